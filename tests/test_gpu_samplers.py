"""Every angular-scattering branch of the collision operator on the GPU
against the oracle (AngularScatteringFunctions.h:27-74; BoltzmannMC.C:
1115-1272): forward, bornDipole, surendra, coulombScreen with option 0
(incident energy) and option 1 (energy after the collision), isotropic, and
the coplanar momentum-conservation (Boeuf) ionization, on conservative and
ionizing processes, for all four energy-sharing operators.

FP64: identical uniform variates give identical channel indices and
post-collision and ejected velocities within 1e-12.  Mixed form: the same
variates give the oracle's scattering angle within FP32 accuracy, and the
cos(chi) it draws follows the reference's angular distributions
(AngularDistributionFunctions.h) by a Kolmogorov-Smirnov test.
"""
import numpy as np
import pytest

import ekmc
import oracle_py as O
from helpers import condition, mixture, speeds, unit_vectors

pytestmark = pytest.mark.gpu

# (type, threshold, scattering, params)
ZOO = [
    (0, 0.0, "isotropic", ()),
    (0, 1.0, "forward", ()),
    (0, 2.5, "bornDipole", ()),
    (0, 4.0, "surendra", ()),
    (0, 6.0, "coulombScreen", (0.0, 0.05)),
    (0, 7.0, "coulombScreen", (1.0, 0.3)),
    (1, 12.0, "momentumConservationIonization", ()),
    (1, 14.0, "bornDipole", ()),
    (1, 16.0, "coulombScreen", (1.0, 0.2)),
    (1, 17.5, "surendra", ()),
    (2, 0.0, "isotropic", ()),
]


def zoo_mixture():
    e = np.logspace(-3, 3, 500)
    procs = []
    for typ, thr, scat, prm in ZOO:
        ee = e[e >= thr] if thr > 0 else e
        if thr > 0:
            ee = np.concatenate([[thr], ee[ee > thr]])
        if typ == 0 and thr == 0:
            sig = 1e-19 * (0.2 + ee / (1 + ee))
        elif typ == 2:
            sig = 3e-22 * np.exp(-((ee - 5.0) / 2.0) ** 2)
        else:
            x = ee / thr
            sig = 3e-21 * np.log(x) / x
        procs.append(dict(type=typ, elastic=(typ == 0 and thr == 0), threshold=thr, energy=ee, sigma=sig,
                          mass_amu=28.0, scattering=scat, params=prm))
    return mixture(procs)


def _forced_rows(cumul, step, P, energies):
    rows = []
    for p in range(P):
        for e0 in energies:
            k = int(e0 / step)
            w1 = (k + 1) - e0 / step
            c = w1 * cumul[k] + (1 - w1) * cumul[k + 1]
            lo = c[p - 1] if p > 0 else 0.0
            if c[p] > lo:
                for f in (0.2, 0.5, 0.8):
                    rows.append((p, e0, lo + f * (c[p] - lo)))
    return rows


def _setup(precision, ioniz):
    mix = zoo_mixture()
    num = ekmc.numerics_default()
    num.precision = precision
    num.ionization_operator = ekmc.IONIZ[ioniz]
    num.n_interp_points = 4000
    cond = condition(en=100.0, n_e=256)
    return mix, cond, num


@pytest.mark.parametrize("ioniz", sorted(ekmc.IONIZ))
def test_every_sampler_matches_oracle_with_identical_variates(ioniz):
    mix, cond, num = _setup(ekmc.FP64, ioniz)
    eng = ekmc.Engine(mix, [cond], num)
    m = O.Model(mix, cond, num)
    maxE = 90.0
    _, cumul, tot, mx, step = m.tables(maxE)
    P = mix.n_processes
    rng = np.random.default_rng(21)
    rows = _forced_rows(cumul, step, P, np.linspace(0.4, 85.0, 60))
    n_forced = len(rows)
    n_rand = 30000
    e = np.concatenate([np.array([r[1] for r in rows]), rng.uniform(0.01, 89.0, n_rand)])
    n = len(e)
    v = unit_vectors(rng, n) * speeds(e)[:, None]
    e = 0.5 * ekmc.EMASS * (v * v).sum(1) / ekmc.CHARGE
    nue = np.full(n, mx[-1] * 1.05)
    var = rng.uniform(0, 1, (n, ekmc.VARIATES_PER_EVENT))
    var[:, [0, 1, 2, 3, 4, 5, 7, 10]] = np.maximum(var[:, [0, 1, 2, 3, 4, 5, 7, 10]], 1e-300)
    var[:n_forced, 5] = np.array([r[2] for r in rows]) * cond.gas_density * np.linalg.norm(v[:n_forced], axis=1) / nue[:n_forced]
    pg, vg, eg, ejvg, ejeg, deg = eng.op_collide(0, maxE, v, e, nue, var)
    st, po, vo, eo, ejvo, ejeo, deo = m.collide(v, e, nue, var)
    assert st == 0
    assert np.array_equal(pg, po)
    want = np.array([r[0] for r in rows])
    assert np.all((po[:n_forced] == want) | (po[:n_forced] == -1))
    counts = np.bincount(po[po >= 0], minlength=P)
    assert np.all(counts[:-1] >= 50) and counts[-1] >= 10, counts  # every branch ran many times
    np.testing.assert_allclose(vg, vo, rtol=1e-12, atol=1e-12 * np.abs(vo).max())
    np.testing.assert_allclose(eg, eo, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(ejvg, ejvo, rtol=1e-12, atol=1e-12 * max(np.abs(ejvo).max(), 1.0))
    np.testing.assert_allclose(ejeg, ejeo, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(deg, deo, rtol=1e-11, atol=1e-12)
    # forward scattering keeps the direction; Boeuf ionization is coplanar:
    # the scattered and ejected electrons and the incident one lie in a plane
    v3, o3 = vo.reshape(n, 3), ejvo.reshape(n, 3)
    fwd = po == 1
    cosf = (v[fwd] * v3[fwd]).sum(1) / np.linalg.norm(v[fwd], axis=1) / np.linalg.norm(v3[fwd], axis=1)
    np.testing.assert_allclose(cosf, 1.0, atol=1e-12)
    bo = (po == 6) & (ejeo > 0) & (eo > 0)
    if bo.any():
        trip = np.einsum("ij,ij->i", v[bo], np.cross(v3[bo], o3[bo]))
        scale = np.linalg.norm(v[bo], axis=1) * np.linalg.norm(v3[bo], axis=1) * np.linalg.norm(o3[bo], axis=1)
        assert np.all(np.abs(trip) <= 1e-10 * scale)
    eng.close()


def _cos_between(a, b):
    return (a * b).sum(1) / np.linalg.norm(a, axis=1) / np.linalg.norm(b, axis=1)


def test_mixed_samplers_follow_oracle_to_fp32_accuracy():
    mix, cond, num = _setup(ekmc.MIXED, "randomUniform")
    eng = ekmc.Engine(mix, [cond], num)
    m = O.Model(mix, cond, num)
    maxE = 90.0
    _, cumul, tot, mx, step = m.tables(maxE)
    P = mix.n_processes
    rng = np.random.default_rng(22)
    rows = _forced_rows(cumul, step, P, np.linspace(0.6, 85.0, 120))
    n = len(rows)
    e = np.array([r[1] for r in rows])
    v = (unit_vectors(rng, n) * speeds(e)[:, None]).astype(np.float32).astype(np.float64)
    e = 0.5 * ekmc.EMASS * (v * v).sum(1) / ekmc.CHARGE
    nue = np.full(n, mx[-1] * 1.05)
    var = rng.uniform(1e-6, 1 - 1e-6, (n, ekmc.VARIATES_PER_EVENT)).astype(np.float32).astype(np.float64)
    var[:, 5] = np.array([r[2] for r in rows]) * cond.gas_density * np.linalg.norm(v, axis=1) / nue
    pg, vg, eg, ejvg, ejeg, deg = eng.op_collide(0, maxE, v, e, nue, var)
    st, po, vo, eo, ejvo, ejeo, deo = m.collide(v, e, nue, var)
    same = (pg == po) & (po >= 0)
    assert same.sum() >= 0.99 * (po >= 0).sum()
    vg3, vo3 = vg.reshape(n, 3)[same], vo.reshape(n, 3)[same]
    cg, co = _cos_between(v[same], vg3), _cos_between(v[same], vo3)
    assert np.all(np.abs(cg - co) <= 2e-4), np.abs(cg - co).max()
    # the mixed form's incident energy comes from FP32 velocities (a few
    # 1e-8 relative); the energy left after a threshold is kept in FP64
    assert np.all(np.abs(eg[same] - eo[same]) <= 2e-5 * eo[same] + 3e-7 * e[same])
    ion = same & (ejeo > 1e-3)
    ce_g = _cos_between(v[ion], ejvg.reshape(n, 3)[ion])
    ce_o = _cos_between(v[ion], ejvo.reshape(n, 3)[ion])
    assert np.all(np.abs(ce_g - ce_o) <= 2e-4)
    for p in range(P):
        assert (po[same] == p).sum() >= 20 or ZOO[p][0] == 2, p
    eng.close()


def _cdf(pdf, x):
    from scipy import integrate
    tot = integrate.quad(pdf, -1, 1, epsabs=1e-13, limit=200)[0]
    return np.array([integrate.quad(pdf, -1, xi, epsabs=1e-13, limit=200)[0] / tot for xi in x])


@pytest.mark.parametrize("p", [2, 3, 4, 5])
def test_mixed_sampler_distributions_follow_reference_pdfs(p):
    """Kolmogorov-Smirnov at the 1 % level on 4000 mixed-form draws at 30 eV
    against the per-solid-angle pdfs of AngularDistributionFunctions.h
    (bornDipole with the process's energy loss; coulombScreen with its
    screening energy over the incident (option 0) or remaining (option 1)
    energy; surendra)."""
    mix, cond, num = _setup(ekmc.MIXED, "randomUniform")
    eng = ekmc.Engine(mix, [cond], num)
    m = O.Model(mix, cond, num)
    maxE = 90.0
    _, cumul, tot, mx, step = m.tables(maxE)
    rng = np.random.default_rng(30 + p)
    e0 = 30.0
    k = int(e0 / step)
    w1 = (k + 1) - e0 / step
    c = w1 * cumul[k] + (1 - w1) * cumul[k + 1]
    n = 4000
    v = (unit_vectors(rng, n) * speeds(np.full(n, e0))[:, None]).astype(np.float32).astype(np.float64)
    e = 0.5 * ekmc.EMASS * (v * v).sum(1) / ekmc.CHARGE
    nue = np.full(n, mx[-1] * 1.05)
    var = rng.uniform(0, 1, (n, ekmc.VARIATES_PER_EVENT))
    var[:, 5] = 0.5 * (c[p - 1] + c[p]) * cond.gas_density * np.linalg.norm(v, axis=1) / nue
    pg, vg, *_ = eng.op_collide(0, maxE, v, e, nue, var)
    assert np.all(pg == p)
    mu = np.sort(_cos_between(v, vg.reshape(n, 3)))
    typ, thr, scat, prm = ZOO[p]
    ea = e0 - thr
    if scat == "bornDipole":
        pdf = lambda x: np.sqrt(e0 * ea) / (ea + e0 - 2 * np.sqrt(e0 * ea) * x)
    elif scat == "surendra":
        pdf = lambda x: e0 / (4 * np.pi * (1 + e0 * (1 - x) / 2) * np.log(1 + e0))
    else:
        xi = prm[1] / (e0 if prm[0] == 0 else ea)
        pdf = lambda x: xi * (xi + 1) / np.pi / (2 * xi + 1 - x) ** 2
    grid = np.linspace(-1, 1, 1601)
    F = np.interp(mu, grid, _cdf(pdf, grid))
    kk = np.arange(1, n + 1) / n
    D = max(np.max(kk - F), np.max(F - (kk - 1.0 / n)))
    assert D < 1.63 / np.sqrt(n), D
    eng.close()
