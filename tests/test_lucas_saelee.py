"""Lucas-Saelee non-conservative model gas vs the published LoKI-MC values."""
import numpy as np
import pytest

import ekmc
from helpers import condition, mixture

pytestmark = pytest.mark.gpu

# Benchmark calculations of LoKI-MC v1.0.0, table 2 (E/N = 10 Td, N = 1e20 m-3,
# M = 1000 m_e): mean energy (eV), flux and bulk drift velocity (1e4 m/s),
# flux and bulk reduced transverse diffusion, flux and bulk reduced
# longitudinal diffusion (1e25 1/(m s)); relative standard deviations in %.
LS_PUBLISHED = {
    (0.0, 0.0, 0.0): ((5.56, 0.05), (7.33, 0.40), (7.33, 0.53), (2.73, 0.59), (2.73, 0.78), (2.66, 0.57), (2.68, 0.86)),
    (0.5, 0.0, 0.0): ((5.22, 0.05), (7.28, 0.40), (8.50, 0.52), (2.54, 0.58), (2.70, 1.35), (2.48, 0.62), (2.85, 1.18)),
    (1.0, 0.0, 0.0): ((4.97, 0.05), (7.36, 0.40), (9.47, 0.53), (2.44, 0.60), (2.75, 1.99), (2.38, 0.62), (2.89, 1.38)),
    (0.0, 5e-4, 0.5): ((5.44, 0.06), (7.35, 0.40), (7.04, 0.68), (2.67, 0.59), (2.61, 1.56), (2.62, 0.59), (2.74, 1.53)),
    (0.0, 2e-3, -0.5): ((5.55, 0.05), (7.34, 0.40), (7.31, 0.47), (2.71, 0.59), (2.77, 1.40), (2.65, 0.60), (2.67, 1.39)),
}


def _procs(F, a, p):
    e = np.concatenate([[0.0], np.logspace(-4, 3, 1401)])
    with np.errstate(divide="ignore"):
        el = np.where(e > 0, 4.0 / np.sqrt(np.maximum(e, 1e-300)), 4.0 / np.sqrt(1e-4)) * 1e-20
    m_amu = 1000.0 * ekmc.EMASS / ekmc.AMU
    procs = [dict(type=0, elastic=True, mass_amu=m_amu, energy=e, sigma=el)]
    eth = np.concatenate([[15.6], e[e > 15.6]])
    if F < 1.0:
        procs.append(dict(type=0, threshold=15.6, mass_amu=m_amu, energy=eth, sigma=0.1 * (1 - F) * (eth - 15.6) * 1e-20))
    if F > 0.0:
        procs.append(dict(type=1, threshold=15.6, w=15.6, mass_amu=m_amu, energy=eth, sigma=0.1 * F * (eth - 15.6) * 1e-20))
    if a > 0.0:
        ea = e[e > 0]
        procs.append(dict(type=2, mass_amu=m_amu, energy=ea, sigma=a * ea ** p * 1e-20))
    return procs


def _digit(v):
    return 0.5 * 10 ** (np.floor(np.log10(abs(v))) - 2)


@pytest.mark.parametrize("case", sorted(LS_PUBLISHED))
def test_lucas_saelee_matches_published_loki_mc_values(case):
    F, a, p = case
    mix = mixture(_procs(F, a, p))
    tg = 300.0
    cond = condition(en=10.0, n_e=100000, tg=tg, p=1e20 * ekmc.BOLTZMANN * tg)
    num = ekmc.numerics_default()
    num.n_integration_points = 1e4
    num.n_integrated_ss_times = 2
    num.ionization_operator = 0
    num.seed = 31
    eng = ekmc.Engine(mix, [cond], num)
    eng.solve()
    r = eng.results(0)
    N = 1e20

    def rel(v, e):
        return e / abs(v)
    ours = [(r.mean_energy, rel(r.mean_energy, r.mean_energy_error)),
            (abs(r.flux_drift_velocity[2]) / 1e4, rel(r.flux_drift_velocity[2], r.flux_drift_velocity_error[2])),
            (abs(r.bulk_drift_velocity[2]) / 1e4, rel(r.bulk_drift_velocity[2], r.bulk_drift_velocity_error[2])),
            (r.flux_red_transv_diff / 1e25, max(rel(r.flux_diffusion[0], r.flux_diffusion_error[0]), rel(r.flux_diffusion[4], r.flux_diffusion_error[4]))),
            (r.bulk_red_transv_diff / 1e25, max(rel(r.bulk_diffusion[0], r.bulk_diffusion_error[0]), rel(r.bulk_diffusion[4], r.bulk_diffusion_error[4]))),
            (r.flux_red_long_diff / 1e25, rel(r.flux_diffusion[8], r.flux_diffusion_error[8])),
            (r.bulk_red_long_diff / 1e25, rel(r.bulk_diffusion[8], r.bulk_diffusion_error[8]))]
    names = ("mean energy", "flux drift", "bulk drift", "flux N D_T", "bulk N D_T", "flux N D_L", "bulk N D_L")
    print(case, ["%s %.4g (pub %.4g)" % (n, v, pv) for n, (v, _), (pv, _) in zip(names, ours, LS_PUBLISHED[case])])
    for name, (v, rl), (pv, prel) in zip(names, ours, LS_PUBLISHED[case]):
        tol = 3.0 * np.hypot(rl * v, prel / 100.0 * pv) + _digit(pv)
        assert abs(v - pv) <= tol, (case, name, v, pv, tol)
    eng.close()
