"""Distributions and run control of the engine, pinned by identities with
independent estimators and by the counts the reference's loop fixes.

Distributions (getTimeDependDistributions, BoltzmannMC.C:1492-1572;
getTimeAverageDistributions, :1574-1605; histogramCount/histogram2DCount,
MathFunctions.C:61-81,106-127).  The EEDF, the (v_r, v_z) EVDF and the
energy-angle distribution (through its first anisotropy) are histograms of the
same sampled ensembles whose means give the mean energy and the flux drift
velocity (calculateMeanDataForSwarmParams, :1410-1482), so their moments must
reproduce those sampled averages up to the binning: mean energy from the EEDF
and from the EVDF, drift velocity from the EVDF's v_z and from the EADF's
first anisotropy.  A half-cell shift of any grid, a wrong cell width or a
wrong normalisation moves these moments by far more than the tolerances.
With no field and thermal targets the EVDF is the Maxwellian at T_g:
<v_z^2> = k T_g / m_e, <v_r^2> = 2 k T_g / m_e, <cos theta> = 0.

Control (BoltzmannMC.C:320-384, 1743-1893, MathFunctions.C:211-231): the time
averages and their errors equal the mean and the 50-batch-means error of the
engine's own sampled series over its integration points; the steady state is
first tested at sampling point 128; K integration points stop the run after
127 + K sampling points; error checks first run at 256 integration points.
"""
import numpy as np
import pytest

import ekmc
from helpers import condition

pytestmark = pytest.mark.gpu

Q_M = ekmc.CHARGE / ekmc.EMASS


def _evdf(eng, c, num):
    L = ekmc.lib()
    nr, na, nc = num.n_radial_velocity_cells, num.n_axial_velocity_cells, num.n_energy_cells
    r, a, f, f1, f2 = np.zeros(nr), np.zeros(na), np.zeros(nr * na), np.zeros(nc), np.zeros(nc)
    ekmc.check(L.ekmc_get_evdf(eng.h, c, ekmc.dptr(r), ekmc.dptr(a), ekmc.dptr(f), ekmc.dptr(f1), ekmc.dptr(f2)))
    return r, a, f.reshape(nr, na), f1, f2


def _moments(eng, c, num):
    cells, eedf = eng.eedf(c)
    de = cells[1] - cells[0]
    p = eedf * np.sqrt(cells) * de
    r, a, f, f1, f2 = _evdf(eng, c, num)
    dr, dz = r[1] - r[0], a[1] - a[0]
    # the reference's normalisation (BoltzmannMC.C:1604): counts over
    # sum_r(row_r dz 2 pi r_r dr), one divisor for every cell, so that
    # sum f 2 pi r dr dz = 1 while f itself stays proportional to the cell
    # counts (the (v_r, v_z) cells already hold the 2 pi v_r Jacobian)
    norm = (f * (2 * np.pi * r[:, None] * dr * dz)).sum()
    w = f / f.sum()
    R, A = np.meshgrid(r, a, indexing="ij")
    speed = np.sqrt(2 * Q_M * cells)
    # the anisotropies are the Legendre coefficients of the energy-angle
    # distribution (BoltzmannMC.C:1583-1600): f1 = 3 f0 <cos theta>,
    # f2 = 5 f0 <P2(cos theta)> per energy cell, f0 the EEDF
    jac = np.sqrt(cells) * de
    return dict(p_sum=p.sum(), e_eedf=(p * cells).sum(), w_sum=norm, vz_evdf=(w * A).sum(),
                e_evdf=(w * 0.5 * (R ** 2 + A ** 2)).sum() / Q_M, vz2=(w * A ** 2).sum(), vr2=(w * R ** 2).sum(),
                vz_eadf=(jac * speed * f1 / 3.0).sum(), cos_mean=(jac * f1 / 3.0).sum(),
                p2_mean=(jac * f2 / 5.0).sum())


@pytest.mark.parametrize("prec", [ekmc.FP64, ekmc.MIXED])
def test_distribution_moments_reproduce_the_sampled_averages(prec):
    mix, _, _ = ekmc.workload("cfg1_atomic40_dc")
    conds = [condition(en=100.0, n_e=20000), condition(en=150.0, n_e=20000)]
    num = ekmc.numerics_default()
    num.precision = prec
    num.seed = 77
    num.initial_elec_temp_over_gas_temp = 38.68
    num.max_collisions_before_ss = 30
    num.n_integration_points = 1500
    eng = ekmc.Engine(mix, conds, num)
    eng.solve()
    for c in range(len(conds)):
        r = eng.results(c)
        m = _moments(eng, c, num)
        vd = r.flux_drift_velocity[2]
        print(c, {k: float(v) for k, v in m.items()}, "mean", r.mean_energy, "vd", vd)
        assert abs(m["p_sum"] - 1) < 1e-12 and abs(m["w_sum"] - 1) < 1e-12
        assert vd > 0  # E along -z: electrons drift along +z
        assert abs(m["e_eedf"] - r.mean_energy) <= 2e-3 * r.mean_energy
        assert abs(m["e_evdf"] - r.mean_energy) <= 3e-3 * r.mean_energy
        assert abs(m["vz_evdf"] - vd) <= 5e-3 * vd
        assert abs(m["vz_eadf"] - vd) <= 2e-2 * vd
    eng.close()


def test_field_free_thermal_evdf_is_the_maxwellian():
    """No field, targets at 300 K (gasTemperatureEffect true): the EVDF's
    second moments are k T_g / m_e along z and 2 k T_g / m_e across, the
    angular distribution is isotropic, within the spread of the samples."""
    from helpers import mixture
    e = np.logspace(-4, 3, 300)
    mix = mixture([dict(type=0, elastic=True, energy=e, sigma=np.full_like(e, 1e-19),
                        mass_amu=20 * ekmc.EMASS / ekmc.AMU)])
    num = ekmc.numerics_default()
    num.gas_temperature_effect = ekmc.GASTEMP["true"]
    num.initial_elec_temp_over_gas_temp = 1.0
    num.n_integration_points = 4000
    num.seed = 8
    T = 300.0
    eng = ekmc.Engine(mix, [condition(en=1e-6, tg=T, n_e=20000)], num)
    eng.solve()
    m = _moments(eng, 0, num)
    kT_m = ekmc.BOLTZMANN * T / ekmc.EMASS
    print(m, kT_m)
    assert abs(m["vz2"] / kT_m - 1) < 6e-3
    assert abs(m["vr2"] / (2 * kT_m) - 1) < 6e-3
    assert abs(m["vz_evdf"]) < 2e-3 * np.sqrt(kT_m)
    assert abs(m["cos_mean"]) < 2e-3 and abs(m["p2_mean"]) < 2e-3
    eng.close()


def _batch_error(x, nb=50):
    """Standard error by 50 batch means of n // 50 points each, divided by
    the number of batches (MathFunctions.C:211-231)."""
    n = len(x) // nb
    means = np.array([x[i * n:(i + 1) * n].mean() for i in range(nb)])
    return np.sqrt(((means - x.mean()) ** 2).sum()) / nb


def _gas():
    from helpers import mixture
    e = np.logspace(-3, 3, 200)
    return mixture([dict(type=0, elastic=True, energy=e, sigma=1e-20 * (0.3 + e / (1 + e)), mass_amu=20.0)])


def test_engine_control_cadence_and_averages_follow_its_series():
    conds = [condition(en=50.0, n_e=2000), condition(en=200.0, n_e=3000)]
    num = ekmc.numerics_default()
    num.max_collisions_before_ss = 1e-9
    num.n_integration_points = 300
    num.initial_elec_temp_over_gas_temp = 38.68
    num.seed = 3
    eng = ekmc.Engine(_gas(), conds, num)
    eng.solve()
    for c in range(2):
        r = eng.results(c)
        s = eng.series(c)
        t = s["times"]
        assert r.n_integration_points == 300 and r.n_sampling_points == 127 + 300 == len(t)
        assert r.steady_state_time == t[127] and r.final_time == t[-1]
        me = s["mean_energy"][127:]
        assert abs(r.mean_energy - me.mean()) <= 1e-13 * r.mean_energy
        assert abs(r.mean_energy_error - _batch_error(me)) <= 1e-12 * r.mean_energy_error
        for key, width, mean, err in (("mean_velocity", 3, r.flux_drift_velocity, r.flux_drift_velocity_error),
                                      ("bulk_velocity", 3, r.bulk_drift_velocity, r.bulk_drift_velocity_error),
                                      ("flux_diffusion", 9, r.flux_diffusion, r.flux_diffusion_error),
                                      ("bulk_diffusion", 9, r.bulk_diffusion, r.bulk_diffusion_error)):
            x = s[key].reshape(-1, width)[127:]
            for k in range(width):
                assert abs(mean[k] - x[:, k].mean()) <= 1e-12 * (abs(x[:, k]).max() + 1e-300), (key, k)
                assert abs(err[k] - _batch_error(x[:, k])) <= 1e-10 * (err[k] + 1e-300), (key, k)
        # bulk quantities are finite differences of the sampled position moments
        pos = s["mean_position"].reshape(-1, 3)
        bv = s["bulk_velocity"].reshape(-1, 3)
        np.testing.assert_allclose(bv[1:], np.diff(pos, axis=0) / np.diff(t)[:, None], rtol=1e-12,
                                   atol=1e-12 * np.abs(bv).max())
    eng.close()


def test_engine_error_checks_stop_at_256_integration_points():
    num = ekmc.numerics_default()
    num.max_collisions_before_ss = 1e-9
    num.n_integration_points = 20
    num.errors_to_be_checked = 1
    num.initial_elec_temp_over_gas_temp = 38.68
    num.seed = 4
    eng = ekmc.Engine(_gas(), [condition(en=50.0, n_e=2000)], num)
    eng.solve()
    r = eng.results(0)
    assert r.good_statistical_errors == 1
    assert r.n_integration_points == 256 and r.n_sampling_points == 127 + 256
    eng.close()
