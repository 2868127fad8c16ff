"""The oracle's whole loop (orc_solve: BoltzmannMC::solve, BoltzmannMC.C:
273-426, with its population control, steady-state and stop control) pinned
on the CPU against the published LoKI-MC results, and its control cadence
against the counts the reference's loop fixes.

Published values: the Reid ramp (table 1) and the Lucas-Saelee model gas
with ionization (F = 0.5, 1) and attachment (a, p) (table 2) of the
reference's benchmark document, within three combined standard deviations
(plus the rounding of the printed value).  The runs use 1e4 electrons (the
published ones 1e5), so their own batch-means errors are larger.

Cadence (BoltzmannMC.C:320-384): the steady state is first tested at the
first sampling point >= 100 that is a multiple of 2^max(log2 n - 11, 6),
i.e. at point 128 (index 127); from there n_integration_points = K stops the
run after exactly K integration points (127 + K sampling points); with error
checks on, errors are first tested at the first multiple of
2^max(log2 n - 9, 7) above 200 integration points, i.e. 256.
"""
import numpy as np
import pytest

import ekmc
import oracle_py as O
from helpers import condition, mixture
from published import (LS_NAMES, LS_PUBLISHED, REID_PUBLISHED, lucas_saelee_processes, ls_values,
                       reid_mixture, reid_values, within)


@pytest.mark.parametrize("en", sorted(REID_PUBLISHED))
def test_oracle_reid_ramp_matches_published_values(en):
    cond = condition(en=en, n_e=10000)
    num = ekmc.numerics_default()
    num.n_integration_points = 2000
    num.n_integrated_ss_times = 2
    num.seed = 5
    r, _, _, _ = O.Model(reid_mixture(), cond, num).solve(n_threads=8)
    for name, (v, rel), (pv, prel) in zip(("mean energy", "bulk drift", "N D_T", "N D_L"),
                                          reid_values(r, cond.gas_density), REID_PUBLISHED[en]):
        ok, tol = within(v, rel, pv, prel, 4)
        assert ok, (en, name, v, pv, tol)


@pytest.mark.parametrize("case", sorted(LS_PUBLISHED))
def test_oracle_lucas_saelee_matches_published_values(case):
    F, a, p = case
    tg = 300.0
    cond = condition(en=10.0, n_e=10000, tg=tg, p=1e20 * ekmc.BOLTZMANN * tg)
    num = ekmc.numerics_default()
    num.n_integration_points = 3000
    num.n_integrated_ss_times = 2
    num.ionization_operator = 0
    num.seed = 31
    r, _, _, _ = O.Model(mixture(lucas_saelee_processes(F, a, p)), cond, num).solve(n_threads=8)
    if F > 0:
        assert r.total_ion_rate_coeff > 0 and r.power_growth < 0
    if a > 0:
        assert r.total_att_rate_coeff > 0
    assert r.power_balance_rel_error < 1e-2
    for name, (v, rel), (pv, prel) in zip(LS_NAMES, ls_values(r), LS_PUBLISHED[case]):
        ok, tol = within(v, rel, pv, prel, 3)
        assert ok, (case, name, v, pv, tol)


def _small_gas():
    e = np.logspace(-3, 3, 200)
    return mixture([dict(type=0, elastic=True, energy=e, sigma=1e-20 * (0.3 + e / (1 + e)), mass_amu=20.0)])


def test_oracle_steady_state_is_first_tested_at_point_128_and_stops_after_k_points():
    num = ekmc.numerics_default()
    num.max_collisions_before_ss = 1e-9  # steady state at the first test
    num.n_integration_points = 300
    num.initial_elec_temp_over_gas_temp = 38.68
    num.seed = 3
    m = O.Model(_small_gas(), condition(en=50.0, n_e=1000), num)
    r, _, _, _ = m.solve(n_threads=4)
    t, e, _ = m.series()
    assert r.n_integration_points == 300
    assert r.n_sampling_points == 127 + 300
    assert r.steady_state_time == t[127]
    assert r.final_time == t[-1]
    # time averages over the integration points [127, 427)
    assert abs(r.mean_energy - e[127:].mean()) <= 1e-14 * r.mean_energy
    assert r.mean_energy_error == O.statistical_error(e[127:])


def test_oracle_error_checks_stop_at_256_integration_points():
    num = ekmc.numerics_default()
    num.max_collisions_before_ss = 1e-9
    num.n_integration_points = 20
    num.errors_to_be_checked = 1  # every limit 1e100: satisfied at the first check
    num.initial_elec_temp_over_gas_temp = 38.68
    num.seed = 4
    r, _, _, _ = O.Model(_small_gas(), condition(en=50.0, n_e=1000), num).solve(n_threads=4)
    assert r.good_statistical_errors == 1
    assert r.n_integration_points == 256
    assert r.n_sampling_points == 127 + 256


def test_oracle_stops_at_max_collisions_after_steady_state():
    num = ekmc.numerics_default()
    num.max_collisions_before_ss = 1e-9
    num.n_integration_points = 1e6
    num.max_collisions_after_ss = 40.0
    num.initial_elec_temp_over_gas_temp = 38.68
    num.seed = 6
    r, _, _, _ = O.Model(_small_gas(), condition(en=50.0, n_e=1000), num).solve(n_threads=4)
    after = r.real_collisions - r.real_collisions_at_ss
    assert r.stopped_by_max_collisions == 1
    assert after >= 40 * 1000
    # one interval earlier it had not been reached: at most one interval of
    # collisions beyond the limit (~0.15 real collisions per electron)
    assert after < 40 * 1000 + 1000
    assert r.n_integration_points > 100
