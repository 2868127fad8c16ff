"""Lucas-Saelee non-conservative model gas vs the published LoKI-MC values
(benchmark calculations of LoKI-MC v1.0.0, table 2): the GPU engine with
1e5 electrons, every published quantity within three combined standard
deviations plus the rounding of the printed value."""
import pytest

import ekmc
from helpers import condition, mixture
from published import LS_NAMES, LS_PUBLISHED, lucas_saelee_processes, ls_values, within

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", sorted(LS_PUBLISHED))
def test_lucas_saelee_matches_published_loki_mc_values(case):
    F, a, p = case
    mix = mixture(lucas_saelee_processes(F, a, p))
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
    ours = ls_values(r)
    print(case, ["%s %.4g (pub %.4g)" % (n, v, pv) for n, (v, _), (pv, _) in zip(LS_NAMES, ours, LS_PUBLISHED[case])])
    for name, (v, rl), (pv, prel) in zip(LS_NAMES, ours, LS_PUBLISHED[case]):
        ok, tol = within(v, rl, pv, prel, 3)
        assert ok, (case, name, v, pv, tol)
    assert eng.status(0).escape_events == 0
    eng.close()
