"""The named workloads at their full sizes (cfg5: one condition of 1e8
electrons and 200 processes; cfg4: 24 AC conditions of 2e5; cfg2: 8 of 1e5;
cfg3: the 64-condition crossed-field scan of 1.56e7 electrons),
a few synchronization intervals each, checked through properties that hold at
any size: every condition keeps exactly its N main electrons, all synchronized
to the interval end with finite energies; about one trial event per electron
and interval (the interval is 1/nu_trial and trial events form a Poisson
process of rate nu_trial); events split into real and null; no flight of any
condition outran the trial frequency; and two engines with the same seed give
the same bits."""
import ctypes as C

import numpy as np
import pytest

import ekmc

pytestmark = pytest.mark.gpu


def _energies_and_times(eng, cond):
    n = eng.conditions[cond].n_electrons
    e, t = np.zeros(n), np.zeros(n)
    ekmc.check(ekmc.lib().ekmc_get_electrons(eng.h, cond, None, None, ekmc.dptr(e), ekmc.dptr(t)))
    return e, t


def _run(wl, prec, intervals, seed=11):
    mix, conds, num = ekmc.workload(wl)
    num.precision = prec
    num.seed = seed
    num.n_integration_points = 1e15  # no condition stops inside the run
    eng = ekmc.Engine(mix, conds, num)
    eng.initialize()
    assert eng.step(intervals) == len(conds)
    return eng, conds


@pytest.mark.parametrize("wl,prec,intervals", [("cfg5_mix5_large", ekmc.MIXED, 8),
                                               ("cfg4_air_ac", ekmc.MIXED, 20),
                                               ("cfg4_air_ac", ekmc.FP64, 10),
                                               ("cfg2_n2like_scan", ekmc.FP64, 20),
                                               ("cfg3_o2like_batched", ekmc.MIXED, 30),
                                               ("cfg3_o2like_batched", ekmc.FP64, 10)])
def test_full_size_intervals_keep_every_condition_whole_and_synchronized(wl, prec, intervals):
    eng, conds = _run(wl, prec, intervals)
    n_total = sum(c.n_electrons for c in conds)
    ev, real = eng.event_count()
    per = ev / (n_total * intervals)
    print(wl, "events per electron and interval %.4f, real fraction %.4f" % (per, real / ev))
    assert 0.9 < per < 1.1  # measured 0.988-0.999
    assert 0 < real < ev
    # the trial frequency bounded the total collision frequency of every
    # flight of every condition (no escape from the null-collision method)
    for c in range(len(conds)):
        assert eng.status(c).escape_events == 0, c
    for c in sorted({0, len(conds) // 2, len(conds) - 1}):
        e, t = _energies_and_times(eng, c)
        assert len(e) == conds[c].n_electrons
        assert np.all(np.isfinite(e)) and np.all(e >= 0)
        assert np.all(t == t[0]) and t[0] > 0
        assert eng.status(c).time == t[0]
        me = eng.series(c)["mean_energy"]
        assert len(me) == intervals + 1 and np.all(np.isfinite(me)) and np.all(me > 0)
        assert abs(me[-1] - e.mean()) <= 1e-9 * me[-1]
    eng.close()


def test_full_size_batch_is_bitwise_reproducible():
    out = []
    for _ in range(2):
        eng, conds = _run("cfg4_air_ac", ekmc.MIXED, 12, seed=23)
        out.append([np.array(eng.event_count(), np.uint64)] +
                   [x for c in (0, 11, 23) for x in _energies_and_times(eng, c)] +
                   [eng.series(c)["mean_energy"] for c in range(len(conds))])
        eng.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
