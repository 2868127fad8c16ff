"""Work-unit autotuner (ekmc_autotune): argument checks on the host, and on
the GPU a tuning pass over a small batch that returns one of its candidates
with a measured rate for every candidate."""
import ctypes as C

import numpy as np
import pytest

import ekmc


def _call(steps, warmup=0, candidates=(256, 1024)):
    mix, conds, num = ekmc.workload("cfg1_atomic40_dc", n_electrons=100)
    arr = (ekmc.Condition * len(conds))(*conds)
    cand = np.asarray(candidates, np.int32)
    best = C.c_int32(-7)
    st = ekmc.lib().ekmc_autotune(C.byref(mix.struct), arr, len(conds), C.byref(num),
                                  cand.ctypes.data_as(C.POINTER(C.c_int32)), len(cand), warmup, steps,
                                  C.byref(best), None)
    return st, best.value


@pytest.mark.parametrize("steps,warmup,candidates", [(0, 0, (256,)), (5, -1, (256,)), (5, 0, (256, 0))])
def test_autotune_rejects_bad_arguments_before_touching_the_device(steps, warmup, candidates):
    st, best = _call(steps, warmup, candidates)
    assert st == 1  # EKMC_ERR_INVALID_ARGUMENT
    assert best == -7
    assert ekmc.lib().ekmc_last_error()


@pytest.mark.gpu
def test_autotune_picks_a_candidate_and_the_pick_runs():
    mix, conds, num = ekmc.workload("cfg3_o2like_batched", n_electrons=20000, conditions=[3, 21, 40, 60])
    num.precision = ekmc.MIXED
    num.seed = 5
    cands = [256, 768, 2048]
    best, rates = ekmc.autotune(mix, conds, num, candidates=cands, warmup=10, steps=10)
    print("autotune", best, rates)
    assert best in cands
    assert sorted(rates) == cands and all(r > 0 for r in rates.values())
    assert rates[best] == max(rates.values())
    num.work_chunk = best
    eng = ekmc.Engine(mix, conds, num)
    eng.initialize()
    assert eng.step(5) == len(conds)
    assert eng.event_count()[0] > 0
    eng.close()
