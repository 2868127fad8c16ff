"""Launch autotuners.  ekmc_autotune (work unit): argument checks on the
host, and on the GPU a tuning pass over a small batch that returns one of its
candidates with a measured rate for every candidate.  ekmc_autotune_launch
(work unit, collision-step lane threshold, envelope coarseness): argument
checks, and on the GPU a pass whose pick runs and gives the same physics as
the defaults within the statistics."""
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


def test_launch_parameters_out_of_range_are_rejected():
    mix, conds, num = ekmc.workload("cfg1_atomic40_dc", n_electrons=100)
    for field, bad in (("park_lanes", -1), ("park_lanes", 65), ("env_coarsen", -1), ("env_coarsen", 4),
                       ("work_chunk", 65536), ("work_chunk", -1), ("evdf3d_cells", -1), ("evdf3d_cells", 257)):
        n = ekmc.Numerics.from_buffer_copy(num)
        setattr(n, field, bad)
        with pytest.raises(ekmc.EngineError, match=field):
            ekmc.Engine(mix, conds, n)


def test_launch_autotune_rejects_bad_arguments():
    mix, conds, num = ekmc.workload("cfg1_atomic40_dc", n_electrons=100)
    arr = (ekmc.Condition * len(conds))(*conds)
    best = ekmc.Numerics()
    best.work_chunk = -7
    L = ekmc.lib()
    assert L.ekmc_autotune_launch(C.byref(mix.struct), arr, len(conds), C.byref(num), 0, 0, C.byref(best), None) == 1
    assert L.ekmc_autotune_launch(C.byref(mix.struct), arr, 0, C.byref(num), 0, 5, C.byref(best), None) == 1
    assert best.work_chunk == -7


@pytest.mark.gpu
def test_launch_autotune_pick_runs_and_keeps_the_physics():
    mix, conds, num = ekmc.workload("cfg3_o2like_batched", n_electrons=20000, conditions=[3, 21, 40, 60])
    num.precision = ekmc.MIXED
    num.seed = 5
    best, rate = ekmc.autotune_launch(mix, conds, num, warmup=10, steps=10)
    print("autotune_launch", best.work_chunk, best.park_lanes, best.env_coarsen, rate)
    assert rate > 0 and best.work_chunk >= 64
    assert best.park_lanes in (16, 24, 32, 40, 48) and best.env_coarsen in (0, 1, 2)
    assert best.seed == num.seed and best.precision == num.precision
    # every launch setting is a schedule of the same per-electron draws: the
    # first interval's events are those of the default launch, for the pick
    # and for settings away from the defaults in every parameter
    other = ekmc.Numerics.from_buffer_copy(num)
    other.work_chunk, other.park_lanes, other.env_coarsen = 512, 16, 2
    counts = []
    for n in (num, best, other):
        eng = ekmc.Engine(mix, conds, n)
        eng.initialize()
        eng.step(1)
        counts.append(eng.event_count())
        eng.close()
    print(counts)
    assert counts[0][1] > 0
    assert counts[1] == counts[0] and counts[2] == counts[0]  # (events, real collisions), bit for bit
