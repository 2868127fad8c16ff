"""The cycle after the flights of an interval -- population control,
sampling, the fixed-order reduction and the download of the per-condition
summaries -- is captured once into a HIP graph and replayed at every later
interval.  Every replay must give the bits the same launches give one by one
(EKMC_GRAPH=0): the runs below compare the two interval by interval, with the
engine's own eager work (flights, uploads, histogram launches) and a stream
synchronization between replays."""
import numpy as np
import pytest

import ekmc
from helpers import condition

pytestmark = pytest.mark.gpu


def _numerics(seed, precision):
    num = ekmc.numerics_default()
    num.precision = precision
    num.seed = seed
    num.initial_elec_temp_over_gas_temp = 38.6817
    num.max_collisions_before_ss = 30
    num.n_integration_points = 600
    return num


def _snapshot(eng, n_cond):
    out = [np.array(eng.event_count(), np.uint64)]
    for c in range(n_cond):
        pos, vel, e, t = eng.electrons(c)
        out += [pos, vel, e, t]
        s = eng.series(c)
        out += [s["mean_energy"], s["mean_position"], s["flux_diffusion"]]
        out += list(eng.process_data(c))
    return out


@pytest.mark.parametrize("precision", [ekmc.FP64, ekmc.MIXED])
def test_graph_replays_equal_eager_launches_every_interval(monkeypatch, precision):
    # 52 Td (attachment: removals refilled by clones), 1000 Td (ionization:
    # newborns moved into holes and marks removed), a crossed-field condition
    mix, conds, _ = ekmc.workload("cfg3_o2like_batched", n_electrons=5000, conditions=[12, 21, 60])
    runs = []
    for graph in ("1", "0"):
        monkeypatch.setenv("EKMC_GRAPH", graph)
        eng = ekmc.Engine(mix, conds, _numerics(17, precision))
        eng.initialize()
        snaps = []
        for _ in range(25):
            assert eng.step(1) == len(conds)
            snaps.append(_snapshot(eng, len(conds)))
        eng.close()
        runs.append(snaps)
    rates, counts = runs[0][-1][8], runs[0][-1][9]
    assert counts.sum() > 0
    for i, (a, b) in enumerate(zip(*runs)):
        assert len(a) == len(b)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), "interval %d differs" % (i + 1)


def test_graph_solve_equals_eager_solve(monkeypatch):
    """Whole runs to their stop criteria, with histograms after steady state
    (counted at load, and by the histogram kernel for the last sample) and an
    AC condition's per-phase sums."""
    mix, _, _ = ekmc.workload("cfg1_atomic40_dc")
    conds = [condition(en=100.0, n_e=6000), condition(en=150.0, angle=60.0, freq=2e9, n_e=6000)]
    res = []
    for graph in ("1", "0"):
        monkeypatch.setenv("EKMC_GRAPH", graph)
        num = _numerics(43, ekmc.FP64)
        num.n_integration_phases = 8
        num.n_integration_points = 300
        eng = ekmc.Engine(mix, conds, num)
        eng.solve()
        out = []
        for c in range(len(conds)):
            d = eng.results(c).as_dict()
            del d["elapsed_time"]  # wall clock
            out += [np.asarray(v, dtype=float).ravel() for v in d.values()]
            out += list(eng.eedf(c))
            out += list(eng.electrons(c))
        res.append(out)
        eng.close()
    assert len(res[0]) == len(res[1])
    for a, b in zip(*res):
        assert np.array_equal(a, b, equal_nan=True)
    # the last condition's entries end with its EEDF (cells, values) and its
    # electrons (positions, velocities, energies, times)
    assert res[0][-5].sum() > 0
