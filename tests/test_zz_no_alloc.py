"""ekmc_op_collide_noalloc, the allocation-free variant of ekmc_op_collide
(its device staging is a static array of the library, one per GPU), against
ekmc_op_collide through the C ABI: identical process ids and bit-identical
velocities, energies, ejected electrons and energy changes on two calls in a
row, the second smaller than the first (the staging keeps the first call's
contents); more than EKMC_NOALLOC_MAX_EVENTS events are rejected."""
import numpy as np
import pytest

import ekmc
from helpers import speeds, unit_vectors

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [ekmc.FP64, ekmc.MIXED])
def test_noalloc_collide_equals_collide_on_two_calls(prec):
    mix, conds, num = ekmc.workload("cfg3_o2like_batched", n_electrons=1024, conditions=[40])
    num.precision = prec
    num.ionization_operator = ekmc.IONIZ["randomUniform"]
    num.n_interp_points = 4000
    eng = ekmc.Engine(mix, conds, num)
    maxE = 80.0
    _, _, _, mx, _ = eng.op_tables(0, maxE)
    rng = np.random.default_rng(5)
    for n in (ekmc.NOALLOC_MAX_EVENTS, 5000):
        e = rng.uniform(0.01, 79.0, n)
        v = unit_vectors(rng, n) * speeds(e)[:, None]
        nue = np.full(n, mx[-1] * 1.05)
        var = np.maximum(rng.uniform(0, 1, (n, ekmc.VARIATES_PER_EVENT)), 1e-300)
        a = eng.op_collide(0, maxE, v, e, nue, var)
        b = eng.op_collide(0, maxE, v, e, nue, var, noalloc=True)
        assert (a[0] >= 0).sum() > n // 20
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    with pytest.raises(ekmc.EngineError) as ei:
        n = ekmc.NOALLOC_MAX_EVENTS + 1
        eng.op_collide(0, maxE, np.ones((n, 3)), np.ones(n), np.ones(n), np.full((n, 11), 0.5), noalloc=True)
    assert ei.value.status == 5 and "at most" in str(ei.value)
    eng.close()
