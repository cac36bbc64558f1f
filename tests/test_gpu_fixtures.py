"""GPU parity on the committed matrices the reference ships and tests with
(1138_bus.mtx, heart2.mtx through its pattern fixture), read with the
reference reader's semantics (every value forced to 1).  These are the
unit-valued inputs that take the reference's special dispatch paths
(cb-spmv-dcu.hpp:1428-1450): 1138_bus the small-direct path, heart2 the
unrolled pattern-reuse kernel.  Both schedules are checked against the
oracle's FP64 SpMV at relative L2 <= 1e-12; the pattern path sums integers'
worth of x entries and must also match exactly when x is integer-valued.
"""
import os

import numpy as np
import pytest

from cbspmv import Plan, read_mtx
from oracle import pyoracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def heart2():
    z = np.load(os.path.join(GOLD, "heart2_pattern.npz"))
    ptr = z["row_ptr"].astype(np.int64)
    m, n = int(z["m"]), int(z["n"])
    r = np.repeat(np.arange(m, dtype=np.int32), np.diff(ptr))
    return m, n, r, z["col"].astype(np.int32), np.ones(r.size)


FIXTURES = {
    "1138_bus": lambda: read_mtx(os.path.join(GOLD, "1138_bus.mtx")),
    "heart2": heart2,
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FIXTURES))
@pytest.mark.parametrize("opts", [dict(), dict(row_owned=False), dict(block_size=32), dict(pattern_wave=False)],
                         ids=["ro", "lpt", "bs32", "nopattern"])
def test_fixture_parity(gpu, name, opts):
    import torch

    m, n, r, c, v = FIXTURES[name]()
    print(f"case {name}/{opts}", flush=True)
    plan = Plan(m, n, r, c, v, **opts)
    for kind in ("uniform", "integer"):
        rng = np.random.default_rng(len(kind))
        x = rng.uniform(-1, 1, n) if kind == "uniform" else rng.integers(-8, 9, n).astype(np.float64)
        y_ref = pyoracle.spmv_seq(m, r, c, v, x)
        xd = torch.tensor(x, dtype=torch.float64, device=gpu)
        for sched in ("stream", "blocked"):
            plan.set_schedule(sched)
            y = plan.spmv(xd).cpu().numpy()
            if kind == "integer":
                assert np.array_equal(y, y_ref), f"{name}/{sched}: integer x not exact"
            else:
                assert pyoracle.rel_l2(y, y_ref) <= TOL, f"{name}/{sched}"
    plan.close()


@pytest.mark.gpu
def test_heart2_takes_pattern_path(gpu):
    """heart2's policy selects the unrolled pattern kernel; its output rows
    cover every row exactly once."""
    m, n, r, c, v = heart2()
    plan = Plan(m, n, r, c, v)
    pol = plan.export("policy").tolist()
    assert pol[:3] == [1, 1, 1]
    rows = plan.export("pattern_output_rows")
    assert np.array_equal(np.sort(rows), np.arange(m))
    plan.close()


@pytest.mark.gpu
def test_bench_on_fixture(gpu):
    m, n, r, c, v = FIXTURES["1138_bus"]()
    plan = Plan(m, n, r, c, v)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    y, runtime, gflops = plan.bench(x)
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v, x)) <= TOL
    assert runtime > 0
    plan.close()
