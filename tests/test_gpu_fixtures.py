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


def repeated_patterns(m=6000, n=5000, patterns=200, seed=3):
    """Unit values, every row a copy of one of `patterns` column sets of
    5-40 columns: 2P < m and fewer than 128 columns per pattern, so the
    reference policy picks the single (not unrolled) pattern-reuse kernel
    (cb-spmv-dcu.hpp:1431-1436, EXEC_PATTERN_REUSE_SINGLE)."""
    rng = np.random.default_rng(seed)
    sets = [np.sort(rng.choice(n, size=int(rng.integers(5, 41)), replace=False)) for _ in range(patterns)]
    pick = rng.integers(0, patterns, size=m)
    r = np.repeat(np.arange(m, dtype=np.int32), [sets[p].size for p in pick])
    c = np.concatenate([sets[p] for p in pick]).astype(np.int32)
    return m, n, r, c, np.ones(r.size)


def test_repeated_patterns_policy_is_single_pattern():
    """CPU: the fixture below takes the single pattern-reuse kernel, as the
    oracle's restatement of the policy says, with P = the distinct row sets."""
    m, n, r, c, v = repeated_patterns()
    plan = Plan(m, n, r, c, v)
    ob = pyoracle.OracleBuild(m, n, r, c, v)
    pol = plan.export("policy").tolist()
    want = ob.policy()
    assert pol[:3] == [1, 1, 0]            # unit values, pattern reuse, not unrolled
    assert want["use_pattern"] == 1 and want["use_unrolled"] == 0
    distinct = {tuple(c[r == i]) for i in range(m)}
    assert pol[7] == len(distinct) and 2 * pol[7] < m
    assert plan.schedule() == "blocked"
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [dict(), dict(row_owned=False), dict(block_size=32)], ids=["ro", "lpt", "bs32"])
def test_single_pattern_kernel_parity(gpu, opts):
    import torch

    m, n, r, c, v = repeated_patterns()
    print(f"case repeated_patterns/{opts}", flush=True)
    plan = Plan(m, n, r, c, v, **opts)
    assert plan.export("policy").tolist()[:3] == [1, 1, 0]
    for kind in ("uniform", "integer"):
        rng = np.random.default_rng(len(kind))
        x = rng.uniform(-1, 1, n) if kind == "uniform" else rng.integers(-8, 9, n).astype(np.float64)
        y_ref = pyoracle.spmv_seq(m, r, c, v, x)
        xd = torch.tensor(x, dtype=torch.float64, device=gpu)
        for sched in ("blocked", "stream"):
            plan.set_schedule(sched)
            y = plan.spmv(xd).cpu().numpy()
            if kind == "integer":
                assert np.array_equal(y, y_ref), f"{sched}: integer x not exact"
            else:
                assert pyoracle.rel_l2(y, y_ref) <= TOL, sched
    plan.close()
