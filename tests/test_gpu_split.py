"""Scattered-entry split on the GPU: y from the stream plan of the other
entries plus the sweep kernel, checked against the oracle's FP64 SpMV of the
matrix (FP64 plans: relative L2 <= 1e-12; FP32 and BF16 plans: of the
matrix with the values rounded as stored, <= 1e-6), on matrices with empty
rows, rows whose
entries are all scattered, a row of 9000 scattered entries (a block alone),
n of 2^25 (128-row blocks), and an all-scattered matrix (empty near plan).
The sweep sums every row in a fixed order: repeated launches are identical.
Replacing values (input or stored order) drops the split and the full
stream carries the new values; a batch of split plans runs their full
streams; a captured graph replays the two launches."""
import numpy as np
import pytest

from cbspmv import Batch, Plan
from oracle import pyoracle
from test_gpu_precision import round_values
from test_split_layout import scattered_matrix


def _empty_and_far_rows(m, n, row, col, val, seed):
    """Drops every entry of some rows (empty rows) and the near-diagonal
    entries of others (rows with scattered entries only)."""
    rng = np.random.default_rng(seed)
    empty = rng.random(m) < 0.05
    far_only = rng.random(m) < 0.05
    near = np.abs(col.astype(np.int64) - row) <= 200
    keep = ~empty[row] & ~(far_only[row] & near)
    return m, n, row[keep], col[keep], val[keep]


def _cases():
    yield "scattered", scattered_matrix(50000, 50000, seed=11)
    yield "empty_and_far_rows", _empty_and_far_rows(*scattered_matrix(50000, 80000, seed=12), seed=13)
    yield "long_row", scattered_matrix(20000, 1500000, seed=14, long_row=9000)
    rng = np.random.default_rng(15)
    m, n = 30000, 1 << 25                       # 25 column bits: blocks of 128 rows
    r = np.sort(rng.integers(0, m, 200000)).astype(np.int32)
    c = rng.integers(0, n, r.size).astype(np.int32)
    yield "n_2e25", (m, n, r, c, rng.uniform(-1, 1, r.size))
    m = n = 40000                               # every entry scattered: the near plan is empty
    r = np.arange(m, dtype=np.int32)
    yield "all_scattered", (m, n, r, ((r.astype(np.int64) * 7919) % n).astype(np.int32), rng.uniform(-1, 1, m))


CASES = dict(_cases())


def _x(n, seed, dtype):
    import torch

    x = np.random.default_rng(seed).uniform(-1, 1, n)
    return x, torch.tensor(x, dtype=torch.float64 if dtype == "f64" else torch.float32, device="cuda:0")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("name", list(CASES))
def test_split_spmv(gpu, name, dtype):
    import torch

    m, n, r, c, v = CASES[name]
    print(f"split {name} {dtype}: m={m} n={n} nnz={r.size}", flush=True)
    p = Plan(m, n, r, c, v, dtype=dtype, split="on")
    st = p.stats()
    assert st["split_entries"] > 0
    x, xd = _x(n, 1, dtype)
    y = p.spmv(xd)
    y2 = p.spmv(xd, torch.empty_like(y))
    torch.cuda.synchronize()
    assert torch.equal(y, y2), "repeated launches differ"
    vv = v if dtype == "f64" else round_values(v, dtype)
    xx = x if dtype == "f64" else x.astype(np.float32).astype(np.float64)
    err = pyoracle.rel_l2(y.cpu().numpy().astype(np.float64), pyoracle.spmv_seq(m, r, c, vv, xx))
    p.close()
    assert err <= (1e-12 if dtype == "f64" else 1e-6), err


@pytest.mark.gpu
def test_split_matches_unsplit_and_values_drop_it(gpu):
    import torch

    m, n, r, c, v = CASES["empty_and_far_rows"]
    p = Plan(m, n, r, c, v, split="on")
    q = Plan(m, n, r, c, v, split="off")
    assert p.stats()["split_entries"] > 0 and q.stats()["split_entries"] == 0
    x, xd = _x(n, 2, "f64")
    ys, yq = p.spmv(xd), q.spmv(xd)
    torch.cuda.synchronize()
    assert pyoracle.rel_l2(ys.cpu().numpy(), yq.cpu().numpy()) <= 1e-14
    v2 = np.random.default_rng(3).uniform(-1, 1, v.size)
    p.set_values(torch.tensor(v2, device=gpu))
    assert p.stats()["split_entries"] == 0
    y = p.spmv(xd)
    torch.cuda.synchronize()
    assert pyoracle.rel_l2(y.cpu().numpy(), pyoracle.spmv_seq(m, r, c, v2, x)) <= 1e-12
    p.close()
    q.close()
    # values in stored order (the transposed plan's) drop the split as well
    p = Plan(m, n, r, c, v, split="on")
    p.build_transpose()
    p.spmv(xd)
    assert p.stats()["split_entries"] > 0
    slot = p.export("grad_slot").astype(np.int64)
    v3 = np.random.default_rng(4).uniform(-1, 1, v.size)
    stored = np.zeros(p.stored_entries())
    np.add.at(stored, slot, v3)
    p.set_values_stored(torch.tensor(stored, device=gpu))
    assert p.stats()["split_entries"] == 0
    y = p.spmv(xd)
    torch.cuda.synchronize()
    assert pyoracle.rel_l2(y.cpu().numpy(), pyoracle.spmv_seq(m, r, c, v3, x)) <= 1e-12
    p.close()


@pytest.mark.gpu
def test_split_plans_in_a_batch_and_a_graph(gpu):
    import torch

    mats = [CASES["scattered"], CASES["long_row"]]
    plans = [Plan(*mat, split="on") for mat in mats]
    assert all(p.stats()["split_entries"] > 0 for p in plans)
    X = [_x(p.n, 4 + i, "f64")[1].reshape(1, -1).repeat(2, 1) for i, p in enumerate(plans)]
    Y = [torch.zeros((2, p.m), dtype=torch.float64, device=gpu) for p in plans]
    b = Batch(plans)
    b.spmv(X, Y)
    torch.cuda.synchronize()
    for (m, n, r, c, v), x, y in zip(mats, X, Y):
        ref = pyoracle.spmv_seq(m, r, c, v, x[0].cpu().numpy())
        for k in range(2):
            assert pyoracle.rel_l2(y[k].cpu().numpy(), ref) <= 1e-12
    b.close()
    # the split plan's two launches in a captured graph, replayed with eager work between
    p = plans[0]
    x = X[0][0].clone()
    eager = p.spmv(x)
    y = torch.zeros_like(eager)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p.spmv(x, y, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p.spmv(x, y)
    for rep in range(3):
        y.fill_(float(rep))
        junk = torch.randn(1000, device=gpu).sum()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), rep
        del junk
    for p in plans:
        p.close()
