"""Batched persistent launch, hipGraph capture/replay and the autotuner
(extension; no reference counterpart).  Every right-hand side of every
matrix is checked against the oracle's FP64 SpMV of the matrix the plan
stores (FP64 plans: the input, relative L2 <= 1e-12; FP32 and BF16 plans:
their stored entries with the stored rounding, <= 1e-6, or 5e-6 on the
matrix with a 30,000-entry row), on uneven matrix sizes including an empty
matrix and a matrix whose long row is cut into pieces, with 1-16
right-hand sides.
Graph replays are checked against the eager result several times, with
other eager work between them."""
import numpy as np
import pytest

from cbspmv import Batch, Plan, autotune, gen
from oracle import pyoracle
from test_gpu_precision import round_values
from test_stream_layout import decode_stream


def stored_matrix(plan, m, n, r, c, v, dtype):
    """The matrix a plan computes with: the input for FP64 plans; for FP32 /
    BF16 plans its stored entries with values rounded as stored, so the
    check measures the kernel, not the storage rounding."""
    if dtype == "f64":
        return m, n, r, c, v
    sr, sc, sv, pad, *_ = decode_stream(plan, dtype)
    keep = ~pad
    return m, n, sr[keep].astype(np.int32), sc[keep].astype(np.int32), round_values(sv[keep], dtype)


def _mats(seed=0, count=9):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(300), np.log(20000))))
        m = n if i % 3 else int(n * rng.uniform(0.5, 1.5)) + 1
        r, c, v = gen.random_coo(m, n, int(rng.integers(2, 30)), seed=seed * 100 + i, dup=bool(i % 2))
        out.append((m, n, r, c, v))
    out.append((50, 40, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0)))
    r, c, v = gen.random_coo(200, 50000, 3, seed=77)
    lc = np.sort(rng.choice(50000, 30000, replace=False)).astype(np.int32)
    at = np.searchsorted(r, 100)
    r = np.concatenate([r[:at], np.full(lc.size, 100, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc, c[at:]])
    v = np.concatenate([v[:at], rng.uniform(-1, 1, lc.size), v[at:]])
    out.append((200, 50000, r, c, v))
    return out


def _check(mats, X, Y, tol):
    """tol applies to matrices whose rows hold at most 4096 entries; FP32
    accumulation over the 30,000-entry row gets 5x that."""
    for (m, n, r, c, v), x, y in zip(mats, X, Y):
        xs, ys = x.cpu().numpy().astype(np.float64), y.cpu().numpy().astype(np.float64)
        long_rows = r.size > 0 and np.bincount(r).max() > 4096
        for k in range(xs.shape[0]):
            ref = pyoracle.spmv_seq(m, r, c, v, xs[k])
            if m == 0:
                continue
            err = pyoracle.rel_l2(ys[k], ref)
            assert err <= (5 * tol if long_rows and tol > 1e-9 else tol), (m, n, k, err)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,comp", [("f64", False), ("f32", False), ("f32", True), ("bf16", True)])
def test_batch_parity(gpu, dtype, comp):
    import torch

    mats = _mats()
    tdt = torch.float64 if dtype == "f64" else torch.float32
    rng = np.random.default_rng(1)
    plans, X, Y = [], [], []
    # tiles of more than 8 items (9, 12, 16) are dealt to the batch kernel's
    # 8 waves in turn, on both the one- and the several-right-hand-side path;
    # 3, 5, 6, 8 and 16 right-hand sides fill one group of eight partly (one
    # or two 16-byte vectors in FP32), exactly, and twice
    waves = [4, 9, 16, 12, 8, 16, 4, 12, 9, 12, 16]
    rhs = [1, 3, 1, 16, 8, 5, 16, 1, 6, 16, 1]
    for i, (m, n, r, c, v) in enumerate(mats):
        opts = dict(dtype=dtype, compensated=comp, tile_waves=waves[i % 11], item_nnz=[0, 256, 0][i % 3])
        plans.append(Plan(m, n, r, c, v, **opts))
        k = rhs[i % 11]
        X.append(torch.tensor(rng.uniform(-1, 1, (k, n)), dtype=tdt, device=gpu))
        Y.append(torch.full((k, m), 7.0, dtype=tdt, device=gpu))
    print(f"batch {dtype}: {len(plans)} plans", flush=True)
    b = Batch(plans)
    b.spmv(X, Y)
    torch.cuda.synchronize()
    tol = 1e-12 if dtype == "f64" else 1e-6
    ref = [stored_matrix(p, *mat, dtype) for p, mat in zip(plans, mats)]
    _check(ref, X, Y, tol)
    # second call with the same pointers (no descriptor upload), new values
    for x in X:
        x.uniform_(-1, 1)
    b.spmv(X, Y)
    torch.cuda.synchronize()
    _check(ref, X, Y, tol)
    b.close()
    for p in plans:
        p.close()


@pytest.mark.gpu
def test_batch_graph_replay(gpu):
    import torch

    mats = _mats(seed=3, count=6)
    rng = np.random.default_rng(4)
    plans = [Plan(m, n, r, c, v, dtype="f32") for (m, n, r, c, v) in mats]
    X = [torch.tensor(rng.uniform(-1, 1, (2, p.n)), dtype=torch.float32, device=gpu) for p in plans]
    Y = [torch.zeros((2, p.m), dtype=torch.float32, device=gpu) for p in plans]
    b = Batch(plans)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.capture(X, Y, s)
    s.synchronize()
    eager = [torch.empty_like(y) for y in Y]
    b.spmv(X, eager)
    torch.cuda.synchronize()
    for e, y in zip(eager, Y):
        assert torch.equal(e, y)
    _check(mats, X, Y, 1e-6)
    for rep in range(4):
        for y in Y:
            y.fill_(float(rep + 1))
        junk = torch.randn(1000, device=gpu).sum()  # eager work between replays
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        for e, y in zip(eager, Y):
            assert torch.equal(e, y), rep
        del junk
    b.close()
    for p in plans:
        p.close()


@pytest.mark.gpu
def test_batch_buffers_follow_rhs_counts(gpu):
    """The interleaved-x buffers grow with the right-hand-side counts of a
    call; a captured graph keeps its own, so eager calls with more
    right-hand sides after a capture leave its replays correct.  The same
    plan may appear several times in one batch (the autotuner times plans
    that way), each copy with its own vectors."""
    import torch

    mats = _mats(seed=5, count=3)[:3]
    plans = [Plan(m, n, r, c, v, dtype="f32") for (m, n, r, c, v) in mats]
    rng = np.random.default_rng(8)

    def vecs(k, plans_, mats_):
        X = [torch.tensor(rng.uniform(-1, 1, (k, p.n)), dtype=torch.float32, device=gpu) for p in plans_]
        Y = [torch.full((k, p.m), 3.0, dtype=torch.float32, device=gpu) for p in plans_]
        return X, Y

    b = Batch(plans)
    for k in (2, 5):   # eager: the buffer grows between calls
        X, Y = vecs(k, plans, mats)
        b.spmv(X, Y)
        torch.cuda.synchronize()
        _check(mats, X, Y, 1e-6)
    s = torch.cuda.Stream()
    Xg, Yg = vecs(3, plans, mats)
    with torch.cuda.stream(s):
        b.capture(Xg, Yg, s)
    s.synchronize()
    _check(mats, Xg, Yg, 1e-6)
    X, Y = vecs(13, plans, mats)   # eager call after the capture, more right-hand sides
    b.spmv(X, Y)
    torch.cuda.synchronize()
    _check(mats, X, Y, 1e-6)
    for y in Yg:
        y.fill_(-1.0)
    torch.cuda.synchronize()
    b.replay(s)
    s.synchronize()
    _check(mats, Xg, Yg, 1e-6)
    b.close()

    same = [plans[0]] * 3
    X, Y = vecs(6, same, [mats[0]] * 3)
    b = Batch(same)
    b.spmv(X, Y)
    torch.cuda.synchronize()
    _check([mats[0]] * 3, X, Y, 1e-6)
    b.close()
    for p in plans:
        p.close()


@pytest.mark.gpu
def test_autotune_returns_a_correct_plan(gpu):
    import torch

    m = n = 20000
    r, c, v = gen.random_coo(m, n, 12, seed=5)
    p, ms = autotune(m, n, r, c, v, nrhs=2, dtype="f32")
    assert ms > 0
    waves, lds = p.export("stream_tile_shape")
    assert (int(waves), int(lds)) == (8, 4096)   # the batch shape with several right-hand sides
    x = np.random.default_rng(6).uniform(-1, 1, n)
    y = p.spmv(torch.tensor(x, dtype=torch.float32, device=gpu)).cpu().numpy().astype(np.float64)
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v, x.astype(np.float32).astype(np.float64))) <= 1e-6
    p.close()
    # explicit candidates and base options are honoured, and the plan runs
    # in a batch
    p, ms = autotune(m, n, r, c, v, nrhs=5, candidates=[(32, 9)], dtype="f32", tile_waves=9)
    assert int(p.export("stream_tile_shape")[0]) == 9
    assert p.stats()["gb_m"] == (m + 31) // 32
    X = torch.tensor(np.random.default_rng(7).uniform(-1, 1, (5, n)), dtype=torch.float32, device=gpu)
    Y = torch.zeros((5, m), dtype=torch.float32, device=gpu)
    b = Batch([p])
    b.spmv([X], [Y])
    torch.cuda.synchronize()
    _check([(m, n, r, c, v)], [X], [Y], 1e-6)
    b.close()
    p.close()
    # without DENSE blocks both block sizes give one stream plan: the second
    # candidate is not timed, whichever would have measured faster (a fixed
    # item size leaves out the item-size stage)
    for first in (16, 32):
        other = 48 - first
        p, _ = autotune(m, n, r, c, v, nrhs=1, candidates=[(first, 9), (other, 9)], dtype="f32", item_nnz=1024)
        assert p.stats()["dense_blocks"] == 0
        assert p.stats()["gb_m"] == (m + first - 1) // first
        p.close()
    # options left open get the batch shape: items of about 2048 entries
    # (the automatic size here would be about 1024), 8 tile waves and a
    # 32 KiB LDS cache with one right-hand side; the plan is correct in a
    # batch with several
    p, _ = autotune(m, n, r, c, v, nrhs=1, candidates=[(16, 16)], dtype="f32")
    items = p.export("stream_items").reshape(-1, 4)
    mean = float((items[:, 1] - items[:, 0]).mean())
    assert 1800 <= mean <= 2600, mean
    assert [int(t) for t in p.export("stream_tile_shape")] == [8, 32768]
    X = torch.tensor(np.random.default_rng(8).uniform(-1, 1, (3, n)), dtype=torch.float32, device=gpu)
    Y = torch.zeros((3, m), dtype=torch.float32, device=gpu)
    b = Batch([p])
    b.spmv([X], [Y])
    torch.cuda.synchronize()
    _check([(m, n, r, c, v)], [X], [Y], 1e-6)
    b.close()
    p.close()
