"""GPU parity of y = A x through the C ABI, both schedules, against the
oracle's sequential FP64 SpMV (cb-spmv-dcu/src/utils.h:98-111).

Bar: relative L2 error <= 1e-12 (FP64), as the target states; the kernels
sum in a different order than the sequential loop, so results are not
bit-identical, but every run of one plan is (checked separately).
"""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle

TOL = 1e-12


def _block_diag_dense(nb=40, bs=16, seed=3):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for b in range(nb):
        r, c = np.meshgrid(np.arange(bs), np.arange(bs), indexing="ij")
        rows.append((b * bs + r).ravel())
        cols.append((b * bs + c).ravel())
        # a few sparse off-diagonal entries
        k = 5
        rows.append(b * bs + rng.integers(0, bs, k))
        cols.append(rng.integers(0, nb * bs, k))
    R, C = np.concatenate(rows), np.concatenate(cols)
    o = np.lexsort((C, R))
    R, C = R[o], C[o]
    key = R * (nb * bs) + C
    keep = np.concatenate([[True], key[1:] != key[:-1]])
    R, C = R[keep], C[keep]
    return nb * bs, nb * bs, R.astype(np.int32), C.astype(np.int32), rng.uniform(-1, 1, R.size)


def _banded(m=2000, bw=6, seed=4):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for r in range(m):
        c = np.arange(max(0, r - bw), min(m, r + bw + 1))
        rows.append(np.full(c.size, r))
        cols.append(c)
    R, C = np.concatenate(rows), np.concatenate(cols)
    return m, m, R.astype(np.int32), C.astype(np.int32), rng.uniform(-1, 1, R.size)


def _uniform_rows(nbr=30, bs=16, n=900, seed=5):
    """Block rows whose 16 rows are identical (uniform-row blocks/segments)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for b in range(nbr):
        c = np.sort(rng.choice(n, size=rng.integers(3, 40), replace=False))
        v = rng.uniform(-1, 1, c.size)
        for r in range(bs):
            rows.append(np.full(c.size, b * bs + r))
            cols.append(c)
            vals.append(v)
    return nbr * bs, n, np.concatenate(rows).astype(np.int32), np.concatenate(cols).astype(np.int32), np.concatenate(vals)


def _long_row(m=300, n=60000, long_len=50000, seed=6):
    rng = np.random.default_rng(seed)
    r, c, v = gen.random_coo(m, n, 5, seed=seed)
    lc = np.sort(rng.choice(n, size=long_len, replace=False))
    at = np.searchsorted(r, 137)
    r = np.concatenate([r[:at], np.full(long_len, 137, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc.astype(np.int32), c[at:]])
    v = np.concatenate([v[:at], rng.uniform(-1, 1, long_len), v[at:]])
    return m, n, r, c, v


def _rand(m, n, per_row, **kw):
    r, c, v = gen.random_coo(m, n, per_row, **kw)
    return m, n, r, c, v


MATRICES = {
    "poisson_4096": lambda: (lambda d: (d["m"], d["n"], d["row"], d["col"], d["val"]))(gen.poisson_4096()),
    "rand_dup": lambda: _rand(500, 700, 3, seed=1, dup=True),
    "unit_dup": lambda: _rand(300, 300, 80, seed=2, dup=True, unit=True),
    "unit_sparse": lambda: _rand(3000, 3000, 4, seed=7, unit=True),
    "empty_rows": lambda: _rand(1001, 777, 6, seed=8, empty_rows=0.3),
    "dense_blocks": _block_diag_dense,
    "banded_csr": _banded,
    "uniform_rows": _uniform_rows,
    "long_row": _long_row,
    "n4095": lambda: _rand(2000, 4095, 3, seed=9),
    "n65535": lambda: _rand(4000, 65535, 20, seed=10),
    "n65536": lambda: _rand(4000, 65536, 20, seed=11),
    "nnz9999": lambda: (lambda m, n, r, c, v: (m, n, r[:9999], c[:9999], v[:9999]))(*_rand(1500, 3000, 8, seed=12)),
    "nnz10000": lambda: (lambda m, n, r, c, v: (m, n, r[:10000], c[:10000], v[:10000]))(*_rand(1500, 3000, 8, seed=12)),
    # 22-bit columns leave 9 bits of row offset in a stream key: items span
    # at most 511 rows
    "wide_4M": lambda: _rand(2000, 1 << 22, 20, seed=13),
    "tall_2M": lambda: _rand(1 << 21, 64, 2, seed=14, empty_rows=0.1),
}

OPTS = {
    "ro_bs16": dict(),
    "lpt_bs16": dict(row_owned=False),
    "ro_bs32": dict(block_size=32),
    "lpt_bs32_w8": dict(block_size=32, waves=8, row_owned=False),
    "ro_bs8": dict(block_size=8),
    "ro_bs64": dict(block_size=64),
    "ro_bs4_w8": dict(block_size=4, waves=8),
    "lpt_bs2": dict(block_size=2, row_owned=False),
    "comp_bs16": dict(builder="comp"),
    "items64": dict(item_nnz=64),
}


def _x_for(n, seed=42):
    return np.random.default_rng(seed).uniform(-1, 1, n)


@pytest.mark.gpu
@pytest.mark.parametrize("mat", sorted(MATRICES))
@pytest.mark.parametrize("opt", sorted(OPTS))
def test_spmv_parity(gpu, mat, opt):
    import torch

    m, n, r, c, v = MATRICES[mat]()
    x = _x_for(n)
    y_ref = pyoracle.spmv_seq(m, r, c, v, x)
    plan = Plan(m, n, r, c, v, **OPTS[opt])
    xd = torch.tensor(x, dtype=torch.float64, device=gpu)
    print(f"case {mat}/{opt}: m={m} n={n} nnz={r.size}", flush=True)
    for sched in ("stream", "blocked"):
        plan.set_schedule(sched)
        y = plan.spmv(xd)
        torch.cuda.synchronize()
        err = pyoracle.rel_l2(y.cpu().numpy(), y_ref)
        assert err <= TOL, f"{mat}/{opt}/{sched}: rel_l2 {err:.3e}"
    plan.close()


@pytest.mark.gpu
def test_repeat_launches_identical(gpu):
    """Same plan, same x: every launch gives the same bits (no atomics in the
    stream schedule's sums; split strips are summed in piece order)."""
    import torch

    m, n, r, c, v = _long_row()
    plan = Plan(m, n, r, c, v, item_nnz=256)
    xd = torch.tensor(_x_for(n), dtype=torch.float64, device=gpu)
    y0 = plan.spmv(xd).clone()
    for _ in range(20):
        y1 = plan.spmv(xd)
        assert torch.equal(y0, y1)
    plan.close()


@pytest.mark.gpu
def test_nnz_zero_and_empty(gpu):
    import torch

    plan = Plan(50, 40, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
    y = plan.spmv(torch.ones(40, dtype=torch.float64, device=gpu), torch.full((50,), 7.0, dtype=torch.float64, device=gpu))
    torch.cuda.synchronize()
    assert torch.count_nonzero(y).item() == 0
    plan.close()


@pytest.mark.gpu
def test_bench_protocol_matches_oracle(gpu):
    """cbspmv_bench (cb_spmv protocol): first-launch y is returned."""
    d = gen.poisson_4096()
    plan = Plan(d["m"], d["n"], d["row"], d["col"], d["val"])
    y, runtime, gflops = plan.bench(d["x"])
    y_ref = pyoracle.spmv_seq(d["m"], d["row"], d["col"], d["val"], d["x"])
    assert pyoracle.rel_l2(y, y_ref) <= TOL
    assert runtime > 0 and gflops > 0
    plan.close()
