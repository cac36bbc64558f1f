"""Reduced-precision storage of A (FP32, BF16) with FP32 vectors, native or
compensated (two-float) accumulation, on the stream schedule.

The reference stores FP64 only (cb-spmv-dcu/src/common.h:7), so these modes
have no reference output; they are checked against the oracle's FP64 SpMV
in two ways:
  * against the exact product of the stored matrix, y* = A_r x_r, computed
    in FP64 from the plan's stored entries (duplicates that the CB format
    merges are summed in FP64 first) rounded as the packer rounds them
    (FP64 -> FP32, then round-to-nearest-even to BF16), and x rounded to
    FP32: this isolates the accumulation error.  Native FP32 accumulation: every element within
    2^-24 (40 + len_i / 128) sum_j |a_ij x_j| (a first-order bound for the
    kernel's summation tree: 8 entries per lane, a 64-lane scan, chunks of
    512 entries and row pieces added in sequence), and relative L2 <= 1e-6
    where no row is longer than 4096 entries.  Compensated: every element
    within 2 ulp of FP32 of y*, plus 2^-44 of sum |a_ij x_j| for rows that
    cancel;
  * against the FP64 result on FP64 inputs: relative L2 <= 1e-6 for FP32
    storage, <= 2^-8 for BF16 storage (8-bit significand).
"""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle
from test_stream_layout import decode_stream


def round_values(v, dtype):
    f = np.asarray(v, np.float64).astype(np.float32)
    if dtype == "f32":
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    h = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32)
    return (h << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def _poisson_4096():
    d = gen.poisson_4096()
    return d["m"], d["n"], d["row"], d["col"], d["val"]


def _rand(m, n, per_row, seed, **kw):
    r, c, v = gen.random_coo(m, n, per_row, seed=seed, **kw)
    return m, n, r, c, v


def _scaled(seed=21):
    """Values spanning 6 decades, one scale per row (the 2,097,152^2 clustered matrix's value law)."""
    m, n, r, c, v = _rand(4000, 4000, 25, seed)
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-3, 3, m)
    return m, n, r, c, rng.standard_normal(r.size) * scale[r]


def _long_row():
    m, n = 300, 60000
    r, c, v = gen.random_coo(m, n, 5, seed=6)
    rng = np.random.default_rng(6)
    lc = np.sort(rng.choice(n, size=40000, replace=False)).astype(np.int32)
    at = np.searchsorted(r, 137)
    r = np.concatenate([r[:at], np.full(lc.size, 137, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc, c[at:]])
    v = np.concatenate([v[:at], rng.uniform(-1, 1, lc.size), v[at:]])
    return m, n, r, c, v


MATRICES = {
    "poisson_4096": _poisson_4096,
    "rand_wide": lambda: _rand(3000, 50000, 12, 22),
    "empty_rows": lambda: _rand(2001, 1777, 9, 23, empty_rows=0.3),
    "scaled": _scaled,
    "long_row": _long_row,
    "unit": lambda: _rand(1500, 1500, 30, 24, unit=True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("mat", sorted(MATRICES))
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("compensated", [False, True], ids=["native", "comp"])
def test_reduced_precision(gpu, mat, dtype, compensated):
    import torch

    m, n, r, c, v = MATRICES[mat]()
    print(f"case {mat}/{dtype}/{'comp' if compensated else 'native'}: m={m} n={n} nnz={r.size}", flush=True)
    x64 = np.random.default_rng(99).uniform(-1, 1, n)
    x32 = x64.astype(np.float32)
    plan = Plan(m, n, r, c, v, dtype=dtype, compensated=compensated)
    sr, sc, sv, pad, *_ = decode_stream(plan, dtype)
    sr, sc, vr = sr[~pad], sc[~pad], round_values(sv[~pad], dtype)
    x32d = x32.astype(np.float64)
    y_star = np.zeros(m)
    np.add.at(y_star, sr, vr * x32d[sc])
    mag = np.zeros(m)
    np.add.at(mag, sr, np.abs(vr * x32d[sc]))
    y_exact = pyoracle.spmv_seq(m, r, c, v, x64)

    assert plan.schedule() == "stream"
    y = plan.spmv(torch.tensor(x32, device=gpu)).cpu().numpy().astype(np.float64)
    plan.close()

    if compensated:
        bound = 2.0 ** -23 * np.abs(y_star) + 2.0 ** -44 * mag
        worst = np.max(np.abs(y - y_star) - bound)
        assert worst <= 0, f"compensated error exceeds 2 ulp by {worst:.3e}"
    else:
        lens = np.bincount(sr, minlength=m)
        bound = 2.0 ** -24 * (40 + lens / 128) * mag
        worst = np.max(np.abs(y - y_star) - bound)
        assert worst <= 0, f"native error exceeds the summation bound by {worst:.3e}"
        if lens.max() <= 4096:
            assert pyoracle.rel_l2(y, y_star) <= 1e-6
    tol = 1e-6 if dtype == "f32" else 2.0 ** -8
    assert pyoracle.rel_l2(y, y_exact) <= tol


@pytest.mark.gpu
def test_compensated_beats_native_on_cancellation(gpu):
    """Rows whose terms cancel: compensated FP32 accumulation keeps the
    result close to FP64, native FP32 does not."""
    import torch

    m, n = 512, 4096
    rng = np.random.default_rng(5)
    rows, cols, vals = [], [], []
    for i in range(m):
        cc = np.sort(rng.choice(n, 400, replace=False))
        big = rng.uniform(1e3, 1e4, 200)
        vv = np.empty(400)
        vv[0::2] = big
        vv[1::2] = -big  # pairs cancel against x[c] ~ x[c'] below
        rows.append(np.full(400, i))
        cols.append(cc)
        vals.append(vv + rng.uniform(-1e-3, 1e-3, 400))
    r = np.concatenate(rows).astype(np.int32)
    c = np.concatenate(cols).astype(np.int32)
    v = np.concatenate(vals)
    x32 = np.ones(n, np.float32)
    vr = round_values(v, "f32")   # no duplicates: one stored entry per input entry
    y_star = pyoracle.spmv_seq(m, r, c, vr, x32.astype(np.float64))
    errs = {}
    for comp in (False, True):
        plan = Plan(m, n, r, c, v, dtype="f32", compensated=comp)
        y = plan.spmv(torch.tensor(x32, device=gpu)).cpu().numpy().astype(np.float64)
        errs[comp] = np.max(np.abs(y - y_star))
        plan.close()
    assert errs[True] <= 2.0 ** -23 * np.max(np.abs(y_star)) + 1e-9
    assert errs[True] < errs[False]


@pytest.mark.gpu
def test_blocked_schedule_rejects_reduced_precision(gpu):
    from cbspmv import CbspmvError

    m, n, r, c, v = _rand(100, 100, 5, 1)
    plan = Plan(m, n, r, c, v, dtype="f32")
    with pytest.raises(CbspmvError):
        plan.set_schedule("blocked")
    plan.close()
