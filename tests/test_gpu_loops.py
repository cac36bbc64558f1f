"""The stream kernel's two loops on the GPU: the ring of buffer loads (x
beyond one XCD's L2, 4 MiB) and the one-ahead loop (x within it), for every
value type, with cut rows and empty rows, and the fused backward (SDDMM over
A^T) on both.

Bars as elsewhere: FP64 within 1e-12 relative L2 of the oracle; FP32/BF16
storage within the native summation bound of the stored matrix (compensated:
2 ulp) and 1e-6 / 2^-8 relative L2 of the FP64 product; dvals exact (FP64)
or the FP32 product of the FP32 inputs (FP32 plans)."""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle
from test_gpu_precision import round_values
from test_stream_layout import decode_stream


def _wide(n, seed):
    """2500 rows over n columns: 20 random columns per row, one row of 30000
    entries (cut into pieces at item_nnz 4096), every 9th row empty."""
    m = 2500
    rng = np.random.default_rng(seed)
    r, c, v = gen.random_coo(m, n, 20, seed=seed)
    keep = r % 9 != 4
    r, c, v = r[keep], c[keep], v[keep]
    lc = np.sort(rng.choice(n, 30000, replace=False)).astype(np.int32)
    at = np.searchsorted(r, 1001)
    r = np.concatenate([r[:at], np.full(lc.size, 1001, np.int32), r[at:]])
    c = np.concatenate([c[:at], lc, c[at:]])
    v = np.concatenate([v[:at], rng.uniform(-1, 1, lc.size), v[at:]])
    return m, n, r, c, v


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(1 << 21) + 5, 200003], ids=["ring", "near"])
@pytest.mark.parametrize("dtype,comp", [("f64", False), ("f32", False), ("f32", True), ("bf16", False),
                                        ("bf16", True)])
def test_loops_all_types(gpu, n, dtype, comp):
    import torch

    m, n, r, c, v = _wide(n, seed=41)
    print(f"loop case n={n} {dtype} comp={comp}", flush=True)
    x64 = np.random.default_rng(5).uniform(-1, 1, n)
    plan = Plan(m, n, r, c, v, dtype=dtype, compensated=comp, item_nnz=4096)
    y_exact = pyoracle.spmv_seq(m, r, c, v, x64)
    if dtype == "f64":
        y = plan.spmv(torch.tensor(x64, device=gpu)).cpu().numpy()
        plan.close()
        assert pyoracle.rel_l2(y, y_exact) <= 1e-12
        return
    x32 = x64.astype(np.float32)
    sr, sc, sv, pad, *_ = decode_stream(plan, dtype)
    sr, sc, vr = sr[~pad], sc[~pad], round_values(sv[~pad], dtype)
    xd = x32.astype(np.float64)
    y_star = np.zeros(m)
    np.add.at(y_star, sr, vr * xd[sc])
    mag = np.zeros(m)
    np.add.at(mag, sr, np.abs(vr * xd[sc]))
    y = plan.spmv(torch.tensor(x32, device=gpu)).cpu().numpy().astype(np.float64)
    plan.close()
    if comp:
        bound = 2.0 ** -23 * np.abs(y_star) + 2.0 ** -44 * mag
    else:
        lens = np.bincount(sr, minlength=m)
        bound = 2.0 ** -24 * (40 + lens / 128) * mag
    assert np.max(np.abs(y - y_star) - bound) <= 0
    assert pyoracle.rel_l2(y, y_exact) <= (1e-6 if dtype == "f32" else 2.0 ** -8)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("m", [(1 << 21) + 3, 100003], ids=["ring", "near"])
def test_loops_backward_fused(gpu, dtype, m):
    """A tall matrix: A^T's x is g, of length m, so m picks the loop of the
    SDDMM launch."""
    import torch

    n = 3000
    r, c, v = gen.random_coo(m, n, 3, seed=43)
    print(f"backward loop case m={m} {dtype}", flush=True)
    p = Plan(m, n, r, c, v, dtype=dtype, backward="fused")
    p.build_transpose()
    p.upload()
    rng = np.random.default_rng(6)
    g, x = rng.uniform(-1, 1, m), rng.uniform(-1, 1, n)
    tdt = torch.float64 if dtype == "f64" else torch.float32
    dx, dv = p.backward(torch.tensor(g, device=gpu, dtype=tdt), torch.tensor(x, device=gpu, dtype=tdt))
    o = np.argsort(c, kind="stable")
    if dtype == "f64":
        dx_ref = pyoracle.spmv_seq(n, c[o], r[o], v[o], g)
        assert pyoracle.rel_l2(dx.cpu().numpy(), dx_ref) <= 1e-12
        assert np.array_equal(dv.cpu().numpy(), g[r] * x[c])
    else:
        g32, x32 = g.astype(np.float32), x.astype(np.float32)
        dx_ref = pyoracle.spmv_seq(n, c[o], r[o], v[o], g32.astype(np.float64))
        assert pyoracle.rel_l2(dx.cpu().numpy().astype(np.float64), dx_ref) <= 1e-6
        assert np.array_equal(dv.cpu().numpy(), g32[r] * x32[c])
    p.close()
