"""GPU parity of the stream schedule with column panels (x split into
column slices gathered on separate XCDs, panel sums added per row by each
tile after the panel kernel).

FP64 against the oracle's sequential SpMV at relative L2 <= 1e-12; FP32 and
BF16 storage against the exact product of the stored matrix with the
summation bound of the precision tests; the same bits on every launch; the
fused backward (SDDMM over a transposed plan that has panels) gives dvals
exactly and dx at 1e-12; cbspmv_set_values reaches the panel entries."""
import numpy as np
import pytest

from cbspmv import Plan
from oracle import pyoracle
from test_gpu_precision import round_values
from test_stream_layout import decode_stream
from test_stream_panels import _mixed

TOL = 1e-12


def _cases():
    yield "mixed_60k", _mixed(60000, 60000, 3, 300, seed=31)
    yield "wide_1M", _mixed(40000, 1 << 20, 4, 200, seed=32)        # 8 MiB of x: panels by default
    yield "odd_n", _mixed(30011, 77777, 5, 100, seed=33)            # panel width not dividing n
    m, n, r, c, v = _mixed(20000, 20000, 2, 50, seed=34)
    rng = np.random.default_rng(34)
    keep = rng.random(r.size) > 0.5 * (r % 7 == 0)                   # sparser rows, some empty
    keep &= r % 11 != 5
    yield "empty_rows", (m, n, r[keep], c[keep], v[keep])


CASES = dict(_cases())


def _t(a, dev, dtype=None):
    import torch

    return torch.tensor(a, device=dev, dtype=dtype or torch.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("panels", [0, 2, 4, 8])
def test_panels_fp64_parity(gpu, case, panels):
    import torch

    m, n, r, c, v = CASES[case]
    print(f"panels case {case} P={panels}: m={m} n={n} nnz={r.size}", flush=True)
    x = np.random.default_rng(1).uniform(-1, 1, n)
    y_ref = pyoracle.spmv_seq(m, r, c, v, x)
    plan = Plan(m, n, r, c, v, panels=panels)
    if panels:
        assert int(plan.export("stream_panels")[0]) == panels
    xd = _t(x, gpu)
    y0 = plan.spmv(xd).clone()
    torch.cuda.synchronize()
    err = pyoracle.rel_l2(y0.cpu().numpy(), y_ref)
    assert err <= TOL, f"rel_l2 {err:.3e}"
    for _ in range(5):   # same bits every time
        assert torch.equal(plan.spmv(xd), y0)
    plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_panels_reduced_precision(gpu, dtype):
    m, n, r, c, v = CASES["mixed_60k"]
    print(f"panels {dtype}", flush=True)
    x32 = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32)
    plan = Plan(m, n, r, c, v, dtype=dtype, panels=4)
    sr, sc, sv, pad, *_ = decode_stream(plan, dtype)
    sr, sc, vr = sr[~pad], sc[~pad], round_values(sv[~pad], dtype)
    xd = x32.astype(np.float64)
    y_star = np.zeros(m)
    np.add.at(y_star, sr, vr * xd[sc])
    mag = np.zeros(m)
    np.add.at(mag, sr, np.abs(vr * xd[sc]))
    y = plan.spmv(_t(x32, gpu, __import__("torch").float32)).cpu().numpy().astype(np.float64)
    plan.close()
    lens = np.bincount(sr, minlength=m)
    bound = 2.0 ** -24 * (40 + lens / 128) * mag
    assert np.max(np.abs(y - y_star) - bound) <= 0
    assert pyoracle.rel_l2(y, y_star) <= 1e-6


@pytest.mark.gpu
def test_panels_backward_fused_and_set_values(gpu):
    import torch

    m, n, r, c, v = CASES["mixed_60k"]
    p = Plan(m, n, r, c, v, panels=4, backward="fused")
    p.build_transpose()
    assert int(p.export("transpose.stream_panels")[0]) == 4
    p.upload()
    rng = np.random.default_rng(3)
    g, x = rng.uniform(-1, 1, m), rng.uniform(-1, 1, n)
    dx, dv = p.backward(_t(g, gpu), _t(x, gpu))
    o = np.argsort(c, kind="stable")
    dx_ref = pyoracle.spmv_seq(n, c[o], r[o], v[o], g)
    assert pyoracle.rel_l2(dx.cpu().numpy(), dx_ref) <= TOL
    assert np.array_equal(dv.cpu().numpy(), g[r] * x[c])
    buf = torch.empty(p.stored_entries(), dtype=torch.float64, device=gpu)
    p.backward_stored(_t(g, gpu), _t(x, gpu), buf)
    gs = p.export("grad_slot").astype(np.int64)
    assert np.array_equal(buf.cpu().numpy()[gs], g[r] * x[c])
    v2 = rng.uniform(-2, 2, r.size)
    p.set_values(_t(v2, gpu))
    y = p.spmv(_t(x, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v2, x)) <= TOL
    dx2 = p.spmv_transpose(_t(g, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(dx2, pyoracle.spmv_seq(n, c[o], r[o], v2[o], g)) <= TOL
    p.close()


@pytest.mark.gpu
def test_panels_match_whole_x(gpu):
    """Panels change only the order of the sums."""
    m, n, r, c, v = CASES["wide_1M"]
    x = _t(np.random.default_rng(4).uniform(-1, 1, n), gpu)
    a = Plan(m, n, r, c, v, panels=1).spmv(x).cpu().numpy()
    b = Plan(m, n, r, c, v, panels=8).spmv(x).cpu().numpy()
    assert pyoracle.rel_l2(a, b) <= 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("panels", [2, 8])
def test_panels_small_cache_and_edges(gpu, panels):
    """A 4 KB LDS cache (31 lines) leaves most entries to the panels; few
    tiles (a grid of 8 workgroups); an empty matrix and one row."""
    import torch

    from cbspmv import gen

    m, n = 3000, 20000
    r, c, v = gen.random_coo(m, n, 7, seed=panels)
    x = np.random.default_rng(9).uniform(-1, 1, n)
    p = Plan(m, n, r, c, v, panels=panels, tile_kib=4)
    assert int(p.export("stream_panels")[0]) == panels
    y = p.spmv(_t(x, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(m, r, c, v, x)) <= TOL
    p.close()
    e = Plan(40, 30, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), panels=panels)
    y = e.spmv(_t(np.ones(30), gpu), _t(np.full(40, 7.0), gpu))
    assert torch.count_nonzero(y).item() == 0
    e.close()
    oc = np.sort(np.random.default_rng(1).choice(n, 50, replace=False)).astype(np.int32)
    ov = np.arange(50, dtype=np.float64)
    one = Plan(1, n, np.zeros(50, np.int32), oc, ov, panels=panels, tile_kib=4)
    y1 = one.spmv(_t(x, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(y1, pyoracle.spmv_seq(1, np.zeros(50, np.int32), oc, ov, x)) <= TOL
    one.close()
