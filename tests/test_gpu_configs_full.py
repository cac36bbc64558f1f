"""GPU parity at the full sizes of the clustered mixed-precision matrix,
the 256-matrix batch and the differentiable power-law matrix, against the
oracle's FP64 SpMV (row-parallel CSR form of utils.h:98-111, built once per
matrix):

  * 2,097,152^2 block-clustered matrix, ~50M entries (64x64 clusters at
    density 0.05 plus 4 uniform entries per row, values N(0,1) 10^U(-3,3) per
    row): FP64 storage at relative L2 <= 1e-12; FP32 and BF16 storage,
    native and compensated accumulation, against the exact product of the
    stored (rounded) matrix with the summation bounds of
    test_gpu_precision.py, and against FP64 at <= 1e-6 (FP32) / 2^-8 (BF16);
  * 256 matrices (n log-uniform in [1e3, 2e5], 258M entries, 1174 FP32
    right-hand sides) in one hipGraph-captured batched launch: every
    right-hand side of every matrix at relative L2 <= 1e-6 against FP64,
    replays bit-identical to the eager launch;
  * 524,288^2 FP32 power-law matrix, ~20M entries: forward y, dx = A^T g and
    dvals = g_i x_j through torch.autograd against FP64 at <= 1e-6.

The generators run on the host (cbspmv/gen.py); each module fixture builds
its matrix once.
"""
import numpy as np
import pytest

from cbspmv import Batch, CBSparse, Plan, cb_spmv_autograd, gen
from oracle import pyoracle
from test_gpu_precision import round_values

THREADS = 16
pytestmark = [pytest.mark.gpu, pytest.mark.slow, pytest.mark.timeout(900)]


# ---- clustered 2M^2, ~50M entries -------------------------------------------

@pytest.fixture(scope="module")
def clustered():
    d = gen.clustered_2m()
    d["csr"] = pyoracle.CSR(d["m"], d["row"], d["col"], d["val"])
    d["y64"] = d["csr"].spmv(d["x"], THREADS)
    return d


def test_clustered_fp64(gpu, clustered):
    import torch

    d = clustered
    print(f"clustered f64: m={d['m']} nnz={d['row'].size}", flush=True)
    p = Plan(d["m"], d["n"], d["row"], d["col"], d["val"])
    st = p.stats()
    assert st["gather_flag"] == 1 and p.schedule() == "stream"
    xd = torch.tensor(d["x"], device=gpu)
    y = p.spmv(xd)
    y2 = p.spmv(xd, torch.empty_like(y))
    assert torch.equal(y, y2), "repeated launches differ"
    err = pyoracle.rel_l2(y.cpu().numpy(), d["y64"])
    p.close()
    assert err <= 1e-12, err


@pytest.mark.parametrize("dtype,comp", [("f32", False), ("f32", True), ("bf16", False), ("bf16", True)],
                         ids=["f32", "f32_comp", "bf16", "bf16_comp"])
def test_clustered_reduced_precision(gpu, clustered, dtype, comp):
    import torch

    d = clustered
    m, n, r, c, v = d["m"], d["n"], d["row"], d["col"], d["val"]
    print(f"clustered {dtype} comp={comp}", flush=True)
    # no duplicate (row, col) in this matrix: one stored entry per input entry
    x32 = d["x"].astype(np.float32)
    x32d = x32.astype(np.float64)
    vr = round_values(v, dtype)
    stored = pyoracle.CSR(m, r, c, vr)
    y_star = stored.spmv(x32d, THREADS)
    mag = pyoracle.CSR(m, r, c, np.abs(vr)).spmv(np.abs(x32d), THREADS)
    p = Plan(m, n, r, c, v, dtype=dtype, compensated=comp)
    y = p.spmv(torch.tensor(x32, device=gpu)).cpu().numpy().astype(np.float64)
    p.close()
    if comp:
        bound = 2.0 ** -23 * np.abs(y_star) + 2.0 ** -44 * mag
    else:
        lens = np.diff(stored.ptr)
        bound = 2.0 ** -24 * (40 + lens / 128) * mag
    worst = float(np.max(np.abs(y - y_star) - bound))
    assert worst <= 0, f"error exceeds the accumulation bound by {worst:.3e}"
    err = pyoracle.rel_l2(y, d["y64"])
    print(f"clustered {dtype} comp={comp}: rel_l2 vs FP64 {err:.3e}", flush=True)
    assert err <= (1e-6 if dtype == "f32" else 2.0 ** -8), err


# ---- 256-matrix batch in one captured launch ---------------------------------

@pytest.fixture(scope="module")
def batch256_mats():
    mats = gen.batch_256()
    print(f"batch: {len(mats)} matrices, {sum(int(d['row'].size) for d in mats)} nnz", flush=True)
    return mats


# default plans, and plans in the shape the autotuner gives a batch (8 tile
# waves, 2048-entry items, 32 or 4 KiB LDS caches: two workgroups per CU)
@pytest.mark.parametrize("shape", ["default", "batch"])
def test_batch_256_graph_replay(gpu, batch256_mats, shape):
    import torch

    mats = batch256_mats
    def opts(d):
        if shape == "default":
            return {}
        return dict(tile_waves=8, item_nnz=2048, tile_kib=4 if d["nrhs"] > 1 else 32)
    plans = [Plan(d["m"], d["n"], d["row"], d["col"], d["val"], dtype="f32", **opts(d)) for d in mats]
    X = [torch.tensor(d["X"], dtype=torch.float32, device=gpu) for d in mats]
    Y = [torch.full((d["nrhs"], d["m"]), float("nan"), dtype=torch.float32, device=gpu) for d in mats]
    eager = [torch.empty_like(y) for y in Y]
    b = Batch(plans)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.capture(X, Y, s)
    s.synchronize()
    b.spmv(X, eager)
    torch.cuda.synchronize()
    for rep in range(2):
        for y in Y:
            y.fill_(float(rep) - 5.0)
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        for e, y in zip(eager, Y):
            assert torch.equal(e, y), f"replay {rep} differs from the eager launch"
    worst, checked = 0.0, 0
    for d, y in zip(mats, Y):
        csr = pyoracle.CSR(d["m"], d["row"], d["col"], d["val"])
        ys = y.double().cpu().numpy()
        for k in range(d["nrhs"]):
            ref = csr.spmv(d["X"][k].astype(np.float64), THREADS)
            worst = max(worst, pyoracle.rel_l2(ys[k], ref))
            checked += 1
    b.close()
    for p in plans:
        p.close()
    print(f"batch ({shape}): {checked} right-hand sides, worst rel_l2 vs FP64 {worst:.3e}", flush=True)
    assert checked == sum(d["nrhs"] for d in mats)
    assert worst <= 1e-6, worst


# ---- differentiable 512K^2 FP32 ------------------------------------------------

@pytest.fixture(scope="module")
def powerlaw():
    d = gen.powerlaw_512k()
    r, c = d["row"], d["col"]
    o = np.argsort(c, kind="stable")
    d["csr"] = pyoracle.CSR(d["m"], r, c, d["val"])
    d["csr_t"] = pyoracle.CSR(d["n"], c[o], r[o], d["val"][o])
    return d


def test_powerlaw_autograd_fp32(gpu, powerlaw):
    import torch

    d = powerlaw
    m, n, r, c, v = d["m"], d["n"], d["row"], d["col"], d["val"]
    print(f"powerlaw autograd f32: m={m} nnz={r.size}", flush=True)
    A = CBSparse(m, n, r, c, v, dtype="f32")
    rng = np.random.default_rng(17)
    g64 = rng.uniform(-1, 1, m)
    x32, g32 = d["x"].astype(np.float32), g64.astype(np.float32)
    vals = torch.tensor(v.astype(np.float32), device=gpu, requires_grad=True)
    x = torch.tensor(x32, device=gpu, requires_grad=True)
    y = cb_spmv_autograd(A, x, vals)
    y.backward(torch.tensor(g32, device=gpu))
    xd, gd = x32.astype(np.float64), g32.astype(np.float64)
    err_y = pyoracle.rel_l2(y.detach().double().cpu().numpy(), d["csr"].spmv(xd, THREADS))
    err_dx = pyoracle.rel_l2(x.grad.double().cpu().numpy(), d["csr_t"].spmv(gd, THREADS))
    err_dv = pyoracle.rel_l2(vals.grad.double().cpu().numpy(), gd[r] * xd[c])
    print(f"powerlaw: y {err_y:.3e} dx {err_dx:.3e} dvals {err_dv:.3e}", flush=True)
    A.close()
    assert err_y <= 1e-6 and err_dx <= 1e-6 and err_dv <= 1e-6
