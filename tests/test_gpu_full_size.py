"""Parity at the benchmark's full size: the 1,048,576^2 matrix with 30.2M
entries (stencil half + power-law half) that bench.py measures.

At this size the oracle's row-parallel CSR SpMV (same per-row order as the
reference's sequential loop, utils.h:98-111) still finishes in well under a
second, so y is compared with it directly (relative L2 <= 1e-12 for FP64;
FP32/BF16 storage as in test_gpu_precision.py).  Size-independent properties
are checked as well: linearity, the column-sum checksum 1^T (A x) =
(1^T A) x, and bit-identical repeated launches.
"""
import numpy as np
import pytest

from cbspmv import Plan, gen
from oracle import pyoracle

THREADS = 16


@pytest.fixture(scope="module")
def stencil_powerlaw():
    d = gen.stencil_powerlaw_1m()
    return d["m"], d["n"], d["row"], d["col"], d["val"], d["x"]


@pytest.fixture(scope="module")
def plan64(stencil_powerlaw):
    m, n, r, c, v, _ = stencil_powerlaw
    p = Plan(m, n, r, c, v)
    yield p
    p.close()


@pytest.mark.gpu
@pytest.mark.slow
def test_full_size_fp64_stream(gpu, stencil_powerlaw, plan64):
    import torch

    m, n, r, c, v, x = stencil_powerlaw
    print(f"stencil_powerlaw stream: m={m} nnz={r.size}", flush=True)
    y_ref = pyoracle.spmv_csr_omp(m, r, c, v, x, THREADS)
    xd = torch.tensor(x, device=gpu)
    plan64.set_schedule("stream")
    y = plan64.spmv(xd)
    y2 = plan64.spmv(xd, torch.empty_like(y))
    assert torch.equal(y, y2), "repeated launches differ"
    err = pyoracle.rel_l2(y.cpu().numpy(), y_ref)
    assert err <= 1e-12, err


@pytest.mark.gpu
@pytest.mark.slow
def test_full_size_fp64_blocked(gpu, stencil_powerlaw, plan64):
    import torch

    m, n, r, c, v, x = stencil_powerlaw
    print("stencil_powerlaw blocked", flush=True)
    y_ref = pyoracle.spmv_csr_omp(m, r, c, v, x, THREADS)
    plan64.set_schedule("blocked")
    y = plan64.spmv(torch.tensor(x, device=gpu)).cpu().numpy()
    plan64.set_schedule("auto")
    assert pyoracle.rel_l2(y, y_ref) <= 1e-12


@pytest.mark.gpu
@pytest.mark.slow
def test_full_size_properties(gpu, stencil_powerlaw, plan64):
    import torch

    m, n, r, c, v, x = stencil_powerlaw
    rng = np.random.default_rng(1)
    x1, x2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    ys = [plan64.spmv(torch.tensor(z, device=gpu)).cpu().numpy() for z in (x1, x2, x1 - 2.0 * x2)]
    assert pyoracle.rel_l2(ys[2], ys[0] - 2.0 * ys[1]) <= 1e-13
    colsum = np.bincount(c, weights=v, minlength=n)
    lhs, rhs = ys[0].sum(), colsum @ x1
    scale = np.abs(colsum) @ np.abs(x1)
    assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.slow
@pytest.mark.parametrize("dtype,comp", [("f32", False), ("f32", True), ("bf16", True)])
def test_full_size_reduced_precision(gpu, stencil_powerlaw, dtype, comp):
    import torch

    m, n, r, c, v, x = stencil_powerlaw
    print(f"stencil_powerlaw {dtype} comp={comp}", flush=True)
    x32 = x.astype(np.float32)
    y_exact = pyoracle.spmv_csr_omp(m, r, c, v, x, THREADS)
    p = Plan(m, n, r, c, v, dtype=dtype, compensated=comp)
    y = p.spmv(torch.tensor(x32, device=gpu)).cpu().numpy().astype(np.float64)
    p.close()
    tol = 1e-6 if dtype == "f32" else 2.0 ** -8
    assert pyoracle.rel_l2(y, y_exact) <= tol


@pytest.mark.gpu
@pytest.mark.slow
def test_full_size_bs32_both_schedules(gpu, stencil_powerlaw):
    """The 32x32 build, where CSR blocks appear (coo2block.h:154-380 selects
    them for blocks of more than 64 entries, here the power-law half's
    near-diagonal blocks), on both schedules: the blocked schedule runs the
    CSR branch of the per-block kernel (cb-spmv-dcu.hpp:881-952)."""
    import torch

    m, n, r, c, v, x = stencil_powerlaw
    p = Plan(m, n, r, c, v, block_size=32)
    st = p.stats()
    print(f"stencil_powerlaw bs32: {st['gb_nnzb']} blocks, COO {st['coo_blocks']} CSR {st['csr_blocks']} "
          f"DENSE {st['dense_blocks']}", flush=True)
    assert st["csr_blocks"] > 0
    y_ref = pyoracle.spmv_csr_omp(m, r, c, v, x, THREADS)
    xd = torch.tensor(x, device=gpu)
    for sched in ("stream", "blocked"):
        p.set_schedule(sched)
        y = p.spmv(xd).cpu().numpy()
        err = pyoracle.rel_l2(y, y_ref)
        print(f"stencil_powerlaw bs32 {sched}: rel_l2 {err:.3e}", flush=True)
        assert err <= 1e-12, (sched, err)
    p.close()
