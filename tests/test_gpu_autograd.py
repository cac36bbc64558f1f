"""Differentiable SpMV on the GPU (extension; the reference has no adjoint).

Checks, on a 2048^2 FP64 instance:
  * dx = A^T g against the oracle's FP64 SpMV of the transposed COO
    (relative L2 <= 1e-12) and dvals[k] = g[row_k] x[col_k] exactly (the
    kernels form the same FP64 product), for both backward modes and the
    stored-order pass;
  * the torch.autograd.Function against central finite differences of a
    nonlinear loss L = sum(sin(A x)) (relative error <= 1e-6 at h = 1e-5) and
    against torch.sparse autograd on the CPU in FP64 (relative L2 <= 1e-12);
  * cbspmv_set_values: the plan then computes with the new values;
  * 50 steps of gradient descent on ||A x - b||^2, b = A x* + 1e-3 N(0,1),
    match the same iteration run with numpy/scipy in FP64 step by step.
FP32 storage is checked against the FP64 gradients at relative L2 <= 1e-6.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from cbspmv import CBSparse, Plan, cb_spmv_autograd, gen
from oracle import pyoracle

M = N = 2048


def _matrix(seed=5, dup=False):
    r, c, v = gen.random_coo(M, N, 10, seed=seed, dup=dup)
    if not dup:  # sorted (row, col): the coalesced order torch.sparse uses
        o = np.lexsort((c, r))
        r, c, v = r[o], c[o], v[o]
    return r, c, v


def _t(a, dev, dtype=None):
    import torch

    return torch.tensor(a, device=dev, dtype=dtype or torch.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fused", "direct"])
@pytest.mark.parametrize("dup", [False, True], ids=["nodup", "dup"])
def test_backward_kernels_fp64(gpu, mode, dup):
    r, c, v = _matrix(dup=dup)
    print(f"backward {mode} dup={dup}", flush=True)
    p = Plan(M, N, r, c, v, backward=mode)
    p.build_transpose()
    p.upload()
    rng = np.random.default_rng(2)
    g, x = rng.uniform(-1, 1, M), rng.uniform(-1, 1, N)
    dx, dv = p.backward(_t(g, gpu), _t(x, gpu))
    o = np.argsort(c, kind="stable")
    dx_ref = pyoracle.spmv_seq(N, c[o], r[o], v[o], g)
    assert pyoracle.rel_l2(dx.cpu().numpy(), dx_ref) <= 1e-12
    assert np.array_equal(dv.cpu().numpy(), g[r] * x[c])
    # dx only / dvals only
    dx2, none = p.backward(_t(g, gpu), _t(x, gpu), want_dvals=False)
    assert none is None and pyoracle.rel_l2(dx2.cpu().numpy(), dx_ref) <= 1e-12
    none, dv2 = p.backward(_t(g, gpu), _t(x, gpu), want_dx=False)
    assert none is None and np.array_equal(dv2.cpu().numpy(), g[r] * x[c])
    # transposed product alone
    assert pyoracle.rel_l2(p.spmv_transpose(_t(g, gpu)).cpu().numpy(), dx_ref) <= 1e-12
    p.close()


@pytest.mark.gpu
def test_backward_stored_order(gpu):
    import torch

    r, c, v = _matrix(dup=True)
    p = Plan(M, N, r, c, v)
    p.build_transpose()
    p.upload()
    rng = np.random.default_rng(3)
    g, x = rng.uniform(-1, 1, M), rng.uniform(-1, 1, N)
    buf = torch.empty(p.stored_entries(), dtype=torch.float64, device=gpu)
    dx = torch.empty(N, dtype=torch.float64, device=gpu)
    p.backward_stored(_t(g, gpu), _t(x, gpu), buf, dx)
    gs = p.export("grad_slot").astype(np.int64)
    assert np.array_equal(buf.cpu().numpy()[gs], g[r] * x[c])
    p.backward_stored(_t(g, gpu), _t(x, gpu), buf)   # without dx
    assert np.array_equal(buf.cpu().numpy()[gs], g[r] * x[c])
    p.close()


@pytest.mark.gpu
def test_set_values(gpu):
    r, c, v = _matrix(dup=True)
    p = Plan(M, N, r, c, v)
    p.build_transpose()
    p.upload()
    v2 = np.random.default_rng(4).uniform(-2, 2, r.size)
    p.set_values(_t(v2, gpu))
    assert p.schedule() == "stream"
    x = np.random.default_rng(5).uniform(-1, 1, N)
    y = p.spmv(_t(x, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(M, r, c, v2, x)) <= 1e-12
    g = np.random.default_rng(6).uniform(-1, 1, M)
    o = np.argsort(c, kind="stable")
    dx = p.spmv_transpose(_t(g, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(dx, pyoracle.spmv_seq(N, c[o], r[o], v2[o], g)) <= 1e-12
    p.close()


@pytest.mark.gpu
def test_autograd_finite_differences(gpu):
    import torch

    r, c, v = _matrix(seed=7)
    A = CBSparse(M, N, r, c, v)
    rng = np.random.default_rng(8)
    x0 = rng.uniform(-1, 1, N)
    vals = _t(v, gpu).requires_grad_(True)
    x = _t(x0, gpu).requires_grad_(True)
    loss = torch.sin(cb_spmv_autograd(A, x, vals)).sum()
    loss.backward()

    def L(vv, xx):
        with torch.no_grad():
            return torch.sin(cb_spmv_autograd(A, _t(xx, gpu), _t(vv, gpu))).sum().item()

    h = 1e-5
    for _ in range(3):
        d = rng.standard_normal(N)
        fd = (L(v, x0 + h * d) - L(v, x0 - h * d)) / (2 * h)
        an = float(x.grad.cpu().numpy() @ d)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
        e = rng.standard_normal(r.size)
        fd = (L(v + h * e, x0) - L(v - h * e, x0)) / (2 * h)
        an = float(vals.grad.cpu().numpy() @ e)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    A.close()


@pytest.mark.gpu
def test_autograd_matches_torch_sparse(gpu):
    import torch

    r, c, v = _matrix(seed=9)
    A = CBSparse(M, N, r, c, v)
    x0 = np.random.default_rng(10).uniform(-1, 1, N)
    vals = _t(v, gpu).requires_grad_(True)
    x = _t(x0, gpu).requires_grad_(True)
    y = cb_spmv_autograd(A, x, vals)
    (torch.sin(y) * torch.arange(M, device=gpu)).sum().backward()

    tv = torch.tensor(v, dtype=torch.float64, requires_grad=True)
    tx = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    S = torch.sparse_coo_tensor(torch.tensor(np.stack([r, c]), dtype=torch.long), tv, (M, N))
    ty = torch.sparse.mm(S, tx.view(-1, 1)).view(-1)
    (torch.sin(ty) * torch.arange(M, dtype=torch.float64)).sum().backward()
    assert pyoracle.rel_l2(y.detach().cpu().numpy(), ty.detach().numpy()) <= 1e-12
    assert pyoracle.rel_l2(x.grad.cpu().numpy(), tx.grad.numpy()) <= 1e-12
    assert pyoracle.rel_l2(vals.grad.cpu().numpy(), tv.grad.numpy()) <= 1e-12
    A.close()


@pytest.mark.gpu
def test_gradient_descent_50_steps(gpu):
    import torch

    r, c, v = _matrix(seed=11)
    A = CBSparse(M, N, r, c, v)
    rng = np.random.default_rng(12)
    S = sp.csr_matrix((v, (r, c)), shape=(M, N))
    x_star = rng.uniform(-1, 1, N)
    b = S @ x_star + 1e-3 * rng.standard_normal(M)
    # step 1 / L with L = 2 ||A||_2^2 (power iteration on A^T A)
    z = rng.standard_normal(N)
    for _ in range(100):
        z = S.T @ (S @ z)
        z /= np.linalg.norm(z)
    lr = 1.0 / (2.0 * np.linalg.norm(S @ z) ** 2)

    x = torch.zeros(N, dtype=torch.float64, device=gpu, requires_grad=True)
    bt = _t(b, gpu)
    x_ref = np.zeros(N)
    losses = []
    for step in range(50):
        x.grad = None
        loss = ((cb_spmv_autograd(A, x) - bt) ** 2).sum()
        loss.backward()
        with torch.no_grad():
            x -= lr * x.grad
        res = S @ x_ref - b
        assert abs(loss.item() - res @ res) <= 1e-10 * (res @ res)
        x_ref = x_ref - lr * 2.0 * (S.T @ res)
        losses.append(loss.item())
        assert pyoracle.rel_l2(x.detach().cpu().numpy(), x_ref) <= 1e-10, step
    assert all(b2 <= a2 for a2, b2 in zip(losses, losses[1:]))
    assert losses[-1] < 0.5 * losses[0]
    A.close()


@pytest.mark.gpu
@pytest.mark.parametrize("compensated", [False, True], ids=["native", "comp"])
def test_fp32_gradients(gpu, compensated):
    import torch

    r, c, v = _matrix(seed=13, dup=True)
    rng = np.random.default_rng(14)
    g, x = rng.uniform(-1, 1, M), rng.uniform(-1, 1, N)
    g32, x32 = g.astype(np.float32), x.astype(np.float32)
    p = Plan(M, N, r, c, v, dtype="f32", compensated=compensated)
    p.build_transpose()
    p.upload()
    dx, dv = p.backward(_t(g32, gpu, torch.float32), _t(x32, gpu, torch.float32))
    o = np.argsort(c, kind="stable")
    dx_ref = pyoracle.spmv_seq(N, c[o], r[o], v[o], g)
    assert pyoracle.rel_l2(dx.cpu().numpy().astype(np.float64), dx_ref) <= 1e-6
    assert pyoracle.rel_l2(dv.cpu().numpy().astype(np.float64), g[r] * x[c]) <= 1e-6
    # FP32 autograd: set_values with float values
    A = CBSparse(M, N, r, c, v, dtype="f32")
    vals = _t(v.astype(np.float32), gpu, torch.float32).requires_grad_(True)
    xt = _t(x32, gpu, torch.float32).requires_grad_(True)
    y = cb_spmv_autograd(A, xt, vals)
    y.backward(_t(g32, gpu, torch.float32))
    assert pyoracle.rel_l2(xt.grad.cpu().numpy().astype(np.float64), dx_ref) <= 1e-6
    p.close()
    A.close()


@pytest.mark.gpu
def test_two_value_sets_share_one_matrix(gpu):
    """Two forwards with different values on one CBSparse before a single
    backward: each node's x.grad must use its own values (A(v1)^T g1 +
    A(v2)^T g2), checked against torch.sparse on the CPU in FP64."""
    import torch

    r, c, v = _matrix(seed=15)
    A = CBSparse(M, N, r, c, v)
    rng = np.random.default_rng(16)
    v1, v2 = rng.uniform(-1, 1, r.size), rng.uniform(-2, 2, r.size)
    x0 = rng.uniform(-1, 1, N)
    x = _t(x0, gpu).requires_grad_(True)
    t1, t2 = _t(v1, gpu).requires_grad_(True), _t(v2, gpu).requires_grad_(True)
    y1 = cb_spmv_autograd(A, x, t1)
    y2 = cb_spmv_autograd(A, x, t2)
    w = torch.arange(M, device=gpu, dtype=torch.float64)
    (torch.sin(y1).sum() + (torch.cos(y2) * w).sum()).backward()

    tx = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    idx = torch.tensor(np.stack([r, c]), dtype=torch.long)
    tv1 = torch.tensor(v1, dtype=torch.float64, requires_grad=True)
    tv2 = torch.tensor(v2, dtype=torch.float64, requires_grad=True)
    ty1 = torch.sparse.mm(torch.sparse_coo_tensor(idx, tv1, (M, N)), tx.view(-1, 1)).view(-1)
    ty2 = torch.sparse.mm(torch.sparse_coo_tensor(idx, tv2, (M, N)), tx.view(-1, 1)).view(-1)
    (torch.sin(ty1).sum() + (torch.cos(ty2) * torch.arange(M, dtype=torch.float64)).sum()).backward()
    assert pyoracle.rel_l2(y1.detach().cpu().numpy(), ty1.detach().numpy()) <= 1e-12
    assert pyoracle.rel_l2(y2.detach().cpu().numpy(), ty2.detach().numpy()) <= 1e-12
    assert pyoracle.rel_l2(x.grad.cpu().numpy(), tx.grad.numpy()) <= 1e-12
    assert pyoracle.rel_l2(t1.grad.cpu().numpy(), tv1.grad.numpy()) <= 1e-12
    assert pyoracle.rel_l2(t2.grad.cpu().numpy(), tv2.grad.numpy()) <= 1e-12
    # an in-place update of saved values between forward and backward is refused
    y3 = cb_spmv_autograd(A, x, t1)
    with torch.no_grad():
        t1.add_(1.0)
    with pytest.raises(RuntimeError):
        y3.sum().backward()
    A.close()


# ---- trainable values in stored order (CBSparse.values) -----------------------

def _sparse_ref(r, c, v, x0, weights):
    """y, x.grad and per-input-entry value gradients of sum(sin(y) * w) by
    torch.sparse on the CPU in FP64."""
    import torch

    tv = torch.tensor(v, dtype=torch.float64, requires_grad=True)
    tx = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    S = torch.sparse_coo_tensor(torch.tensor(np.stack([r, c]), dtype=torch.long), tv, (M, N))
    ty = torch.sparse.mm(S, tx.view(-1, 1)).view(-1)
    (torch.sin(ty) * torch.tensor(weights)).sum().backward()
    return ty.detach().numpy(), tx.grad.numpy(), tv.grad.numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("dup", [False, True], ids=["nodup", "dup"])
def test_stored_values_match_torch_sparse(gpu, dup):
    """cb_spmv(A, x) differentiates A.values (stored order) with the fused
    pass alone; mapped to input order the gradient is g_i x_j per entry."""
    import torch

    r, c, v = _matrix(seed=21, dup=dup)
    A = CBSparse(M, N, r, c, v, requires_grad=True)
    assert A.values.numel() == A.stored
    x0 = np.random.default_rng(22).uniform(-1, 1, N)
    w = np.arange(M, dtype=np.float64)
    x = _t(x0, gpu).requires_grad_(True)
    y = cb_spmv_autograd(A, x)
    (torch.sin(y) * _t(w, gpu)).sum().backward()
    ty, txg, tvg = _sparse_ref(r, c, v, x0, w)
    assert pyoracle.rel_l2(y.detach().cpu().numpy(), ty) <= 1e-12
    assert pyoracle.rel_l2(x.grad.cpu().numpy(), txg) <= 1e-12
    # each input entry's gradient is its stored entry's (duplicates share one)
    assert pyoracle.rel_l2(A.to_input(A.values.grad).cpu().numpy(), tvg) <= 1e-12
    # stored entries holding no input entry (padding) get a zero gradient
    live = np.zeros(A.stored, bool)
    live[A.plan.export("grad_slot")] = True
    assert np.all(A.values.grad.cpu().numpy()[~live] == 0)
    A.close()


@pytest.mark.gpu
def test_stored_values_finite_differences(gpu):
    import torch

    r, c, v = _matrix(seed=23)
    A = CBSparse(M, N, r, c, v, requires_grad=True)
    x0 = np.random.default_rng(24).uniform(-1, 1, N)
    x = _t(x0, gpu)
    loss = torch.sin(cb_spmv_autograd(A, x)).sum()
    loss.backward()
    grad = A.values.grad.cpu().numpy()
    base = A.values.detach().clone()

    def L(wv):
        A.values = wv
        with torch.no_grad():
            return torch.sin(cb_spmv_autograd(A, x)).sum().item()

    rng = np.random.default_rng(25)
    h = 1e-5
    live = np.zeros(A.stored, bool)
    live[A.plan.export("grad_slot")] = True
    for _ in range(3):
        e = rng.standard_normal(A.stored) * live
        et = _t(e, gpu)
        fd = (L(base + h * et) - L(base - h * et)) / (2 * h)
        an = float(grad @ e)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    A.close()


@pytest.mark.gpu
def test_stored_values_training(gpu):
    """Gradient descent on A.values for ||A x - b||^2 with a fixed x: each
    step's loss and values match the same iteration in FP64 numpy; the next
    forward pushes the updated values (cbspmv_set_values_stored)."""
    import torch

    r, c, v = _matrix(seed=27)      # no duplicates: stored order is a permutation of input order
    A = CBSparse(M, N, r, c, v, requires_grad=True)
    rng = np.random.default_rng(28)
    x0 = rng.uniform(-1, 1, N)
    v_star = v + 0.1 * rng.standard_normal(v.size)
    b = sp.csr_matrix((v_star, (r, c)), shape=(M, N)) @ x0
    x, bt = _t(x0, gpu), _t(b, gpu)
    opt = torch.optim.SGD([A.values], lr=0.05)
    vals = v.copy()
    slot = A.plan.export("grad_slot").astype(np.int64)
    losses = []
    for step in range(20):
        opt.zero_grad()
        loss = ((cb_spmv_autograd(A, x) - bt) ** 2).sum()
        loss.backward()
        opt.step()
        res = sp.csr_matrix((vals, (r, c)), shape=(M, N)) @ x0 - b
        assert abs(loss.item() - res @ res) <= 1e-10 * max(res @ res, 1e-30), step
        vals = vals - 0.05 * 2.0 * res[r] * x0[c]
        got = A.values.detach().cpu().numpy()[slot]
        assert pyoracle.rel_l2(got, vals) <= 1e-12, step
        losses.append(loss.item())
    assert losses[-1] < 0.5 * losses[0]
    # after pushing the last step's values the plans compute with the trained
    # values (forward and transpose)
    A.sync_stored(A.values)
    y = A.plan.spmv(x).cpu().numpy()
    assert pyoracle.rel_l2(y, pyoracle.spmv_seq(M, r, c, vals, x0)) <= 1e-12
    g = rng.uniform(-1, 1, M)
    o = np.argsort(c, kind="stable")
    dx = A.plan.spmv_transpose(_t(g, gpu)).cpu().numpy()
    assert pyoracle.rel_l2(dx, pyoracle.spmv_seq(N, c[o], r[o], vals[o], g)) <= 1e-12
    A.close()


@pytest.mark.gpu
def test_stored_values_dense_fill_masked(gpu):
    """A matrix with dense blocks: A^T's zero fill entries get no gradient,
    and the trained values stay consistent between A and A^T."""
    import torch

    rng = np.random.default_rng(29)
    D = (rng.random((96, 96)) < 0.1) * rng.uniform(-1, 1, (96, 96))
    D[16:32, 32:48] = rng.uniform(-1, 1, (16, 16))
    D[20, 40] = 0.0                      # a hole: zero fill in the dense block
    r, c = np.nonzero(D)
    r, c = r.astype(np.int32), c.astype(np.int32)
    v = D[r, c]
    A = CBSparse(96, 96, r, c, v, requires_grad=True, builder="comp")
    assert A._mask is not None
    x0 = rng.uniform(-1, 1, 96)
    x = _t(x0, gpu).requires_grad_(True)
    y = cb_spmv_autograd(A, x)
    (y * y).sum().backward()
    live = A._mask.cpu().numpy().astype(bool)
    grad = A.values.grad.cpu().numpy()
    assert np.all(grad[~live] == 0)
    gy = 2 * (D @ x0)
    assert pyoracle.rel_l2(A.to_input(A.values.grad).cpu().numpy(), gy[r] * x0[c]) <= 1e-12
    assert pyoracle.rel_l2(x.grad.cpu().numpy(), D.T @ gy) <= 1e-12
    A.close()
