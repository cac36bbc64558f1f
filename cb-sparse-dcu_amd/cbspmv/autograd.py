"""torch.autograd wrapper over the C-ABI adjoint (extension: the reference
computes y = A x only and has no gradients).

    A = CBSparse(m, n, row, col, val, dtype="f32")      # builds A and A^T once
    vals = torch.tensor(val, device="cuda", requires_grad=True)
    y = cb_spmv(A, x, vals)          # y = A(vals) x
    loss(y).backward()               # x.grad = A^T g, vals.grad[k] = g[row_k] x[col_k]

The backward runs hand-written kernels only (cbspmv_backward): dx over the
transposed CB plan, and vals.grad in input COO order, the order of `vals`,
either by a pass over the input coordinates (backward="direct", the
default) or by the fused SDDMM over A^T's stored pattern followed by a
permutation to input order (backward="fused").  A caller that keeps its
own values in A^T's stored order can take the fused pass alone
(Plan.backward_stored), which is the fastest: about 1.3x the forward time
for dx and dvals together, against about 2.7x for the direct mode.
Values are pushed to the device plans (cbspmv_set_values) when `vals` is
another tensor than the one pushed last, or was updated in place since.
"""
import weakref

import torch

from .plan import Plan


class CBSparse:
    """A sparse matrix held as a forward CB plan plus its transposed plan."""

    def __init__(self, m, n, row, col, val, device=None, **opts):
        self.plan = Plan(m, n, row, col, val, **opts)
        self.plan.build_transpose()
        self.m, self.n, self.nnz = self.plan.m, self.plan.n, self.plan.nnz
        self.dtype = self.plan.vector_dtype
        if device is not None:
            with torch.cuda.device(device):
                self.plan.upload()
        else:
            self.plan.upload()
        self._synced = (None, -1)   # (weak reference to the tensor, its version)

    def sync_values(self, vals):
        """Pushes vals to the device plans unless this very tensor, at this
        version, was pushed last (a new tensor may reuse freed storage, so
        the storage address alone does not identify it)."""
        if vals.dtype != self.dtype or vals.numel() != self.nnz or vals.device.type != "cuda":
            raise ValueError(f"vals must be a CUDA {self.dtype} tensor of {self.nnz} elements")
        ref, ver = self._synced
        if ref is not None and ref() is vals and ver == vals._version:
            return
        self.plan.set_values(vals.detach().contiguous())
        self._synced = (weakref.ref(vals), vals._version)

    def close(self):
        self.plan.close()


class _SpMV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, x, mat):
        if vals is not None:
            mat.sync_values(vals)
        x = x.contiguous()
        y = mat.plan.spmv(x)
        ctx.mat = mat
        ctx.has_vals = vals is not None
        # saving vals lets autograd's version check reject an in-place update
        # between forward and backward; backward pushes them again in case
        # another forward pushed other values to the same matrix meanwhile
        if vals is not None:
            ctx.save_for_backward(x, vals)
        else:
            ctx.save_for_backward(x)
        return y

    @staticmethod
    def backward(ctx, g):
        saved = ctx.saved_tensors
        x = saved[0]
        want_vals, want_x = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if ctx.has_vals and want_x:
            ctx.mat.sync_values(saved[1])   # dx = A(vals)^T g needs these values on the device
        dx, dvals = ctx.mat.plan.backward(g.contiguous(), x, want_dx=want_x, want_dvals=want_vals)
        return dvals, dx, None


def cb_spmv(mat, x, vals=None):
    """y = A x, differentiable in x and (when given) in A's values."""
    if x.dim() != 1 or x.numel() != mat.n:
        raise ValueError(f"x must be a vector of {mat.n} elements")
    return _SpMV.apply(vals, x, mat)
