"""torch.autograd wrapper over the C-ABI adjoint (extension: the reference
computes y = A x only and has no gradients).

    A = CBSparse(m, n, row, col, val, dtype="f32", requires_grad=True)
    y = cb_spmv(A, x)               # y = A x with A's trainable values
    loss(y).backward()              # x.grad = A^T g, A.values.grad = g_i x_j

A.values holds the matrix's values in stored order: one per entry of the
transposed plan's stream (duplicates of one (row, col) merged into one
value), the order in which the fused backward pass over A^T produces the
gradients.  So the wrapper's backward is that one pass (dx and dvals
together, cbspmv_backward_stored) with no permutation after it.
A.to_stored(v) / A.to_input(w) convert between input COO order and stored
order.  An optimizer step on A.values changes them in place; the next
forward pushes them to both device plans (cbspmv_set_values_stored).

Values in input COO order remain possible:

    vals = torch.tensor(val, device="cuda", requires_grad=True)
    y = cb_spmv(A, x, vals)         # vals.grad[k] = g[row_k] x[col_k]

Then the backward runs dx over A^T and dvals by a pass over the input
coordinates (backward="direct", the default) or the fused pass plus a
permutation to input order (backward="fused").
"""
import weakref

import numpy as np
import torch

from .plan import Plan


class CBSparse:
    """A sparse matrix held as a forward CB plan plus its transposed plan,
    with its values as a stored-order tensor (`values`)."""

    def __init__(self, m, n, row, col, val, device=None, requires_grad=False, **opts):
        # its values are replaced through the full stream, which drops a
        # scattered-entry split: none is built
        opts.setdefault("split", "off")
        self.plan = Plan(m, n, row, col, val, **opts)
        self.plan.build_transpose()
        self.m, self.n, self.nnz = self.plan.m, self.plan.n, self.plan.nnz
        self.dtype = self.plan.vector_dtype
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        with torch.cuda.device(dev):
            self.plan.upload()
        self.stored = self.plan.stored_entries()
        self._grad_slot = self.plan.export("grad_slot").astype(np.int64)
        self._slot_dev = None   # device copy of grad_slot, made by to_stored / to_input
        live = np.zeros(self.stored, dtype=bool)
        live[self._grad_slot] = True
        # stored entries without an input entry are padding (their dvals are
        # exactly 0) or zero fill of dense blocks of A^T (their dvals are
        # g_i x_j, not a gradient of anything): masked when dense blocks exist
        dense = self.plan.export("transpose.type_per_blk")
        self._mask = None
        if np.any((dense & 0x7F) == 3):
            self._mask = torch.tensor(live, dtype=self.dtype, device=dev)
        v = np.asarray(val, dtype=np.float64)
        vs = np.bincount(self._grad_slot, weights=v, minlength=self.stored) if self.nnz else np.zeros(self.stored)
        self.values = torch.tensor(vs, dtype=self.dtype, device=dev).requires_grad_(requires_grad)
        # the build-time values are on the device already
        self._synced = ("stored", weakref.ref(self.values), self.values._version)

    # ---- orders
    def _slots(self, device):
        if self._slot_dev is None or self._slot_dev.device != device:
            self._slot_dev = torch.tensor(self._grad_slot, device=device)
        return self._slot_dev

    def to_stored(self, vals):
        """Input-order values (tensor of nnz) -> stored order (sums of
        duplicates, 0 where no input entry)."""
        out = torch.zeros(self.stored, dtype=vals.dtype, device=vals.device)
        return out.index_add_(0, self._slots(vals.device), vals)

    def to_input(self, w):
        """Stored-order tensor -> input order (each input entry reads its
        stored entry)."""
        return w.index_select(0, self._slots(w.device))

    # ---- value pushes
    def _is_synced(self, kind, vals):
        k, ref, ver = self._synced
        return k == kind and ref is not None and ref() is vals and ver == vals._version

    def sync_values(self, vals):
        """Pushes input-order vals to the device plans unless this very
        tensor, at this version, was pushed last (a new tensor may reuse
        freed storage, so the storage address alone does not identify it)."""
        if vals.dtype != self.dtype or vals.numel() != self.nnz or vals.device.type != "cuda":
            raise ValueError(f"vals must be a CUDA {self.dtype} tensor of {self.nnz} elements")
        if self._is_synced("input", vals):
            return
        self.plan.set_values(vals.detach().contiguous())
        self._synced = ("input", weakref.ref(vals), vals._version)

    def sync_stored(self, vals):
        """Pushes stored-order vals (e.g. self.values after an optimizer
        step) unless this tensor at this version was pushed last."""
        if vals.dtype != self.dtype or vals.numel() != self.stored or vals.device.type != "cuda":
            raise ValueError(f"stored values must be a CUDA {self.dtype} tensor of {self.stored} elements")
        if self._is_synced("stored", vals):
            return
        self.plan.set_values_stored(vals.detach().contiguous())
        self._synced = ("stored", weakref.ref(vals), vals._version)

    def close(self):
        self.plan.close()


class _SpMV(torch.autograd.Function):
    """y = A(vals) x with vals in input COO order."""

    @staticmethod
    def forward(ctx, vals, x, mat):
        mat.sync_values(vals)
        x = x.contiguous()
        y = mat.plan.spmv(x)
        ctx.mat = mat
        # saving vals lets autograd's version check reject an in-place update
        # between forward and backward; backward pushes them again in case
        # another forward pushed other values to the same matrix meanwhile
        ctx.save_for_backward(x, vals)
        return y

    @staticmethod
    def backward(ctx, g):
        x, vals = ctx.saved_tensors
        want_vals, want_x = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if want_x:
            ctx.mat.sync_values(vals)   # dx = A(vals)^T g needs these values on the device
        dx, dvals = ctx.mat.plan.backward(g.contiguous(), x, want_dx=want_x, want_dvals=want_vals)
        return dvals, dx, None


class _SpMVStored(torch.autograd.Function):
    """y = A(w) x with w in stored order: the backward is the fused pass
    over A^T alone."""

    @staticmethod
    def forward(ctx, w, x, mat):
        mat.sync_stored(w)
        x = x.contiguous()
        y = mat.plan.spmv(x)
        ctx.mat = mat
        ctx.save_for_backward(x, w)
        return y

    @staticmethod
    def backward(ctx, g):
        x, w = ctx.saved_tensors
        mat = ctx.mat
        want_w, want_x = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        g = g.contiguous()
        if want_x:
            mat.sync_stored(w)
        if not want_w:
            return None, mat.plan.spmv_transpose(g) if want_x else None, None
        dw = torch.empty(mat.stored, dtype=mat.dtype, device=g.device)
        dx = torch.empty(mat.n, dtype=mat.dtype, device=g.device) if want_x else None
        mat.plan.backward_stored(g, x, dw, dx)
        if mat._mask is not None:
            dw.mul_(mat._mask)
        return dw, dx, None


def cb_spmv(mat, x, vals=None):
    """y = A x, differentiable in x and in A's values: A.values (stored
    order) by default, or `vals` (input COO order) when given."""
    if x.dim() != 1 or x.numel() != mat.n:
        raise ValueError(f"x must be a vector of {mat.n} elements")
    if vals is None:
        return _SpMVStored.apply(mat.values, x, mat)
    return _SpMV.apply(vals, x, mat)
