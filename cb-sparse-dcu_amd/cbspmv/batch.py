"""Batched persistent launch, hipGraph capture and the autotuner (extension:
the reference runs one matrix per process with block size and waves fixed
at compile time, cb-spmv-dcu/src/macros.h:38-82)."""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .plan import Plan, _as_i32, _ptr, _stream_handle, make_opts

_sigs_done = False


def _sigs():
    global _sigs_done
    if _sigs_done:
        return
    vp = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    _lib.bind("cbspmv_batch_create", ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp), vp])
    _lib.bind("cbspmv_batch_spmv", ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ip, vp])
    _lib.bind("cbspmv_batch_capture", ctypes.c_int, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ip, vp])
    _lib.bind("cbspmv_batch_replay", ctypes.c_int, [vp, vp])
    _lib.bind("cbspmv_batch_destroy", None, [vp])
    _lib.bind("cbspmv_autotune", ctypes.c_int,
              [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, ctypes.POINTER(ctypes.c_double),
               ctypes.POINTER(_lib.Opts), ip, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp),
               ctypes.POINTER(ctypes.c_double), vp])
    _sigs_done = True


class Batch:
    """Plans of one value type run by one persistent launch.  X[i] and Y[i]
    are CUDA tensors of shape (nrhs_i, n_i) / (nrhs_i, m_i) (each row one
    right-hand side, i.e. column-major n_i x nrhs_i)."""

    def __init__(self, plans, stream=None):
        _sigs()
        self.plans = list(plans)
        arr = (ctypes.c_void_p * len(self.plans))(*[p.handle.value for p in self.plans])
        h = ctypes.c_void_p()
        check(lib().cbspmv_batch_create(arr, len(self.plans), ctypes.byref(h), _stream_handle(stream)))
        self._h = h
        for p in self.plans:
            p._uploaded = True

    def _tables(self, X, Y):
        n = len(self.plans)
        if len(X) != n or len(Y) != n:
            raise ValueError("one X and one Y per plan")
        for p, x, y in zip(self.plans, X, Y):
            if x.dim() != 2 or x.shape[1] != p.n or not x.is_contiguous() or x.dtype != p.vector_dtype:
                raise ValueError("X[i] must be a contiguous (nrhs, n_i) tensor of the plan's vector type")
            if y.dim() != 2 or y.shape[1] != p.m or y.shape[0] != x.shape[0] or not y.is_contiguous() \
                    or y.dtype != p.vector_dtype:
                raise ValueError("Y[i] must be a contiguous (nrhs, m_i) tensor like X[i]")
        xs = (ctypes.c_void_p * n)(*[x.data_ptr() for x in X])
        ys = (ctypes.c_void_p * n)(*[y.data_ptr() for y in Y])
        nr = (ctypes.c_int * n)(*[int(x.shape[0]) for x in X])
        return xs, ys, nr

    def spmv(self, X, Y, stream=None):
        xs, ys, nr = self._tables(X, Y)
        check(lib().cbspmv_batch_spmv(self._h, xs, ys, nr, _stream_handle(stream)))
        return Y

    def capture(self, X, Y, stream):
        """Captures the launch into a graph (on a created stream) and runs it.
        The graph holds the tensors' device addresses, so the batch keeps
        X and Y alive until the next capture or close()."""
        xs, ys, nr = self._tables(X, Y)
        check(lib().cbspmv_batch_capture(self._h, xs, ys, nr, _stream_handle(stream)))
        self._captured = (list(X), list(Y))

    def replay(self, stream):
        check(lib().cbspmv_batch_replay(self._h, _stream_handle(stream)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().cbspmv_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        self._captured = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(m, n, row, col, val, nrhs=1, candidates=None, stream=None, **opts):
    """Returns (Plan, ms): the fastest of the candidate (block_size, waves)
    builds for this matrix, timed on the device."""
    _sigs()
    row, col = _as_i32(row), _as_i32(col)
    val = np.ascontiguousarray(np.asarray(val, dtype=np.float64))
    o = make_opts(**opts)
    if candidates is None:
        cand, nc = None, 0
    else:
        flat = [int(v) for pair in candidates for v in pair]
        cand, nc = (ctypes.c_int * len(flat))(*flat), len(candidates)
    h = ctypes.c_void_p()
    ms = ctypes.c_double()
    check(lib().cbspmv_autotune(int(m), int(n), int(row.size), _ptr(row, ctypes.c_int), _ptr(col, ctypes.c_int),
                                _ptr(val, ctypes.c_double), ctypes.byref(o), cand, nc, int(nrhs), ctypes.byref(h),
                                ctypes.byref(ms), _stream_handle(stream)))
    p = Plan.__new__(Plan)
    p.m, p.n, p.nnz = int(m), int(n), int(row.size)
    p.dtype = opts.get("dtype", "f64")
    p._h = h
    p._uploaded = True
    return p, ms.value


__all__ = ["Batch", "autotune"]
