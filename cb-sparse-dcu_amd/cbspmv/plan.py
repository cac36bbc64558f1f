"""Host-side plan object over the C ABI.

A Plan is the GLOBAL_BLOCK of the reference (cb-spmv-dcu/src/format.h:62)
plus its device slabs: `Plan(m, n, row, col, val)` runs the builder
(coo2block_gather, cb-spmv-dcu/src/coo2block.h:628), `plan.spmv(x, y)` is one
device SpMV on torch tensors, and `plan.bench(x)` follows cb_spmv's protocol
(cb-spmv-dcu/src/cb-spmv-dcu.hpp:1088).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

_EXPORT_DTYPES = {
    "type_per_blk": np.uint8,
    "segment_direct": np.uint8,
    "segment_uniform": np.uint8,
    "block_row_uniform": np.uint8,
    "image": np.uint8,
    "pl_val": np.float64,
    "stream_key": np.uint32,
    "stream_val": np.float64,
    "stream_items": np.uint32,
    "policy": np.int32,
    "stream_slab": np.uint8,
    "stream_tiles": np.uint32,
    "stream_tile_lines": np.uint32,
    "stream_empty_rows": np.uint32,
    "stream_panel_items": np.uint32,
    "stream_panel_mask": np.uint8,
    "grad_slot": np.uint32,
    "stream_value_pos": np.int64,
    "slot_of": np.uint32,
    "split_key": np.uint32,
    "split_val": np.float64,
    "split_src": np.uint32,
    "split_wave_off": np.uint32,
    "split_shape": np.uint32,
}
ARRAY_NAMES = [
    "gb_row_idx", "gb_col_idx", "gb_row_ptr", "nnz_per_blk", "type_per_blk",
    "segment_rows", "segment_begin", "segment_end", "segment_direct", "segment_uniform",
    "block_row_uniform", "flat_row_ptr", "flat_col_idx", "pattern_row_ptr",
    "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows", "block_cols",
    "block_cols_offset", "bia_nnz", "bia_origin", "bia_end", "pl_idx", "pl_val",
]

_DTYPE_CODES = {"f64": _lib.F64, "f32": _lib.F32, "bf16": _lib.BF16}
_SCHEDULES = {"auto": _lib.SCHED_AUTO, "blocked": _lib.SCHED_BLOCKED, "stream": _lib.SCHED_STREAM}


def make_opts(block_size=16, waves=4, row_owned=True, pattern_wave=True, force_coo64=False,
              dense_sixteenths=15, builder="gather", dtype="f64", compensated=False,
              schedule="auto", item_nnz=0, tile_waves=0, tile_kib=0, backward="direct", panels=0, tiles=0,
              split="auto"):
    o = _lib.Opts()
    lib().cbspmv_default_opts(ctypes.byref(o))
    o.block_size = int(block_size)
    o.waves_per_block = int(waves)
    o.row_owned = int(bool(row_owned))
    o.pattern_wave = int(bool(pattern_wave))
    o.force_coo64 = int(bool(force_coo64))
    o.dense_sixteenths = int(dense_sixteenths)
    o.builder = _lib.BUILDER_COMP if builder == "comp" else _lib.BUILDER_GATHER
    o.value_dtype = _DTYPE_CODES[dtype]
    o.accumulate = _lib.ACC_COMPENSATED if compensated else _lib.ACC_NATIVE
    o.schedule = _SCHEDULES[schedule]
    o.stream_item_nnz = int(item_nnz)
    o.stream_tile_waves = int(tile_waves)
    o.stream_tile_kib = int(tile_kib)
    o.backward_mode = {"fused": _lib.BACKWARD_FUSED, "direct": _lib.BACKWARD_DIRECT}[backward]
    o.stream_panels = int(panels)
    o.stream_tiles = int(tiles)
    o.stream_split = {"auto": _lib.SPLIT_AUTO, "off": _lib.SPLIT_OFF, "on": _lib.SPLIT_ON}[split]
    return o


def _as_i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Plan:
    """A CB plan built from a row-major COO matrix (m x n)."""

    def __init__(self, m, n, row, col, val, **opts):
        row, col = _as_i32(row), _as_i32(col)
        val = np.ascontiguousarray(np.asarray(val, dtype=np.float64))
        if not (row.shape == col.shape == val.shape) or row.ndim != 1:
            raise ValueError("row, col and val must be 1-D arrays of equal length")
        self.m, self.n, self.nnz = int(m), int(n), int(row.size)
        self.dtype = opts.get("dtype", "f64")
        o = make_opts(**opts)
        h = ctypes.c_void_p()
        check(lib().cbspmv_build(self.m, self.n, self.nnz, _ptr(row, ctypes.c_int), _ptr(col, ctypes.c_int),
                                 _ptr(val, ctypes.c_double), ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._uploaded = False

    # ---- lifecycle
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().cbspmv_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- host-side inspection
    def stats(self):
        s = _lib.Stats()
        check(lib().cbspmv_plan_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def export(self, name):
        """Returns the named plan array as a numpy array (see cbspmv.h)."""
        cnt = ctypes.c_size_t(0)
        check(lib().cbspmv_export(self._h, name.encode(), None, ctypes.byref(cnt)))
        dt = _EXPORT_DTYPES.get(name.split(".", 1)[-1], np.int32)
        out = np.empty(cnt.value, dtype=dt)
        if cnt.value:
            check(lib().cbspmv_export(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt)))
        return out

    def layout_offset(self, name):
        return int(lib().cbspmv_layout_offset(self._h, name.encode()))

    # ---- device execution
    @property
    def vector_dtype(self):
        return torch.float64 if self.dtype == "f64" else torch.float32

    def upload(self, stream=None):
        check(lib().cbspmv_upload(self._h, _stream_handle(stream)))
        self._uploaded = True

    def set_schedule(self, name):
        check(lib().cbspmv_set_schedule(self._h, _SCHEDULES[name]))

    def schedule(self):
        code = lib().cbspmv_get_schedule(self._h)
        return {v: k for k, v in _SCHEDULES.items()}[code]

    def _check_vec(self, t, length, name):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise TypeError(f"{name} must be a CUDA (HIP) tensor")
        if t.dtype != self.vector_dtype or not t.is_contiguous() or t.numel() < length:
            raise ValueError(f"{name} must be a contiguous {self.vector_dtype} tensor of >= {length} elements")

    def spmv(self, x, y=None, stream=None):
        """y = A x on the device (torch tensors); returns y."""
        if not self._uploaded:
            self.upload(stream)
        self._check_vec(x, self.n, "x")
        if y is None:
            y = torch.empty(self.m, dtype=self.vector_dtype, device=x.device)
        self._check_vec(y, self.m, "y")
        check(lib().cbspmv_spmv(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                _stream_handle(stream)))
        return y

    # ---- adjoint (extension: no reference counterpart)
    def build_transpose(self):
        """Builds A^T once (host); device copies follow on first use."""
        check(lib().cbspmv_build_transpose(self._h))

    def stored_entries(self):
        # fixed once the transposed plan exists (-1 before): asked once then
        n = getattr(self, "_stored", None)
        if n is None:
            n = int(lib().cbspmv_stored_entries(self._h))
            if n >= 0:
                self._stored = n
        return n

    def spmv_transpose(self, g, dx=None, stream=None):
        """dx = A^T g on the device."""
        self._check_vec(g, self.m, "g")
        if dx is None:
            dx = torch.empty(self.n, dtype=self.vector_dtype, device=g.device)
        self._check_vec(dx, self.n, "dx")
        check(lib().cbspmv_spmv_transpose(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(dx.data_ptr()),
                                          _stream_handle(stream)))
        return dx

    def backward(self, g, x, want_dx=True, want_dvals=True, stream=None):
        """(dx, dvals) of y = A x for upstream gradient g; dvals in input COO
        order.  Either may be None when not wanted."""
        self._check_vec(g, self.m, "g")
        self._check_vec(x, self.n, "x")
        dx = torch.empty(self.n, dtype=self.vector_dtype, device=g.device) if want_dx else None
        dv = torch.empty(self.nnz, dtype=self.vector_dtype, device=g.device) if want_dvals else None
        check(lib().cbspmv_backward(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                    ctypes.c_void_p(dx.data_ptr() if dx is not None else 0),
                                    ctypes.c_void_p(dv.data_ptr() if dv is not None else 0), _stream_handle(stream)))
        return dx, dv

    def backward_stored(self, g, x, dvals, dx=None, stream=None):
        """Fused pass: dx (optional) and dvals in stored order."""
        self._check_vec(g, self.m, "g")
        self._check_vec(x, self.n, "x")
        self._check_vec(dvals, self.stored_entries(), "dvals")
        check(lib().cbspmv_backward_stored(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(dx.data_ptr() if dx is not None else 0),
                                           ctypes.c_void_p(dvals.data_ptr()), _stream_handle(stream)))
        return dx, dvals

    def set_values(self, vals, stream=None):
        """Replaces A's values (device tensor, input COO order)."""
        if not self._uploaded:
            self.upload(stream)
        self._check_vec(vals, self.nnz, "vals")
        check(lib().cbspmv_set_values(self._h, ctypes.c_void_p(vals.data_ptr()), _stream_handle(stream)))

    def set_values_stored(self, vals, stream=None):
        """Replaces A's values from stored order (stored_entries() values,
        the order of backward_stored's dvals)."""
        if not self._uploaded:
            self.upload(stream)
        self._check_vec(vals, self.stored_entries(), "vals")
        f = _lib.bind("cbspmv_set_values_stored", ctypes.c_int, [ctypes.c_void_p] * 3)
        check(f(self._h, ctypes.c_void_p(vals.data_ptr()), _stream_handle(stream)))

    # ---- persistence
    def save(self, path):
        """Writes the built plan to a file (cbspmv_save)."""
        f = _lib.bind("cbspmv_save", ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p])
        check(f(self._h, str(path).encode()))

    @classmethod
    def load(cls, path):
        """A plan read back from cbspmv_save's file, without rebuilding."""
        f = _lib.bind("cbspmv_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)])
        h = ctypes.c_void_p()
        check(f(str(path).encode(), ctypes.byref(h)))
        p = cls.__new__(cls)
        p._h = h
        p._uploaded = False
        st = p.stats()
        p.m, p.n, p.nnz = st["m"], st["n"], st["nnz"]
        p.dtype = {_lib.F64: "f64", _lib.F32: "f32", _lib.BF16: "bf16"}[int(p.export("plan_dtype")[0])]
        return p

    def bench(self, x):
        """cb_spmv() protocol on host FP64 x: returns (y, runtime_ms, gflops)."""
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
        y = np.zeros(self.m, dtype=np.float64)
        met = _lib.Metrics()
        check(lib().cbspmv_bench(self._h, _ptr(x, ctypes.c_double), _ptr(y, ctypes.c_double), ctypes.byref(met)))
        self._uploaded = True
        return y, met.runtime, met.gflops


def read_mtx(path):
    """Matrix Market file -> (m, n, row, col, val) with the reference
    reader's semantics (values forced to 1, symmetric expansion, row-major
    COO; cb-spmv-dcu/src/io/mmio_highlevel.h:62-403)."""
    L = lib()
    f = L.cbspmv_read_mtx
    f.restype = ctypes.c_int
    ip = ctypes.POINTER(ctypes.c_int)
    f.argtypes = [ctypes.c_char_p, ip, ip, ip, ctypes.POINTER(ip), ctypes.POINTER(ip),
                  ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]
    L.cbspmv_free.argtypes = [ctypes.c_void_p]
    m, n, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    r, c, v = ip(), ip(), ctypes.POINTER(ctypes.c_double)()
    check(f(str(path).encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(r),
            ctypes.byref(c), ctypes.byref(v)))
    try:
        k = nnz.value
        row = np.ctypeslib.as_array(r, shape=(k,)).copy() if k else np.zeros(0, np.int32)
        col = np.ctypeslib.as_array(c, shape=(k,)).copy() if k else np.zeros(0, np.int32)
        val = np.ctypeslib.as_array(v, shape=(k,)).copy() if k else np.zeros(0)
    finally:
        L.cbspmv_free(ctypes.cast(r, ctypes.c_void_p))
        L.cbspmv_free(ctypes.cast(c, ctypes.c_void_p))
        L.cbspmv_free(ctypes.cast(v, ctypes.c_void_p))
    return m.value, n.value, row.astype(np.int32), col.astype(np.int32), val
