"""ctypes binding of libcbspmv.so (the C ABI declared in include/cbspmv.h).

torch is imported before the library is loaded so that the process has a
single HIP runtime: torch's bundled libamdhip64 and the system one share the
soname libamdhip64.so.7, and whichever loads first serves both.
"""
import ctypes
import os

import torch  # noqa: F401  (load torch's HIP runtime first, see above)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libcbspmv.so")

CBSPMV_OK = 0
ERR_ARG, ERR_RANGE, ERR_ORDER, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6, -7
ERRORS = {
    -1: "CBSPMV_ERR_ARG",
    -2: "CBSPMV_ERR_RANGE",
    -3: "CBSPMV_ERR_ORDER",
    -4: "CBSPMV_ERR_HIP",
    -5: "CBSPMV_ERR_NOMEM",
    -6: "CBSPMV_ERR_STATE",
    -7: "CBSPMV_ERR_UNSUPPORTED",
}

BUILDER_GATHER, BUILDER_COMP = 0, 1
F64, F32, BF16 = 0, 1, 2
ACC_NATIVE, ACC_COMPENSATED = 0, 1
SCHED_AUTO, SCHED_BLOCKED, SCHED_STREAM = 0, 1, 2
BACKWARD_DIRECT, BACKWARD_FUSED = 0, 1
SPLIT_AUTO, SPLIT_OFF, SPLIT_ON = 0, 1, 2


class CbspmvError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [
        ("block_size", ctypes.c_int),
        ("waves_per_block", ctypes.c_int),
        ("row_owned", ctypes.c_int),
        ("pattern_wave", ctypes.c_int),
        ("force_coo64", ctypes.c_int),
        ("dense_sixteenths", ctypes.c_int),
        ("builder", ctypes.c_int),
        ("value_dtype", ctypes.c_int),
        ("accumulate", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("stream_item_nnz", ctypes.c_int),
        ("stream_tile_waves", ctypes.c_int),
        ("stream_tile_kib", ctypes.c_int),
        ("backward_mode", ctypes.c_int),
        ("stream_panels", ctypes.c_int),
        ("stream_tiles", ctypes.c_int),
        ("stream_split", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int), ("n", ctypes.c_int), ("nnz", ctypes.c_int),
        ("gb_m", ctypes.c_int), ("gb_n", ctypes.c_int), ("gb_nnzb", ctypes.c_int),
        ("gather_flag", ctypes.c_int), ("direct_coo_columns", ctypes.c_int),
        ("compact_coo_indices", ctypes.c_int),
        ("super_sparse_blocks", ctypes.c_int), ("analyzed_blocks", ctypes.c_int),
        ("super_sparse_ratio", ctypes.c_double),
        ("load_balanced", ctypes.c_int),
        ("load_stddev_before", ctypes.c_double), ("load_stddev_after", ctypes.c_double),
        ("coo_blocks", ctypes.c_int), ("csr_blocks", ctypes.c_int),
        ("dense_blocks", ctypes.c_int), ("uniform_blocks", ctypes.c_int),
        ("segments", ctypes.c_int), ("patterns", ctypes.c_int),
        ("unit_values", ctypes.c_int),
        ("slab_bytes", ctypes.c_longlong), ("stream_bytes", ctypes.c_longlong),
        ("restore_entries", ctypes.c_longlong),
        ("build_seconds", ctypes.c_double),
        ("split_entries", ctypes.c_longlong),
        ("split_blocks", ctypes.c_int),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Metrics(ctypes.Structure):
    _fields_ = [("runtime", ctypes.c_double), ("gflops", ctypes.c_double)]


_lib = None


def lib():
    """Loads libcbspmv.so once.  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    sig = {
        "cbspmv_default_opts": (None, [ctypes.POINTER(Opts)]),
        "cbspmv_build": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, dp,
                                        ctypes.POINTER(Opts), ctypes.POINTER(vp)]),
        "cbspmv_destroy": (None, [vp]),
        "cbspmv_plan_stats": (ctypes.c_int, [vp, ctypes.POINTER(Stats)]),
        "cbspmv_export": (ctypes.c_int, [vp, ctypes.c_char_p, vp, ctypes.POINTER(ctypes.c_size_t)]),
        "cbspmv_layout_offset": (ctypes.c_longlong, [vp, ctypes.c_char_p]),
        "cbspmv_upload": (ctypes.c_int, [vp, vp]),
        "cbspmv_spmv": (ctypes.c_int, [vp, vp, vp, vp]),
        "cbspmv_set_schedule": (ctypes.c_int, [vp, ctypes.c_int]),
        "cbspmv_get_schedule": (ctypes.c_int, [vp]),
        "cbspmv_bench": (ctypes.c_int, [vp, dp, dp, ctypes.POINTER(Metrics)]),
        "cbspmv_last_error": (ctypes.c_char_p, []),
        "cbspmv_version": (ctypes.c_char_p, []),
        "cbspmv_last_kernel_ms": (ctypes.c_double, [vp]),
        "cbspmv_build_transpose": (ctypes.c_int, [vp]),
        "cbspmv_stored_entries": (ctypes.c_longlong, [vp]),
        "cbspmv_spmv_transpose": (ctypes.c_int, [vp, vp, vp, vp]),
        "cbspmv_backward": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "cbspmv_backward_stored": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "cbspmv_set_values": (ctypes.c_int, [vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def bind(name, restype, argtypes):
    """Declares the signature of an additional C-ABI function and returns it."""
    f = getattr(lib(), name)
    f.restype = restype
    f.argtypes = argtypes
    return f


def check(rc):
    if rc != CBSPMV_OK:
        raise CbspmvError(rc, lib().cbspmv_last_error().decode(errors="replace"))
    return rc
