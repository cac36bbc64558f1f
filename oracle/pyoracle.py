"""Python access to the oracles (test infrastructure only).

  liboracle.so      CPU restatement of the reference path (cb_oracle.cpp)
  _ref/ref_*.so     the reference's own host code compiled from its sources
                    (ref_shim.cpp, built by oracle/Makefile when the reference
                    checkout is present)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the oracle.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

ARRAY_NAMES = [
    "gb_row_idx", "gb_col_idx", "gb_row_ptr", "nnz_per_blk", "type_per_blk",
    "segment_rows", "segment_begin", "segment_end", "segment_direct", "segment_uniform",
    "block_row_uniform", "flat_row_ptr", "flat_col_idx", "pattern_row_ptr",
    "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows", "block_cols",
    "block_cols_offset", "bia_nnz", "bia_origin", "bia_end", "pl_idx", "pl_val",
]
SCALAR_NAMES = [
    "gb_m", "gb_n", "gb_nnzb", "gather_flag", "direct_coo_columns", "compact_coo_indices",
    "super_sparse_blocks", "analyzed_blocks", "super_sparse_ratio", "load_balanced",
    "load_stddev_before", "load_stddev_after",
]
_DT = {"type_per_blk": np.uint8, "segment_direct": np.uint8, "segment_uniform": np.uint8,
       "block_row_uniform": np.uint8, "pl_val": np.float64, "image": np.uint8}


class _OracleOpts(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in
                ("bs", "waves", "row_owned", "force_coo64", "dense_sixteenths", "pattern_wave")]


_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_LIB)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.oracle_build.restype = vp
        L.oracle_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, dp,
                                   ctypes.POINTER(_OracleOpts), ctypes.c_int]
        L.oracle_free.argtypes = [vp]
        L.oracle_scalar.restype = ctypes.c_double
        L.oracle_scalar.argtypes = [vp, ctypes.c_char_p]
        L.oracle_array.restype = ctypes.c_long
        L.oracle_array.argtypes = [vp, ctypes.c_char_p, vp]
        L.oracle_layout.restype = ctypes.c_long
        L.oracle_layout.argtypes = [vp, ctypes.c_char_p]
        L.oracle_policy.argtypes = [vp, ctypes.POINTER(ctypes.c_long)]
        L.oracle_spmv_seq.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int, dp, dp]
        L.oracle_spmv_csr_omp.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), ip, dp, dp, dp, ctypes.c_int]
        L.oracle_time_csr_omp.restype = ctypes.c_double
        L.oracle_time_csr_omp.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), ip, dp, dp, dp, ctypes.c_int,
                                          ctypes.c_int]
        L.oracle_host_threads.argtypes = [ip, ip, ip]
        _oracle = L
    return _oracle


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class _Built:
    """A built CB structure (oracle or reference) with array access by name."""

    def __init__(self, lib, handle, array_fn, scalar_fn, free_fn):
        self._lib, self._h = lib, handle
        self._array, self._scalar, self._free = array_fn, scalar_fn, free_fn

    def array(self, name):
        n = self._array(self._h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.empty(n, dtype=_DT.get(name, np.int32))
        if n:
            self._array(self._h, name.encode(), out.ctypes.data_as(ctypes.c_void_p))
        return out

    def scalar(self, name):
        return self._scalar(self._h, name.encode())

    def close(self):
        if self._h:
            self._free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class OracleBuild(_Built):
    """The oracle's restatement of coo2block_gather / coo2block_comp."""

    def __init__(self, m, n, row, col, val, bs=16, waves=4, row_owned=True, force_coo64=False,
                 dense_sixteenths=15, pattern_wave=True, comp=False):
        L = oracle_lib()
        row, col, val = _i32(row), _i32(col), _f64(val)
        o = _OracleOpts(bs, waves, int(row_owned), int(force_coo64), dense_sixteenths, int(pattern_wave))
        h = L.oracle_build(int(m), int(n), int(row.size), _p(row, ctypes.c_int), _p(col, ctypes.c_int),
                           _p(val, ctypes.c_double), ctypes.byref(o), int(comp))
        super().__init__(L, h, L.oracle_array, L.oracle_scalar, L.oracle_free)

    def layout(self, name):
        return int(self._lib.oracle_layout(self._h, name.encode()))

    def policy(self):
        out = (ctypes.c_long * 8)()
        self._lib.oracle_policy(self._h, out)
        keys = ["unit_values", "use_pattern", "use_unrolled", "use_small_direct", "launch_waves",
                "grid", "aggregate_coo", "pattern_count"]
        return dict(zip(keys, list(out)))


_ref_libs = {}


def ref_lib_path(bs=16, waves=4, row_owned=True, force_coo64=False):
    return os.path.join(REF_DIR, f"ref_bs{bs}_w{waves}_ro{int(row_owned)}_f{int(force_coo64)}.so")


def ref_available(bs=16, waves=4, row_owned=True, force_coo64=False):
    return os.path.exists(ref_lib_path(bs, waves, row_owned, force_coo64))


def ref_lib(bs=16, waves=4, row_owned=True, force_coo64=False):
    path = ref_lib_path(bs, waves, row_owned, force_coo64)
    if path not in _ref_libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: build it with `make -C oracle ref`")
        L = ctypes.CDLL(path)
        vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        L.ref_build.restype = vp
        L.ref_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip, dp, ctypes.c_int]
        L.ref_free.argtypes = [vp]
        L.ref_scalar.restype = ctypes.c_double
        L.ref_scalar.argtypes = [vp, ctypes.c_char_p]
        L.ref_array.restype = ctypes.c_long
        L.ref_array.argtypes = [vp, ctypes.c_char_p, vp]
        L.ref_check.restype = ctypes.c_int
        L.ref_check.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.ref_read_mtx.restype = ctypes.c_int
        L.ref_read_mtx.argtypes = [ctypes.c_char_p, ip, ip, ip, ctypes.POINTER(ip), ctypes.POINTER(ip),
                                   ctypes.POINTER(dp)]
        L.ref_free_buf.argtypes = [vp]
        L.ref_time_check.restype = ctypes.c_double
        L.ref_time_check.argtypes = [ctypes.c_int, ip, ip, dp, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
        _ref_libs[path] = L
    return _ref_libs[path]


class RefBuild(_Built):
    """The reference builder itself (oracle/_ref)."""

    def __init__(self, m, n, row, col, val, bs=16, waves=4, row_owned=True, force_coo64=False, comp=False):
        L = ref_lib(bs, waves, row_owned, force_coo64)
        row, col, val = _i32(row), _i32(col), _f64(val)
        h = L.ref_build(int(m), int(n), int(row.size), _p(row, ctypes.c_int), _p(col, ctypes.c_int),
                        _p(val, ctypes.c_double), int(comp))
        super().__init__(L, h, L.ref_array, L.ref_scalar, L.ref_free)


def spmv_seq(m, row, col, val, x):
    """y = A x by the reference's sequential loop (utils.h:98-111), FP64."""
    row, col, val, x = _i32(row), _i32(col), _f64(val), _f64(x)
    y = np.zeros(int(m), dtype=np.float64)
    oracle_lib().oracle_spmv_seq(int(row.size), _p(row, ctypes.c_int), _p(col, ctypes.c_int),
                                 _p(val, ctypes.c_double), int(m), _p(x, ctypes.c_double),
                                 _p(y, ctypes.c_double))
    return y


def csr_row_ptr(m, row):
    """int64 row pointer of a row-major COO row array (rows non-decreasing)."""
    ptr = np.zeros(int(m) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum(np.bincount(np.asarray(row, dtype=np.int64), minlength=int(m)))
    return ptr


class CSR:
    """The CSR form of a row-major COO matrix, built once, for repeated
    row-parallel SpMVs (same per-row order as utils.h:98-111)."""

    def __init__(self, m, row, col, val):
        self.m = int(m)
        self.ptr = csr_row_ptr(m, row)
        self.col, self.val = _i32(col), _f64(val)

    def spmv(self, x, threads=16):
        x = _f64(x)
        y = np.zeros(self.m, dtype=np.float64)
        oracle_lib().oracle_spmv_csr_omp(self.m, _p(self.ptr, ctypes.c_long), _p(self.col, ctypes.c_int),
                                         _p(self.val, ctypes.c_double), _p(x, ctypes.c_double),
                                         _p(y, ctypes.c_double), int(threads))
        return y


    def time_spmv(self, x, threads, reps):
        """Median seconds per SpMV over `reps` runs after one warm-up, timed
        in C++ around the OpenMP loop alone."""
        x = _f64(x)
        y = np.zeros(self.m, dtype=np.float64)
        return oracle_lib().oracle_time_csr_omp(self.m, _p(self.ptr, ctypes.c_long), _p(self.col, ctypes.c_int),
                                                _p(self.val, ctypes.c_double), _p(x, ctypes.c_double),
                                                _p(y, ctypes.c_double), int(threads), int(reps))


def host_threads():
    """{omp_max_threads, hardware_concurrency, affinity_cpus} of this process."""
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    oracle_lib().oracle_host_threads(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return {"omp_max_threads": a.value, "hardware_concurrency": b.value, "affinity_cpus": c.value}


def spmv_csr_omp(m, row, col, val, x, threads):
    """Row-parallel CSR form of the same sum (same per-row order)."""
    return CSR(m, row, col, val).spmv(x, threads)


def ref_check(m, n, row, col, val, x, y):
    """check_results_consistency() of the reference (utils.h:85-136)."""
    L = ref_lib()
    row, col, val, x, y = _i32(row), _i32(col), _f64(val), _f64(x), _f64(y)
    return bool(L.ref_check(int(row.size), _p(row, ctypes.c_int), _p(col, ctypes.c_int), _p(val, ctypes.c_double),
                            int(m), int(n), _p(x, ctypes.c_double), _p(y, ctypes.c_double)))


def ref_read_mtx(path):
    """The reference reader (read_matrix_coo_data, mmio_highlevel.h:313)."""
    L = ref_lib()
    ip = ctypes.POINTER(ctypes.c_int)
    m, n, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    r, c, v = ip(), ip(), ctypes.POINTER(ctypes.c_double)()
    rc = L.ref_read_mtx(str(path).encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(r),
                        ctypes.byref(c), ctypes.byref(v))
    if rc != 0:
        raise RuntimeError(f"reference reader returned {rc}")
    k = nnz.value
    out = (m.value, n.value, np.ctypeslib.as_array(r, shape=(k,)).copy(),
           np.ctypeslib.as_array(c, shape=(k,)).copy(), np.ctypeslib.as_array(v, shape=(k,)).copy())
    for ptr in (r, c, v):
        L.ref_free_buf(ctypes.cast(ptr, ctypes.c_void_p))
    return out


def rel_l2(y, y_ref):
    y, y_ref = np.asarray(y, dtype=np.float64), np.asarray(y_ref, dtype=np.float64)
    den = np.linalg.norm(y_ref)
    num = np.linalg.norm(y - y_ref)
    return float(num / den) if den > 0 else float(num)
