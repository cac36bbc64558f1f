"""The rocSPARSE comparators bench.py times beside the CB-SpMV kernels
(libvendor_csr.so: rocsparse_spmv on CSR, rocsparse_bsrmv at block dims 2, 4,
8, 16 after rocsparse_csr2bsr, standing for the reference's hipSPARSE CSR and
BSR baseline programs).  Their y is checked against the oracle's FP64 SpMV
(relative L2 <= 1e-12 for FP64; 1e-5 for FP32 against the oracle on the
FP32-rounded matrix and x), on matrices whose sizes are not multiples of
the block dims, so that padded x and y blocks are exercised."""
import ctypes
import os

import numpy as np
import pytest

from cbspmv import gen, read_mtx
from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "cb-sparse-dcu_amd", "libvendor_csr.so")
GOLD = os.path.join(ROOT, "tests", "golden")
CSR_ALGS = {"csr_adaptive": 2, "csr_rowsplit": 3, "csr_lrb": 7}


def _lib():
    if not os.path.exists(LIB):
        pytest.fail("libvendor_csr.so is not built")
    L = ctypes.CDLL(LIB)
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.vendor_csr_spmv.restype = ctypes.c_int
    L.vendor_csr_spmv.argtypes = [ctypes.c_int] * 3 + [ip, ip, dp, dp, dp] + [ctypes.c_int] * 4 + [dp]
    L.vendor_bsr_spmv.restype = ctypes.c_int
    L.vendor_bsr_spmv.argtypes = ([ctypes.c_int] * 3 + [ip, ip, dp, dp, dp] + [ctypes.c_int] * 4 +
                                  [dp, ctypes.POINTER(ctypes.c_longlong)])
    return L, ip, dp


def _matrices():
    m, n, r, c, _ = read_mtx(os.path.join(GOLD, "1138_bus.mtx"))
    v = np.random.default_rng(1).uniform(-1, 1, r.size)
    yield "1138_bus", (m, n, r, c, v)
    r, c, v = gen.random_coo(3001, 2777, 9, seed=4, dup=False)
    yield "rect_3001x2777", (3001, 2777, r, c, v)


@pytest.mark.gpu
@pytest.mark.parametrize("f32", [0, 1])
def test_vendor_comparators_match_oracle(gpu, f32):
    L, ip, dp = _lib()
    for name, (m, n, r, c, v) in _matrices():
        ptr = np.zeros(m + 1, np.int32)
        np.cumsum(np.bincount(r, minlength=m), out=ptr[1:])
        order = np.lexsort((c, r))   # rocSPARSE's CSR: columns ascending within a row
        col = np.ascontiguousarray(c[order], np.int32)
        val = np.ascontiguousarray(v[order], np.float64)
        x = np.random.default_rng(2).uniform(-1, 1, n)
        if f32:
            ref = pyoracle.spmv_seq(m, r, c, v.astype(np.float32).astype(np.float64),
                                    x.astype(np.float32).astype(np.float64))
            tol = 1e-5
        else:
            ref = pyoracle.spmv_seq(m, r, c, v, x)
            tol = 1e-12
        args = (m, n, int(r.size), ptr.ctypes.data_as(ip), col.ctypes.data_as(ip), val.ctypes.data_as(dp),
                x.ctypes.data_as(dp))
        for alg_name, alg in CSR_ALGS.items():
            y = np.full(m, 7.0)
            ms = ctypes.c_double()
            print(f"vendor csr {name} {alg_name} f32={f32}", flush=True)
            assert L.vendor_csr_spmv(*args, y.ctypes.data_as(dp), alg, f32, 2, 3, ctypes.byref(ms)) == 0
            assert ms.value > 0
            assert pyoracle.rel_l2(y, ref) <= tol, (name, alg_name)
        for bd in (2, 4, 8, 16):
            y = np.full(m, 7.0)
            ms, nnzb = ctypes.c_double(), ctypes.c_longlong()
            print(f"vendor bsr {name} bd={bd} f32={f32}", flush=True)
            assert L.vendor_bsr_spmv(*args, y.ctypes.data_as(dp), bd, f32, 2, 3, ctypes.byref(ms),
                                     ctypes.byref(nnzb)) == 0
            # block count: distinct (row // bd, col // bd) pairs
            want = np.unique((r // bd).astype(np.int64) * ((n + bd - 1) // bd) + c // bd).size
            assert nnzb.value == want, (name, bd)
            assert pyoracle.rel_l2(y, ref) <= tol, (name, bd)
