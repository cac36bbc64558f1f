"""CB-SpMV for AMD Instinct MI355X (gfx950): Python host side.

The compute path is libcbspmv.so (hand-written HIP kernels behind a C ABI,
include/cbspmv.h).  This package binds it with ctypes and uses torch only for
device memory and streams.
"""
from ._lib import CbspmvError, LIB_PATH, lib  # noqa: F401
from .plan import ARRAY_NAMES, Plan, make_opts, read_mtx  # noqa: F401
from .autograd import CBSparse, cb_spmv as cb_spmv_autograd  # noqa: F401
from .batch import Batch, autotune  # noqa: F401
from .reference_api import (  # noqa: F401
    GLOBAL_BLOCK,
    PROGRAM_METRICS,
    cb_spmv,
    coo2block_comp,
    coo2block_gather,
    benchmark_report_dict,
    print_benchmark_report,
    writeCSV,
)
