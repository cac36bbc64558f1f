"""The reference's host interface for this path, with the same names and
argument meaning, over the MI355X library:

  GLOBAL_BLOCK(m, n, nnz)                    cb-spmv-dcu/src/format.h:106
  coo2block_gather(row, col, val, gb)        cb-spmv-dcu/src/coo2block.h:628
  coo2block_comp(row, col, val, gb)          cb-spmv-dcu/src/coo2block.h:605
  cb_spmv(gb, x, y) -> PROGRAM_METRICS       cb-spmv-dcu/src/cb-spmv-dcu.hpp:1088
  print_benchmark_report(...)                cb-spmv-dcu/src/utils.h:147

The compile-time macros of the reference (CB_BLOCK_SIZE, CB_WAVES_PER_BLOCK,
CB_ROW_OWNED, CB_PATTERN_WAVE, ...) become keyword options of GLOBAL_BLOCK.
Errors raise CbspmvError instead of exiting the process.
"""
import sys
from dataclasses import dataclass

import numpy as np

from .plan import ARRAY_NAMES, Plan


@dataclass
class PROGRAM_METRICS:
    """format.h:185-188: runtime in ms per launch and GFLOPS = 2 nnz / t."""
    runtime: float = 0.0
    gflops: float = 0.0


class GLOBAL_BLOCK:
    """The CB structure of one matrix; filled by coo2block_gather/_comp."""

    def __init__(self, mtx_m, mtx_n, mtx_nnz, block_size=16, waves_per_block=4, row_owned=True,
                 pattern_wave=True, force_coo64=False, dense_sixteenths=15, **device_opts):
        self.mtx_m, self.mtx_n, self.mtx_nnz = int(mtx_m), int(mtx_n), int(mtx_nnz)
        self.block_size = int(block_size)
        self.gb_m = (self.mtx_m + self.block_size - 1) // self.block_size
        self.gb_n = (self.mtx_n + self.block_size - 1) // self.block_size
        self.gb_nnzb = 0
        self._opts = dict(block_size=block_size, waves=waves_per_block, row_owned=row_owned,
                          pattern_wave=pattern_wave, force_coo64=force_coo64,
                          dense_sixteenths=dense_sixteenths, **device_opts)
        self.plan = None
        self._stats = {}

    def _build(self, row_idx, col_idx, val, builder):
        row = np.asarray(row_idx)
        if row.size != self.mtx_nnz:
            raise ValueError(f"GLOBAL_BLOCK was sized for {self.mtx_nnz} nonzeros, got {row.size}")
        if self.plan is not None:
            self.plan.close()
        self.plan = Plan(self.mtx_m, self.mtx_n, row, col_idx, val, builder=builder, **self._opts)
        self._stats = self.plan.stats()
        self.gb_nnzb = self._stats["gb_nnzb"]

    def __getattr__(self, name):
        # GLOBAL_BLOCK scalar fields (format.h:63-77) and arrays (:79-104)
        if name.startswith("_") or name == "plan":
            raise AttributeError(name)
        if self.plan is None:
            raise AttributeError(f"{name}: GLOBAL_BLOCK not built yet")
        scalars = {"gather_flag", "direct_coo_columns", "compact_coo_indices", "super_sparse_blocks",
                   "analyzed_blocks", "super_sparse_ratio", "load_balanced", "load_stddev_before",
                   "load_stddev_after"}
        if name in scalars:
            v = self._stats[name]
            return bool(v) if name in {"gather_flag", "direct_coo_columns", "compact_coo_indices",
                                        "load_balanced"} else v
        if name in ARRAY_NAMES:
            return self.plan.export(name)
        raise AttributeError(name)


def coo2block_gather(row_idx, col_idx, val, gb):
    """Builds gb with column gathering when super-sparse blocks dominate."""
    gb._build(row_idx, col_idx, val, "gather")


def coo2block_comp(row_idx, col_idx, val, gb):
    """Builds gb without column gathering."""
    gb._build(row_idx, col_idx, val, "comp")


def cb_spmv(gb, x, y):
    """y[:] = A x from the first device launch; returns the timing of 1000
    launches after 100 warm-ups (cb-spmv-dcu.hpp:1625-1660)."""
    if gb.plan is None:
        raise RuntimeError("cb_spmv: GLOBAL_BLOCK not built")
    y_out, runtime, gflops = gb.plan.bench(x)
    y[: gb.mtx_m] = y_out
    return PROGRAM_METRICS(runtime, gflops)


def print_benchmark_report(matrix_name, rows, columns, nonzeros, block, metrics, result_is_correct,
                           file=sys.stdout):
    """Same report layout as utils.h:147-201 (without terminal colours)."""
    div, sep = "=" * 64, "-" * 64
    w = file.write
    w(f"\n{div}\n                       CB-SpMV Benchmark\n{div}\n")
    w(f"{'Matrix':<24}: {matrix_name}\n{'Dimensions':<24}: {rows} x {columns}\n")
    w(f"{'Nonzeros':<24}: {nonzeros}\n")
    w(f"{'Sparse blocks (nnz<=8)':<24}: {block.super_sparse_blocks} / {block.analyzed_blocks} "
      f"({block.super_sparse_ratio * 100.0:.6f}%)\n")
    w(f"{'Column gather':<24}: {'YES' if block.gather_flag else 'NO'}\n")
    w(f"{'Load balance':<24}: {'YES' if block.load_balanced else 'NO'}\n")
    if block.load_balanced:
        w(f"{'Load stddev':<24}: {block.load_stddev_before:.6f} -> {block.load_stddev_after:.6f}\n")
    else:
        w(f"{'Load stddev':<24}: N/A -> N/A\n")
    w(f"{sep}\n{'Method':<14}{'Time (ms)':>25}{'GFLOPS':>25}\n{sep}\n")
    w(f"{'CB-SpMV-MI355X':<14}{metrics.runtime:>25.6f}{metrics.gflops:>25.6f}\n{sep}\n")
    w(f"{'Validation':<14}: {'PASS' if result_is_correct else 'FAIL'}\n{div}\n")


def writeCSV(filename, metrics, mtx_name, nnz):
    """utils.h:256-268: appends "matrix,nnz,runtime_ms,gflops" to a CSV file."""
    with open(filename, "a") as fh:
        fh.write(f"{mtx_name},{nnz},{metrics.runtime:g},{metrics.gflops:g}\n")


def benchmark_report_dict(matrix_name, rows, columns, nonzeros, block, metrics, result_is_correct):
    """The fields of print_benchmark_report (utils.h:147-201) as a dict, for
    JSON output."""
    return {
        "matrix": matrix_name, "rows": int(rows), "columns": int(columns), "nonzeros": int(nonzeros),
        "super_sparse_blocks": int(block.super_sparse_blocks), "analyzed_blocks": int(block.analyzed_blocks),
        "super_sparse_ratio": float(block.super_sparse_ratio), "column_gather": bool(block.gather_flag),
        "load_balanced": bool(block.load_balanced),
        "load_stddev": [float(block.load_stddev_before), float(block.load_stddev_after)] if block.load_balanced else None,
        "runtime_ms": float(metrics.runtime), "gflops": float(metrics.gflops), "validation": bool(result_is_correct),
    }
