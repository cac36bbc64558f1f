// cb_spmv.hpp - source-level drop-in for the reference's C++ call surface,
// header-only over the C ABI of libcbspmv.so (cbspmv.h).
//
// A reference caller (cb-spmv-dcu/src/main.hip.cpp:46-75) builds and runs a
// matrix with
//     GLOBAL_BLOCK gb(m, n, nnz);                       // format.h:106
//     coo2block_gather(row, col, val, gb);              // coo2block.h:628
//     PROGRAM_METRICS met = cb_spmv(gb, x, y);           // cb-spmv-dcu.hpp:1088
//     print_benchmark_report(name, m, n, nnz, gb, met, ok);   // utils.h:147
// Including this header instead of format.h / coo2block.h /
// cb-spmv-dcu.hpp keeps those lines unchanged.  The reference's compile-time
// options are read from the same macros (CB_BLOCK_SIZE, CB_WAVES_PER_BLOCK,
// CB_ROW_OWNED, CB_PATTERN_WAVE, CB_FORCE_COO64, CB_DENSE_SIXTEENTHS) when
// defined, else from the reference Makefile's build (16, 4, 1, 1, 0, 15);
// a GLOBAL_BLOCK can also be given a cbspmv_opts at run time.
//
// Differences from the reference, by design:
//   * errors throw cbspmv_error (the reference prints and exits);
//   * GLOBAL_BLOCK owns a device plan that persists across cb_spmv calls (the
//     reference uploads and frees a slab on every call); its arrays are
//     filled from the plan on construction and are read-only copies;
//   * the payload maps (data_per_*_blk) are not materialised; the packed
//     payload streams are available through cbspmv_export ("pl_idx",
//     "pl_val").
#ifndef CB_SPMV_HPP
#define CB_SPMV_HPP

#include <cstdint>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cbspmv.h"

struct cbspmv_error : std::runtime_error {
    int code;
    cbspmv_error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void cbspmv_check(int rc) {
    if (rc != CBSPMV_OK) throw cbspmv_error(rc, cbspmv_last_error());
}

inline cbspmv_opts cbspmv_macro_opts() {
    cbspmv_opts o;
    cbspmv_default_opts(&o);
#ifdef CB_BLOCK_SIZE
    o.block_size = CB_BLOCK_SIZE;
#endif
#ifdef CB_WAVES_PER_BLOCK
    o.waves_per_block = CB_WAVES_PER_BLOCK;
#endif
#ifdef CB_ROW_OWNED
    o.row_owned = CB_ROW_OWNED;
#endif
#ifdef CB_PATTERN_WAVE
    o.pattern_wave = CB_PATTERN_WAVE;
#endif
#ifdef CB_FORCE_COO64
    o.force_coo64 = CB_FORCE_COO64;
#endif
#ifdef CB_DENSE_SIXTEENTHS
    o.dense_sixteenths = CB_DENSE_SIXTEENTHS;
#endif
    return o;
}

// format.h:185-188
struct PROGRAM_METRICS {
    double runtime{0.0};   // ms per launch
    double gflops{0.0};    // 2 nnz / time
};

// format.h:62-115: the scalar fields and index arrays of the CB structure.
struct GLOBAL_BLOCK {
    int mtx_m, mtx_n, mtx_nnz;
    bool gather_flag = false;
    bool direct_coo_columns = false;
    bool compact_coo_indices = false;
    int super_sparse_blocks = 0;
    int analyzed_blocks = 0;
    double super_sparse_ratio = 0.0;
    bool load_balanced = false;
    double load_stddev_before = 0.0;
    double load_stddev_after = 0.0;
    int gb_m, gb_n, gb_nnzb = 0;

    std::vector<int> gb_row_idx, gb_col_idx, gb_row_ptr;
    std::vector<int> segment_rows, segment_begin, segment_end;
    std::vector<uint8_t> segment_direct, segment_uniform, block_row_uniform;
    std::vector<int> flat_row_ptr, flat_col_idx;
    std::vector<int> pattern_row_ptr, pattern_col_idx, pattern_output_ptr, pattern_output_rows;
    std::vector<int> nnz_per_blk;
    std::vector<uint8_t> type_per_blk;
    std::vector<int> block_cols_offset, block_cols;

    cbspmv_opts opts;
    cbspmv_plan* plan = nullptr;

    GLOBAL_BLOCK(int m, int n, int nnz, const cbspmv_opts& o = cbspmv_macro_opts())
        : mtx_m(m), mtx_n(n), mtx_nnz(nnz), opts(o) {
        gb_m = (m + o.block_size - 1) / o.block_size;
        gb_n = (n + o.block_size - 1) / o.block_size;
    }
    GLOBAL_BLOCK(const GLOBAL_BLOCK&) = delete;
    GLOBAL_BLOCK& operator=(const GLOBAL_BLOCK&) = delete;
    ~GLOBAL_BLOCK() { cbspmv_destroy(plan); }

    void build(const int* row, const int* col, const double* val, int builder) {
        cbspmv_destroy(plan);
        plan = nullptr;
        cbspmv_opts o = opts;
        o.builder = builder;
        cbspmv_check(cbspmv_build(mtx_m, mtx_n, mtx_nnz, row, col, val, &o, &plan));
        cbspmv_stats s;
        cbspmv_check(cbspmv_plan_stats(plan, &s));
        gather_flag = s.gather_flag;
        direct_coo_columns = s.direct_coo_columns;
        compact_coo_indices = s.compact_coo_indices;
        super_sparse_blocks = s.super_sparse_blocks;
        analyzed_blocks = s.analyzed_blocks;
        super_sparse_ratio = s.super_sparse_ratio;
        load_balanced = s.load_balanced;
        load_stddev_before = s.load_stddev_before;
        load_stddev_after = s.load_stddev_after;
        gb_m = s.gb_m;
        gb_n = s.gb_n;
        gb_nnzb = s.gb_nnzb;
        fetch("gb_row_idx", gb_row_idx);
        fetch("gb_col_idx", gb_col_idx);
        fetch("gb_row_ptr", gb_row_ptr);
        fetch("segment_rows", segment_rows);
        fetch("segment_begin", segment_begin);
        fetch("segment_end", segment_end);
        fetch("segment_direct", segment_direct);
        fetch("segment_uniform", segment_uniform);
        fetch("block_row_uniform", block_row_uniform);
        fetch("flat_row_ptr", flat_row_ptr);
        fetch("flat_col_idx", flat_col_idx);
        fetch("pattern_row_ptr", pattern_row_ptr);
        fetch("pattern_col_idx", pattern_col_idx);
        fetch("pattern_output_ptr", pattern_output_ptr);
        fetch("pattern_output_rows", pattern_output_rows);
        fetch("nnz_per_blk", nnz_per_blk);
        fetch("type_per_blk", type_per_blk);
        fetch("block_cols_offset", block_cols_offset);
        fetch("block_cols", block_cols);
    }

   private:
    template <typename T>
    void fetch(const char* name, std::vector<T>& v) {
        size_t n = 0;
        cbspmv_check(cbspmv_export(plan, name, nullptr, &n));
        v.resize(n);
        if (n) cbspmv_check(cbspmv_export(plan, name, v.data(), &n));
    }
};

// coo2block.h:628: CB build with column gathering when super-sparse blocks
// dominate (bit-exact with the reference builder for the same options).
inline void coo2block_gather(const int* row_idx, const int* col_idx, const double* val, GLOBAL_BLOCK& gb) {
    gb.build(row_idx, col_idx, val, CBSPMV_BUILDER_GATHER);
}

// coo2block.h:605: CB build that never gathers columns.
inline void coo2block_comp(const int* row_idx, const int* col_idx, const double* val, GLOBAL_BLOCK& gb) {
    gb.build(row_idx, col_idx, val, CBSPMV_BUILDER_COMP);
}

// cb-spmv-dcu.hpp:1088: y = A x on the GPU with the reference's protocol
// (validated first launch copied to host y, 100 warm-up and 1000 timed
// launches).  x and y are host arrays of n and m doubles.
inline PROGRAM_METRICS cb_spmv(const GLOBAL_BLOCK& gb, const double* x, double* y) {
    if (!gb.plan) throw cbspmv_error(CBSPMV_ERR_STATE, "cb_spmv: GLOBAL_BLOCK not built");
    cbspmv_metrics met{};
    cbspmv_check(cbspmv_bench(gb.plan, x, y, &met));
    PROGRAM_METRICS out;
    out.runtime = met.runtime;
    out.gflops = met.gflops;
    return out;
}

// utils.h:147-201: the same report text.
inline void print_benchmark_report(const std::string& matrix_name, int rows, int columns, int nonzeros,
                                   const GLOBAL_BLOCK& block, const PROGRAM_METRICS& metrics,
                                   bool result_is_correct) {
    const std::string divider(64, '='), separator(64, '-');
    std::ostream& o = std::cout;
    o << '\n' << divider << '\n' << "                       CB-SpMV Benchmark\n" << divider << '\n' << std::left
      << std::setw(24) << "Matrix" << ": " << matrix_name << '\n'
      << std::setw(24) << "Dimensions" << ": " << rows << " x " << columns << '\n'
      << std::setw(24) << "Nonzeros" << ": " << nonzeros << '\n'
      << std::setw(24) << "Sparse blocks (nnz<=8)" << ": " << block.super_sparse_blocks << " / "
      << block.analyzed_blocks << " (" << std::fixed << std::setprecision(6) << block.super_sparse_ratio * 100.0
      << "%)\n"
      << std::setw(24) << "Column gather" << ": " << (block.gather_flag ? "YES" : "NO") << '\n'
      << std::setw(24) << "Load balance" << ": " << (block.load_balanced ? "YES" : "NO") << '\n'
      << std::setw(24) << "Load stddev" << ": ";
    if (block.load_balanced) o << block.load_stddev_before << " -> " << block.load_stddev_after << '\n';
    else o << "N/A -> N/A\n";
    o << separator << '\n' << std::left << std::setw(14) << "Method" << std::right << std::setw(25) << "Time (ms)"
      << std::setw(25) << "GFLOPS" << '\n' << separator << '\n' << std::left << std::setw(14) << "CB-SpMV-MI355X"
      << std::right << std::fixed << std::setprecision(6) << std::setw(25) << metrics.runtime << std::setw(25)
      << metrics.gflops << '\n' << separator << '\n' << std::left << std::setw(14) << "Validation" << ": "
      << (result_is_correct ? "\033[32m" : "\033[31m") << (result_is_correct ? "PASS" : "FAIL") << "\033[0m" << '\n'
      << divider << '\n';
}

// utils.h:256-268: appends "matrix,nnz,runtime_ms,gflops" to a CSV file.
inline bool writeCSV(const std::string& filename, const PROGRAM_METRICS& metrics, const std::string& mtx_name,
                     int nnz) {
    FILE* f = std::fopen(filename.c_str(), "a");
    if (!f) return false;
    std::fprintf(f, "%s,%d,%g,%g\n", mtx_name.c_str(), nnz, metrics.runtime, metrics.gflops);
    return std::fclose(f) == 0;
}

#endif  // CB_SPMV_HPP
