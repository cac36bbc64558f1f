// Oracle-side shim around the reference's own host code, compiled from the
// reference sources where they lie (see oracle/Makefile; nothing is copied).
// The resulting oracle/_ref/*.so is test infrastructure: tests compare the
// oracle's restatement and the product builder against it, and bench.py's
// cpu_baseline leg may time check_results_consistency() from it.
//
// Wrapped reference entry points:
//   read_matrix_coo_data         cb-spmv-dcu/src/io/mmio_highlevel.h:313
//   GLOBAL_BLOCK ctor            cb-spmv-dcu/src/format.h:106
//   coo2block_gather / _comp     cb-spmv-dcu/src/coo2block.h:628 / :605
//   check_results_consistency    cb-spmv-dcu/src/utils.h:85
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "coo2block.h"
#include "utils.h"
#include "io/mmio_highlevel.h"

namespace {
template <typename T>
long copy_out(const std::vector<T>& v, void* dst) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
    return (long)v.size();
}
}  // namespace

extern "C" {

// Reads a .mtx file with the reference reader; returns its status and the
// row-major COO (caller frees the three arrays with ref_free_buf).
int ref_read_mtx(const char* path, int* m, int* n, int* nnz, int** row, int** col, double** val) {
    std::string type;
    int M = 0, N = 0, L = 0;
    int rc = read_matrix_info(path, type, M, N, L);
    if (rc != 0) return rc;
    std::vector<int> I, J;
    std::vector<double> V;
    rc = read_matrix_coo_data(path, I, J, V);
    if (rc != 0) return rc;
    *m = M; *n = N; *nnz = (int)I.size();
    *row = (int*)malloc(sizeof(int) * (I.size() + 1));
    *col = (int*)malloc(sizeof(int) * (I.size() + 1));
    *val = (double*)malloc(sizeof(double) * (I.size() + 1));
    std::copy(I.begin(), I.end(), *row);
    std::copy(J.begin(), J.end(), *col);
    std::copy(V.begin(), V.end(), *val);
    return 0;
}
void ref_free_buf(void* p) { free(p); }

int ref_block_size() { return BLOCK_SIZE; }
int ref_waves() { return WARP_PER_BLOCK; }
int ref_row_owned() { return CB_ROW_OWNED; }

void* ref_build(int m, int n, int nnz, const int* row, const int* col, const double* val, int comp) {
    GLOBAL_BLOCK* gb = new GLOBAL_BLOCK(m, n, nnz);
    if (comp) coo2block_comp(row, col, val, *gb);
    else coo2block_gather(row, col, val, *gb);
    return gb;
}

void ref_free(void* h) { delete static_cast<GLOBAL_BLOCK*>(h); }

double ref_scalar(void* h, const char* name) {
    const GLOBAL_BLOCK& g = *static_cast<GLOBAL_BLOCK*>(h);
    const std::string s(name);
    if (s == "gb_m") return g.gb_m;
    if (s == "gb_n") return g.gb_n;
    if (s == "gb_nnzb") return g.gb_nnzb;
    if (s == "gather_flag") return g.gather_flag;
    if (s == "direct_coo_columns") return g.direct_coo_columns;
    if (s == "compact_coo_indices") return g.compact_coo_indices;
    if (s == "super_sparse_blocks") return g.super_sparse_blocks;
    if (s == "analyzed_blocks") return g.analyzed_blocks;
    if (s == "super_sparse_ratio") return g.super_sparse_ratio;
    if (s == "load_balanced") return g.load_balanced;
    if (s == "load_stddev_before") return g.load_stddev_before;
    if (s == "load_stddev_after") return g.load_stddev_after;
    return -1.0;
}

// Same array names and payload-stream encoding as the oracle (oracle_array).
long ref_array(void* h, const char* name, void* dst) {
    GLOBAL_BLOCK& g = *static_cast<GLOBAL_BLOCK*>(h);
    const std::string s(name);
    if (s == "gb_row_idx") return copy_out(g.gb_row_idx, dst);
    if (s == "gb_col_idx") return copy_out(g.gb_col_idx, dst);
    if (s == "gb_row_ptr") return copy_out(g.gb_row_ptr, dst);
    if (s == "nnz_per_blk") return copy_out(g.nnz_per_blk, dst);
    if (s == "type_per_blk") return copy_out(g.type_per_blk, dst);
    if (s == "segment_rows") return copy_out(g.segment_rows, dst);
    if (s == "segment_begin") return copy_out(g.segment_begin, dst);
    if (s == "segment_end") return copy_out(g.segment_end, dst);
    if (s == "segment_direct") return copy_out(g.segment_direct, dst);
    if (s == "segment_uniform") return copy_out(g.segment_uniform, dst);
    if (s == "block_row_uniform") return copy_out(g.block_row_uniform, dst);
    if (s == "flat_row_ptr") return copy_out(g.flat_row_ptr, dst);
    if (s == "flat_col_idx") return copy_out(g.flat_col_idx, dst);
    if (s == "pattern_row_ptr") return copy_out(g.pattern_row_ptr, dst);
    if (s == "pattern_col_idx") return copy_out(g.pattern_col_idx, dst);
    if (s == "pattern_output_ptr") return copy_out(g.pattern_output_ptr, dst);
    if (s == "pattern_output_rows") return copy_out(g.pattern_output_rows, dst);
    if (s == "block_cols") return copy_out(g.block_cols, dst);
    if (s == "block_cols_offset") return copy_out(g.block_cols_offset, dst);
    if (s == "bia_nnz" || s == "bia_origin" || s == "bia_end") {
        std::vector<int> v;
        for (const BLOCK_IDX& b : g.block_idx_array)
            v.push_back(s == "bia_nnz" ? b.blk_nnz : s == "bia_origin" ? b.origin_idx : b.end_idx);
        return copy_out(v, dst);
    }
    if (s == "pl_idx" || s == "pl_val") {
        std::vector<int32_t> idx;
        std::vector<double> val;
        for (int i = 0; i < g.gb_nnzb; ++i) {
            if (g.nnz_per_blk[i] <= 0) continue;
            const int o = g.block_idx_array[i].origin_idx;
            const int t = g.type_per_blk[i] & BLOCK_TYPE_MASK;
            if (t == COO) {
                const COO_BLOCK_COMP& b = g.data_per_coo_blk_comp.at(o);
                for (IndexTypeComp w : b.idx) idx.push_back((int32_t)w);
                val.insert(val.end(), b.val.begin(), b.val.end());
            } else if (t == CSR) {
                const CSR_BLOCK& b = g.data_per_csr_blk.at(o);
                for (IndexTypeOrig w : b.row_ptr) idx.push_back((int32_t)w);
                for (IndexTypeOrig w : b.col_idx) idx.push_back((int32_t)w);
                val.insert(val.end(), b.val.begin(), b.val.end());
            } else {
                const DENSE_BLOCK& b = g.data_per_dense_blk.at(o);
                val.insert(val.end(), b.val.begin(), b.val.end());
            }
        }
        return s == "pl_idx" ? copy_out(idx, dst) : copy_out(val, dst);
    }
    return -1;
}

// The reference CPU path: check_results_consistency() recomputes y with the
// sequential FP64 loop and compares it with y_test.  Returns 1 if it passes.
int ref_check(int nnz, const int* row, const int* col, const double* val,
              int m, int n, const double* x, const double* y_test) {
    std::vector<int> r(row, row + nnz), c(col, col + nnz);
    std::vector<double> v(val, val + nnz), xv(x, x + n), yv(y_test, y_test + m);
    return check_results_consistency(r, c, v, xv, yv) ? 1 : 0;
}

// Times `reps` calls of check_results_consistency() on vectors built once;
// returns the median seconds per call.
double ref_time_check(int nnz, const int* row, const int* col, const double* val,
                      int m, int n, const double* x, const double* y_test, int reps) {
    std::vector<int> r(row, row + nnz), c(col, col + nnz);
    std::vector<double> v(val, val + nnz), xv(x, x + n), yv(y_test, y_test + m);
    std::vector<double> t;
    for (int k = 0; k < reps; ++k) {
        const auto a = std::chrono::steady_clock::now();
        volatile bool ok = check_results_consistency(r, c, v, xv, yv);
        (void)ok;
        const auto b = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(b - a).count());
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? 0.0 : t[t.size() / 2];
}

}  // extern "C"
