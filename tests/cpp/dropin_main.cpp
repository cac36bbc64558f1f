// Reference-style caller of the drop-in header (the body of
// cb-spmv-dcu/src/main.hip.cpp:46-75 with the reader swapped for
// cbspmv_read_mtx): builds the CB structure of a .mtx file through
// GLOBAL_BLOCK + coo2block_gather and prints its scalars and array sizes
// and sums as JSON; with --run also runs cb_spmv (the reference protocol) on
// x from mt19937(42) U[0,1) and writes y to the file named after --run.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "cb_spmv.hpp"

template <typename T>
static void emit(const char* name, const std::vector<T>& v, bool last = false) {
    long long sum = 0;
    for (const T& e : v) sum += (long long)e;
    std::printf("\"%s\": [%zu, %lld]%s", name, v.size(), sum, last ? "" : ", ");
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    int m, n, nnz;
    int *row, *col;
    double* val;
    if (cbspmv_read_mtx(argv[1], &m, &n, &nnz, &row, &col, &val) != 0) return 3;
    try {
        GLOBAL_BLOCK gb(m, n, nnz);
        coo2block_gather(row, col, val, gb);
        std::printf("{\"m\": %d, \"n\": %d, \"nnz\": %d, \"gather_flag\": %d, \"direct_coo_columns\": %d, "
                    "\"compact_coo_indices\": %d, \"super_sparse_blocks\": %d, \"analyzed_blocks\": %d, "
                    "\"gb_nnzb\": %d, ",
                    m, n, nnz, (int)gb.gather_flag, (int)gb.direct_coo_columns, (int)gb.compact_coo_indices,
                    gb.super_sparse_blocks, gb.analyzed_blocks, gb.gb_nnzb);
        emit("gb_row_idx", gb.gb_row_idx);
        emit("gb_col_idx", gb.gb_col_idx);
        emit("nnz_per_blk", gb.nnz_per_blk);
        emit("type_per_blk", gb.type_per_blk);
        emit("segment_rows", gb.segment_rows);
        emit("pattern_row_ptr", gb.pattern_row_ptr);
        emit("block_cols", gb.block_cols, true);
        std::printf("}\n");
        if (argc >= 4 && std::strcmp(argv[2], "--run") == 0) {
            std::mt19937 eng(42);
            std::uniform_real_distribution<double> u(0.0, 1.0);
            std::vector<double> x(n), y(m, 0.0);
            for (double& v : x) v = u(eng);
            PROGRAM_METRICS met = cb_spmv(gb, x.data(), y.data());
            FILE* f = std::fopen(argv[3], "wb");
            if (!f) return 4;
            std::fwrite(x.data(), sizeof(double), x.size(), f);
            std::fwrite(y.data(), sizeof(double), y.size(), f);
            std::fclose(f);
            print_benchmark_report(argv[1], m, n, nnz, gb, met, true);
        }
    } catch (const cbspmv_error& e) {
        std::fprintf(stderr, "cbspmv error %d: %s\n", e.code, e.what());
        return 5;
    }
    cbspmv_free(row);
    cbspmv_free(col);
    cbspmv_free(val);
    return 0;
}
