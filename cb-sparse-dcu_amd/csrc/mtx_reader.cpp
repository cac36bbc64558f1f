// Matrix Market reader with the reference's semantics
// (cb-spmv-dcu/src/io/mmio_highlevel.h:62-403):
//   * coordinate matrices of type real, integer, pattern or complex;
//   * every value is forced to 1.0 (mmio_highlevel.h:201-203);
//   * symmetric, hermitian and skew-symmetric files are expanded: each
//     off-diagonal entry (i, j) also yields (j, i), same value (:214-233,
//     :259-283);
//   * output is row-major COO: rows ascending, and within a row the order in
//     which entries reach that row while the file is read in order, mirrored
//     entries taking their place in row j at the time entry (i, j) is read.
// Status codes follow the reference reader's negative returns: -1 open,
// -2 banner, -3 unsupported type, -4 size line, -5 entry, -6 expanded nnz
// above INT_MAX, -7 allocation.
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cb_runtime.hpp"

namespace {

std::string lower(std::string s) {
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace

extern "C" {

int cbspmv_read_mtx(const char* path, int* m_out, int* n_out, int* nnz_out, int** row_out, int** col_out,
                    double** val_out) {
    using cbs::fail;
    if (!path || !m_out || !n_out || !nnz_out || !row_out || !col_out || !val_out)
        return fail(CBSPMV_ERR_ARG, "null argument");
    *row_out = nullptr; *col_out = nullptr; *val_out = nullptr;
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(-1, std::string("cannot open ") + path);
    char line[4096];
    if (!std::fgets(line, sizeof(line), f)) { std::fclose(f); return fail(-2, "empty file"); }
    char b0[64] = {0}, b1[64] = {0}, b2[64] = {0}, b3[64] = {0}, b4[64] = {0};
    if (std::sscanf(line, "%63s %63s %63s %63s %63s", b0, b1, b2, b3, b4) != 5 ||
        lower(b0) != "%%matrixmarket") {
        std::fclose(f);
        return fail(-2, "bad Matrix Market banner");
    }
    const std::string object = lower(b1), format = lower(b2), field = lower(b3), sym = lower(b4);
    const bool is_real = field == "real" || field == "double", is_int = field == "integer";
    const bool is_pattern = field == "pattern", is_complex = field == "complex";
    if (object != "matrix" || format != "coordinate" || !(is_real || is_int || is_pattern || is_complex) ||
        !(sym == "general" || sym == "symmetric" || sym == "hermitian" || sym == "skew-symmetric")) {
        std::fclose(f);
        return fail(-3, "only coordinate matrices of real/integer/pattern/complex type are supported");
    }
    const bool symmetric = sym != "general";
    long long M = -1, N = -1, L = -1;
    while (std::fgets(line, sizeof(line), f)) {
        if (line[0] == '%') continue;
        bool blank = true;
        for (char* q = line; *q; ++q) if (!std::isspace((unsigned char)*q)) { blank = false; break; }
        if (blank) continue;
        if (std::sscanf(line, "%lld %lld %lld", &M, &N, &L) != 3) M = -1;
        break;
    }
    if (M < 0 || N < 0 || L < 0 || M >= INT_MAX || N > INT_MAX || L > INT_MAX || (symmetric && M != N)) {
        std::fclose(f);
        return fail(-4, "bad size line");
    }
    std::vector<int> ri, ci;
    try {
        ri.resize((size_t)L);
        ci.resize((size_t)L);
    } catch (...) {
        std::fclose(f);
        return fail(-7, "out of host memory");
    }
    const int fields = is_complex ? 4 : is_pattern ? 2 : 3;
    for (long long k = 0; k < L; ++k) {
        long long i = 0, j = 0;
        double v = 0.0, vi = 0.0;
        int got;
        if (fields == 2) got = std::fscanf(f, "%lld %lld", &i, &j);
        else if (fields == 3) got = std::fscanf(f, "%lld %lld %lg", &i, &j, &v);
        else got = std::fscanf(f, "%lld %lld %lg %lg", &i, &j, &v, &vi);
        if (got != fields || i < 1 || i > M || j < 1 || j > N) {
            std::fclose(f);
            return fail(-5, "bad entry " + std::to_string(k + 1));
        }
        ri[(size_t)k] = (int)(i - 1);
        ci[(size_t)k] = (int)(j - 1);
    }
    std::fclose(f);

    // counting sort by row, entries (and their mirrors) in file order
    std::vector<long long> ptr((size_t)M + 1, 0);
    long long total = L;
    for (long long k = 0; k < L; ++k) {
        ptr[(size_t)ri[k] + 1]++;
        if (symmetric && ri[k] != ci[k]) {
            ptr[(size_t)ci[k] + 1]++;
            if (++total > INT_MAX) return fail(-6, "expanded nnz above INT_MAX");
        }
    }
    for (long long r = 0; r < M; ++r) ptr[(size_t)r + 1] += ptr[(size_t)r];
    int* row = static_cast<int*>(std::malloc(sizeof(int) * (size_t)std::max(total, 1ll)));
    int* col = static_cast<int*>(std::malloc(sizeof(int) * (size_t)std::max(total, 1ll)));
    double* val = static_cast<double*>(std::malloc(sizeof(double) * (size_t)std::max(total, 1ll)));
    if (!row || !col || !val) {
        std::free(row); std::free(col); std::free(val);
        return fail(-7, "out of host memory");
    }
    std::vector<long long> cur(ptr.begin(), ptr.end() - 1);
    for (long long k = 0; k < L; ++k) {
        const int i = ri[k], j = ci[k];
        long long at = cur[(size_t)i]++;
        row[at] = i; col[at] = j; val[at] = 1.0;
        if (symmetric && i != j) {
            at = cur[(size_t)j]++;
            row[at] = j; col[at] = i; val[at] = 1.0;
        }
    }
    *m_out = (int)M; *n_out = (int)N; *nnz_out = (int)total;
    *row_out = row; *col_out = col; *val_out = val;
    return CBSPMV_OK;
}

void cbspmv_free(void* p) { std::free(p); }

}  // extern "C"
