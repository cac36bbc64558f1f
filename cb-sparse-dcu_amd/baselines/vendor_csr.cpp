// Vendor-library comparators: rocSPARSE CSR and BSR SpMV on the same matrix
// and GPU, timed with the reference protocol (warm-up launches, then timed
// launches between hipEvents).  They play the role of the reference's
// hipSPARSE baseline programs (cb-spmv-dcu/benchmark/hipSparse-csr/
// main.hip.cpp:104-181, hipSparse-bsr/main.hip.cpp:99-187), written here
// against rocSPARSE's public API; a measurement aid for bench.py, not part
// of the CB-SpMV library.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdio>
#include <vector>

namespace {
struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
};
// rocSPARSE's typed BSR entry points for double and float
rocsparse_status csr2bsr(rocsparse_handle h, rocsparse_int m, rocsparse_int n, const rocsparse_mat_descr dc,
                         const double* v, const rocsparse_int* rp, const rocsparse_int* ci, rocsparse_int bd,
                         const rocsparse_mat_descr db, double* bv, rocsparse_int* brp, rocsparse_int* bci) {
    return rocsparse_dcsr2bsr(h, rocsparse_direction_row, m, n, dc, v, rp, ci, bd, db, bv, brp, bci);
}
rocsparse_status csr2bsr(rocsparse_handle h, rocsparse_int m, rocsparse_int n, const rocsparse_mat_descr dc,
                         const float* v, const rocsparse_int* rp, const rocsparse_int* ci, rocsparse_int bd,
                         const rocsparse_mat_descr db, float* bv, rocsparse_int* brp, rocsparse_int* bci) {
    return rocsparse_scsr2bsr(h, rocsparse_direction_row, m, n, dc, v, rp, ci, bd, db, bv, brp, bci);
}
rocsparse_status bsrmv_analysis(rocsparse_handle h, rocsparse_int mb, rocsparse_int nb, rocsparse_int nnzb,
                                const rocsparse_mat_descr d, const double* v, const rocsparse_int* rp,
                                const rocsparse_int* ci, rocsparse_int bd, rocsparse_mat_info info) {
    return rocsparse_dbsrmv_analysis(h, rocsparse_direction_row, rocsparse_operation_none, mb, nb, nnzb, d, v, rp, ci,
                                     bd, info);
}
rocsparse_status bsrmv_analysis(rocsparse_handle h, rocsparse_int mb, rocsparse_int nb, rocsparse_int nnzb,
                                const rocsparse_mat_descr d, const float* v, const rocsparse_int* rp,
                                const rocsparse_int* ci, rocsparse_int bd, rocsparse_mat_info info) {
    return rocsparse_sbsrmv_analysis(h, rocsparse_direction_row, rocsparse_operation_none, mb, nb, nnzb, d, v, rp, ci,
                                     bd, info);
}
rocsparse_status bsrmv(rocsparse_handle h, rocsparse_int mb, rocsparse_int nb, rocsparse_int nnzb,
                       const rocsparse_mat_descr d, const double* v, const rocsparse_int* rp, const rocsparse_int* ci,
                       rocsparse_int bd, rocsparse_mat_info info, const double* x, double* y) {
    const double one = 1.0, zero = 0.0;
    return rocsparse_dbsrmv(h, rocsparse_direction_row, rocsparse_operation_none, mb, nb, nnzb, &one, d, v, rp, ci,
                            bd, info, x, &zero, y);
}
rocsparse_status bsrmv(rocsparse_handle h, rocsparse_int mb, rocsparse_int nb, rocsparse_int nnzb,
                       const rocsparse_mat_descr d, const float* v, const rocsparse_int* rp, const rocsparse_int* ci,
                       rocsparse_int bd, rocsparse_mat_info info, const float* x, float* y) {
    const float one = 1.0f, zero = 0.0f;
    return rocsparse_sbsrmv(h, rocsparse_direction_row, rocsparse_operation_none, mb, nb, nnzb, &one, d, v, rp, ci,
                            bd, info, x, &zero, y);
}

// BSR comparator body for value type T (see vendor_bsr_spmv).
template <typename T>
int bsr_spmv(int m, int n, int nnz, const int* row_ptr, const int* col, const double* val, const double* x, double* y,
             int bd, int warmup, int reps, double* ms, long long* nnzb_out) {
    auto H = [](hipError_t e) { return e == hipSuccess ? 0 : -(int)e; };
    if (bd < 1 || m < 1 || n < 1) return (int)rocsparse_status_invalid_size;
    const int mb = (m + bd - 1) / bd, nb = (n + bd - 1) / bd;
    const size_t bb = (size_t)bd * bd;
    Buf drp, dcol, dval, dx, dy, brp, bci, bval;
    int rc;
    if ((rc = H(hipMalloc(&drp.p, (size_t)(m + 1) * 4))) || (rc = H(hipMalloc(&dcol.p, (size_t)nnz * 4 + 4))) ||
        (rc = H(hipMalloc(&dval.p, (size_t)nnz * sizeof(T) + 8))) ||
        (rc = H(hipMalloc(&dx.p, (size_t)nb * bd * sizeof(T)))) ||
        (rc = H(hipMalloc(&dy.p, (size_t)mb * bd * sizeof(T)))) || (rc = H(hipMalloc(&brp.p, (size_t)(mb + 1) * 4))))
        return rc;
    std::vector<T> vt(val, val + nnz), xt((size_t)nb * bd, T(0));
    for (int i = 0; i < n; ++i) xt[i] = (T)x[i];
    (void)hipMemcpy(drp.p, row_ptr, (size_t)(m + 1) * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(dcol.p, col, (size_t)nnz * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(dval.p, vt.data(), (size_t)nnz * sizeof(T), hipMemcpyHostToDevice);
    (void)hipMemcpy(dx.p, xt.data(), xt.size() * sizeof(T), hipMemcpyHostToDevice);

    rocsparse_handle h = nullptr;
    rocsparse_mat_descr dc = nullptr, db = nullptr;
    rocsparse_mat_info info = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = 0;
    auto S = [&](rocsparse_status st) { if (!rc && st != rocsparse_status_success) rc = (int)st; };
    (void)hipStreamCreate(&s);
    S(rocsparse_create_handle(&h));
    S(rocsparse_set_stream(h, s));
    S(rocsparse_create_mat_descr(&dc));
    S(rocsparse_create_mat_descr(&db));
    S(rocsparse_create_mat_info(&info));
    rocsparse_int nnzb = 0;
    S(rocsparse_csr2bsr_nnz(h, rocsparse_direction_row, m, n, dc, (const rocsparse_int*)drp.p,
                            (const rocsparse_int*)dcol.p, bd, db, (rocsparse_int*)brp.p, &nnzb));
    if (!rc) (void)hipStreamSynchronize(s);
    if (!rc) rc = H(hipMalloc(&bci.p, (size_t)nnzb * 4 + 4));
    if (!rc) rc = H(hipMalloc(&bval.p, (size_t)nnzb * bb * sizeof(T) + 8));
    T* BV = static_cast<T*>(bval.p);
    const rocsparse_int* BR = static_cast<const rocsparse_int*>(brp.p);
    const rocsparse_int* BC = static_cast<const rocsparse_int*>(bci.p);
    S(csr2bsr(h, m, n, dc, static_cast<const T*>(dval.p), (const rocsparse_int*)drp.p, (const rocsparse_int*)dcol.p,
              bd, db, BV, (rocsparse_int*)brp.p, (rocsparse_int*)bci.p));
    S(bsrmv_analysis(h, mb, nb, nnzb, db, BV, BR, BC, bd, info));
    const T* X = static_cast<const T*>(dx.p);
    T* Y = static_cast<T*>(dy.p);
    for (int i = 0; !rc && i < warmup; ++i) S(bsrmv(h, mb, nb, nnzb, db, BV, BR, BC, bd, info, X, Y));
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    for (int i = 0; !rc && i < reps; ++i) S(bsrmv(h, mb, nb, nnzb, db, BV, BR, BC, bd, info, X, Y));
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = reps > 0 ? t / reps : 0.0;
    *nnzb_out = nnzb;
    if (!rc) {
        std::vector<T> yt((size_t)mb * bd);
        (void)hipMemcpy(yt.data(), dy.p, yt.size() * sizeof(T), hipMemcpyDeviceToHost);
        for (int i = 0; i < m; ++i) y[i] = (double)yt[i];
    }
    if (info) rocsparse_destroy_mat_info(info);
    if (dc) rocsparse_destroy_mat_descr(dc);
    if (db) rocsparse_destroy_mat_descr(db);
    if (h) rocsparse_destroy_handle(h);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}
}  // namespace

extern "C" {

// y = A x with rocsparse_{d,s}bsrmv after rocsparse_csr2bsr with square
// blocks of bd x bd (row-major inside a block, as the reference's
// hipsparseDcsr2bsr / hipsparseDbsrmv calls; bd in {2, 4, 8, 16} there).
// CSR input as for vendor_csr_spmv; x is zero-padded to whole blocks.
// Writes the mean ms per bsrmv launch, the BSR block count and y (m
// doubles) from the last launch.  Returns 0, or the failing rocSPARSE / HIP
// status (negated HIP status).
int vendor_bsr_spmv(int m, int n, int nnz, const int* row_ptr, const int* col, const double* val, const double* x,
                    double* y, int bd, int f32, int warmup, int reps, double* ms, long long* nnzb) {
    return f32 ? bsr_spmv<float>(m, n, nnz, row_ptr, col, val, x, y, bd, warmup, reps, ms, nnzb)
               : bsr_spmv<double>(m, n, nnz, row_ptr, col, val, x, y, bd, warmup, reps, ms, nnzb);
}

// y = A x with rocsparse_spmv on CSR (row_ptr[m+1], col[nnz], val[nnz] on the
// host; FP64, or FP32 when f32 != 0, values and vectors converted here).
// alg: a rocsparse_spmv_alg value.  Writes the mean ms per compute launch to
// *ms and y (m doubles) from the last launch.  Returns 0, or the failing
// rocSPARSE / HIP status (negated HIP status).
int vendor_csr_spmv(int m, int n, int nnz, const int* row_ptr, const int* col, const double* val, const double* x,
                    double* y, int alg, int f32, int warmup, int reps, double* ms) {
    const size_t vb = f32 ? 4 : 8;
    Buf drp, dcol, dval, dx, dy, work;
    auto H = [](hipError_t e) { return e == hipSuccess ? 0 : -(int)e; };
    int rc;
    if ((rc = H(hipMalloc(&drp.p, (size_t)(m + 1) * 4))) || (rc = H(hipMalloc(&dcol.p, (size_t)nnz * 4 + 4))) ||
        (rc = H(hipMalloc(&dval.p, (size_t)nnz * vb + 8))) || (rc = H(hipMalloc(&dx.p, (size_t)n * vb + 8))) ||
        (rc = H(hipMalloc(&dy.p, (size_t)m * vb + 8))))
        return rc;
    std::vector<float> vf, xf;
    if (f32) {
        vf.assign(val, val + nnz);
        xf.assign(x, x + n);
    }
    (void)hipMemcpy(drp.p, row_ptr, (size_t)(m + 1) * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(dcol.p, col, (size_t)nnz * 4, hipMemcpyHostToDevice);
    (void)hipMemcpy(dval.p, f32 ? (const void*)vf.data() : (const void*)val, (size_t)nnz * vb, hipMemcpyHostToDevice);
    (void)hipMemcpy(dx.p, f32 ? (const void*)xf.data() : (const void*)x, (size_t)n * vb, hipMemcpyHostToDevice);

    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnvec_descr X = nullptr, Y = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const rocsparse_datatype dt = f32 ? rocsparse_datatype_f32_r : rocsparse_datatype_f64_r;
    const double a64 = 1.0, b64 = 0.0;
    const float a32 = 1.0f, b32 = 0.0f;
    const void* alpha = f32 ? (const void*)&a32 : (const void*)&a64;
    const void* beta = f32 ? (const void*)&b32 : (const void*)&b64;
    size_t bytes = 0;
    rc = 0;
    auto S = [&](rocsparse_status st) { if (!rc && st != rocsparse_status_success) rc = (int)st; };
    (void)hipStreamCreate(&s);
    S(rocsparse_create_handle(&h));
    S(rocsparse_set_stream(h, s));
    S(rocsparse_create_csr_descr(&A, m, n, nnz, drp.p, dcol.p, dval.p, rocsparse_indextype_i32, rocsparse_indextype_i32,
                                 rocsparse_index_base_zero, dt));
    S(rocsparse_create_dnvec_descr(&X, n, dx.p, dt));
    S(rocsparse_create_dnvec_descr(&Y, m, dy.p, dt));
    const auto algo = (rocsparse_spmv_alg)alg;
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdeprecated-declarations"
    S(rocsparse_spmv(h, rocsparse_operation_none, alpha, A, X, beta, Y, dt, algo, rocsparse_spmv_stage_buffer_size,
                     &bytes, nullptr));
    if (!rc && bytes) rc = H(hipMalloc(&work.p, bytes));
    S(rocsparse_spmv(h, rocsparse_operation_none, alpha, A, X, beta, Y, dt, algo, rocsparse_spmv_stage_preprocess,
                     &bytes, work.p));
    for (int i = 0; !rc && i < warmup; ++i)
        S(rocsparse_spmv(h, rocsparse_operation_none, alpha, A, X, beta, Y, dt, algo, rocsparse_spmv_stage_compute,
                         &bytes, work.p));
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    for (int i = 0; !rc && i < reps; ++i)
        S(rocsparse_spmv(h, rocsparse_operation_none, alpha, A, X, beta, Y, dt, algo, rocsparse_spmv_stage_compute,
                         &bytes, work.p));
#pragma GCC diagnostic pop
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = reps > 0 ? t / reps : 0.0;
    if (!rc) {
        if (f32) {
            std::vector<float> yf(m);
            (void)hipMemcpy(yf.data(), dy.p, (size_t)m * 4, hipMemcpyDeviceToHost);
            for (int i = 0; i < m; ++i) y[i] = yf[i];
        } else {
            (void)hipMemcpy(y, dy.p, (size_t)m * 8, hipMemcpyDeviceToHost);
        }
    }
    if (X) rocsparse_destroy_dnvec_descr(X);
    if (Y) rocsparse_destroy_dnvec_descr(Y);
    if (A) rocsparse_destroy_spmat_descr(A);
    if (h) rocsparse_destroy_handle(h);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

}  // extern "C"
