// Vendor-library comparator: rocSPARSE CSR SpMV on the same matrix and GPU,
// timed with the reference protocol (warm-up launches, then timed launches
// between hipEvents).  It plays the role of the reference's hipSPARSE CSR
// baseline program (cb-spmv-dcu/benchmark/hipSparse-csr/main.hip.cpp:104-181),
// written here against rocSPARSE's public generic API; it is a measurement
// aid for bench.py, not part of the CB-SpMV library.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdio>
#include <vector>

namespace {
struct Buf {
    void* p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
};
}  // namespace

extern "C" {

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
