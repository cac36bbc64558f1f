// Sweep kernel of a split plan: y += the scattered entries' products, after
// the stream plan of the other entries has written y.  The layout is built
// by split.cpp (SweepPart).
//
// One workgroup of kSweepWaves waves per row block (at most kSweepMaxRows
// rows).  Each wave owns a contiguous run of the block's rows and walks
// their entries in column order, 64 at a time: the 64 are ordered by row,
// so one DPP segmented scan gives each row's sum over them, which the row's
// last lane adds to the block's row tile in LDS.  A row belongs to one wave
// and a wave's chunks run in order, so every row's sum is formed in one
// fixed order (results are identical from launch to launch).  At the end the
// workgroup adds its row tile to y.  Because every wave moves through x from
// its first column to its last at about the same pace, the waves of an XCD
// gather from a narrow window of x at any time and its lines stay in the
// XCD's L2.  A wave keeps one chunk's gathers in flight (64 lanes) and
// loads the next chunk's keys and values meanwhile: with more chunks in
// flight the waves drift apart and the window outgrows L2 (the 1,048,576^2
// stencil / power-law matrix, FP64, split: 127.4 us per SpMV with one chunk,
// 134.1 with two, 139.7 with four; 135.9 us unsplit).
#include <hip/hip_runtime.h>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"
#include "kernels.hpp"

namespace cbs {
namespace dev {

constexpr int kSweepThreads = kSweepWaves * kWave;
constexpr int kSweepChunks = 1;   // 64-entry chunks per wave per step

// A stored value as the accumulation type: FP64 and FP32 as they are, BF16
// (its 16 bits) widened to FP32.
template <typename A, typename V>
__device__ __forceinline__ A sweep_value(const V* v, uint32_t k) { return A(gptr(v)[k]); }
template <>
__device__ __forceinline__ float sweep_value<float, uint16_t>(const uint16_t* v, uint32_t k) {
    return __uint_as_float((uint32_t)gptr(v)[k] << 16);
}

template <typename A, typename V, bool PACED>
__global__ __launch_bounds__(kSweepThreads) void sweep_spmv(const A* __restrict__ x, const uint32_t* __restrict__ key,
                                                            const V* __restrict__ val,
                                                            const uint32_t* __restrict__ wave_off,
                                                            const int32_t* __restrict__ blk_row,
                                                            const int32_t* __restrict__ blk_span, uint32_t col_bits,
                                                            A* __restrict__ y) {
    constexpr int U = kSweepChunks;
    __shared__ A yt[kSweepMaxRows];
    const uint32_t b = blockIdx.x;
    const int lane = lane_id(), wave = (int)(threadIdx.x / kWave);
    const uint32_t cmask = (1u << col_bits) - 1u;
    for (int i = (int)threadIdx.x; i < kSweepMaxRows; i += kSweepThreads) yt[i] = A(0);
    const uint32_t e0 = gptr(wave_off)[b * kSweepWaves + wave], e1 = gptr(wave_off)[b * kSweepWaves + wave + 1];
    const int r0 = gptr(blk_row)[b], span = gptr(blk_span)[b];   // span >= 1
    uint32_t kn[U];
    A vn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t k = e0 + u * kWave + lane;
        kn[u] = k < e1 ? gptr(key)[k] : 0u;
        vn[u] = k < e1 ? sweep_value<A>(val, k) : A(0);
    }
    __syncthreads();
    // Paced (x beyond 8 MiB): a wave's priority starts at 3 and falls by one
    // each quarter of its entries, which keeps the waves of a CU at one pace
    // through x (the 2,097,152^2 block-clustered matrix, FP64, x 16 MiB:
    // 194.0 -> 186.9 us per split SpMV; the 1,048,576^2 stencil / power-law
    // one, x 8 MiB: 124.0 -> 127.7 us, so not there).
    [[maybe_unused]] const uint32_t quarter = (e1 - e0) / 4;
    [[maybe_unused]] int pace = 0;
    if constexpr (PACED) __builtin_amdgcn_s_setprio(3);
    for (uint32_t k0 = e0; k0 < e1; k0 += U * kWave) {
        if constexpr (PACED) {
            if (pace == 0 && k0 - e0 >= quarter) {
                __builtin_amdgcn_s_setprio(2);
                pace = 1;
            } else if (pace == 1 && k0 - e0 >= 2 * quarter) {
                __builtin_amdgcn_s_setprio(1);
                pace = 2;
            } else if (pace == 2 && k0 - e0 >= 3 * quarter) {
                __builtin_amdgcn_s_setprio(0);
                pace = 3;
            }
        }
        uint32_t kk[U];
        A v[U], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            kk[u] = kn[u];
            v[u] = vn[u];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xv[u] = k0 + u * kWave + lane < e1 ? gptr(x)[kk[u] & cmask] : A(0);
        const uint32_t k1 = k0 + U * kWave;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = k1 + u * kWave + lane;
            kn[u] = k < e1 ? gptr(key)[k] : 0u;
            vn[u] = k < e1 ? sweep_value<A>(val, k) : A(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool valid = k0 + u * kWave + lane < e1;
            const int row = valid ? (int)(kk[u] >> col_bits) : 0x7ffffffe;   // rows ascend within the chunk
            const A s = seg_scan_dpp(v[u] * xv[u], row);
            const int nrow = shfl_down(row, 1);
            if (valid && (lane == kWave - 1 || nrow != row)) yt[row] += s;
        }
    }
    if constexpr (PACED) __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    // y[rows] += the row tile: every y load of the thread is issued first
    // (indices clamped to the block's last row), and each value passes an
    // empty asm, which keeps the compiler from sinking a load into its
    // store's branch; a loop of load, wait and store would pay one round
    // trip per 256 rows.
    constexpr int R = kSweepMaxRows / kSweepThreads;
    A yv[R];
#pragma unroll
    for (int k = 0; k < R; ++k) yv[k] = gptr(y)[r0 + min((int)threadIdx.x + k * kSweepThreads, span - 1)];
#pragma unroll
    for (int k = 0; k < R; ++k) __asm__ volatile("" : "+v"(yv[k]));
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = (int)threadIdx.x + k * kSweepThreads;
        if (i < span) gptr(y)[r0 + i] = yv[k] + yt[i];
    }
}

}  // namespace dev

int run_sweep(cbspmv_plan& p, const void* x, void* y, hipStream_t s) {
    const SweepPart& sw = p.sweep;
    const DevicePlan& d = p.dev;
    const uint32_t blocks = (uint32_t)sw.blocks();
    if (blocks == 0) return CBSPMV_OK;
    if (sw.max_rows > kSweepMaxRows || sw.col_bits < 1 || sw.col_bits > 31 || !d.sw_key.ptr || !d.sw_val.ptr)
        return fail(CBSPMV_ERR_STATE, "scattered-entry split not uploaded");
    const auto* key = static_cast<const uint32_t*>(d.sw_key.ptr);
    const auto* wo = static_cast<const uint32_t*>(d.sw_wave_off.ptr);
    const auto* br = static_cast<const int32_t*>(d.sw_blk_row.ptr);
    const auto* bs = static_cast<const int32_t*>(d.sw_blk_span.ptr);
    const size_t xb = (size_t)p.host.n * (p.host.o.dtype == CBSPMV_F64 ? 8 : 4);
    const bool paced = xb > (size_t(8) << 20);
    if (p.host.o.dtype == CBSPMV_F64)
        hipLaunchKernelGGL((paced ? dev::sweep_spmv<double, double, true> : dev::sweep_spmv<double, double, false>), dim3(blocks), dim3(dev::kSweepThreads), 0, s,
                           static_cast<const double*>(x), key, static_cast<const double*>(d.sw_val.ptr), wo, br, bs,
                           (uint32_t)sw.col_bits, static_cast<double*>(y));
    else if (p.host.o.dtype == CBSPMV_F32)
        hipLaunchKernelGGL((paced ? dev::sweep_spmv<float, float, true> : dev::sweep_spmv<float, float, false>), dim3(blocks), dim3(dev::kSweepThreads), 0, s,
                           static_cast<const float*>(x), key, static_cast<const float*>(d.sw_val.ptr), wo, br, bs,
                           (uint32_t)sw.col_bits, static_cast<float*>(y));
    else   // BF16 values, FP32 vectors
        hipLaunchKernelGGL((paced ? dev::sweep_spmv<float, uint16_t, true> : dev::sweep_spmv<float, uint16_t, false>), dim3(blocks), dim3(dev::kSweepThreads), 0, s,
                           static_cast<const float*>(x), key, static_cast<const uint16_t*>(d.sw_val.ptr), wo, br, bs,
                           (uint32_t)sw.col_bits, static_cast<float*>(y));
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? CBSPMV_OK : hip_fail(e, "sweep launch");
}

}  // namespace cbs
