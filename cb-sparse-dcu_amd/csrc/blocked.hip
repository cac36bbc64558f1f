// Blocked schedule: the reference's compute + gather kernel pair re-derived
// for gfx950, running on the canonical packed slab exactly as the reference
// packs it (cb-spmv-dcu/src/cb-spmv-dcu.hpp:1130-1348).
//
// One workgroup per row-owned segment (or per group of W blocks in LPT order),
// one wave per block, per-block format dispatch:
//   COO   one lane per stored entry; rows are sorted inside a block, so a
//         segmented scan keyed on the row (or gated by the encoded in-row
//         position in the direct / 32x32 encodings, with the long-row flag
//         choosing 3 or 5 scan steps) leaves each row's sum in its last lane
//   CSR   64/bs lanes per row striding the row, summed with DPP row_shr
//   DENSE lane = (row, column chunk), chunks summed with xor shuffles
//   uniform-row blocks compute row 0 once and broadcast it
// x comes from a per-wave LDS tile of the block column (non-gathered plans)
// or through the restore map / direct column (gathered plans).  Partial rows
// of a segment meet in LDS; a direct segment stores its rows, other segments
// add them atomically to rows zeroed by a fill launch first (this replaces the
// reference's inter-workgroup epoch spin, cb-spmv-dcu.hpp:236-283).
// Unit-valued plans with many repeated row patterns run the pattern kernel:
// one wave per distinct row pattern, its sum scattered to all rows that share
// it (cb-spmv-dcu.hpp:172-220, policy :1428-1450).
#include "blocked.hpp"
#include "kernels.hpp"

namespace cbs {
namespace dev {

template <int BS>
struct OrigT { using type = uint8_t; };
template <>
struct OrigT<32> { using type = uint16_t; };
template <>
struct OrigT<64> { using type = uint16_t; };

template <int BS>
struct MaskT { using type = uint8_t; };
template <>
struct MaskT<16> { using type = uint16_t; };
template <>
struct MaskT<32> { using type = uint32_t; };
template <>
struct MaskT<64> { using type = uint64_t; };

// Sum of lanes [16g, 16g+16) for BS-wide groups, delivered to every lane of
// the first group (the uniform-row reduction).
template <int BS>
__device__ __forceinline__ double uniform_row_total(double v) {
#pragma unroll
    for (int d = BS / 2; d > 0; d >>= 1) v += __shfl_down(v, d, BS);
    return __shfl(v, 0, kWave);
}

template <int BS>
__global__ void blocked_spmv(BlockedArgs a) {
    using Orig = typename OrigT<BS>::type;
    using Mask = typename MaskT<BS>::type;
    constexpr int LPR = kWave / BS > 0 ? kWave / BS : 1;   // lanes per row (CSR)
    constexpr int BCB = BS == 2 ? 1 : BS == 4 ? 2 : BS == 8 ? 3 : BS == 16 ? 4 : BS == 32 ? 5 : 6;
    __shared__ double s_y[kMaxWaves][BS];
    __shared__ double s_x[kMaxWaves][BS];
    const int lane = lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    const int W = a.waves;

    int b_first, b_end, owned_row = -1;
    bool direct_out = false, owned_uniform = false;
    if (a.row_owned) {
        const int s = blockIdx.x;
        owned_row = a.seg_rows[s];
        b_first = a.seg_begin[s];
        b_end = a.seg_end[s];
        direct_out = a.seg_direct[s] != 0;
        owned_uniform = a.seg_uniform[s] != 0;
    } else {
        b_first = blockIdx.x * W;
        b_end = min(b_first + W, a.nnzb);
    }
    if (lane < BS) s_y[wave][lane] = 0.0;
    __builtin_amdgcn_wave_barrier();

    for (int g = b_first + wave; g < b_end; g += W) {
        const int k = a.nnzpb[g];
        if (k <= 0) continue;
        const uint8_t enc = a.type[g];
        const int t = enc & kTypeMask;
        const bool uni = (enc & kUniform) != 0;
        const uint8_t* base = a.slab + a.addr[g];
        const int br = a.brow[g], bc = a.bcol[g];
        const int goff = a.gather ? a.block_cols_off[br] : 0;
        const int gcount = a.gather ? a.block_cols_off[br + 1] - goff : a.n;
        if (!a.gather) {
            // LDS tile of x for this block column
            if (lane < BS) {
                const int c = bc * BS + lane;
                s_x[wave][lane] = c < a.n ? a.x[c] : 0.0;
            }
            __builtin_amdgcn_wave_barrier();
        }
        // x at block-local column c (gathered: through the restore map)
        auto xcol = [&](int c) -> double {
            if (!a.gather) return s_x[wave][c];
            const int gc = bc * BS + c;
            return gc < gcount ? a.x[a.block_cols[goff + gc]] : 0.0;
        };
        // where one block's row sums go
        auto emit = [&](int row, double v) {
            if (a.row_owned) {
                s_y[wave][row] += v;
            } else {
                const int64_t yr = (int64_t)br * BS + row;
                if (yr < a.m) atomicAdd(a.y + yr, v);
            }
        };

        if (t == kCOO) {
            const int isz = a.compact ? 2 : 4;
            const double* vals = reinterpret_cast<const double*>(base + ((k * isz + 7) & ~7));
            if (lane < k) {
                const uint32_t w = a.compact ? reinterpret_cast<const uint16_t*>(base)[lane]
                                             : reinterpret_cast<const uint32_t*>(base)[lane];
                int row;
                double xv;
                if (a.compact) {
                    row = (int)((w >> 12) & (BS - 1));
                    xv = a.x[w & 0xfff];
                } else if (a.direct) {
                    row = (int)((w >> 16) & (BS - 1));
                    xv = a.x[w & 0xffff];
                } else {
                    row = (int)((w >> BCB) & (BS - 1));
                    xv = xcol((int)(w & (BS - 1)));
                }
                double v = (owned_uniform && row != 0) ? 0.0 : (a.unit ? xv : vals[lane] * xv);
                if (a.aggregate || a.row_owned) {
                    bool last;
                    const bool positional = BS != 64 && ((a.direct && !a.compact) || BS == 32);
                    if (positional) {
                        // scan gated by the encoded in-row position (5 bits)
                        const int pshift = a.direct ? 16 + BCB : 10;
                        const int pos = (int)((w >> pshift) & 0x1f);
                        const bool long_rows = (w & (a.direct ? (1u << (21 + BCB)) : 0x8000u)) != 0;
#pragma unroll
                        for (int d = 1; d < 8; d <<= 1) {
                            const double pv = shfl_up(v, d);
                            if (pos >= d) v += pv;
                        }
                        if (long_rows) {
#pragma unroll
                            for (int d = 8; d < BS; d <<= 1) {
                                const double pv = shfl_up(v, d);
                                if (pos >= d) v += pv;
                            }
                        }
                        const int npos = shfl_down(pos, 1);
                        last = lane == k - 1 || npos == 0;
                    } else {
                        // scan keyed on the row (rows are sorted in a block)
#pragma unroll
                        for (int d = 1; d < BS && d < kWave; d <<= 1) {
                            const double pv = shfl_up(v, d);
                            const int pr = shfl_up(row, d);
                            if (lane >= d && pr == row) v += pv;
                        }
                        const int nrow = shfl_down(row, 1);
                        last = lane == k - 1 || nrow != row;
                    }
                    if (last) emit(row, v);
                } else {
                    emit(row, v);
                }
            }
        } else if (t == kCSR) {
            const Orig* rp = reinterpret_cast<const Orig*>(base);
            const Orig* cc = rp + BS + 1;
            const double* vals = reinterpret_cast<const double*>(base + (((BS + 1 + k) * sizeof(Orig) + 7) & ~(size_t)7));
            double sum = 0.0;
            if (uni) {
                // row 0 once: lanes 0..BS-1 stride row 0 by BS
                for (int q = (int)rp[0] + lane; q < (int)rp[1]; q += BS)
                    sum += a.unit ? xcol(cc[q]) : xcol(cc[q]) * vals[q];
                sum = uniform_row_total<BS>(sum);
                if (owned_uniform) { if (lane == 0 && sum != 0.0) emit(0, sum); }
                else if (lane < BS && sum != 0.0) emit(lane, sum);
            } else {
                const int row = lane / LPR, off = lane % LPR;
                if (row < BS)
                    for (int q = (int)rp[row] + off; q < (int)rp[row + 1]; q += LPR)
                        sum += a.unit ? xcol(cc[q]) : xcol(cc[q]) * vals[q];
                sum = group_sum_to_last<LPR>(sum);
                if (off == LPR - 1 && row < BS && sum != 0.0) emit(row, sum);
            }
        } else {  // DENSE
            double sum = 0.0;
            if (uni) {
                if (lane < BS) {
                    if (a.unit) {
                        const Mask mask = reinterpret_cast<const Mask*>(base)[0];
                        if ((mask >> lane) & 1) sum = xcol(lane);
                    } else {
                        sum = reinterpret_cast<const double*>(base)[lane] * xcol(lane);
                    }
                }
                sum = uniform_row_total<BS>(sum);
                if (owned_uniform) { if (lane == 0 && sum != 0.0) emit(0, sum); }
                else if (lane < BS && sum != 0.0) emit(lane, sum);
            } else {
                const int row = lane % BS, chunk = lane / BS;
                constexpr int CPL = LPR <= BS ? BS / LPR : 1;     // columns per lane
                const int c0 = LPR <= BS ? chunk * CPL : chunk;
                const int c1 = LPR <= BS ? c0 + CPL : BS;
                constexpr int CS = LPR <= BS ? 1 : LPR;
                if (a.unit) {
                    const Mask mask = reinterpret_cast<const Mask*>(base)[row];
                    for (int c = c0; c < c1; c += CS)
                        if ((mask >> c) & 1) sum += xcol(c);
                } else {
                    const double* dv = reinterpret_cast<const double*>(base);
                    for (int c = c0; c < c1; c += CS) sum += dv[row * BS + c] * xcol(c);
                }
#pragma unroll
                for (int d = BS; d < kWave; d <<= 1) sum += shfl_xor(sum, d);
                if (lane < BS) emit(row, sum);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }

    if (!a.row_owned) return;
    __syncthreads();
    if (wave == 0 && lane < BS) {
        double total = 0.0;
        for (int w = 0; w < W; ++w) total += s_y[w][owned_uniform ? 0 : lane];
        const int64_t yr = (int64_t)owned_row * BS + lane;
        if (yr < a.m) {
            if (direct_out) a.y[yr] = total;
            else atomicAdd(a.y + yr, total);
        }
    }
}

// Zero the rows of the listed block rows (or all m rows when list is null).
__global__ void zero_rows(double* y, int m, const uint32_t* strips, uint32_t n_strips, int bs) {
    const size_t total = strips ? (size_t)n_strips * bs : (size_t)m;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const int64_t row = strips ? (int64_t)strips[t / bs] * bs + (int64_t)(t % bs) : (int64_t)t;
        if (row < m) y[row] = 0.0;
    }
}

// Zero `words` 32-bit words (a kernel rather than hipMemsetAsync so that the
// call can sit inside a captured graph).
__global__ void zero_words(uint32_t* p, size_t words) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < words; t += (size_t)gridDim.x * blockDim.x)
        p[t] = 0u;
}

// One wave per distinct row pattern of a unit-valued matrix.
template <bool UNROLLED>
__global__ void pattern_spmv(PatternArgs a) {
    const int lane = lane_id();
    const int p = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    if (p >= a.patterns) return;
    double sum = 0.0;
    const int b = a.row_ptr[p], e = a.row_ptr[p + 1];
    if (UNROLLED) {
        int q = b + lane;
        for (; q + kWave < e; q += 2 * kWave) sum += a.x[a.col[q]] + a.x[a.col[q + kWave]];
        if (q < e) sum += a.x[a.col[q]];
    } else {
        for (int q = b + lane; q < e; q += kWave) sum += a.x[a.col[q]];
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, kWave);
    sum = __shfl(sum, 0, kWave);
    for (int q = a.out_ptr[p] + lane; q < a.out_ptr[p + 1]; q += kWave) a.y[a.out_rows[q]] = sum;
}

}  // namespace dev

hipError_t launch_blocked(int bs, const BlockedArgs& a, int grid, hipStream_t s) {
    const dim3 g(grid), b(a.waves * dev::kWave);
    switch (bs) {
        case 2: hipLaunchKernelGGL(dev::blocked_spmv<2>, g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL(dev::blocked_spmv<4>, g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL(dev::blocked_spmv<8>, g, b, 0, s, a); break;
        case 16: hipLaunchKernelGGL(dev::blocked_spmv<16>, g, b, 0, s, a); break;
        case 32: hipLaunchKernelGGL(dev::blocked_spmv<32>, g, b, 0, s, a); break;
        case 64: hipLaunchKernelGGL(dev::blocked_spmv<64>, g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero_rows(double* y, int m, const uint32_t* strips, uint32_t n, int bs, hipStream_t s) {
    const size_t total = strips ? (size_t)n * bs : (size_t)m;
    if (total == 0) return hipSuccess;
    const int grid = (int)std::min<size_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::zero_rows, dim3(grid), dim3(256), 0, s, y, m, strips, n, bs);
    return hipGetLastError();
}

hipError_t launch_zero_words(void* p, size_t words, hipStream_t s) {
    if (words == 0) return hipSuccess;
    const int grid = (int)std::min<size_t>((words + 255) / 256, 4096);
    hipLaunchKernelGGL(dev::zero_words, dim3(grid), dim3(256), 0, s, static_cast<uint32_t*>(p), words);
    return hipGetLastError();
}

hipError_t launch_pattern(const PatternArgs& a, bool unrolled, int waves, hipStream_t s) {
    const int grid = (a.patterns + waves - 1) / waves;
    if (grid == 0) return hipSuccess;
    if (unrolled) hipLaunchKernelGGL(dev::pattern_spmv<true>, dim3(grid), dim3(waves * 64), 0, s, a);
    else hipLaunchKernelGGL(dev::pattern_spmv<false>, dim3(grid), dim3(waves * 64), 0, s, a);
    return hipGetLastError();
}

}  // namespace cbs
