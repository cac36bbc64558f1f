// gfx950 (CDNA4, wave64) device code shared by the CB-SpMV schedules.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cb_plan.hpp"

namespace cbs {
namespace dev {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// The XCD (XCC) this wave runs on, from the hardware register XCC_ID
// (bits 3:0; gfx940 and later).
__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xfu;
}

// Lanes 0..lane inclusive.
__device__ __forceinline__ uint64_t lanes_le(int lane) {
    return lane == 63 ? ~0ull : ((2ull << lane) - 1);
}

// ---- global-memory pointers ----------------------------------------------------
// A pointer the compiler cannot prove global (one read from a descriptor in
// device memory) makes its loads and stores flat instructions, and a flat
// load counts against the LDS wait counter too, so every wait for an LDS
// read would also wait for the gathers in flight.  Every vector this code
// addresses is device memory: gptr() states that.
template <typename T>
using gbl = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ const gbl<T>* gptr(const T* p) { return (const gbl<T>*)p; }
template <typename T>
__device__ __forceinline__ gbl<T>* gptr(T* p) { return (gbl<T>*)p; }

// ---- value types ------------------------------------------------------------
// BF16 values are stored as the upper half of an FP32 bit pattern.
struct bf16_t { uint16_t bits; };

__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ float widen(bf16_t v) { return __uint_as_float((uint32_t)v.bits << 16); }

template <typename T>
__device__ __forceinline__ T stream_load(const T* p) { return __builtin_nontemporal_load(p); }
template <>
__device__ __forceinline__ bf16_t stream_load(const bf16_t* p) {
    bf16_t r;
    r.bits = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
    return r;
}

// ---- wave shuffles (ds_bpermute) and DPP -------------------------------------
template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, d, kWave); }
template <typename T>
__device__ __forceinline__ T shfl_down(T v, int d) { return __shfl_down(v, d, kWave); }
template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int d) { return __shfl_xor(v, d, kWave); }

// DPP move of a 32-bit lane value; lanes without a source (outside the row,
// or in rows not selected by ROWMASK) read `old`.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ uint32_t dpp32(uint32_t v, uint32_t old = 0) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROWMASK, 0xf, false);
}
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = dpp32<CTRL, ROWMASK>((uint32_t)b), hi = dpp32<CTRL, ROWMASK>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ float dpp(float v) { return __uint_as_float(dpp32<CTRL, ROWMASK>(__float_as_uint(v))); }

// DPP controls (GFX9): row_shr:n = 0x110 + n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143.
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

// DPP move for a value that a key comparison masks out where it has no
// source: lanes without a source read 0 (bound_ctrl) and lanes in rows that
// ROWMASK leaves out keep whatever the register held, so no old value has
// to be written first.
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ uint32_t dpp32_masked(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, ROWMASK, 0xf, true);
}
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp_masked(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = dpp32_masked<CTRL, ROWMASK>((uint32_t)b), hi = dpp32_masked<CTRL, ROWMASK>((uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ float dpp_masked(float v) {
    return __uint_as_float(dpp32_masked<CTRL, ROWMASK>(__float_as_uint(v)));
}
// Key of the source lane for a segmented scan step, keys stored as key + 1
// (never 0): lanes without a source, and lanes in rows ROWMASK leaves out,
// read 0, which matches no key.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_key(int key1) {
    if constexpr (ROWMASK == 0xf) return (int)dpp32_masked<CTRL, 0xf>((uint32_t)key1);
    else return (int)dpp32<CTRL, ROWMASK>((uint32_t)key1, 0u);
}

// One step of a segmented inclusive scan: add the value DPP brings in when it
// comes from the same segment.  `seg1` (key + 1) is non-decreasing along the
// wave, so an equal key means every lane in between belongs to the same run.
template <int CTRL, int ROWMASK, typename T>
__device__ __forceinline__ void seg_step(T& v, int seg1) {
    const T pv = dpp_masked<CTRL, ROWMASK>(v);
    if (dpp_key<CTRL, ROWMASK>(seg1) == seg1) v += pv;
}

// Segmented inclusive scan over the 64 lanes for non-decreasing keys (below
// 0x7fffffff + 1), all on DPP: row_shr 1, 2, 4, 8 inside each 16-lane row,
// then row_bcast:15 into rows 1 and 3 and row_bcast:31 into rows 2 and 3.
// Run-end lanes end with their run's sum.
template <typename T>
__device__ __forceinline__ T seg_scan_dpp(T v, int seg) {
    const int seg1 = (int)((uint32_t)seg + 1u);
    seg_step<kRowShr1, 0xf>(v, seg1);
    seg_step<kRowShr2, 0xf>(v, seg1);
    seg_step<kRowShr4, 0xf>(v, seg1);
    seg_step<kRowShr8, 0xf>(v, seg1);
    seg_step<kRowBcast15, 0xa>(v, seg1);
    seg_step<kRowBcast31, 0xc>(v, seg1);
    return v;
}

// Sum over groups of G consecutive lanes (G = 2, 4, 8 via DPP row_shr, larger
// via shuffles); the total lands in the group's LAST lane.
template <int G, typename T>
__device__ __forceinline__ T group_sum_to_last(T v) {
    if constexpr (G >= 2) v += dpp<kRowShr1>(v);
    if constexpr (G >= 4) v += dpp<kRowShr2>(v);
    if constexpr (G >= 8) v += dpp<kRowShr4>(v);
    if constexpr (G >= 16) v += dpp<kRowShr8>(v);
    if constexpr (G >= 32) {
        // across 16-lane rows: lanes shifted by 16 (shuffle; DPP rows are 16 wide)
        const T o = shfl_up(v, 16);
        if ((lane_id() & 31) >= 16) v += o;
    }
    return v;
}

// Segmented inclusive scan over the wave for non-decreasing segment keys:
// every lane ends with the sum of its run prefix; run-end lanes hold run sums.
template <typename T>
__device__ __forceinline__ T seg_scan_sorted(T v, int seg) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const T pv = shfl_up(v, d);
        const int ps = shfl_up(seg, d);
        if (lane >= d && ps == seg) v += pv;
    }
    return v;
}

// Error-free transforms for the compensated (two-float) accumulation:
// Knuth's TwoSum and the FMA TwoProduct (Ogita, Rump, Oishi, "Accurate Sum
// and Dot Product", SIAM J. Sci. Comput. 26(6), 2005).
struct f2 { float hi, lo; };
__device__ __forceinline__ f2 two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    const float e = (a - (s - bb)) + (b - bb);
    return {s, e};
}
__device__ __forceinline__ f2 add_f2(f2 a, f2 b) {
    const f2 s = two_sum(a.hi, b.hi);
    return {s.hi, s.lo + (a.lo + b.lo)};
}
__device__ __forceinline__ f2 two_prod(float a, float b) {
    const float p = a * b;
    return {p, __builtin_fmaf(a, b, -p)};
}

}  // namespace dev
}  // namespace cbs
