// Stream schedule: y = A x over the transcoded strip stream (cb_plan.hpp).
//
// Each wave takes one item (about item_nnz entries, at most kMaxItemStrips
// strips) at a time.  Keys and values stream in with non-temporal loads, U
// 64-entry chunks in flight per lane; x is gathered through L2/MALL.  Because
// a strip's entries are ordered by row, (strip-in-item, row) is non-decreasing
// along the item, so each 64-lane chunk is reduced with one segmented scan and
// each run's sum is added to a per-wave LDS accumulator [strip][row] by a
// single lane (no atomics).  Whole strips are stored to y directly; a strip cut
// into several items leaves partial rows in a small buffer and the last item
// to finish it (an atomic ticket per split strip) sums the pieces in piece
// order, so results are deterministic run to run.  Strips without entries are
// zeroed by the same launch.
#include "kernels.hpp"
#include "stream.hpp"

namespace cbs {
namespace dev {

template <typename A>
__device__ __forceinline__ A load_partial(const A* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int BS, typename V, typename A, bool COMP>
__global__ __launch_bounds__(kStreamThreads) void stream_spmv(StreamArgs<V, A> a) {
    constexpr int WPB = kStreamThreads / kWave;
    constexpr int U = sizeof(V) >= 8 ? 4 : 8;          // 64-entry chunks in flight per lane
    constexpr int SLOTS = kMaxItemStrips * BS;
    using Acc = typename std::conditional<COMP, f2, A>::type;
    __shared__ Acc acc_all[WPB][SLOTS];

    const int lane = lane_id();
    const int wave = (int)(threadIdx.x >> 6);
    // XCD-aware order: consecutive logical blocks (hence consecutive items and
    // strips) run on the same XCD, so x windows stay in that XCD's L2.
    uint32_t blk = blockIdx.x;
    if ((gridDim.x & 7u) == 0) blk = (blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const uint32_t gwave = blk * WPB + wave;
    const uint32_t nwaves = gridDim.x * WPB;

    for (uint32_t t = gwave * kWave + lane; t < a.n_empty * BS; t += nwaves * kWave) {
        const int64_t row = (int64_t)a.empty[t / BS] * BS + (t % BS);
        if (row < a.m) a.y[row] = A(0);
    }

    Acc* acc = acc_all[wave];
    const uint32_t cmask = (1u << a.col_bits) - 1u;
    for (uint32_t item = gwave; item < a.n_items; item += nwaves) {
        const StreamItem it = a.items[item];
        const int ns = it.n_strips;
        for (int t = lane; t < ns * BS; t += kWave) {
            if constexpr (COMP) acc[t] = f2{0.f, 0.f};
            else acc[t] = A(0);
        }
        __builtin_amdgcn_wave_barrier();
        int kbase = it.first_partial ? 0 : -1;
        const uint32_t e_last = it.e_end - 1;
        for (uint32_t e0 = it.e_begin; e0 < it.e_end; e0 += kWave * U) {
            // all loads unconditional (index clamped to the item) so that
            // the U key/value loads and then the U x gathers issue together
            uint32_t key[U];
            V val[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t e = min(e0 + u * kWave + lane, e_last);
                key[u] = stream_load(a.key + e);
                val[u] = stream_load(a.val + e);
            }
            A xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = a.x[key[u] & cmask];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool valid = e0 + u * kWave + lane <= e_last;
                const uint64_t starts = __ballot(valid && (key[u] >> 31));
                const int k = kbase + (int)__popcll(starts & lanes_le(lane));
                kbase += (int)__popcll(starts);
                const int r = (int)((key[u] >> a.col_bits) & (BS - 1));
                const int seg = valid ? k * BS + r : 0x7fffffff;
                Acc v;
                if constexpr (COMP) {
                    // segmented scan on (hi, lo) pairs with TwoSum
                    v = valid ? two_prod(widen(val[u]), xv[u]) : f2{0.f, 0.f};
#pragma unroll
                    for (int d = 1; d < kWave; d <<= 1) {
                        const float ph = shfl_up(v.hi, d), pl = shfl_up(v.lo, d);
                        const int ps = shfl_up(seg, d);
                        if (lane >= d && ps == seg) v = add_f2(v, f2{ph, pl});
                    }
                } else {
                    v = seg_scan_dpp(valid ? A(widen(val[u])) * xv[u] : A(0), seg);
                }
                const int nseg = shfl_down(seg, 1);
                const bool end = valid && (lane == kWave - 1 || nseg != seg);
                if (end) {
                    if constexpr (COMP) acc[seg] = add_f2(acc[seg], v);
                    else atomicAdd(acc + seg, v);   // ds_add: run ends are unique per chunk
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // ---- write rows: whole strips to y, cut strips to the partial buffer
        const uint32_t split_first = it.first_partial ? a.item_split_first[item] : 0xffffffffu;
        const uint32_t split_last = it.last_partial ? a.item_split[item] : 0xffffffffu;
        for (int t = lane; t < ns * BS; t += kWave) {
            const int k = t / BS, r = t % BS;
            A v;
            if constexpr (COMP) v = acc[t].hi + acc[t].lo;
            else v = acc[t];
            const bool cut_first = k == 0 && it.first_partial;
            const bool cut_last = k == ns - 1 && it.last_partial;
            if (cut_first || cut_last) {
                const StreamSplit sp = a.splits[cut_first ? split_first : split_last];
                a.partial[(size_t)(sp.partial_base + (item - sp.first_item)) * BS + r] = v;
            } else {
                const int64_t row = (int64_t)a.strip_ids[it.s_begin + k] * BS + r;
                if (row < a.m) a.y[row] = v;
            }
        }
        // ---- last piece of a cut strip sums all pieces in order
        for (int which = 0; which < 2; ++which) {
            const uint32_t sid = which == 0 ? split_first : split_last;
            if (sid == 0xffffffffu) continue;
            if (which == 1 && sid == split_first) continue;   // interior piece: once
            __threadfence();
            uint32_t ticket = 0;
            if (lane == 0) ticket = atomicAdd(a.counters + sid, 1u);
            ticket = __shfl(ticket, 0, kWave);
            const StreamSplit sp = a.splits[sid];
            if (ticket == sp.pieces - 1) {
                __threadfence();
                for (int r = lane; r < BS; r += kWave) {
                    A s = A(0);
                    for (uint32_t q = 0; q < sp.pieces; ++q)
                        s += load_partial(a.partial + (size_t)(sp.partial_base + q) * BS + r);
                    const int64_t row = (int64_t)sp.strip * BS + r;
                    if (row < a.m) a.y[row] = s;
                }
                if (lane == 0) atomicExch(a.counters + sid, 0u);
            }
        }
    }
}

}  // namespace dev

template <typename V, typename A, bool COMP>
hipError_t launch_stream_t(int bs, const StreamArgs<V, A>& a, int grid, hipStream_t s) {
    const dim3 g(grid), b(kStreamThreads);
    switch (bs) {
        case 2: hipLaunchKernelGGL((dev::stream_spmv<2, V, A, COMP>), g, b, 0, s, a); break;
        case 4: hipLaunchKernelGGL((dev::stream_spmv<4, V, A, COMP>), g, b, 0, s, a); break;
        case 8: hipLaunchKernelGGL((dev::stream_spmv<8, V, A, COMP>), g, b, 0, s, a); break;
        case 16: hipLaunchKernelGGL((dev::stream_spmv<16, V, A, COMP>), g, b, 0, s, a); break;
        case 32: hipLaunchKernelGGL((dev::stream_spmv<32, V, A, COMP>), g, b, 0, s, a); break;
        case 64: hipLaunchKernelGGL((dev::stream_spmv<64, V, A, COMP>), g, b, 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stream(int bs, const StreamArgs<double, double>& a, int grid, hipStream_t s) {
    return launch_stream_t<double, double, false>(bs, a, grid, s);
}
hipError_t launch_stream(int bs, const StreamArgs<float, float>& a, int grid, bool comp, hipStream_t s) {
    return comp ? launch_stream_t<float, float, true>(bs, a, grid, s)
                : launch_stream_t<float, float, false>(bs, a, grid, s);
}
hipError_t launch_stream(int bs, const StreamArgs<dev::bf16_t, float>& a, int grid, bool comp, hipStream_t s) {
    return comp ? launch_stream_t<dev::bf16_t, float, true>(bs, a, grid, s)
                : launch_stream_t<dev::bf16_t, float, false>(bs, a, grid, s);
}

}  // namespace cbs
