// Stream schedule: y = A x over the transcoded strip stream (cb_plan.hpp).
//
// One workgroup (tile_items waves) per tile.  The workgroup first copies
// the tile's hot x cache lines into LDS with coalesced loads, so entries whose
// column is staged read LDS instead of issuing a divergent gather to L2 (one
// L2 request per lane), which is what bounds a gather-heavy SpMV here.  Then
// each wave streams its item in chunks of 64 lanes' entries, non-temporal:
// with x in L2 one chunk ahead (stream_item_near), with x beyond L2 a ring of
// chunks in flight through buffer instructions (stream_item), chosen per
// launch by the size of x.  Rows are contiguous runs inside an item, so a chunk is
// reduced with one DPP segmented scan, each finished row is stored to y by its
// last lane, and the row still open at the end of a chunk is carried into
// the next one.  Items are cut at row boundaries, so rows have one writer and
// need no atomics; only a row longer than four items is cut, and its pieces
// meet in a partial buffer where the last piece to arrive (atomic ticket) sums
// them in piece order.  Results are the same bits on every launch.
//
// The SDDMM variant runs on the transposed plan of a backward pass: besides
// dx = A^T g it writes, for every stream entry (row j of A^T, column i),
// dvals = g[i] * x[j] in stream order, reusing the g[i] the product gathers
// anyway; lanes own consecutive entries, so those stores are coalesced.
//
// With column panels (cb_plan.hpp) every workgroup also runs, on as many
// extra waves, items of its XCD's panel into per-panel row sums, and
// panel_sums adds those to y in a second launch.
#include <algorithm>
#include <type_traits>

#include "kernels.hpp"
#include "stream.hpp"

namespace cbs {
namespace dev {

constexpr int kMultiRhs = kRhsGroup;   // right-hand sides reduced per chunk load

template <typename A>
__device__ __forceinline__ A load_partial(const A* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A stream item's 16-byte record, read as one global load.
__device__ __forceinline__ StreamItem load_item(const StreamItem* items, uint32_t i) {
    typedef uint32_t w4 __attribute__((ext_vector_type(4)));
    const w4 w = gptr(reinterpret_cast<const w4*>(items))[i];
    return __builtin_bit_cast(StreamItem, w);
}

__device__ __forceinline__ StreamSplit load_split(const StreamSplit* splits, uint32_t i) {
    typedef uint32_t w4 __attribute__((ext_vector_type(4)));
    const w4 w = gptr(reinterpret_cast<const w4*>(splits))[i];
    return __builtin_bit_cast(StreamSplit, w);
}

// (hi, lo) pairs for the compensated accumulation
__device__ __forceinline__ f2& operator+=(f2& a, const f2& b) {
    a = add_f2(a, b);
    return a;
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ f2 dpp_f2(f2 v) { return f2{dpp<CTRL, ROWMASK>(v.hi), dpp<CTRL, ROWMASK>(v.lo)}; }

template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_step2(f2& v, int seg1) {
    const f2 pv = f2{dpp_masked<CTRL, ROWMASK>(v.hi), dpp_masked<CTRL, ROWMASK>(v.lo)};
    if (dpp_key<CTRL, ROWMASK>(seg1) == seg1) v += pv;
}
__device__ __forceinline__ f2 seg_scan_dpp(f2 v, int seg) {
    const int seg1 = (int)((uint32_t)seg + 1u);
    seg_step2<kRowShr1, 0xf>(v, seg1);
    seg_step2<kRowShr2, 0xf>(v, seg1);
    seg_step2<kRowShr4, 0xf>(v, seg1);
    seg_step2<kRowShr8, 0xf>(v, seg1);
    seg_step2<kRowBcast15, 0xa>(v, seg1);
    seg_step2<kRowBcast31, 0xc>(v, seg1);
    return v;
}

// value of lane `src` (wave-uniform src) with v_readlane
__device__ __forceinline__ float read_lane(float v, int src) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)__float_as_uint(v), src));
}
__device__ __forceinline__ double read_lane(double v, int src) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ f2 read_lane(f2 v, int src) { return f2{read_lane(v.hi, src), read_lane(v.lo, src)}; }

template <typename A>
__device__ __forceinline__ A finish(A v) { return v; }
__device__ __forceinline__ float finish(f2 v) { return v.hi + v.lo; }

// Chunk loads: 16-byte lane planes (cb_plan.hpp), non-temporal because the
// slab is read once per SpMV and should not displace x from L2.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename V, int EPL>
struct Chunk {
    static constexpr int NK = EPL * 4 / 16;             // key planes
    static constexpr int NV = EPL * (int)sizeof(V) / 16; // value planes
    uint32_t k[EPL];
    V v[EPL];
    // chunk starting at entry e0 with G lanes; lanes >= G re-read lane G-1
    __device__ __forceinline__ void load(const uint8_t* slab, uint32_t e0, uint32_t G, int lane) {
        const uint8_t* base = slab + (size_t)e0 * (4 + sizeof(V));
        const uint32_t l = min((uint32_t)lane, G - 1) * 16;
#pragma unroll
        for (int p = 0; p < NK; ++p) {
            const u32x4 w = __builtin_nontemporal_load(gptr(reinterpret_cast<const u32x4*>(base + p * G * 16 + l)));
            k[4 * p] = w.x; k[4 * p + 1] = w.y; k[4 * p + 2] = w.z; k[4 * p + 3] = w.w;
        }
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const u32x4 w = __builtin_nontemporal_load(gptr(reinterpret_cast<const u32x4*>(base + (NK + q) * G * 16 + l)));
            uint32_t ws[4] = {w.x, w.y, w.z, w.w};
            __builtin_memcpy(&v[q * (16 / sizeof(V))], ws, 16);
        }
    }
};

// ---- buffer instructions (raw buffer resources, gfx9 layout).  A lane whose
// offset is at or past the resource's byte size reads zeros and writes
// nothing, without a memory request, so a lane that has nothing to load or
// store passes kOutOfRange instead of branching around the instruction.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;       // dword 3 of a raw buffer resource
constexpr uint32_t kOutOfRange = 0x80000000u; // every resource here is < 2 GiB
constexpr int kNonTemporal = 2;               // cache policy bit: nt
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
template <typename A>
__device__ __forceinline__ A buffer_load(rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ double buffer_load<double>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ float buffer_load<float>(rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <typename A>
__device__ __forceinline__ void buffer_store(A v, rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ void buffer_store<double>(double v, rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
}
template <>
__device__ __forceinline__ void buffer_store<float>(float v, rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, off, 0, 0);
}

// One chunk of an item through its buffer resource: r0 is the chunk's first
// entry relative to the item (ne entries); a chunk past the end reads zeros.
template <typename V, int EPL>
struct LaneChunk {
    static constexpr int NK = EPL * 4 / 16;
    static constexpr int NV = EPL * (int)sizeof(V) / 16;
    uint32_t k[EPL];
    V v[EPL];
    __device__ __forceinline__ void load(rsrc_t rs, uint32_t r0, uint32_t ne, int lane) {
        const uint32_t G = r0 < ne ? min(64u, (ne - r0) / EPL) : 64u;
        const uint32_t base = r0 < ne ? r0 * (4 + (uint32_t)sizeof(V)) + min((uint32_t)lane, G - 1) * 16 : kOutOfRange;
#pragma unroll
        for (int p = 0; p < NK; ++p) {
            const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rs, base + p * G * 16, 0, kNonTemporal);
            k[4 * p] = w.x; k[4 * p + 1] = w.y; k[4 * p + 2] = w.z; k[4 * p + 3] = w.w;
        }
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(rs, base + (NK + q) * G * 16, 0, kNonTemporal);
            __builtin_memcpy(&v[q * (16 / sizeof(V))], &w, 16);
        }
    }
};

// DPP whole-wave shifts (GFX9): wave_shl:1 (lane i reads i+1), wave_shr:1
// (lane i reads i-1); the lane without a source reads `old`.
template <int CTRL>
__device__ __forceinline__ int wave_shift(int v, int old) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xf, 0xf, false);
}
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;
__device__ __forceinline__ double wave_shr1(double v) { return dpp<kWaveShr1, 0xf>(v); }
__device__ __forceinline__ float wave_shr1(float v) { return dpp<kWaveShr1, 0xf>(v); }
__device__ __forceinline__ f2 wave_shr1(f2 v) { return f2{dpp<kWaveShr1, 0xf>(v.hi), dpp<kWaveShr1, 0xf>(v.lo)}; }

// Stages a tile's hot x lines in LDS, PER slots per thread per pass with all
// line-id loads, then all x loads, issued back to back (16 consecutive
// threads read one line): two dependent round trips per pass, one pass for
// the default cache; the last slot stays 0 for the padding entries.
// mid() runs once, between the first pass's line-id loads and its x loads
// (or alone when the tile stages nothing): the caller's wave issues its own
// first loads there, so that their round trips overlap the staging's.
// Every thread of the workgroup calls this; it ends with a workgroup barrier.
template <typename V, typename A, typename Mid>
__device__ __forceinline__ void stage_tile(const StreamArgs<V, A>& a, uint32_t tile, A* xs, Mid&& mid) {
    // one pass for the default 96 KB cache and 16 waves: up to 12,288 FP64
    // or 24,576 FP32 slots over 1024 threads
    constexpr int PER = 96 * 1024 / (int)sizeof(A) / kMaxStreamThreads;
    const uint32_t nthreads = blockDim.x;
    const uint32_t lb = a.tiles[2 * tile], cnt = a.tiles[2 * tile + 1] * kLineElems;
    const uint32_t kZero = a.tile_slots - 1;   // stays 0: read by the padding entries
    bool pending = true;   // mid() not yet run (wave-uniform)
    for (uint32_t p0 = 0; p0 < cnt; p0 += PER * nthreads) {
        const uint32_t s0 = p0 + threadIdx.x;
        // Every load is issued by every lane, from an index clamped into its
        // array (a slot past the tile's reads its last line, a column past
        // x reads x's last element; both are dropped), so that all of them
        // are in flight together: behind a branch each would wait for the last.
        uint32_t lid[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t s = min(s0 + j * nthreads, cnt - 1);
            lid[j] = gptr(a.tile_lines)[lb + s / kLineElems];
        }
        if (pending) {
            mid();
            pending = false;
        }
        A xv[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t c = lid[j] * kLineElems + ((s0 + j * nthreads) % kLineElems);
            xv[j] = gptr(a.x)[min(c, (uint32_t)a.n - 1)];
        }
        // every lane stores too (a slot past the tile writes the zero slot,
        // past every staged one), else the compiler sinks a load into the store's branch
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint32_t s = s0 + j * nthreads;
            const uint32_t c = lid[j] * kLineElems + (s % kLineElems);
            xs[s < cnt ? s : kZero] = s < cnt && c < (uint32_t)a.n ? xv[j] : A(0);
        }
    }
    if (pending) mid();
    if (threadIdx.x == 0) xs[kZero] = A(0);
    __syncthreads();
}

// stage_tile's mid() for a wave without loads of its own to overlap
struct NoLoads {
    __device__ __forceinline__ void operator()() const {}
};

// The stage() of an item that has no tile to stage (or one staged before):
// its first loads are issued at once.
struct RunNow {
    template <typename F>
    __device__ __forceinline__ void operator()(F&& f) const { f(); }
};
constexpr RunNow run_now{};

// Pieces of a cut row: each piece's wave takes a ticket once its partial
// sum is stored; the last piece to arrive sums them in piece order.
template <typename V, typename A, bool COMP>
__device__ __forceinline__ void join_pieces(const StreamArgs<V, A>& a, uint32_t split_first, uint32_t split_last,
                                            A* yout) {
    using Acc = typename std::conditional<COMP, f2, A>::type;
    constexpr int kParts = COMP ? 2 : 1;
    const int lane = lane_id();
    for (int which = 0; which < 2; ++which) {
        const uint32_t sid = which == 0 ? split_first : split_last;
        if (sid == 0xffffffffu || (which == 1 && sid == split_first)) continue;
        __threadfence();
        uint32_t ticket = 0;
        if (lane == 0) ticket = atomicAdd(a.counters + sid, 1u);
        ticket = __shfl(ticket, 0, kWave);
        const StreamSplit sp = load_split(a.splits, sid);
        if (ticket == sp.pieces - 1) {
            __threadfence();
            if (lane == 0) {
                Acc sum{};
                for (uint32_t q = 0; q < sp.pieces; ++q) {
                    const uint32_t slot = kParts * (sp.partial_base + q);
                    if constexpr (COMP) sum += f2{load_partial(a.partial + slot), load_partial(a.partial + slot + 1)};
                    else sum += load_partial(a.partial + slot);
                }
                yout[sp.row] = finish(sum);
                atomicExch(a.counters + sid, 0u);
            }
        }
    }
}

// SDDMM products dv[j] = xc[j] * xrow[row of entry j] for one chunk.  A
// chunk's rows are few and consecutive: lane l loads xrow[r0 + l] once and
// entries take their row's value by a lane shuffle.  A lane's entries
// usually span one or two rows, its first entry's and its last entry's, so
// two shuffles serve them all; a lane with an entry in a row in between, or
// in a row 64 or more past r0, redoes its entries one shuffle (or load)
// each.  Every shuffle is issued by the whole wave.
template <typename A, int EPL>
__device__ __forceinline__ void row_products(const A* xrow, uint32_t m, uint32_t row_base, const int (&off)[EPL],
                                             bool valid, const A (&xc)[EPL], A (&dv)[EPL]) {
    const int lane = lane_id();
    const uint32_t r0 = row_base + (uint32_t)__builtin_amdgcn_readfirstlane(off[0]);
    const A xw = gptr(xrow)[min(r0 + (uint32_t)lane, m - 1u)];
    const uint32_t dlo = row_base + (uint32_t)off[0] - r0, dhi = row_base + (uint32_t)off[EPL - 1] - r0;
    const A xa = __shfl(xw, valid && dlo < 64u ? (int)dlo : 0, kWave);
    const A xb = __shfl(xw, valid && dhi < 64u ? (int)dhi : 0, kWave);
    bool slow = valid && (dlo >= 64u || dhi >= 64u);
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
        const bool lo = off[j] == off[0];
        dv[j] = xc[j] * (lo ? xa : xb);
        slow = slow || (valid && !lo && off[j] != off[EPL - 1]);
    }
    if (__any(slow)) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const uint32_t d = row_base + (uint32_t)off[j] - r0;
            const A v = __shfl(xw, valid && d < 64u ? (int)d : 0, kWave);
            if (slow) dv[j] = xc[j] * (d < 64u ? v : gptr(xrow)[min(row_base + (uint32_t)off[j], m - 1u)]);
        }
    }
}

// Chunks in flight per wave (the ring of stream_item): 5 when a chunk holds
// 12 registers (FP64, BF16), 4 when it holds 16 (FP32); 3 where registers
// are short (the batched kernel, FP32 backward).  The ring pays off when x is
// gathered from beyond L2; with x in L2 the one-ahead loop of
// stream_item_near (ring size 2) is faster.
template <typename V, int EPL>
constexpr int deep_ring() { return EPL * (4 + (int)sizeof(V)) / 4 <= 12 ? 5 : 4; }
constexpr int kNearRing = 2;

// One item with x in L2: chunk c + 1's keys and values are loaded (non-
// temporal) while chunk c is reduced, then chunk c + 1's x is gathered; lanes
// without a gather or a finished run skip the instruction.
template <typename V, typename A, bool COMP, int EPL, bool SDDMM, bool LDS, typename Stage>
__device__ __forceinline__ void stream_item_near(const StreamArgs<V, A>& a, uint32_t item, const A* xs, A* yout,
                                                 Stage&& stage) {
    using Acc = typename std::conditional<COMP, f2, A>::type;
    const uint32_t kZero = a.tile_slots - 1;
    const int lane = lane_id();
    const StreamItem it = load_item(a.items, item);
    const int cb = a.col_bits;
    const uint32_t cmask = (1u << cb) - 1u;
    const uint32_t omask = (1u << (31 - cb)) - 1u;
    // The cut-row tables are read at once when the item has no tile to stage
    // (the registers are shorter then), else after the staging, which must
    // not wait for this item's descriptor.
    constexpr bool kLate = !std::is_same<std::decay_t<Stage>, RunNow>::value;
    uint32_t split_first = 0xffffffffu, split_last = 0xffffffffu;
    if constexpr (!kLate) {
        if (it.first_partial) split_first = gptr(a.item_split_first)[item];
        if (it.last_partial) split_last = gptr(a.item_split)[item];
    }
    auto put_partial = [&](uint32_t slot, Acc v) {
        if constexpr (COMP) {
            gptr(a.partial)[2 * slot] = v.hi;
            gptr(a.partial)[2 * slot + 1] = v.lo;
        } else {
            gptr(a.partial)[slot] = v;
        }
    };
    // a finished run: to y, or to the partial buffer when it is a piece of
    // a cut row (the item's first run, offset 0, or its last run)
    auto emit = [&](bool go, int off, Acc v, bool final_run) {
        if (!go) return;
        if (it.first_partial && off == 0) {
            const StreamSplit sp = load_split(a.splits, split_first);
            put_partial(sp.partial_base + (item - sp.first_item), v);
        } else if (final_run && it.last_partial) {
            const StreamSplit sp = load_split(a.splits, split_last);
            put_partial(sp.partial_base + (item - sp.first_item), v);
        } else {
            gptr(yout)[it.row_base + (uint32_t)off] = finish(v);
        }
    };
    auto gather = [&](const uint32_t* kk, A* xx) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const uint32_t c = kk[j] & cmask;
            const bool cached = (kk[j] & kKeyCached) != 0;
            A g = A(0);
            if (!cached) g = gptr(a.x)[c];
            A l = A(0);
            if constexpr (LDS) l = xs[min(c, kZero)];
            xx[j] = cached ? l : g;
        }
    };
    constexpr uint32_t kChunk = kWave * EPL;
    auto lanes_of = [&](uint32_t e0) { return min(64u, (it.e_end - e0) / EPL); };

    int carry_off = -1;
    Acc carry{};
    Chunk<V, EPL> cur, nxt;
    A xc[EPL];
    stage([&] { cur.load(a.slab, it.e_begin, lanes_of(it.e_begin), lane); });
    if constexpr (kLate) {
        if (it.first_partial) split_first = gptr(a.item_split_first)[item];
        if (it.last_partial) split_last = gptr(a.item_split)[item];
    }
    gather(cur.k, xc);
    for (uint32_t e0 = it.e_begin; e0 < it.e_end; e0 += kChunk) {
        const bool more = e0 + kChunk < it.e_end;
        if (more) nxt.load(a.slab, e0 + kChunk, lanes_of(e0 + kChunk), lane);
        const int L = (int)min(63u, (it.e_end - e0) / EPL - 1);   // last lane with entries
        const bool valid = lane <= L;
        int off[EPL];
        Acc s[EPL];
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            off[j] = valid ? (int)((cur.k[j] >> cb) & omask) : 0x7fffffff;
            if constexpr (COMP) s[j] = two_prod(widen(cur.v[j]), xc[j]);
            else s[j] = A(widen(cur.v[j])) * xc[j];
        }
        if constexpr (SDDMM) {
            A dv[EPL];
            row_products<A, EPL>(a.xrow, (uint32_t)a.m, it.row_base, off, valid, xc, dv);
            if (valid) {
                gbl<u32x4>* dst = gptr(reinterpret_cast<u32x4*>(a.dvals + e0 + (uint32_t)lane * EPL));
#pragma unroll
                for (int q = 0; q < EPL * (int)sizeof(A) / 16; ++q) {
                    u32x4 w;
                    __builtin_memcpy(&w, reinterpret_cast<const unsigned char*>(dv) + 16 * q, 16);
                    __builtin_nontemporal_store(w, dst + q);
                }
            }
        }
        // the run left open by the previous chunk
        if (carry_off >= 0) {
            const int o0 = __builtin_amdgcn_readfirstlane(off[0]);
            if (o0 == carry_off) {
                if (lane == 0) s[0] += carry;
            } else {
                emit(lane == 0, carry_off, carry, false);
            }
        }
        // runs inside the lane, then across lanes (segmented scan)
        bool ends[EPL];
#pragma unroll
        for (int j = 1; j < EPL; ++j) {
            const bool same = off[j] == off[j - 1];
            ends[j - 1] = !same;
            if (same) s[j] += s[j - 1];
        }
        const Acc T = seg_scan_dpp(s[EPL - 1], off[EPL - 1]);
        const int prev_off = wave_shift<kWaveShr1>(off[EPL - 1], -1);
        const Acc prev_T = wave_shr1(T);
        const int next_off = wave_shift<kWaveShl1>(off[0], -1);
        bool first = true;   // still in the lane's first run
#pragma unroll
        for (int j = 0; j < EPL - 1; ++j) {
            Acc v = s[j];
            if (first && prev_off == off[0]) v += prev_T;
            emit(valid && ends[j], off[j], v, false);
            first = first && !ends[j];
        }
        emit(valid && lane < L && next_off != off[EPL - 1], off[EPL - 1], T, false);
        carry = read_lane(T, L);
        carry_off = __builtin_amdgcn_readlane(off[EPL - 1], L);
        if (more) {
            cur = nxt;
            gather(cur.k, xc);
        }
    }
    emit(lane == 0, carry_off, carry, true);
    join_pieces<V, A, COMP>(a, split_first, split_last, yout);
}

// One item, streamed by one wave: y rows go to yout (y, or a panel's partial
// sums).  With LDS, keys flagged as staged read x from the tile's LDS cache
// xs; without (panel items), such keys are padding and read 0.
template <typename V, typename A, bool COMP, int EPL, bool SDDMM, bool LDS, int RING, typename Stage>
__device__ __forceinline__ void stream_item(const StreamArgs<V, A>& a, uint32_t item, const A* xs, A* yout,
                                            Stage&& stage) {
    if constexpr (RING == kNearRing) {
        stream_item_near<V, A, COMP, EPL, SDDMM, LDS>(a, item, xs, yout, stage);
        return;
    }
    using Acc = typename std::conditional<COMP, f2, A>::type;
    const uint32_t kZero = a.tile_slots - 1;
    const int lane = lane_id();
    const StreamItem it = load_item(a.items, item);
    const int cb = a.col_bits;
    const uint32_t cmask = (1u << cb) - 1u;
    const uint32_t omask = (1u << (31 - cb)) - 1u;
    // The cut-row tables are read at once when the item has no tile to stage
    // (the registers are shorter then), else after the staging, which must
    // not wait for this item's descriptor.
    constexpr bool kLate = !std::is_same<std::decay_t<Stage>, RunNow>::value;
    uint32_t split_first = 0xffffffffu, split_last = 0xffffffffu;
    if constexpr (!kLate) {
        if (it.first_partial) split_first = gptr(a.item_split_first)[item];
        if (it.last_partial) split_last = gptr(a.item_split)[item];
    }

    // A finished run (row offset `off`) goes to y, or to the partial buffer
    // when it is a piece of a cut row (the item's first run, offset 0, or its
    // last run).  Pieces keep their full accumulator (hi and lo of a
    // compensated sum).
    auto put_partial = [&](uint32_t slot, Acc v) {
        if constexpr (COMP) {
            a.partial[2 * slot] = v.hi;
            a.partial[2 * slot + 1] = v.lo;
        } else {
            a.partial[slot] = v;
        }
    };

    // Every memory instruction of the loop below is issued by all lanes and on
    // every pass (buffer instructions whose out-of-range lanes neither read nor
    // write), so the compiler can count them and wait for exactly the load a
    // value comes from: chunk c + RING - 1 is loaded and the x of chunk c + 1
    // gathered while chunk c is reduced.  The gather of chunk c + 1 waits for
    // its keys, loaded RING - 2 chunks earlier: that distance, not the wave
    // count, hides the HBM latency.
    static_assert(RING == kNearRing || RING >= 3, "the ring holds the reduced, the gathered and a loading chunk");
    constexpr uint32_t EB = 4 + sizeof(V);
    constexpr uint32_t kChunk = kWave * EPL;
    // Set by stage()'s hook (nothing that waits for the item's descriptor
    // may come before the staging's first loads) and after it.
    uint32_t ne = 0;
    rsrc_t rs = make_rsrc(a.slab, 0);
    const rsrc_t rx = make_rsrc(a.x, (uint32_t)a.n * (uint32_t)sizeof(A));
    rsrc_t ry = make_rsrc(yout, 0);
    bool first_piece = false;

    auto gather = [&](const LaneChunk<V, EPL>& c, A* g, bool live) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const bool cached = (c.k[j] & kKeyCached) != 0;
            g[j] = buffer_load<A>(rx, live && !cached ? (c.k[j] & cmask) * (uint32_t)sizeof(A) : kOutOfRange);
        }
    };

    int carry_off = -1;
    Acc carry{};
    Acc held{};            // the item's first run when it is a piece of a cut row
    bool has_held = false;
    auto reduce = [&](const LaneChunk<V, EPL>& c, const A* g, uint32_t r0) {
        const int L = (int)min(63u, (ne - r0) / EPL - 1);   // last lane with entries
        const bool valid = lane <= L;
        int off[EPL];
        Acc s[EPL];
        A xc[EPL];
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            A l = A(0);
            if constexpr (LDS) l = xs[min(c.k[j] & cmask, kZero)];
            xc[j] = (c.k[j] & kKeyCached) != 0 ? l : g[j];
            off[j] = valid ? (int)((c.k[j] >> cb) & omask) : 0x7fffffff;
            if constexpr (COMP) s[j] = two_prod(widen(c.v[j]), xc[j]);
            else s[j] = A(widen(c.v[j])) * xc[j];
        }
        if constexpr (SDDMM) {
            A dv[EPL];
            row_products<A, EPL>(a.xrow, (uint32_t)a.m, it.row_base, off, valid, xc, dv);
            const rsrc_t rd = make_rsrc(a.dvals + it.e_begin, ne * (uint32_t)sizeof(A));
            const uint32_t at = valid ? (r0 + (uint32_t)lane * EPL) * (uint32_t)sizeof(A) : kOutOfRange;
#pragma unroll
            for (int q = 0; q < EPL * (int)sizeof(A) / 16; ++q) {
                u32x4 w;
                __builtin_memcpy(&w, reinterpret_cast<const unsigned char*>(dv) + 16 * q, 16);
                __builtin_amdgcn_raw_buffer_store_b128(w, rd, at + 16 * q, 0, kNonTemporal);
            }
        }
        // a finished run inside the loop: to y, or held when it is the first
        // piece of a cut row
        auto emit = [&](bool go, int o, Acc v) {
            const bool to_partial = first_piece && o == 0;
            if (go && to_partial) {
                held = v;
                has_held = true;
            }
            buffer_store<A>(finish(v), ry, go && !to_partial ? (uint32_t)o * (uint32_t)sizeof(A) : kOutOfRange);
        };
        // the run left open by the previous chunk
        const int o0 = __builtin_amdgcn_readfirstlane(off[0]);
        const bool cont = carry_off >= 0 && o0 == carry_off;
        if (lane == 0 && cont) s[0] += carry;
        emit(lane == 0 && carry_off >= 0 && !cont, carry_off, carry);
        // runs inside the lane: s[j] becomes the partial sum of the run
        // through entry j; ends[j] marks a run that stops at entry j
        bool ends[EPL];
#pragma unroll
        for (int j = 1; j < EPL; ++j) {
            const bool same = off[j] == off[j - 1];
            ends[j - 1] = !same;
            if (same) s[j] += s[j - 1];
        }
        // runs across lanes: segmented scan of the lanes' last runs
        const Acc T = seg_scan_dpp(s[EPL - 1], off[EPL - 1]);
        const int prev_off = wave_shift<kWaveShr1>(off[EPL - 1], -1);
        const Acc prev_T = wave_shr1(T);
        const int next_off = wave_shift<kWaveShl1>(off[0], -1);
        bool first = true;   // still in the lane's first run
#pragma unroll
        for (int j = 0; j < EPL - 1; ++j) {
            Acc v = s[j];
            if (first && prev_off == off[0]) v += prev_T;
            emit(valid && ends[j], off[j], v);
            first = first && !ends[j];
        }
        emit(valid && lane < L && next_off != off[EPL - 1], off[EPL - 1], T);
        carry = read_lane(T, L);
        carry_off = __builtin_amdgcn_readlane(off[EPL - 1], L);
    };

    LaneChunk<V, EPL> c[RING];
    A g[RING][EPL];
    // The first RING - 1 chunks are loaded while the tile's x lines are staged.
    stage([&] {
        ne = it.e_end - it.e_begin;
        rs = make_rsrc(a.slab + (size_t)it.e_begin * EB, ne * EB);
#pragma unroll
        for (int d = 0; d < RING - 1; ++d) c[d].load(rs, d * kChunk, ne, lane);
    });
    ry = make_rsrc(yout + it.row_base, min((uint32_t)a.m - it.row_base, omask + 1u) * (uint32_t)sizeof(A));
    first_piece = it.first_partial != 0;
    if constexpr (kLate) {
        if (it.first_partial) split_first = gptr(a.item_split_first)[item];
        if (it.last_partial) split_last = gptr(a.item_split)[item];
    }
    gather(c[0], g[0], true);
    uint32_t r0 = 0;
    // Paced plans (StreamArgs::paced): priority 3 for the item's first
    // quarter, down to 0 for its last.  Measured on the 1,048,576^2 stencil /
    // power-law matrix's other entries (99.5% of them staged): 126.0 -> 124.2
    // us per split SpMV; on a gather-heavy plan it costs (219.8 -> 224.8 us
    // for the 2,097,152^2 block-clustered one unsplit), hence the flag.
    const bool paced = a.paced != 0;
    const uint32_t quarter = ne / 4;
    int pace = 0;
    if (paced) __builtin_amdgcn_s_setprio(3);
    for (bool more = true; more;) {
#pragma unroll
        for (int u = 0; u < RING; ++u) {
            if (r0 >= ne) {
                more = false;
                break;
            }
            gather(c[(u + 1) % RING], g[(u + 1) % RING], r0 + kChunk < ne);
            c[(u + RING - 1) % RING].load(rs, r0 + (RING - 1) * kChunk, ne, lane);
            reduce(c[u], g[u], r0);
            r0 += kChunk;
        }
        if (paced) {
            if (pace == 0 && r0 >= quarter) {
                __builtin_amdgcn_s_setprio(2);
                pace = 1;
            } else if (pace == 1 && r0 >= 2 * quarter) {
                __builtin_amdgcn_s_setprio(1);
                pace = 2;
            } else if (pace == 2 && r0 >= 3 * quarter) {
                __builtin_amdgcn_s_setprio(0);
                pace = 3;
            }
        }
    }
    if (paced) __builtin_amdgcn_s_setprio(0);
    // the first piece of a cut row, if it ended inside the loop
    if (first_piece) {
        const uint64_t hm = __ballot(has_held);
        if (hm != 0) {
            const Acc v = read_lane(held, (int)__builtin_ctzll(hm));
            if (lane == 0) {
                const StreamSplit sp = load_split(a.splits, split_first);
                put_partial(sp.partial_base + (item - sp.first_item), v);
            }
        }
    }
    // the last run: a piece of a cut row (first and last piece when the item
    // holds one run), or a whole row
    if (lane == 0 && carry_off >= 0) {
        if (first_piece && carry_off == 0) {
            const StreamSplit sp = load_split(a.splits, split_first);
            put_partial(sp.partial_base + (item - sp.first_item), carry);
        } else if (it.last_partial) {
            const StreamSplit sp = load_split(a.splits, split_last);
            put_partial(sp.partial_base + (item - sp.first_item), carry);
        } else {
            yout[it.row_base + (uint32_t)carry_off] = finish(carry);
        }
    }

    join_pieces<V, A, COMP>(a, split_first, split_last, yout);
}

// One tile of the stream: stage its hot x lines in LDS, then each of its
// first tile_items waves streams one item.  With LOOP the workgroup may have
// fewer waves than the tile has items: wave w streams items w, w + waves, ...
// Every thread of the workgroup calls this.
template <typename V, typename A, bool COMP, int EPL, bool SDDMM, int RING, bool LOOP = false>
__device__ __forceinline__ void stream_tile(const StreamArgs<V, A>& a, uint32_t tile, A* xs) {
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));   // wave-uniform
    if constexpr (LOOP) {
        stage_tile(a, tile, xs, NoLoads{});
        const uint32_t waves = blockDim.x >> 6;
        for (uint32_t w = wave; w < a.tile_items; w += waves) {
            const uint32_t item = tile * a.tile_items + w;
            if (item >= a.n_items) break;
            stream_item<V, A, COMP, EPL, SDDMM, true, RING>(a, item, xs, a.y, run_now);
        }
    } else {
        // A wave with an item stages the tile from inside the item, after
        // its first slab loads are issued; the others stage it here.  Every
        // wave reaches the staging's barrier once.  Not with 8-entry lanes
        // in the ring, whose registers are short already: their waves stage
        // the tile first.
        const uint32_t item = tile * a.tile_items + wave;
        const bool has_item = wave < a.tile_items && item < a.n_items;
        if constexpr (RING == kNearRing || EPL == 4) {
            if (has_item)
                stream_item<V, A, COMP, EPL, SDDMM, true, RING>(a, item, xs, a.y,
                                                                [&](auto&& mid) { stage_tile(a, tile, xs, mid); });
            else
                stage_tile(a, tile, xs, NoLoads{});
        } else {
            stage_tile(a, tile, xs, NoLoads{});
            if (has_item) stream_item<V, A, COMP, EPL, SDDMM, true, RING>(a, item, xs, a.y, run_now);
        }
    }
}

// The R values x[col] of R right-hand sides, from x interleaved by column
// (p = xg + col * R; 16-byte aligned): 16-byte loads issued back to back
// over one column's R values, which the vector L1 serves as about one
// gather (a probe of random columns: one 16-byte load 265-290 G/s, two of
// one 32-byte column 227-238 G/s, two of separate columns 133-142 G/s).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
// Only the vectors holding one of the nr (wave-uniform) right-hand sides in
// use are loaded; the others stay 0.
template <int R>
__device__ __forceinline__ void load_rhs(const float* p, float (&v)[R], uint32_t nr) {
#pragma unroll
    for (int q = 0; q < R / 4; ++q) {
        f32x4 w = {0.f, 0.f, 0.f, 0.f};
        if ((uint32_t)(4 * q) < nr) w = *gptr(reinterpret_cast<const f32x4*>(p + 4 * q));
        v[4 * q] = w.x; v[4 * q + 1] = w.y; v[4 * q + 2] = w.z; v[4 * q + 3] = w.w;
    }
}
template <int R>
__device__ __forceinline__ void load_rhs(const double* p, double (&v)[R], uint32_t nr) {
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
        f64x2 w = {0.0, 0.0};
        if ((uint32_t)(2 * q) < nr) w = *gptr(reinterpret_cast<const f64x2*>(p + 2 * q));
        v[2 * q] = w.x; v[2 * q + 1] = w.y;
    }
}

// Several right-hand sides at once: each chunk of keys and values is loaded
// once and reduced for every right-hand side r < nr (at most R), so the slab
// is read once per tile instead of once per right-hand side.  x comes
// interleaved (a.x[col * gw + r], written by batch_prologue) and is gathered
// from global memory (L2) for every entry, four right-hand sides at a time
// (load_rhs); LDS holds the tile's line table to
// decode the column of an entry whose key names an LDS slot.  RHS r writes
// y + r*ldy; pieces of cut rows go to partial + r*pstride and share one
// ticket per split.  multi_item runs one item with one wave;
// stream_tile_multi deals the tile's items to the workgroup's waves.
template <typename V, typename A, bool COMP, int EPL, int R>
__device__ __forceinline__ void multi_item(const StreamArgs<V, A>& a, uint32_t item, const uint32_t* lines,
                                           uint64_t ldy, uint64_t pstride, uint32_t nr, uint32_t gw) {
    using Acc = typename std::conditional<COMP, f2, A>::type;
    const uint32_t kZero = a.tile_slots - 1;
    const int lane = lane_id();
    const StreamItem it = load_item(a.items, item);
    const int cb = a.col_bits;
    const uint32_t cmask = (1u << cb) - 1u;
    const uint32_t omask = (1u << (31 - cb)) - 1u;
    const uint32_t split_first = it.first_partial ? gptr(a.item_split_first)[item] : 0xffffffffu;
    const uint32_t split_last = it.last_partial ? gptr(a.item_split)[item] : 0xffffffffu;
    constexpr int kParts = COMP ? 2 : 1;
    const uint32_t item_rows = min((uint32_t)a.m - it.row_base, omask + 1u);
    // the first piece of a cut row: to the partial buffer of right-hand side r
    auto put_partial_r = [&](uint32_t r, Acc v) {
        const StreamSplit sp = load_split(a.splits, split_first);
        const uint32_t slot = sp.partial_base + (item - sp.first_item);
        gbl<A>* p = gptr(a.partial + (size_t)r * pstride);
        if constexpr (COMP) {
            p[2 * slot] = v.hi;
            p[2 * slot + 1] = v.lo;
        } else {
            p[slot] = v;
        }
    };

    auto emit = [&](bool go, int off, Acc v, bool final_run, uint32_t r) {
        if (!go) return;
        uint32_t slot = 0xffffffffu;
        if (it.first_partial && off == 0) {
            const StreamSplit sp = load_split(a.splits, split_first);
            slot = sp.partial_base + (item - sp.first_item);
        } else if (final_run && it.last_partial) {
            const StreamSplit sp = load_split(a.splits, split_last);
            slot = sp.partial_base + (item - sp.first_item);
        }
        if (slot != 0xffffffffu) {
            gbl<A>* p = gptr(a.partial + (size_t)r * pstride);
            if constexpr (COMP) {
                p[2 * slot] = v.hi;
                p[2 * slot + 1] = v.lo;
            } else {
                p[slot] = v;
            }
        } else {
            gptr(a.y)[(size_t)r * ldy + it.row_base + (uint32_t)off] = finish(v);
        }
    };
    // column of every entry: padding reads column 0 with value 0
    auto columns = [&](const uint32_t* kk, uint32_t* cc) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const uint32_t low = kk[j] & cmask;
            const bool cached = (kk[j] & kKeyCached) != 0;
            uint32_t c = low;
            if (cached) c = low == kZero ? 0u : lines[low / kLineElems] * kLineElems + low % kLineElems;
            cc[j] = c;
        }
    };
    constexpr uint32_t kChunk = kWave * EPL;
    auto lanes_of = [&](uint32_t e0) { return min(64u, (it.e_end - e0) / EPL); };

    int carry_off = -1;
    Acc carry[R];
#pragma unroll
    for (int r = 0; r < R; ++r) carry[r] = Acc{};
    Chunk<V, EPL> cur, nxt;
    uint32_t col[EPL];
    cur.load(a.slab, it.e_begin, lanes_of(it.e_begin), lane);
    for (uint32_t e0 = it.e_begin; e0 < it.e_end; e0 += kChunk) {
        const bool more = e0 + kChunk < it.e_end;
        if (more) nxt.load(a.slab, e0 + kChunk, lanes_of(e0 + kChunk), lane);
        const int L = (int)min(63u, (it.e_end - e0) / EPL - 1);
        const bool valid = lane <= L;
        columns(cur.k, col);
        int off[EPL];
        bool ends[EPL];
#pragma unroll
        for (int j = 0; j < EPL; ++j) off[j] = valid ? (int)((cur.k[j] >> cb) & omask) : 0x7fffffff;
#pragma unroll
        for (int j = 1; j < EPL; ++j) ends[j - 1] = off[j] != off[j - 1];
        const int o0 = __builtin_amdgcn_readfirstlane(off[0]);
        const int prev_off = wave_shift<kWaveShr1>(off[EPL - 1], -1);
        const int next_off = wave_shift<kWaveShl1>(off[0], -1);
        const bool carried = carry_off >= 0 && o0 == carry_off;
        const bool carry_emit = carry_off >= 0 && !carried;   // the previous chunk's open run ended
        // byte offsets of the rows whose runs end at each entry slot; the
        // first piece of a cut row (offset 0 of an item with first_partial)
        // goes to the partial buffer instead (row0_end)
        uint32_t so[EPL];
        bool row0_end[EPL];
        bool any0 = false;
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
            const bool end = j < EPL - 1 ? valid && ends[j] : valid && lane < L && next_off != off[EPL - 1];
            row0_end[j] = end && it.first_partial && off[j] == 0;
            any0 = any0 || row0_end[j];
            so[j] = end && !row0_end[j] ? (uint32_t)off[j] * (uint32_t)sizeof(A) : kOutOfRange;
        }
        const uint32_t so_carry = lane == 0 && carry_emit && !(it.first_partial && carry_off == 0)
                                      ? (uint32_t)carry_off * (uint32_t)sizeof(A) : kOutOfRange;
        // wave-uniform: the first piece of a cut row ends in this chunk
        const bool row0 = it.first_partial && (__ballot(any0) != 0 || (carry_emit && carry_off == 0));
        // four right-hand sides at a time (a 16-byte load per entry for FP32):
        // the x of eight for a chunk would hold 64 VGPRs, and with at most
        // 128 per wave the kernel spilled; the second four read the same
        // cache lines as the first
#pragma unroll
        for (int h = 0; h < R; h += kRhsGroupNarrow) {
          if ((uint32_t)h >= nr) break;
          A xr[EPL][kRhsGroupNarrow];
#pragma unroll
          for (int j = 0; j < EPL; ++j) load_rhs<kRhsGroupNarrow>(a.x + (size_t)col[j] * gw + h, xr[j], nr - h);
#pragma unroll
          for (int q = 0; q < kRhsGroupNarrow; ++q) {
            const int r = h + q;
            if ((uint32_t)r >= nr) break;
            Acc s[EPL];
#pragma unroll
            for (int j = 0; j < EPL; ++j) {
                const A xv = xr[j][q];
                if constexpr (COMP) s[j] = two_prod(widen(cur.v[j]), xv);
                else s[j] = A(widen(cur.v[j])) * xv;
            }
            if (carried && lane == 0) s[0] += carry[r];
#pragma unroll
            for (int j = 1; j < EPL; ++j)
                if (!ends[j - 1]) s[j] += s[j - 1];
            const Acc T = seg_scan_dpp(s[EPL - 1], off[EPL - 1]);
            const Acc prev_T = wave_shr1(T);
            // runs that end in this chunk, and the run the previous chunk left
            // open when it ends here: one buffer store per entry slot, lanes
            // without a run end out of range
            const uint32_t yb = (uint32_t)r * (uint32_t)ldy;
            const rsrc_t ry = make_rsrc(a.y + yb + it.row_base, item_rows * (uint32_t)sizeof(A));
            buffer_store<A>(finish(carry[r]), ry, so_carry);
            bool first = true;
#pragma unroll
            for (int j = 0; j < EPL; ++j) {
                Acc v = j == EPL - 1 ? T : s[j];
                if (j < EPL - 1 && first && prev_off == off[0]) v += prev_T;
                buffer_store<A>(finish(v), ry, so[j]);
                if (row0 && row0_end[j]) put_partial_r(r, v);   // the first piece of a cut row
                if (j < EPL - 1) first = first && !ends[j];
            }
            if (row0 && carry_emit && carry_off == 0 && lane == 0) put_partial_r(r, carry[r]);
            carry[r] = read_lane(T, L);
          }
        }
        carry_off = __builtin_amdgcn_readlane(off[EPL - 1], L);
        if (more) cur = nxt;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if ((uint32_t)r < nr) emit(lane == 0, carry_off, carry[r], true, r);

    for (int which = 0; which < 2; ++which) {
        const uint32_t sid = which == 0 ? split_first : split_last;
        if (sid == 0xffffffffu || (which == 1 && sid == split_first)) continue;
        __threadfence();
        uint32_t ticket = 0;
        if (lane == 0) ticket = atomicAdd(a.counters + sid, 1u);
        ticket = __shfl(ticket, 0, kWave);
        const StreamSplit sp = load_split(a.splits, sid);
        if (ticket == sp.pieces - 1) {
            __threadfence();
            if (lane == 0) {
                for (uint32_t r = 0; r < nr; ++r) {
                    const A* p = a.partial + (size_t)r * pstride;
                    Acc sum{};
                    for (uint32_t q = 0; q < sp.pieces; ++q) {
                        const uint32_t slot = kParts * (sp.partial_base + q);
                        if constexpr (COMP) sum += f2{load_partial(p + slot), load_partial(p + slot + 1)};
                        else sum += load_partial(p + slot);
                    }
                    a.y[(size_t)r * ldy + sp.row] = finish(sum);
                }
                atomicExch(a.counters + sid, 0u);
            }
        }
    }
}

// The workgroup's LDS holds the tile's line table (to decode the columns of
// keys that name LDS slots); x is gathered from the interleaved copy for
// every entry.  Staging the lines' interleaved x in LDS as well, when all
// of a tile's lines fit, made the 256-matrix batch slower (2.22 against
// 2.02 ms): a banded matrix's gathers already hit the vector L1, and the
// staging delays every unit's start.
template <typename V, typename A, bool COMP, int EPL, int R>
__device__ __forceinline__ void stream_tile_multi(const StreamArgs<V, A>& a, uint32_t tile, unsigned char* lds,
                                                  uint64_t ldy, uint64_t pstride, uint32_t nr, uint32_t gw) {
    uint32_t* lines = reinterpret_cast<uint32_t*>(lds);
    const uint32_t lb = a.tiles[2 * tile], cnt = a.tiles[2 * tile + 1];
    for (uint32_t s = threadIdx.x; s < cnt; s += blockDim.x) lines[s] = gptr(a.tile_lines)[lb + s];
    __syncthreads();
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));   // wave-uniform
    const uint32_t waves = blockDim.x >> 6;
    for (uint32_t w = wave; w < a.tile_items; w += waves) {
        const uint32_t item = tile * a.tile_items + w;
        if (item >= a.n_items) break;
        multi_item<V, A, COMP, EPL, R>(a, item, lines, ldy, pstride, nr, gw);
    }
}

template <typename V, typename A, bool COMP, int EPL, bool SDDMM, bool NEAR>
__global__ __launch_bounds__(kMaxStreamThreads) void stream_spmv(StreamArgs<V, A> a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    A* xs = reinterpret_cast<A*>(lds_raw);   // tile_slots elements (dynamic LDS)
    // Workgroup b streams tile b.  The hardware deals workgroups round-robin
    // over the 8 XCDs, so neighbouring tiles (rows) land on different XCDs
    // and a matrix whose costly rows are contiguous (a 1M-row matrix with a
    // stencil half and a power-law half with scattered columns) spreads its
    // gathers over every XCD's L2; contiguous tile ranges per XCD (for
    // shared x windows) were 3.6% slower there and no faster on banded
    // matrices.
    const uint32_t tile = blockIdx.x;
    const uint32_t nthreads = blockDim.x;
    for (uint32_t t = blockIdx.x * nthreads + threadIdx.x; t < a.n_empty_rows; t += gridDim.x * nthreads)
        a.y[a.empty_rows[t]] = A(0);
    constexpr int RING = NEAR ? kNearRing : SDDMM && EPL == 8 ? 3 : deep_ring<V, EPL>();
    if (tile < a.n_tiles) stream_tile<V, A, COMP, EPL, SDDMM, RING>(a, tile, xs);
    if (a.panels <= 1) return;
    // Column panels: waves [tile_items, 2 tile_items) run items of panel
    // (XCD % P) (the grid is a multiple of 8: workgroup b is on XCD b % 8),
    // dealt over the panel waves of the panel's XCDs, into panel q's sums.
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < a.tile_items) return;
    const uint32_t P = a.panels, per = kXcds / P;
    const uint32_t xcd = blockIdx.x % kXcds, q = xcd % P;
    const uint32_t rank = (blockIdx.x / kXcds) * per + xcd / P;   // among panel q's workgroups
    const uint32_t stride = (gridDim.x / kXcds) * per * a.tile_items;
    A* out = a.panel_out + (size_t)q * (uint32_t)a.m;
    for (uint32_t i = a.panel_item_ptr[q] + rank * a.tile_items + (wave - a.tile_items); i < a.panel_item_ptr[q + 1];
         i += stride)
        stream_item<V, A, false, EPL, SDDMM, false, RING>(a, i, nullptr, out, run_now);
}

// y[r] += the panel sums of row r (in panel order), for rows with panel
// entries; other rows are not touched.
template <typename A>
__global__ __launch_bounds__(256) void panel_sums(A* y, const A* out, const uint8_t* mask, uint32_t m, uint32_t P) {
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < m; r += gridDim.x * blockDim.x) {
        const uint32_t mk = mask[r];
        if (mk == 0) continue;
        A t = A(0);
        for (uint32_t q = 0; q < P; ++q)
            if ((mk >> q) & 1u) t += out[(size_t)q * m + r];
        y[r] = y[r] + t;
    }
}

// Batched persistent launch: work units are (plan, tile) pairs.  Each XCD
// owns the contiguous unit range split.begin[xcd] .. split.begin[xcd + 1]
// (ranges of equal estimated cost, so a plan's tiles share one L2), and its
// workgroups take the next unit of that range from queue[xcd] (zeroed by
// batch_prologue before the launch).  A workgroup whose XCD's range is done
// goes on with the other XCDs' queues in turn, so an error of the cost
// estimate leaves no XCD idle; a unit is taken once whoever takes it.  A
// unit runs its tile for every right-hand side of its plan in turn.  Unit 0
// of a plan also zeroes the plan's empty rows.
template <typename V, typename A, bool COMP, int EPL>
__global__ __launch_bounds__(kMaxBatchThreads) void stream_batch(const BatchPlan<V, A>* __restrict__ plans, uint32_t nplans,
                                                                  XcdSplit split, uint32_t* queue) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ uint32_t next_unit;
    A* xs = reinterpret_cast<A*>(lds_raw);
    const uint32_t nq = split.nq;
    uint32_t q = xcc_id() % nq;   // start with this XCD's own queue
    uint32_t queues_done = 0;
    for (;;) {
        if (threadIdx.x == 0) next_unit = split.begin[q] + atomicAdd(queue + q, 1u);
        __syncthreads();
        const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)next_unit);
        __syncthreads();   // next_unit is rewritten at the next unit
        if (u >= split.begin[q + 1]) {
            if (++queues_done == nq) break;
            q = q + 1u == nq ? 0u : q + 1u;
            continue;
        }
        // plan of unit u: last plan with unit_begin <= u (binary search)
        uint32_t lo = 0, hi = nplans;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (plans[mid].unit_begin <= u) lo = mid;
            else hi = mid;
        }
        const BatchPlan<V, A>& bp = plans[lo];
        const uint32_t tile = u - bp.unit_begin;
        if (tile == 0)
            for (uint32_t r = 0; r < bp.nrhs; ++r)
                for (uint32_t t = threadIdx.x; t < bp.args.n_empty_rows; t += blockDim.x)
                    gptr(bp.y)[(size_t)r * bp.ldy + gptr(bp.args.empty_rows)[t]] = A(0);
        if (tile >= bp.args.n_tiles) continue;
        if (bp.nrhs == 1) {   // one right-hand side: the LDS x cache
            StreamArgs<V, A> a = bp.args;
            a.x = bp.x;
            a.y = bp.y;
            a.partial = bp.partial;
            a.counters = bp.counters;
            stream_tile<V, A, COMP, EPL, false, kNearRing, true>(a, tile, xs);
            __syncthreads();   // LDS is restaged by the next unit
            continue;
        }
        // several: groups of up to gw share each chunk load and each x
        // gather (x interleaved per group by batch_prologue)
        const uint32_t gw = bp.gw;
        for (uint32_t r0 = 0; r0 < bp.nrhs; r0 += gw) {
            StreamArgs<V, A> a = bp.args;
            a.x = bp.xg + (size_t)r0 * bp.args.n;
            a.y = bp.y + (size_t)r0 * bp.ldy;
            a.partial = bp.partial + (size_t)r0 * bp.partial_stride;
            a.counters = bp.counters + (size_t)r0 * bp.counter_stride;
            stream_tile_multi<V, A, COMP, EPL, kMultiRhs>(a, tile, lds_raw, bp.ldy, bp.partial_stride,
                                                          min(bp.nrhs - r0, gw), gw);
            __syncthreads();
        }
    }
}

// batch_prologue writes the x of every plan with several right-hand sides
// interleaved in groups of gw (the plan's group width) for stream_tile_multi:
// xg[(g*n + col)*gw + j] = x[(g*gw + j)*ldx + col].  Reads are
// coalesced over col; each thread writes the 16-byte vectors of one column
// that hold a right-hand side (entries past the last one are 0; vectors
// holding none are neither written nor loaded by load_rhs).
template <int R>
__device__ __forceinline__ void store_rhs(float* p, const float (&v)[R], uint32_t nr) {
    static_assert(R % 4 == 0, "whole 16-byte vectors");
#pragma unroll
    for (int q = 0; q < R / 4; ++q)
        if ((uint32_t)(4 * q) < nr)
            reinterpret_cast<f32x4*>(p)[q] = f32x4{v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]};
}
template <int R>
__device__ __forceinline__ void store_rhs(double* p, const double (&v)[R], uint32_t nr) {
    static_assert(R % 2 == 0, "whole 16-byte vectors");
#pragma unroll
    for (int q = 0; q < R / 2; ++q)
        if ((uint32_t)(2 * q) < nr) reinterpret_cast<f64x2*>(p)[q] = f64x2{v[2 * q], v[2 * q + 1]};
}

// Runs before every stream_batch launch: zeroes the per-XCD unit queues
// and interleaves x.  Block b works on plan block_plan[b] (a table the host
// appends to the descriptor), on that plan's interleaved columns
// (b - xg_block0) * kPrologueCols + [0, kPrologueCols); the plans' column
// ranges are padded to whole blocks, so every block has one plan.
template <typename V, typename A>
__global__ __launch_bounds__(256) void batch_prologue(const BatchPlan<V, A>* plans, const uint32_t* block_plan,
                                                      uint32_t nblocks, uint32_t* queue) {
    if (blockIdx.x == 0 && threadIdx.x < kMaxXcds) queue[threadIdx.x] = 0;
    if (blockIdx.x >= nblocks) return;   // the one block launched when no plan needs x interleaved
    const BatchPlan<V, A>& bp = plans[block_plan[blockIdx.x]];
    const uint32_t n = (uint32_t)bp.args.n;
    const uint32_t gw = bp.gw;
    const uint32_t cols = (bp.nrhs + gw - 1) / gw * n;
    const uint32_t c0 = (blockIdx.x - bp.xg_block0) * kPrologueCols + threadIdx.x;
    // all loads of the thread's columns first (a column past the end reads
    // the last one and is not stored), then the stores
    constexpr uint32_t K = kPrologueCols / 256;
    A v[K][kMultiRhs];
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t c = min(c0 + k * 256, cols - 1);
        const uint32_t g = c / n, col = c - g * n;
#pragma unroll
        for (int j = 0; j < kMultiRhs; ++j) {
            const uint32_t r = g * gw + (uint32_t)j;
            v[k][j] = (uint32_t)j < gw && r < bp.nrhs ? gptr(bp.x)[(size_t)r * bp.ldx + col] : A(0);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t c = c0 + k * 256;
        if (c < cols) {
            const uint32_t g = c / n;
            store_rhs<kMultiRhs>(bp.xg + (size_t)c * gw, v[k], min(bp.nrhs - g * gw, gw));
        }
    }
}

}  // namespace dev

template <typename V, typename A, bool COMP, bool SDDMM, bool NEAR>
hipError_t launch_stream_t(const StreamArgs<V, A>& a, int grid, hipStream_t s) {
    constexpr int EPL = sizeof(V) >= 8 ? 4 : 8;
    auto kern = dev::stream_spmv<V, A, COMP, EPL, SDDMM, NEAR>;
    const size_t lds = (size_t)a.tile_slots * sizeof(A);
    static bool attr_set = false;   // allow more than 64 KB of dynamic LDS
    if (!attr_set) {
        const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxTileBytes);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    // with column panels, as many panel waves as tile waves
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * a.tile_items * (a.panels > 1 ? 2 : 1)), lds, s, a);
    return hipGetLastError();
}

// x within one XCD's L2 (kPanelL2Bytes): the one-ahead loop, else the ring
template <typename V, typename A, bool COMP, bool SDDMM>
hipError_t launch_stream_x(const StreamArgs<V, A>& a, int grid, hipStream_t s) {
    const bool near = (int64_t)a.n * (int64_t)sizeof(A) <= kPanelL2Bytes;
    return near ? launch_stream_t<V, A, COMP, SDDMM, true>(a, grid, s) : launch_stream_t<V, A, COMP, SDDMM, false>(a, grid, s);
}

template <typename V, typename A>
hipError_t launch_stream_any(const StreamArgs<V, A>& a, int grid, bool comp, bool sddmm, hipStream_t s) {
    if (comp) return sddmm ? launch_stream_x<V, A, true, true>(a, grid, s) : launch_stream_x<V, A, true, false>(a, grid, s);
    return sddmm ? launch_stream_x<V, A, false, true>(a, grid, s) : launch_stream_x<V, A, false, false>(a, grid, s);
}

template <typename V, typename A, bool COMP>
hipError_t launch_batch_t(const BatchPlan<V, A>* plans, uint32_t nplans, const XcdSplit& split, uint32_t* queue,
                          int grid, int threads, size_t lds, hipStream_t s) {
    constexpr int EPL = sizeof(V) >= 8 ? 4 : 8;
    auto kern = dev::stream_batch<V, A, COMP, EPL>;
    static bool attr_set = false;
    if (!attr_set) {
        const hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxTileBytes);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, s, plans, nplans, split, queue);
    return hipGetLastError();
}

int batch_blocks_per_cu(int dtype, bool comp, int threads, size_t lds) {
    int n = 0;
    hipError_t e;
    if (dtype == CBSPMV_F64)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::stream_batch<double, double, false, 4>, threads, lds);
    else if (dtype == CBSPMV_F32)
        e = comp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::stream_batch<float, float, true, 8>, threads, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::stream_batch<float, float, false, 8>, threads, lds);
    else
        e = comp ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::stream_batch<dev::bf16_t, float, true, 8>, threads, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, dev::stream_batch<dev::bf16_t, float, false, 8>, threads, lds);
    return e == hipSuccess ? n : 0;
}

template <typename V, typename A>
hipError_t launch_prologue_t(const BatchPlan<V, A>* p, const uint32_t* block_plan, uint32_t nblocks, uint32_t* queue,
                             hipStream_t s) {
    hipLaunchKernelGGL((dev::batch_prologue<V, A>), dim3(std::max(nblocks, 1u)), dim3(256), 0, s, p, block_plan,
                       nblocks, queue);
    return hipGetLastError();
}
hipError_t launch_prologue(const BatchPlan<double, double>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s) {
    return launch_prologue_t(p, block_plan, nblocks, queue, s);
}
hipError_t launch_prologue(const BatchPlan<float, float>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s) {
    return launch_prologue_t(p, block_plan, nblocks, queue, s);
}
hipError_t launch_prologue(const BatchPlan<dev::bf16_t, float>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s) {
    return launch_prologue_t(p, block_plan, nblocks, queue, s);
}

hipError_t launch_batch(const BatchPlan<double, double>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s) {
    (void)comp;
    return launch_batch_t<double, double, false>(p, np, split, queue, grid, threads, lds, s);
}
hipError_t launch_batch(const BatchPlan<float, float>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s) {
    return comp ? launch_batch_t<float, float, true>(p, np, split, queue, grid, threads, lds, s)
                : launch_batch_t<float, float, false>(p, np, split, queue, grid, threads, lds, s);
}
hipError_t launch_batch(const BatchPlan<dev::bf16_t, float>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s) {
    return comp ? launch_batch_t<dev::bf16_t, float, true>(p, np, split, queue, grid, threads, lds, s)
                : launch_batch_t<dev::bf16_t, float, false>(p, np, split, queue, grid, threads, lds, s);
}

template <typename A>
hipError_t launch_panel_sums_t(A* y, const A* out, const uint8_t* mask, uint32_t m, uint32_t panels, hipStream_t s) {
    const uint32_t grid = std::min<uint32_t>(4096u, (m + 255u) / 256u);
    hipLaunchKernelGGL(dev::panel_sums<A>, dim3(std::max(grid, 1u)), dim3(256), 0, s, y, out, mask, m, panels);
    return hipGetLastError();
}
hipError_t launch_panel_sums(double* y, const double* out, const uint8_t* mask, uint32_t m, uint32_t panels,
                             hipStream_t s) {
    return launch_panel_sums_t(y, out, mask, m, panels, s);
}
hipError_t launch_panel_sums(float* y, const float* out, const uint8_t* mask, uint32_t m, uint32_t panels,
                             hipStream_t s) {
    return launch_panel_sums_t(y, out, mask, m, panels, s);
}

hipError_t launch_stream(const StreamArgs<double, double>& a, int grid, bool sddmm, hipStream_t s) {
    return sddmm ? launch_stream_x<double, double, false, true>(a, grid, s)
                 : launch_stream_x<double, double, false, false>(a, grid, s);
}
hipError_t launch_stream(const StreamArgs<float, float>& a, int grid, bool comp, bool sddmm, hipStream_t s) {
    return launch_stream_any(a, grid, comp, sddmm, s);
}
hipError_t launch_stream(const StreamArgs<dev::bf16_t, float>& a, int grid, bool comp, bool sddmm, hipStream_t s) {
    return launch_stream_any(a, grid, comp, sddmm, s);
}

}  // namespace cbs
