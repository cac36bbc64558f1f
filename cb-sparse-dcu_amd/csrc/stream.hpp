// Launch interface of the stream schedule (stream.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cb_plan.hpp"

namespace cbs {

namespace dev { struct bf16_t; }

constexpr int kMaxStreamThreads = 64 * kMaxTileItems;   // one wave per tile item
// The batched kernel has at most 16 waves per workgroup (at most 128 VGPRs
// each, one workgroup per CU); the items of a tile with more are dealt to its
// waves in turn.
constexpr int kMaxBatchTileItems = 16;
// Right-hand sides of one pass of the batched kernel over a tile at most, and
// the widest group x is interleaved in for it (BatchPlan::xg); a plan's
// group width is 4 or 8 (BatchPlan::gw).
constexpr int kRhsGroup = 8;
constexpr int kRhsGroupNarrow = 4;
constexpr int kMaxBatchThreads = 64 * kMaxBatchTileItems;

template <typename V, typename A>
struct StreamArgs {
    const uint8_t* slab;          // chunk lane planes (cb_plan.hpp)
    const StreamItem* items;
    const uint32_t* item_split;
    const uint32_t* item_split_first;
    const StreamSplit* splits;
    const uint32_t* tiles;        // (first line, line count) per tile
    const uint32_t* tile_lines;
    const uint32_t* empty_rows;
    A* partial;
    uint32_t* counters;
    uint32_t n_empty_rows;
    uint32_t n_items;
    uint32_t n_tiles;
    uint32_t tile_items;          // waves per workgroup
    uint32_t tile_slots;          // LDS x-cache elements
    int col_bits;
    int m, n;
    const A* x;
    A* y;
    // backward over the transposed plan: dvals[e] = x[col(e)] * xrow[row(e)]
    // for every padded stream entry e (SDDMM launches only)
    const A* xrow;
    A* dvals;
    // column panels (panels > 1; cb_plan.hpp): n_items counts the tiles'
    // items only, the panel items follow them
    uint32_t panels;
    const uint32_t* panel_item_ptr;  // panel q's items: [ptr[q], ptr[q + 1])
    A* panel_out;                    // panels x m partial sums
    // 1 on the other entries' plan of a scattered-entry split: a wave's issue
    // priority falls as its item progresses (ring loop), so the tile's
    // lagging waves go first and its waves end closer together
    uint32_t paced;
};

// One plan of a batched launch: its stream arguments (x, y, partial and
// counters are set per right-hand side from the fields below) and the first
// work unit (tile) it owns in the batch.
template <typename V, typename A>
struct BatchPlan {
    StreamArgs<V, A> args;
    uint32_t unit_begin;
    uint32_t nrhs;
    const A* x;               // column-major n x nrhs
    A* y;                     // column-major m x nrhs
    uint64_t ldx, ldy;
    A* xg;                    // nrhs > 1: x interleaved in groups of gw right-hand sides
    uint32_t gw;              // group width: 4 or 8 (kRhsGroup)
    uint32_t xg_block0;       // the plan's first block of the interleave (batch_prologue)
    A* partial;               // nrhs x partial_stride
    uint32_t* counters;       // nrhs x counter_stride (zero between launches)
    uint64_t partial_stride, counter_stride;
};

// The work units (plan, tile) of XCD k in a batched launch:
// [begin[k], begin[k + 1]).
struct XcdSplit {
    uint32_t nq;                   // queues: the device's XCD count (1..kMaxXcds)
    uint32_t begin[kMaxXcds + 1];
};

// Interleaved columns (groups * n) per block of batch_prologue.
constexpr uint32_t kPrologueCols = 1024;

// Runs before every launch_batch, on the same stream: zeroes the kMaxXcds unit
// queue counters and fills BatchPlan::xg of every plan with nrhs > 1 from
// its x.  block_plan[b] (device memory, nblocks entries, 0 if no plan has
// several right-hand sides) is the plan of block b; plan p's blocks start
// at p.xg_block0 and cover its groups * n columns kPrologueCols at a time.
hipError_t launch_prologue(const BatchPlan<double, double>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s);
hipError_t launch_prologue(const BatchPlan<float, float>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s);
hipError_t launch_prologue(const BatchPlan<dev::bf16_t, float>* p, const uint32_t* block_plan, uint32_t nblocks,
                           uint32_t* queue, hipStream_t s);

// Resident workgroups per CU of the batched kernel for this shape (0 = unknown).
int batch_blocks_per_cu(int dtype, bool comp, int threads, size_t lds);
// grid: a multiple of kXcds
hipError_t launch_batch(const BatchPlan<double, double>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s);
hipError_t launch_batch(const BatchPlan<float, float>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s);
hipError_t launch_batch(const BatchPlan<dev::bf16_t, float>* p, uint32_t np, const XcdSplit& split, uint32_t* queue,
                        int grid, int threads, size_t lds, bool comp, hipStream_t s);

// y[r] += the panel sums of row r, for every row with panel entries (after
// a stream launch of a plan with column panels, on the same stream).
hipError_t launch_panel_sums(double* y, const double* panel_out, const uint8_t* mask, uint32_t m, uint32_t panels,
                             hipStream_t s);
hipError_t launch_panel_sums(float* y, const float* panel_out, const uint8_t* mask, uint32_t m, uint32_t panels,
                             hipStream_t s);

// sddmm: also write dvals (a.xrow, a.dvals set)
hipError_t launch_stream(const StreamArgs<double, double>& a, int grid, bool sddmm, hipStream_t s);
hipError_t launch_stream(const StreamArgs<float, float>& a, int grid, bool comp, bool sddmm, hipStream_t s);
hipError_t launch_stream(const StreamArgs<dev::bf16_t, float>& a, int grid, bool comp, bool sddmm, hipStream_t s);

}  // namespace cbs
