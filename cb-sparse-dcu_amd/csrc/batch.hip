// Batched persistent launch, its hipGraph capture, and the autotuner (no
// reference counterpart: the reference runs one matrix per process, with
// block size and waves fixed at compile time, cb-spmv-dcu/src/macros.h:38-82).
//
// A batch is a set of uploaded plans of one value type.  One launch of
// stream_batch covers every (plan, tile, right-hand side): the work units
// (plan, tile) are dealt to a grid sized to the resident capacity of the
// GPU, and each unit runs its tile once per right-hand side.  The per-plan
// descriptor (stream arguments, vector pointers, partial-sum buffers per
// right-hand side) lives in device memory; it is re-uploaded only when the
// vector pointers or RHS counts change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "cb_runtime.hpp"
#include "kernels.hpp"
#include "stream.hpp"

namespace cbs {
constexpr int kBatchMaxRhs = 16;
// autotuner: copies of a candidate per timed batch (work units to fill)
constexpr uint32_t kAutotuneUnits = 2048;
constexpr uint32_t kAutotuneMaxCopies = 64;
constexpr size_t kAutotuneYBytes = size_t(1) << 30;   // bound on the copies' y
// autotuner: batch shape when the base options leave it open
constexpr int kBatchTileWaves = 8;        // stream tile waves
constexpr int kBatchItemEntries = 2048;   // entries per item
constexpr int kBatchTileKib = 32;         // LDS x cache, one right-hand side
constexpr int kBatchMultiTileKib = 4;     // LDS x cache, several right-hand sides
// autotuner: a later candidate replaces the one kept only when it is this
// much faster.  A candidate is timed against copies of itself, which
// predicts its cost inside a mixed batch only roughly: on a batch of 256
// matrices (1e3 to 2e5 rows, 258M entries in all), when candidates still
// chose the stream tile waves, keeping
// every one that timed faster at all picked 9 waves for 67 matrices and
// made the batch slower (1.63 ms) than 16 waves for all of them (1.56 ms).
constexpr double kAutotuneMargin = 0.05;
}  // namespace cbs

struct cbspmv_batch {
    std::vector<cbspmv_plan*> plans;
    int dtype = CBSPMV_F64;
    bool comp = false;
    uint32_t units = 0;
    int xcds = 8;     // XCDs of the device the batch was created on (queues)
    int threads = 64;
    size_t lds = 0;
    int grid = 8;
    cbs::DeviceBuf desc, partial, counters;
    std::vector<uint8_t> host_desc;   // last uploaded descriptor bytes
    // stream entries per tile of every plan (the unit cost estimate), and
    // the per-XCD unit queue counters (zeroed by the prologue kernel)
    std::vector<std::vector<uint32_t>> tile_entries;
    cbs::DeviceBuf queue;
    cbs::XcdSplit split{}, graph_split{};
    // x interleaved per group of right-hand sides (plans with several) and
    // the interleave's block count, for eager calls
    cbs::DeviceBuf xg;
    uint32_t xg_blocks = 0;
    // captured graph, the descriptor and interleave buffer it reads
    cbs::DeviceBuf graph_desc, graph_xg;
    uint32_t graph_xg_blocks = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace cbs {
namespace {

template <typename V, typename A>
void fill_args(const cbspmv_plan& p, StreamArgs<V, A>& a) {
    const Plan& h = p.host;
    const DevicePlan& d = p.dev;
    a = StreamArgs<V, A>{};
    a.slab = static_cast<const uint8_t*>(d.key.ptr);
    a.items = static_cast<const StreamItem*>(d.items.ptr);
    a.item_split = static_cast<const uint32_t*>(d.item_split.ptr);
    a.item_split_first = static_cast<const uint32_t*>(d.item_split_first.ptr);
    a.splits = static_cast<const StreamSplit*>(d.splits.ptr);
    a.tiles = static_cast<const uint32_t*>(d.tiles.ptr);
    a.tile_lines = static_cast<const uint32_t*>(d.tile_lines.ptr);
    a.empty_rows = static_cast<const uint32_t*>(d.empty.ptr);
    a.n_empty_rows = (uint32_t)h.empty_rows.size();
    a.n_items = h.tile_items_count;
    a.n_tiles = (uint32_t)(h.tiles.size() / 2);
    a.tile_items = (uint32_t)h.o.tile_items;
    a.tile_slots = (uint32_t)tile_slots(h.o.tile_bytes, h.o.dtype);
    a.col_bits = h.col_bits;
    a.m = h.m;
    a.n = h.n;
}

// Group width of a plan with several right-hand sides: 4 when they fit one
// group of 4 (16 bytes per FP32 column, so half the cache lines a group of
// 8 would touch), else 8.  Splitting 5-8 right-hand sides of the largest
// plans into two groups of 4, so that one group's interleaved x stays in
// an XCD's L2, measured slower on the 256-matrix batch (2.13 against
// 2.05 ms): the second pass over their slab costs more than the gathers
// it takes off the fabric.
uint32_t group_width(const Plan& h, int nrhs) {
    (void)h;
    return nrhs <= kRhsGroupNarrow ? (uint32_t)kRhsGroupNarrow : (uint32_t)kRhsGroup;
}

// Estimated cost of one work unit, in stream entries: a tile with one
// right-hand side streams its entries once with its LDS x cache; with
// several, once per group of gw, gathering every entry's x values (about
// 1.5x the one-RHS cost per entry); plus a fixed cost per unit.
double unit_cost(uint32_t entries, int nrhs, uint32_t gw) {
    constexpr double kUnitFixed = 2048.0, kMultiWeight = 1.5;
    if (nrhs <= 0) return kUnitFixed;
    if (nrhs == 1) return kUnitFixed + entries;
    return kUnitFixed + kMultiWeight * entries * (double)(((uint32_t)nrhs + gw - 1) / gw);
}

// Contiguous unit ranges of (about) equal estimated cost, one per XCD of the
// current device.
XcdSplit split_units(const cbspmv_batch& b, const int* nrhs) {
    std::vector<double> cost;
    cost.reserve(b.units);
    for (size_t i = 0; i < b.plans.size(); ++i) {
        const std::vector<uint32_t>& te = b.tile_entries[i];
        const uint32_t gw = group_width(b.plans[i]->host, nrhs[i]);
        if (te.empty()) cost.push_back(unit_cost(0, nrhs[i], gw));
        for (uint32_t e : te) cost.push_back(unit_cost(e, nrhs[i], gw));
    }
    double total = 0.0;
    for (double c : cost) total += c;
    XcdSplit sp{};
    const uint32_t nq = (uint32_t)b.xcds;
    sp.nq = nq;
    double acc = 0.0;
    uint32_t u = 0, k = 1;
    for (; u < (uint32_t)cost.size() && k < nq; ++u) {
        acc += cost[u];
        while (k < nq && acc >= total * k / nq) sp.begin[k++] = u + 1;
    }
    for (; k <= nq; ++k) sp.begin[k] = (uint32_t)cost.size();
    return sp;
}

// Descriptor bytes for one call: the BatchPlan array (x/y pointers, RHS
// counts), then the interleave's block -> plan table (xg_blocks entries).
// Plans with several right-hand sides get a slice of xg (grown here when
// too small) for their interleaved x.
template <typename V, typename A>
int make_desc(cbspmv_batch& b, const void* const* x, void* const* y, const int* nrhs, DeviceBuf& xg,
              uint32_t& xg_blocks, std::vector<uint8_t>& out, hipStream_t s) {
    const size_t np = b.plans.size();
    size_t xg_elems = 0;
    uint64_t cols_sum = 0;
    for (size_t i = 0; i < np; ++i) {
        const Plan& h = b.plans[i]->host;
        if (nrhs[i] < 0 || nrhs[i] > kBatchMaxRhs) return fail(CBSPMV_ERR_ARG, "nrhs must be in 0..16");
        if (nrhs[i] > 0 && ((h.m > 0 && !y[i]) || (h.nnz > 0 && !x[i]))) return fail(CBSPMV_ERR_ARG, "null vector");
        if (nrhs[i] > 1) {
            const uint32_t gw = group_width(h, nrhs[i]);
            const uint64_t cols = ((uint64_t)nrhs[i] + gw - 1) / gw * (uint64_t)h.n;
            xg_elems += (size_t)cols * gw;
            cols_sum += cols;
        }
    }
    if (cols_sum >= (1ull << 31))
        return fail(CBSPMV_ERR_UNSUPPORTED, "batch: sum over plans of n x (right-hand-side groups) must stay below 2^31");
    std::vector<uint32_t> block_plan;
    if (xg.bytes < xg_elems * sizeof(A)) {
        const int rc = upload_bytes(xg, nullptr, xg_elems * sizeof(A), s);
        if (rc) return rc;
    }
    std::vector<BatchPlan<V, A>> d(np);
    A* part = static_cast<A*>(b.partial.ptr);
    uint32_t* cnt = static_cast<uint32_t*>(b.counters.ptr);
    size_t xg_off = 0;   // elements of xg handed out so far
    uint32_t unit = 0;
    for (size_t i = 0; i < np; ++i) {
        const cbspmv_plan& p = *b.plans[i];
        const Plan& h = p.host;
        BatchPlan<V, A>& e = d[i];
        fill_args(p, e.args);
        e.unit_begin = unit;
        unit += std::max<uint32_t>(1u, e.args.n_tiles);
        e.nrhs = (uint32_t)nrhs[i];
        e.x = static_cast<const A*>(x[i]);
        e.y = static_cast<A*>(y[i]);
        e.ldx = (uint64_t)h.n;
        e.ldy = (uint64_t)h.m;
        const size_t parts = b.comp ? 2 : 1;   // (hi, lo) per row piece
        e.partial_stride = (uint64_t)h.partial_slots * parts;
        e.counter_stride = h.splits.size();
        e.partial = part;
        e.counters = cnt;
        e.xg = nullptr;
        e.gw = group_width(h, nrhs[i]);
        e.xg_block0 = (uint32_t)block_plan.size();
        if (nrhs[i] > 1) {
            const size_t cols = (size_t)((uint32_t)nrhs[i] + e.gw - 1) / e.gw * (size_t)h.n;
            e.xg = static_cast<A*>(xg.ptr) + xg_off;
            xg_off += cols * e.gw;
            block_plan.insert(block_plan.end(), (cols + kPrologueCols - 1) / kPrologueCols, (uint32_t)i);
        }
        part += (size_t)h.partial_slots * parts * kBatchMaxRhs;
        cnt += h.splits.size() * kBatchMaxRhs;
    }
    xg_blocks = (uint32_t)block_plan.size();
    out.resize(np * sizeof(BatchPlan<V, A>) + block_plan.size() * sizeof(uint32_t));
    if (np) std::memcpy(out.data(), d.data(), np * sizeof(BatchPlan<V, A>));
    if (!block_plan.empty())
        std::memcpy(out.data() + np * sizeof(BatchPlan<V, A>), block_plan.data(), block_plan.size() * sizeof(uint32_t));
    return CBSPMV_OK;
}

int build_desc(cbspmv_batch& b, const void* const* x, void* const* y, const int* nrhs, DeviceBuf& xg,
               uint32_t& xg_blocks, XcdSplit& split, std::vector<uint8_t>& out, hipStream_t s) {
    int rc;
    if (b.dtype == CBSPMV_F64) rc = make_desc<double, double>(b, x, y, nrhs, xg, xg_blocks, out, s);
    else if (b.dtype == CBSPMV_F32) rc = make_desc<float, float>(b, x, y, nrhs, xg, xg_blocks, out, s);
    else rc = make_desc<dev::bf16_t, float>(b, x, y, nrhs, xg, xg_blocks, out, s);
    if (rc == CBSPMV_OK) split = split_units(b, nrhs);
    return rc;
}

int upload_desc(DeviceBuf& buf, const std::vector<uint8_t>& bytes, hipStream_t s) {
    if (buf.bytes < bytes.size()) {
        const int rc = upload_bytes(buf, nullptr, bytes.size(), s);
        if (rc) return rc;
    }
    if (!bytes.empty()) {
        CBS_HIP(hipMemcpyAsync(buf.ptr, bytes.data(), bytes.size(), hipMemcpyHostToDevice, s));
        CBS_HIP(hipStreamSynchronize(s));   // the host bytes may change after return
    }
    return CBSPMV_OK;
}

// The prologue (queue reset, x interleave), then the batched launch, on
// stream s.
int launch(cbspmv_batch& b, const void* desc, uint32_t xg_blocks, const XcdSplit& split, hipStream_t s) {
    if (b.plans.empty()) return CBSPMV_OK;
    hipError_t e;
    const uint32_t np = (uint32_t)b.plans.size();
    uint32_t* q = static_cast<uint32_t*>(b.queue.ptr);
    // the block -> plan table follows the np BatchPlan entries
    auto table = [&](size_t plan_bytes) {
        return reinterpret_cast<const uint32_t*>(static_cast<const uint8_t*>(desc) + np * plan_bytes);
    };
    if (b.dtype == CBSPMV_F64) {
        const auto* d = static_cast<const BatchPlan<double, double>*>(desc);
        e = launch_prologue(d, table(sizeof(*d)), xg_blocks, q, s);
        if (e == hipSuccess) e = launch_batch(d, np, split, q, b.grid, b.threads, b.lds, false, s);
    } else if (b.dtype == CBSPMV_F32) {
        const auto* d = static_cast<const BatchPlan<float, float>*>(desc);
        e = launch_prologue(d, table(sizeof(*d)), xg_blocks, q, s);
        if (e == hipSuccess) e = launch_batch(d, np, split, q, b.grid, b.threads, b.lds, b.comp, s);
    } else {
        const auto* d = static_cast<const BatchPlan<dev::bf16_t, float>*>(desc);
        e = launch_prologue(d, table(sizeof(*d)), xg_blocks, q, s);
        if (e == hipSuccess) e = launch_batch(d, np, split, q, b.grid, b.threads, b.lds, b.comp, s);
    }
    return e == hipSuccess ? CBSPMV_OK : hip_fail(e, "batched launch");
}

void drop_graph(cbspmv_batch& b) {
    if (b.exec) (void)hipGraphExecDestroy(b.exec);
    if (b.graph) (void)hipGraphDestroy(b.graph);
    b.exec = nullptr;
    b.graph = nullptr;
}

}  // namespace
}  // namespace cbs

using cbs::fail;

extern "C" {

int cbspmv_batch_create(cbspmv_plan* const* plans, int count, cbspmv_batch** batch, void* stream) {
    if (!batch || count < 0 || (count > 0 && !plans)) return fail(CBSPMV_ERR_ARG, "null argument");
    *batch = nullptr;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cbspmv_batch* b = new (std::nothrow) cbspmv_batch();
    if (!b) return fail(CBSPMV_ERR_NOMEM, "out of host memory");
    size_t partial_elems = 0, counter_elems = 0;
    int max_items = 1, max_slots = 0;
    for (int i = 0; i < count; ++i) {
        cbspmv_plan* p = plans[i];
        if (!p) { delete b; return fail(CBSPMV_ERR_ARG, "null plan in batch"); }
        const cbs::Plan& h = p->host;
        if (i == 0) { b->dtype = h.o.dtype; b->comp = h.o.accum == CBSPMV_ACC_COMPENSATED; }
        if (h.o.dtype != b->dtype || (h.o.accum == CBSPMV_ACC_COMPENSATED) != b->comp) {
            delete b;
            return fail(CBSPMV_ERR_ARG, "plans of one batch must share value type and accumulation");
        }
        if (h.panels > 1) {
            delete b;
            return fail(CBSPMV_ERR_UNSUPPORTED, "plans of a batch need stream_panels = 1");
        }
        if (!cbs::stream_vectors_fit(h)) {
            delete b;
            return fail(CBSPMV_ERR_UNSUPPORTED, "the stream schedule needs x and y below 2 GiB each");
        }
        const int rc = cbs::upload_plan(*p, s);
        if (rc) { delete b; return rc; }
        b->plans.push_back(p);
        b->units += std::max<uint32_t>(1u, (uint32_t)(h.tiles.size() / 2));
        {
            std::vector<uint32_t> te(h.tiles.size() / 2, 0u);
            const size_t ti = (size_t)h.o.tile_items;
            for (size_t it = 0; it < (size_t)h.tile_items_count; ++it)
                te[it / ti] += h.items[it].e_end - h.items[it].e_begin;
            b->tile_entries.push_back(std::move(te));
        }
        partial_elems += (size_t)h.partial_slots * cbs::kBatchMaxRhs * 2;   // room for (hi, lo)
        counter_elems += h.splits.size() * cbs::kBatchMaxRhs;
        // waves per workgroup: a tile with more items than waves deals them round-robin
        max_items = std::max(max_items, std::min(h.o.tile_items, cbs::kMaxBatchTileItems));
        max_slots = std::max(max_slots, cbs::tile_slots(h.o.tile_bytes, h.o.dtype));
    }
    const size_t eb = b->dtype == CBSPMV_F64 ? 8 : 4;
    b->threads = 64 * max_items;
    b->lds = (size_t)max_slots * eb;
    // a persistent grid: as many workgroups as stay resident at once
    int per_cu = cbs::batch_blocks_per_cu(b->dtype, b->comp, b->threads, b->lds);
    if (per_cu <= 0) {   // occupancy query unavailable: bound by LDS and 32 waves per CU
        const int by_lds = b->lds ? (int)std::max<size_t>(1, (size_t)(160 * 1024) / b->lds) : 8;
        per_cu = std::max(1, std::min(by_lds, 32 / max_items));
    }
    const cbs::DeviceShape shape = cbs::device_shape();
    b->xcds = shape.xcds;
    const uint32_t cap = (uint32_t)(shape.cus * per_cu);
    b->grid = (int)((std::min<uint32_t>(std::max<uint32_t>(b->units, 1u), cap) + 7u) / 8u * 8u);
    int rc;
    if ((rc = cbs::upload_bytes(b->queue, nullptr, cbs::kMaxXcds * sizeof(uint32_t), s)) ||
        (rc = cbs::upload_bytes(b->partial, nullptr, std::max<size_t>(partial_elems, 1) * eb, s)) ||
        (rc = cbs::upload_bytes(b->counters, nullptr, std::max<size_t>(counter_elems, 1) * 4, s))) {
        cbspmv_batch_destroy(b);
        return rc;
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        cbspmv_batch_destroy(b);
        return fail(CBSPMV_ERR_HIP, "batch create: stream synchronize");
    }
    *batch = b;
    return CBSPMV_OK;
}

int cbspmv_batch_spmv(cbspmv_batch* batch, const void* const* d_x, void* const* d_y, const int* nrhs,
                      void* stream) {
    if (!batch) return fail(CBSPMV_ERR_ARG, "null batch");
    if (batch->plans.empty()) return CBSPMV_OK;
    if (!d_x || !d_y || !nrhs) return fail(CBSPMV_ERR_ARG, "null argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    std::vector<uint8_t> bytes;
    int rc = cbs::build_desc(*batch, d_x, d_y, nrhs, batch->xg, batch->xg_blocks, batch->split, bytes, s);
    if (rc) return rc;
    if (bytes != batch->host_desc) {
        if ((rc = cbs::upload_desc(batch->desc, bytes, s))) return rc;
        batch->host_desc.swap(bytes);
    }
    return cbs::launch(*batch, batch->desc.ptr, batch->xg_blocks, batch->split, s);
}

int cbspmv_batch_capture(cbspmv_batch* batch, const void* const* d_x, void* const* d_y, const int* nrhs,
                         void* stream) {
    if (!batch) return fail(CBSPMV_ERR_ARG, "null batch");
    if (!stream) return fail(CBSPMV_ERR_ARG, "graph capture needs a created (non-null) stream");
    if (!d_x || !d_y || !nrhs) return fail(CBSPMV_ERR_ARG, "null argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    cbs::drop_graph(*batch);   // before graph_xg may be reallocated
    std::vector<uint8_t> bytes;
    int rc = cbs::build_desc(*batch, d_x, d_y, nrhs, batch->graph_xg, batch->graph_xg_blocks, batch->graph_split, bytes,
                             s);
    if (rc) return rc;
    if ((rc = cbs::upload_desc(batch->graph_desc, bytes, s))) return rc;
    CBS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = cbs::launch(*batch, batch->graph_desc.ptr, batch->graph_xg_blocks, batch->graph_split, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return cbs::hip_fail(e, "hipStreamEndCapture");
    batch->graph = g;
    CBS_HIP(hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0));
    CBS_HIP(hipGraphLaunch(batch->exec, s));
    return CBSPMV_OK;
}

int cbspmv_batch_replay(cbspmv_batch* batch, void* stream) {
    if (!batch) return fail(CBSPMV_ERR_ARG, "null batch");
    if (!batch->exec) return fail(CBSPMV_ERR_STATE, "nothing captured (call cbspmv_batch_capture first)");
    CBS_HIP(hipGraphLaunch(batch->exec, static_cast<hipStream_t>(stream)));
    return CBSPMV_OK;
}

void cbspmv_batch_destroy(cbspmv_batch* batch) {
    if (!batch) return;
    cbs::drop_graph(*batch);
    for (cbs::DeviceBuf* d : {&batch->desc, &batch->partial, &batch->counters, &batch->graph_desc, &batch->xg,
                              &batch->graph_xg, &batch->queue})
        cbs::release(*d);
    delete batch;
}

int cbspmv_autotune(int m, int n, int nnz, const int* row, const int* col, const double* val,
                    const cbspmv_opts* base, const int* candidates, int ncandidates, int nrhs,
                    cbspmv_plan** best, double* best_ms, void* stream) {
    if (!best) return fail(CBSPMV_ERR_ARG, "null output");
    *best = nullptr;
    if (nrhs < 1 || nrhs > cbs::kBatchMaxRhs) return fail(CBSPMV_ERR_ARG, "nrhs must be in 1..16");
    static const int kDefault[] = {16, 0, 32, 0, 16, 4, 16, 9, 32, 4, 32, 9};
    if (!candidates) { candidates = kDefault; ncandidates = 6; }
    if (ncandidates < 1) return fail(CBSPMV_ERR_ARG, "no candidates");
    cbspmv_opts o;
    if (base) o = *base;
    else cbspmv_default_opts(&o);
    // the plans run in batches, which read the full stream: no scattered-entry split
    o.stream_split = CBSPMV_SPLIT_OFF;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t eb = o.value_dtype == CBSPMV_F64 ? 8 : 4;
    // one x and y block for all candidates, filled with a fixed pattern
    void* dx = nullptr;
    void* dy = nullptr;
    CBS_HIP(hipMalloc(&dx, std::max<size_t>(1, (size_t)n * nrhs * eb)));
    if (hipMalloc(&dy, std::max<size_t>(1, (size_t)m * nrhs * eb)) != hipSuccess) {
        (void)hipFree(dx);
        return fail(CBSPMV_ERR_NOMEM, "autotune: hipMalloc");
    }
    {
        std::vector<uint8_t> hx((size_t)n * nrhs * eb);
        std::mt19937 gen(12345);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        for (size_t i = 0; i < (size_t)n * nrhs; ++i) {
            if (eb == 8) { const double v = u(gen); std::memcpy(&hx[i * 8], &v, 8); }
            else { const float v = (float)u(gen); std::memcpy(&hx[i * 4], &v, 4); }
        }
        if (!hx.empty()) (void)hipMemcpy(dx, hx.data(), hx.size(), hipMemcpyHostToDevice);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int rc = CBSPMV_OK;
    double best_t = 1e300;
    // Times an uploaded plan the way a batch runs it: K copies of the plan
    // in one batch (K sized so that the copies' work units fill the GPU
    // several times over), costed at 1/K of that launch, i.e. its throughput
    // beside other matrices rather than its latency alone on the GPU.  Plans
    // a batch does not take: one launch per right-hand side.
    auto time_plan = [&](cbspmv_plan* p) -> double {
        const uint32_t units = std::max<uint32_t>(1u, (uint32_t)(p->host.tiles.size() / 2));
        const size_t y_bytes = std::max<size_t>(1, (size_t)m * nrhs * eb);
        const int K = (int)std::max<size_t>(
            1, std::min<size_t>({(size_t)cbs::kAutotuneMaxCopies, (cbs::kAutotuneUnits + units - 1) / units,
                                 cbs::kAutotuneYBytes / y_bytes}));
        cbspmv_batch* one = nullptr;
        void* dyk = nullptr;
        std::vector<const void*> bx((size_t)K, dx);
        std::vector<void*> by((size_t)K, nullptr);
        std::vector<int> bn((size_t)K, nrhs);
        if (hipMalloc(&dyk, (size_t)K * y_bytes) == hipSuccess) {
            for (int k = 0; k < K; ++k) by[k] = static_cast<uint8_t*>(dyk) + (size_t)k * y_bytes;
            std::vector<cbspmv_plan*> copies((size_t)K, p);
            if (cbspmv_batch_create(copies.data(), K, &one, stream) != CBSPMV_OK) one = nullptr;
        }
        const double per_run = one ? (double)K : 1.0;
        auto run = [&]() {
            if (one) {
                if (rc == CBSPMV_OK) rc = cbspmv_batch_spmv(one, bx.data(), by.data(), bn.data(), stream);
                return;
            }
            for (int r = 0; r < nrhs && rc == CBSPMV_OK; ++r)
                rc = cbs::run_spmv(*p, static_cast<uint8_t*>(dx) + (size_t)r * n * eb,
                                   static_cast<uint8_t*>(dy) + (size_t)r * m * eb, s);
        };
        for (int w = 0; w < 3; ++w) run();
        constexpr int kReps = 20;
        (void)hipEventRecord(e0, s);
        for (int k = 0; k < kReps; ++k) run();
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (one) cbspmv_batch_destroy(one);
        if (dyk) (void)hipFree(dyk);
        return ms / kReps / per_run;
    };
    // Builds, times and keeps or destroys one candidate; `timed` lists the
    // stream plan shapes timed so far.
    auto try_candidate = [&](const cbspmv_opts& oc, std::vector<cbs::StreamPlanShape>& timed) {
        cbspmv_plan* p = nullptr;
        rc = cbspmv_build(m, n, nnz, row, col, val, &oc, &p);
        if (rc) return;
        if (p->schedule == CBSPMV_SCHED_STREAM) {
            const cbs::StreamPlanShape sh = cbs::stream_plan_shape(p->host);
            if (!sh.dense) {
                if (std::find(timed.begin(), timed.end(), sh) != timed.end()) {
                    cbspmv_destroy(p);
                    return;
                }
                timed.push_back(sh);
            }
        }
        if ((rc = cbs::upload_plan(*p, s))) {
            cbspmv_destroy(p);
            return;
        }
        const double t = time_plan(p);
        if (rc == CBSPMV_OK && t < (*best ? best_t * (1.0 - cbs::kAutotuneMargin) : best_t)) {
            if (*best) cbspmv_destroy(*best);
            *best = p;
            best_t = t;
        } else {
            cbspmv_destroy(p);
        }
    };
    // The stream schedule reads the strip entries, not the blocks: without
    // DENSE blocks (whose zero fill depends on the block size) two block
    // sizes give the same entries, item cuts and staged lines, so a
    // candidate whose stream plan has the shape of one already timed is
    // skipped (the first one stays).
    std::vector<cbs::StreamPlanShape> timed;
    for (int c = 0; c < ncandidates && rc == CBSPMV_OK; ++c) {
        cbspmv_opts oc = o;
        oc.block_size = candidates[2 * c];
        if (candidates[2 * c + 1] > 0) oc.waves_per_block = candidates[2 * c + 1];
        // Batch shape, where `base` leaves it open.  The batched launch runs
        // every unit with the widest tile of the batch and as much LDS as the
        // largest cache, so the stream tile shape is one choice for all of
        // its plans, not a per-matrix one (a 9- or 16-wave plan among 8-wave
        // ones leaves waves idle or halves the workgroups per CU):
        //  * 8 tile waves with a cache of at most 32 KiB, so that two
        //    workgroups share a CU and a unit's slowest wave holds up only
        //    half of it (a batch of 256 matrices, 258M entries: 1.557 -> 1.466 ms against
        //    16 waves; 4 waves 1.475, 6 1.541, 10 1.685 ms);
        //  * with several right-hand sides x is gathered, not staged, so the
        //    cache only decodes staged keys: 4 KiB; with one, 32 KiB (4 to
        //    32 KiB measured alike on that batch);
        //  * items of kBatchItemEntries entries: the automatic size fills the
        //    GPU with one matrix's tiles (down to 1024 entries per item, up
        //    to whatever one round of tiles needs), short items pay their
        //    start-up for little work and long ones leave a unit's waves
        //    waiting for its slowest (1024: 1.521, 1536: 1.481, 2048:
        //    1.466, 2560: 1.500 ms; 2048 or the one-round size, if larger:
        //    1.515 ms).  A matrix with more than 8 x 2048 entries per CU
        //    then has more tiles than a launch of its own would run at once,
        //    which a batch's unit queues absorb.
        if (o.stream_tile_waves == 0) oc.stream_tile_waves = cbs::kBatchTileWaves;
        if (o.stream_tile_kib == 0) oc.stream_tile_kib = nrhs > 1 ? cbs::kBatchMultiTileKib : cbs::kBatchTileKib;
        if (o.stream_item_nnz == 0) oc.stream_item_nnz = cbs::kBatchItemEntries;
        try_candidate(oc, timed);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dx);
    (void)hipFree(dy);
    if (rc) {
        if (*best) cbspmv_destroy(*best);
        *best = nullptr;
        return rc;
    }
    if (best_ms) *best_ms = best_t;
    return CBSPMV_OK;
}

}  // extern "C"
