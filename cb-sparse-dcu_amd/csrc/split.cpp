// Scattered-entry split of a stream plan (host half): when a plan splits,
// the stream plan of the other entries, and dropping the split.  Which
// entries are scattered and their sweep layout: split_layout.cpp; the
// kernel and its launch: sweep.hip.
//
// The stream schedule serves an entry's x from its tile's LDS cache when the
// tile reads the entry's x line often enough to earn one of the cache's
// lines; an entry whose line the tile reads only a few times is a random
// gather: one vector-L1 miss, and one 128-byte L2 miss once x outgrows one
// XCD's 4 MB L2 (every tile of an XCD draws its own random lines).  The
// split runs those entries in a second kernel after the stream plan of the
// others.  There every wave walks its rows' scattered entries in column
// order, so the waves of an XCD sweep x from its first column to its last
// together and fetch each x line into the XCD's L2 about once for all of
// them.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"

namespace cbs {
namespace {

// Automatic split: x larger than one XCD's L2 (below that the gathers hit
// L2 anyway and the split only adds a launch), at most 16 MiB (the largest
// x measured), and enough scattered entries for a second launch to pay.
// Measured (per SpMV, unsplit -> split): the 1,048,576^2 stencil /
// power-law matrix, FP64 (x 8 MiB, 7.7M of 30.2M entries scattered), 136.1
// -> 126.7 us; the 2,097,152^2 block-clustered matrix, FP64 (x 16 MiB, 8.2M
// of 50M), 218.3 -> 196.2 us.
constexpr size_t kSplitMinXBytes = 4u << 20;
constexpr size_t kSplitMaxXBytes = 16u << 20;
constexpr int64_t kSplitMinEntries = 1 << 20;
constexpr double kSplitMinShare = 0.15;
}  // namespace

int build_split(cbspmv_plan& p, int nnz, const int* row, const int* col, const double* val) {
    const Plan& h = p.host;
    const int mode = p.opts.stream_split;
    if (mode == CBSPMV_SPLIT_OFF || nnz == 0 || h.m == 0) return CBSPMV_OK;
    if (h.o.accum != CBSPMV_ACC_NATIVE || h.o.panels > 1 ||
        p.schedule != CBSPMV_SCHED_STREAM)
        return CBSPMV_OK;
    const size_t xb = vec_elem_bytes(h);
    if (mode == CBSPMV_SPLIT_AUTO && ((size_t)h.n * xb <= kSplitMinXBytes || (size_t)h.n * xb > kSplitMaxXBytes))
        return CBSPMV_OK;
    int cb = 1;
    while ((1ll << cb) < (long long)h.n) ++cb;
    const int max_rows = cb >= 26 ? 0 : (int)std::min<long long>(kSweepMaxRows, 1ll << (32 - cb));
    if (max_rows < 64) return CBSPMV_OK;   // too few key bits beside the column
    const std::vector<uint8_t> far = scattered_entries(nnz, row, col, h.o.tiles, (uint32_t)(128 / xb));
    std::vector<uint32_t> fe;
    for (int e = 0; e < nnz; ++e)
        if (far[e]) fe.push_back((uint32_t)e);
    const int64_t nf = (int64_t)fe.size();
    if (nf == 0) return CBSPMV_OK;
    if (mode == CBSPMV_SPLIT_AUTO && (nf < kSplitMinEntries || (double)nf < kSplitMinShare * nnz)) return CBSPMV_OK;
    std::vector<int> nr, nc;
    std::vector<double> nv;
    nr.reserve((size_t)(nnz - nf));
    nc.reserve((size_t)(nnz - nf));
    nv.reserve((size_t)(nnz - nf));
    for (int e = 0; e < nnz; ++e)
        if (!far[e]) { nr.push_back(row[e]); nc.push_back(col[e]); nv.push_back(val[e]); }
    cbspmv_opts o = p.opts;
    o.stream_split = CBSPMV_SPLIT_OFF;
    cbspmv_plan* near = nullptr;
    const int rc = build_plan_object(o, h.m, h.n, (int)nr.size(), nr.data(), nc.data(), nv.data(), false, &near);
    if (rc != CBSPMV_OK) return rc;
    near->paced = true;
    p.near = near;
    p.sweep = sweep_layout(fe, row, col, val, cb, max_rows);
    return CBSPMV_OK;
}

void drop_split(cbspmv_plan& p) {
    if (p.near) {
        cbspmv_destroy(p.near);
        p.near = nullptr;
    }
    p.sweep = SweepPart{};
    for (DeviceBuf* b : {&p.dev.sw_key, &p.dev.sw_val, &p.dev.sw_wave_off, &p.dev.sw_blk_row, &p.dev.sw_blk_span})
        release(*b);
}

}  // namespace cbs
