// Scattered-entry split of a stream plan (host half): which entries are
// scattered, the stream plan of the others, and the sweep kernel's layout.
// The kernel and its launch are in sweep.hip.
//
// The stream schedule serves an entry's x from its tile's LDS cache when the
// tile reads the entry's x line often enough to earn one of the cache's
// lines; an entry whose line the tile reads only a few times is a random
// gather: one vector-L1 miss, and one 128-byte L2 miss once x outgrows one
// XCD's 4 MB L2 (every tile of an XCD draws its own random lines).  The
// split runs those entries in a second kernel after the stream plan of the
// others.  There every wave walks its
// rows' scattered entries in column order, so the waves of an XCD sweep x
// from its first column to its last together and fetch each x line into the
// XCD's L2 about once for all of them.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <utility>
#include <vector>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"

namespace cbs {
namespace {

// Automatic split: x larger than one XCD's L2 but at most twice it (a
// sweep's window of x has to stay in L2 while the waves drift apart), and
// enough scattered entries for a second launch to pay.  Measured (per SpMV,
// unsplit -> split): the 1,048,576^2 stencil / power-law matrix, FP64 (x 8
// MiB, 7.7M of 30.2M entries scattered), 135.9 -> 127.3 us; the 2,097,152^2
// block-clustered matrix, FP32 (x 8 MiB, 8.1M of 50M), 154.0 -> 153.6 us,
// and FP64 (x 16 MiB) 217.8 -> 233.2 us, which the x bound leaves unsplit.
constexpr size_t kSplitMinXBytes = 4u << 20;
constexpr size_t kSplitMaxXBytes = 8u << 20;
constexpr int64_t kSplitMinEntries = 1 << 20;
constexpr double kSplitMinShare = 0.15;
// Entries of one sweep block at most (a row with more is a block alone).
constexpr int64_t kSweepBlockEntries = 8000;
// A line read this many times or fewer in a group is a scattered line.
// Uniformly scattered columns meet in a line by chance: on the 1,048,576^2
// stencil / power-law matrix ~42% of them share their line with another one
// in their group, and almost none with three others; a line the stream
// stages (the diagonal band) is read hundreds of times.
constexpr size_t kScatteredReads = 3;

// 1 for the scattered entries of a row-major COO: rows are cut into groups
// of about nnz / tiles entries at row boundaries (the stream schedule's
// tiles), and an entry is scattered when at most kScatteredReads entries of
// its group read its x line (line_elems consecutive columns).
std::vector<uint8_t> classify(int nnz, const int* row, const int* col, int tiles, uint32_t line_elems) {
    std::vector<uint8_t> far((size_t)nnz, 0);
    if (nnz == 0) return far;
    const int64_t target = std::max<int64_t>(1, (int64_t)nnz / std::max(1, tiles));
    std::vector<int64_t> cut{0};
    for (int64_t e = target; e < nnz; e += target) {
        while (e < nnz && row[e] == row[e - 1]) ++e;
        if (e >= nnz) break;
        cut.push_back(e);
    }
    cut.push_back(nnz);
    const long long groups = (long long)cut.size() - 1;
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, uint32_t>> ln;   // (line, entry)
#pragma omp for schedule(dynamic, 1)
        for (long long g = 0; g < groups; ++g) {
            ln.clear();
            for (int64_t e = cut[g]; e < cut[g + 1]; ++e) ln.push_back({(uint32_t)col[e] / line_elems, (uint32_t)e});
            std::sort(ln.begin(), ln.end());
            for (size_t q = 0; q < ln.size();) {
                size_t r = q + 1;
                while (r < ln.size() && ln[r].first == ln[q].first) ++r;
                if (r - q <= kScatteredReads)
                    for (size_t t = q; t < r; ++t) far[ln[t].second] = 1;
                q = r;
            }
        }
    }
    return far;
}

// The sweep layout of the scattered entries fe (input indices, row-major).
// Block and wave boundaries are found in one pass; the waves' entries are
// then ordered block by block in parallel, each into its own output range.
SweepPart lay_out(const std::vector<uint32_t>& fe, const int* row, const int* col, const double* val, int col_bits,
                  int max_rows) {
    SweepPart sw;
    sw.col_bits = col_bits;
    sw.max_rows = max_rows;
    const int64_t nf = (int64_t)fe.size();
    auto next_row = [&](int64_t k) {   // first entry after k's row
        const int r = row[fe[k]];
        while (k < nf && row[fe[k]] == r) ++k;
        return k;
    };
    std::vector<int64_t> cuts;   // per block: kSweepWaves + 1 boundaries into fe (shared between blocks)
    for (int64_t i = 0; i < nf;) {
        const int r0 = row[fe[i]];
        int64_t j = i;
        while (j < nf && row[fe[j]] - r0 < max_rows) {
            const int64_t k = next_row(j);
            if (j > i && k - i > kSweepBlockEntries) break;
            j = k;
        }
        // block: entries [i, j), rows [r0, row of entry j - 1]; its waves cut
        // at row boundaries near equal shares of its entries
        sw.blk_row.push_back(r0);
        sw.blk_span.push_back(row[fe[j - 1]] - r0 + 1);
        int64_t a = i;
        for (int w = 0; w < kSweepWaves; ++w) {
            cuts.push_back(a);
            int64_t b = j;
            if (w + 1 < kSweepWaves) {
                b = std::max(a, i + (j - i) * (w + 1) / kSweepWaves);
                if (b > i && b < j && row[fe[b]] == row[fe[b - 1]]) b = next_row(b);
                b = std::min(b, j);
            }
            a = b;
        }
        i = j;
    }
    cuts.push_back(nf);
    const long long waves = (long long)cuts.size() - 1;
    sw.wave_off.assign(cuts.begin(), cuts.end());
    sw.key.resize((size_t)nf);
    sw.val.resize((size_t)nf);
    sw.src.resize((size_t)nf);
#pragma omp parallel
    {
        std::vector<uint32_t> grp;
#pragma omp for schedule(dynamic, 16)
        for (long long w = 0; w < waves; ++w) {
            const int64_t a = cuts[w], b = cuts[w + 1];
            const int r0 = sw.blk_row[(size_t)(w / kSweepWaves)];
            grp.assign(fe.begin() + a, fe.begin() + b);
            std::stable_sort(grp.begin(), grp.end(), [&](uint32_t x, uint32_t y) {
                return col[x] != col[y] ? col[x] < col[y] : row[x] < row[y];
            });
            for (size_t q = 0; q < grp.size(); q += 64)
                std::stable_sort(grp.begin() + q, grp.begin() + std::min(grp.size(), q + 64),
                                 [&](uint32_t x, uint32_t y) { return row[x] < row[y]; });
            for (size_t t = 0; t < grp.size(); ++t) {
                const uint32_t e = grp[t];
                sw.key[(size_t)a + t] = ((uint32_t)(row[e] - r0) << col_bits) | (uint32_t)col[e];
                sw.val[(size_t)a + t] = val[e];
                sw.src[(size_t)a + t] = e;
            }
        }
    }
    return sw;
}

}  // namespace

int build_split(cbspmv_plan& p, int nnz, const int* row, const int* col, const double* val) {
    const Plan& h = p.host;
    const int mode = p.opts.stream_split;
    if (mode == CBSPMV_SPLIT_OFF || nnz == 0 || h.m == 0) return CBSPMV_OK;
    if ((h.o.dtype != CBSPMV_F64 && h.o.dtype != CBSPMV_F32) || h.o.accum != CBSPMV_ACC_NATIVE || h.o.panels > 1 ||
        p.schedule != CBSPMV_SCHED_STREAM)
        return CBSPMV_OK;
    const size_t xb = vec_elem_bytes(h);
    if (mode == CBSPMV_SPLIT_AUTO && ((size_t)h.n * xb <= kSplitMinXBytes || (size_t)h.n * xb > kSplitMaxXBytes))
        return CBSPMV_OK;
    int cb = 1;
    while ((1ll << cb) < (long long)h.n) ++cb;
    const int max_rows = cb >= 26 ? 0 : (int)std::min<long long>(kSweepMaxRows, 1ll << (32 - cb));
    if (max_rows < 64) return CBSPMV_OK;   // too few key bits beside the column
    const std::vector<uint8_t> far = classify(nnz, row, col, h.o.tiles, (uint32_t)(128 / xb));
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
    p.near = near;
    p.sweep = lay_out(fe, row, col, val, cb, max_rows);
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
