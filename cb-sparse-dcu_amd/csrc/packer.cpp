// Packers: the canonical slab (the reference's device layout, byte for byte)
// and the stream slab (the transcoded layout of the stream schedule).
#include <algorithm>
#include <cstring>
#include <numeric>

#include <omp.h>

#include "cb_plan.hpp"

namespace cbs {

// ---- canonical slab: cb-spmv-dcu/src/cb-spmv-dcu.hpp:1100-1348 -------------
// Regions in order, each start aligned to its element type: block rows/cols,
// RO segment arrays, nnz and type per block, pattern arrays, gather restore
// map, block address table, per-block payloads (final block order), x, y, RO
// output epochs.  The address table holds byte offsets from the slab base
// (0 = no payload) where the reference stores device pointers, so the image
// does not depend on where it is uploaded.
void pack_canonical(Plan& p) {
    if (p.image_ready) return;
    const int bs = p.o.bs;
    bool unit = true;
    for (int g = 0; g < p.nnzb && unit; ++g) {
        const int o = p.bia_origin[g];
        if (p.nnzpb[g] <= 0) continue;
        const int t = p.type[g] & kTypeMask;
        if (t == kCOO) {
            for (int q = 0; q < p.coo_count[o]; ++q) unit = unit && p.ent_val[p.pl_start[o] + q] == 1.0;
        } else if (t == kCSR) {
            for (int q = 0; q < p.raw_nnz[o]; ++q) unit = unit && p.ent_val[p.pl_start[o] + q] == 1.0;
        } else {
            const double* d = p.dense_val.data() + (size_t)p.dense_slot[o] * bs * bs;
            for (int q = 0; q < bs * bs; ++q) unit = unit && (d[q] == 0.0 || d[q] == 1.0);
        }
    }
    p.unit_values = unit;

    size_t cur = 0;
    auto take = [&cur](size_t bytes, size_t al) { cur = align_up(cur, al); const size_t at = cur; cur += bytes; return at; };
    const size_t nb = (size_t)p.nnzb, S = p.seg_rows.size();
    auto& off = p.off;
    off.clear();
    off["gb_row_idx"] = take(nb * 4, 4);
    off["gb_col_idx"] = take(nb * 4, 4);
    if (p.o.row_owned) {
        off["segment_rows"] = take(S * 4, 4);
        off["segment_begin"] = take(S * 4, 4);
        off["segment_end"] = take(S * 4, 4);
        off["segment_direct"] = take(S, 1);
        off["segment_uniform"] = take(S, 1);
    }
    off["nnz_per_blk"] = take(nb * 4, 4);
    off["type_per_blk"] = take(nb, 1);
    if (p.o.pattern_wave) {
        off["pattern_row_ptr"] = take(p.pat_row_ptr.size() * 4, 4);
        off["pattern_col_idx"] = take(p.pat_col.size() * 4, 4);
        off["pattern_output_ptr"] = take(p.pat_out_ptr.size() * 4, 4);
        off["pattern_output_rows"] = take(p.pat_out_rows.size() * 4, 4);
    }
    if (p.gather) {
        off["block_cols"] = take(p.block_cols.size() * 4, 4);
        off["block_cols_offset"] = take(p.block_cols_off.size() * 4, 4);
    }
    off["block_addr"] = take(nb * 8, 8);
    const size_t ob = orig_index_bytes(bs), mb = dense_mask_bytes(bs);
    std::vector<size_t> boff(nb, 0);
    for (size_t g = 0; g < nb; ++g) {
        if (p.nnzpb[g] <= 0) continue;
        const int o = p.bia_origin[g];
        const int t = p.type[g] & kTypeMask;
        if (t == kCOO) {
            const size_t k = (size_t)p.coo_count[o];
            const size_t ib = k * (p.compact ? 2 : 4);
            boff[g] = take(unit ? ib : align_up(ib, 8) + k * 8, unit ? 4 : 8);
        } else if (t == kCSR) {
            const size_t k = (size_t)p.raw_nnz[o];
            const size_t ib = (bs + 1) * ob + k * ob;
            boff[g] = take(unit ? ib : align_up(ib, 8) + k * 8, unit ? ob : 8);
        } else {
            boff[g] = take(unit ? bs * mb : (size_t)bs * bs * 8, unit ? mb : 8);
        }
    }
    off["x"] = take((size_t)p.gb_n * bs * 8, 8);
    off["y"] = take((size_t)p.gb_m * bs * 8, 8);
    if (p.o.row_owned) off["output_epochs"] = take((size_t)p.gb_m * 4, 4);
    off["total"] = cur;

    p.image.assign(off["y"], 0);
    auto put = [&p](size_t at, const void* src, size_t bytes) { if (bytes) std::memcpy(p.image.data() + at, src, bytes); };
    put(off["gb_row_idx"], p.brow.data(), nb * 4);
    put(off["gb_col_idx"], p.bcol.data(), nb * 4);
    if (p.o.row_owned) {
        put(off["segment_rows"], p.seg_rows.data(), S * 4);
        put(off["segment_begin"], p.seg_begin.data(), S * 4);
        put(off["segment_end"], p.seg_end.data(), S * 4);
        put(off["segment_direct"], p.seg_direct.data(), S);
        put(off["segment_uniform"], p.seg_uniform.data(), S);
    }
    put(off["nnz_per_blk"], p.nnzpb.data(), nb * 4);
    put(off["type_per_blk"], p.type.data(), nb);
    if (p.o.pattern_wave) {
        put(off["pattern_row_ptr"], p.pat_row_ptr.data(), p.pat_row_ptr.size() * 4);
        put(off["pattern_col_idx"], p.pat_col.data(), p.pat_col.size() * 4);
        put(off["pattern_output_ptr"], p.pat_out_ptr.data(), p.pat_out_ptr.size() * 4);
        put(off["pattern_output_rows"], p.pat_out_rows.data(), p.pat_out_rows.size() * 4);
    }
    if (p.gather) {
        put(off["block_cols"], p.block_cols.data(), p.block_cols.size() * 4);
        put(off["block_cols_offset"], p.block_cols_off.data(), p.block_cols_off.size() * 4);
    }
    std::vector<uint64_t> addr(nb, 0);
    for (size_t g = 0; g < nb; ++g) if (p.nnzpb[g] > 0) addr[g] = boff[g];
    put(off["block_addr"], addr.data(), nb * 8);
#pragma omp parallel for schedule(dynamic, 1024)
    for (long long gg = 0; gg < (long long)nb; ++gg) {
        const size_t g = (size_t)gg;
        if (p.nnzpb[g] <= 0) continue;
        const int o = p.bia_origin[g];
        const int t = p.type[g] & kTypeMask;
        uint8_t* dst = p.image.data() + boff[g];
        if (t == kCOO) {
            const size_t k = (size_t)p.coo_count[o];
            const uint32_t* idx = p.coo_idx.data() + p.pl_start[o];
            if (p.compact) {
                for (size_t q = 0; q < k; ++q) { const uint16_t w = (uint16_t)idx[q]; std::memcpy(dst + 2 * q, &w, 2); }
            } else {
                std::memcpy(dst, idx, k * 4);
            }
            if (!unit) std::memcpy(dst + align_up(k * (p.compact ? 2 : 4), 8), p.ent_val.data() + p.pl_start[o], k * 8);
        } else if (t == kCSR) {
            const size_t k = (size_t)p.raw_nnz[o];
            const int* rp = p.csr_rp.data() + p.csr_rp_start[o];
            const int* cc = p.csr_col.data() + p.pl_start[o];
            auto put_orig = [&](size_t pos, int v) {
                if (ob == 1) dst[pos] = (uint8_t)v;
                else { const uint16_t w = (uint16_t)v; std::memcpy(dst + pos * 2, &w, 2); }
            };
            for (int r = 0; r <= bs; ++r) put_orig((size_t)r, rp[r]);
            for (size_t q = 0; q < k; ++q) put_orig(bs + 1 + q, cc[q]);
            if (!unit) std::memcpy(dst + align_up((bs + 1 + k) * ob, 8), p.ent_val.data() + p.pl_start[o], k * 8);
        } else {
            const double* d = p.dense_val.data() + (size_t)p.dense_slot[o] * bs * bs;
            if (unit) {
                for (int r = 0; r < bs; ++r) {
                    uint64_t mask = 0;
                    for (int c = 0; c < bs; ++c) if (d[r * bs + c] != 0.0) mask |= 1ull << c;
                    std::memcpy(dst + (size_t)r * mb, &mask, mb);
                }
            } else {
                std::memcpy(dst, d, (size_t)bs * bs * 8);
            }
        }
    }
    p.image_ready = true;
}

// ---- stream slab -------------------------------------------------------------
// Every block of a strip (block row) contributes entries (column, row, value):
// COO blocks their merged entries, CSR blocks their entries, DENSE blocks
// every position inside the matrix.  Columns are original columns: a gathered
// block's local column goes through the restore map once, here, instead of
// once per SpMV.  Inside a strip, entries are ordered by row, then block
// order, then in-block order (a stable counting sort), so every row is one
// contiguous run.
namespace {

// Per-entry cost weights of the automatic item cut (integers, so that item
// targets stay integral): an entry whose x comes from LDS or is read again
// in its tile costs kStagedWeight, a random gather the other.  Per-workgroup
// end times on the 1,048,576^2 stencil / power-law matrix (30.2M entries)
// with items cut by entry count put stencil entries at 0.66 ns (FP64) /
// 0.38 ns (FP32) of one CU's time and the power-law half's scattered ones at
// 2.3 / 1.7 ns.  Swept on the same matrix (gathered weight over staged,
// launch time): FP64 3.5x 135.7 us, 4.5x 136.5, 5.5x 136.2, 7x 145.5
// (137.0 by entry count); FP32 3.5x 96.7, 4.5x 96.3, 5.5x 94.8, 7x 94.4, 9x
// 104.0 us (99.7 by entry count).  Heavier weights let a few tiles with many
// scattered entries run long.
constexpr int64_t kStagedWeight = 2;
constexpr int64_t kGatheredWeightF64 = 7;
constexpr int64_t kGatheredWeightF32 = 12;

// Stream entries in row order.  zero marks an entry that only makes its row a
// run (value 0, padding key).
struct Entries {
    std::vector<uint32_t> row, col;
    std::vector<double> val;
    std::vector<uint8_t> zero;
    int64_t size() const { return (int64_t)row.size(); }
    void reserve(size_t n) { row.reserve(n); col.reserve(n); val.reserve(n); zero.reserve(n); }
    void push(uint32_t r, uint32_t c, double v, uint8_t z) { row.push_back(r); col.push_back(c); val.push_back(v); zero.push_back(z); }
};

struct Piece { int64_t b, e; uint8_t fp, lp; uint32_t split_first, split_last; };

struct StreamShape {
    int epl, cb, max_lines;
    uint32_t omask, zero_slot;
    size_t tile_items;
};

StreamShape stream_shape(const Plan& p) {
    StreamShape s{};
    s.epl = stream_lane_entries(p.o.dtype);
    const int slots = tile_slots(p.o.tile_bytes, p.o.dtype);
    s.max_lines = tile_lines(p.o.tile_bytes, p.o.dtype);
    s.tile_items = (size_t)p.o.tile_items;
    s.zero_slot = (uint32_t)slots - 1;
    s.cb = 1;
    while ((1ll << s.cb) < (long long)std::max(p.n, slots)) ++s.cb;   // columns and LDS slots
    s.omask = (1u << (31 - s.cb)) - 1u;                                // row-offset field
    return s;
}

// Entries [b0, e0) of E in items of about `target` cost units, cut at row
// boundaries, an item's rows spanning at most omask + 1 row offsets.  The
// cost of entries [a, b) is wp[b] - wp[a] (prefix sums of per-entry
// weights), or b - a without wp.  With `split`, a row costing more than 4
// target is cut into pieces of about target (by entry count) that meet in
// the partial buffer (splits; first_item counts from item_base); without,
// such a row is one item.
// Starts of the row runs of E in [b0, e0), then e0.
std::vector<int64_t> row_runs(const Entries& E, int64_t b0, int64_t e0) {
    std::vector<int64_t> runs;
    for (int64_t r = b0; r < e0; ++r)
        if (r == b0 || E.row[r] != E.row[r - 1]) runs.push_back(r);
    runs.push_back(e0);
    return runs;
}

// `runs`: row_runs(E, b0, e0), so that the repeated cuts of a bisection
// step over rows rather than entries.
void cut_items(const Entries& E, const std::vector<int64_t>& runs, int64_t target, bool split, uint32_t omask,
               size_t item_base, std::vector<Piece>& pieces, std::vector<StreamSplit>& splits,
               uint32_t& partial_slots, const std::vector<int64_t>* wp = nullptr) {
    auto cost = [&](int64_t a, int64_t b) { return wp ? (*wp)[b] - (*wp)[a] : b - a; };
    const int64_t split_len = 4 * target;
    const int64_t b0 = runs.front(), e0 = runs.back();
    int64_t cur = b0;
    for (size_t k = 0; k + 1 < runs.size(); ++k) {
        const int64_t r = runs[k], re = runs[k + 1];
        const int64_t len = re - r, len_c = cost(r, re);
        if (split && len_c > split_len) {
            if (cur < r) pieces.push_back({cur, r, 0, 0, ~0u, ~0u});
            const int64_t np = (len_c + target - 1) / target;
            StreamSplit sp{};
            sp.row = E.row[r];
            sp.first_item = (uint32_t)(item_base + pieces.size());
            sp.pieces = (uint32_t)np;
            sp.partial_base = partial_slots;
            partial_slots += (uint32_t)np;
            const uint32_t id = (uint32_t)splits.size();
            splits.push_back(sp);
            for (int64_t q = 0; q < np; ++q)
                pieces.push_back({r + len * q / np, r + len * (q + 1) / np, (uint8_t)(q > 0), (uint8_t)(q + 1 < np),
                                  q > 0 ? id : ~0u, q + 1 < np ? id : ~0u});
            cur = re;
        } else if (r > cur && (cost(cur, r) + len_c > target || E.row[re - 1] - E.row[cur] > omask)) {
            pieces.push_back({cur, r, 0, 0, ~0u, ~0u});
            cur = r;
        }
    }
    if (cur < e0) pieces.push_back({cur, e0, 0, 0, ~0u, ~0u});
}

// Appends the pieces as items, padded to the lane width: fills p.items, the
// split ids, p.s_val and the per-position column / row offset / padding flag;
// src[o] is the entry of E at padded position o (-1 for padding).
void append_items(Plan& p, const Entries& E, const std::vector<Piece>& pieces, int epl, std::vector<uint32_t>& p_col,
                  std::vector<uint32_t>& p_off, std::vector<uint8_t>& p_zero, std::vector<int64_t>* src) {
    const size_t I0 = p.items.size(), NI = pieces.size();
    int64_t at = (int64_t)p_col.size();
    std::vector<int64_t> out_at(NI + 1, at);
    for (size_t i = 0; i < NI; ++i) {
        const int64_t len = pieces[i].e - pieces[i].b;
        out_at[i + 1] = out_at[i] + (len + epl - 1) / epl * epl;
    }
    const size_t padded = (size_t)out_at[NI];
    p_col.resize(padded);
    p_off.resize(padded);
    p_zero.resize(padded, 0);
    p.s_val.resize(padded, 0.0);
    if (src) src->resize(padded, -1);
    p.items.resize(I0 + NI);
    p.item_split.resize(I0 + NI, ~0u);
    p.item_split_first.resize(I0 + NI, ~0u);
    for (size_t i = 0; i < NI; ++i) {
        const Piece& pc = pieces[i];
        StreamItem& it = p.items[I0 + i];
        it = StreamItem{};
        it.e_begin = (uint32_t)out_at[i];
        it.e_end = (uint32_t)out_at[i + 1];
        it.row_base = E.row[pc.b];
        it.first_partial = pc.fp;
        it.last_partial = pc.lp;
        p.item_split_first[I0 + i] = pc.split_first;
        p.item_split[I0 + i] = pc.split_last;
        int64_t o = out_at[i];
        for (int64_t e = pc.b; e < pc.e; ++e, ++o) {
            p_col[o] = E.col[e];
            p_off[o] = E.row[e] - it.row_base;
            p.s_val[o] = E.zero[e] ? 0.0 : E.val[e];
            p_zero[o] = E.zero[e];
            if (src) (*src)[o] = E.zero[e] ? -1 : e;
        }
        for (; o < out_at[i + 1]; ++o) {   // padding: extends the last run by 0 * 0
            p_col[o] = 0;
            p_off[o] = E.row[pc.e - 1] - it.row_base;
            p_zero[o] = 1;
        }
    }
}

// Tile items of the entries E: items sized so that the tiles (tile_items
// items each) number at most kStreamCUs by default (every workgroup of a
// launch runs in one round on a 256-CU MI355X; an explicit item_nnz is used
// as given), then each tile's hot x lines (16 elements, hit at least twice;
// the most hit ones when more than fit) and the keys.  Resets the stream
// tables of p.  src receives the entry of E at each padded position.
void layout_tiles(Plan& p, const StreamShape& sh, const Entries& E, std::vector<uint32_t>& p_col,
                  std::vector<uint32_t>& p_off, std::vector<uint8_t>& p_zero, std::vector<int64_t>& src,
                  const std::vector<int64_t>* wp = nullptr) {
    const int64_t total = E.size();
    const int64_t total_c = wp ? wp->back() : total;   // cost of all entries
    p.items.clear();
    p.item_split.clear();
    p.item_split_first.clear();
    p.splits.clear();
    p.partial_slots = 0;
    p.s_val.clear();
    p_col.clear(); p_off.clear(); p_zero.clear(); src.clear();
    std::vector<Piece> pieces;
    const std::vector<int64_t> runs = row_runs(E, 0, total);
    auto cut = [&](int64_t target) {
        pieces.clear();
        p.splits.clear();
        p.partial_slots = 0;
        cut_items(E, runs, target, true, sh.omask, 0, pieces, p.splits, p.partial_slots, wp);
    };
    if (p.o.item_nnz > 0) {
        cut(std::max<int64_t>(64, std::min<int64_t>(p.o.item_nnz, 1 << 22)));
    } else {
        // the smallest item size whose cut fits one round of tiles (one per
        // CU): cutting at row boundaries adds items, so grow the target until
        // the cut fits, then bisect back towards the smallest that fits, so
        // that few CUs are left without a tile (30M entries in FP64: 256
        // tiles instead of 249)
        // (in cost units with weights: at least kMinItemEntries entries of
        // the average cost)
        const int64_t max_items = (int64_t)p.o.tiles * p.o.tile_items;
        const int64_t min_item = total > 0 ? std::max<int64_t>(1, kMinItemEntries * total_c / total) : kMinItemEntries;
        int64_t lo = std::max<int64_t>(min_item, total_c / max_items + 1), hi = lo;
        cut(hi);
        bool fits = (int64_t)pieces.size() <= max_items;
        for (int round = 0; round < 64 && !fits && hi < ((int64_t)1 << 40); ++round) {
            lo = hi + 1;
            hi = hi + hi / 16 + 1;
            cut(hi);
            fits = (int64_t)pieces.size() <= max_items;
        }
        if (fits && lo < hi) {
            while (lo < hi) {   // hi fits, lo - 1 did not
                const int64_t mid = lo + (hi - lo) / 2;
                cut(mid);
                if ((int64_t)pieces.size() <= max_items) hi = mid;
                else lo = mid + 1;
            }
            cut(hi);
        }
    }
    append_items(p, E, pieces, sh.epl, p_col, p_off, p_zero, &src);
    const size_t NI = p.items.size();
    const int64_t padded = (int64_t)p_col.size();
    p.s_key.assign((size_t)padded, 0);
    const size_t NT = (NI + sh.tile_items - 1) / sh.tile_items;
    std::vector<std::vector<uint32_t>> lines(NT);
#pragma omp parallel
    {
        std::vector<uint32_t> ln;
        std::vector<std::pair<uint32_t, uint32_t>> cnt;   // (count, line)
#pragma omp for schedule(dynamic, 4)
        for (long long t = 0; t < (long long)NT; ++t) {
            const int64_t b = p.items[t * sh.tile_items].e_begin;
            const int64_t e = p.items[std::min(NI, (size_t)(t + 1) * sh.tile_items) - 1].e_end;
            ln.clear();
            for (int64_t q = b; q < e; ++q) if (!p_zero[q]) ln.push_back(p_col[q] / kLineElems);
            std::sort(ln.begin(), ln.end());
            cnt.clear();
            for (size_t q = 0; q < ln.size();) {
                size_t r = q + 1;
                while (r < ln.size() && ln[r] == ln[q]) ++r;
                if (r - q >= 2) cnt.push_back({(uint32_t)(r - q), ln[q]});
                q = r;
            }
            if ((int)cnt.size() > sh.max_lines) {
                std::nth_element(cnt.begin(), cnt.begin() + sh.max_lines, cnt.end(),
                                 [](const std::pair<uint32_t, uint32_t>& x, const std::pair<uint32_t, uint32_t>& y) {
                                     return x.first != y.first ? x.first > y.first : x.second < y.second;
                                 });
                cnt.resize(sh.max_lines);
            }
            std::vector<uint32_t>& mine = lines[t];
            mine.clear();
            for (auto& c : cnt) mine.push_back(c.second);
            std::sort(mine.begin(), mine.end());
            for (int64_t q = b; q < e; ++q) {
                const uint32_t off = (p_off[q] & sh.omask) << sh.cb;
                if (p_zero[q]) { p.s_key[q] = kKeyCached | off | sh.zero_slot; continue; }
                const uint32_t col = p_col[q];
                auto it = std::lower_bound(mine.begin(), mine.end(), col / kLineElems);
                if (it != mine.end() && *it == col / kLineElems)
                    p.s_key[q] = kKeyCached | off | ((uint32_t)(it - mine.begin()) * kLineElems + col % kLineElems);
                else
                    p.s_key[q] = off | col;
            }
        }
    }
    p.tiles.assign(2 * NT, 0);
    p.tile_lines.clear();
    for (size_t t = 0; t < NT; ++t) {
        p.tiles[2 * t] = (uint32_t)p.tile_lines.size();
        p.tiles[2 * t + 1] = (uint32_t)lines[t].size();
        p.tile_lines.insert(p.tile_lines.end(), lines[t].begin(), lines[t].end());
    }
}

}  // namespace

int stream_panels(const Plan& p) {
    if (p.o.panels > 0) return p.o.panels;
    if (!kPanelsByDefault || p.o.accum == CBSPMV_ACC_COMPENSATED || p.o.tile_items > kMaxTileItems / 2) return 1;
    const int64_t x_bytes = (int64_t)p.n * (p.o.dtype == CBSPMV_F64 ? 8 : 4);
    if (x_bytes <= kPanelL2Bytes) return 1;
    int P = 2;
    while (P < kXcds && (int64_t)P * kPanelBytes < x_bytes) P *= 2;
    return P;
}

void build_stream(Plan& p) {
    if (p.stream_ready) return;
    const int bs = p.o.bs, bcb = coord_bits(bs);
    const StreamShape sh = stream_shape(p);
    const int epl = sh.epl;
    p.col_bits = sh.cb;

    // blocks of each strip in final block order (pad blocks skipped)
    std::vector<int> sb_ptr(p.gb_m + 1, 0), sb_list;
    for (int g = 0; g < p.nnzb; ++g) if (p.nnzpb[g] > 0) sb_ptr[p.brow[g] + 1]++;
    for (int b = 0; b < p.gb_m; ++b) sb_ptr[b + 1] += sb_ptr[b];
    sb_list.assign(sb_ptr[p.gb_m], 0);
    {
        std::vector<int> c(sb_ptr.begin(), sb_ptr.end() - 1);
        for (int g = 0; g < p.nnzb; ++g) if (p.nnzpb[g] > 0) sb_list[c[p.brow[g]]++] = g;
    }
    // usable columns of a DENSE block: positions past the gathered column
    // count (or past n) hold no entry (cb-spmv-dcu.hpp:962, :1018, :1034)
    auto block_cols_used = [&](int g) {
        const int b = p.brow[g];
        const int ncols = p.gather ? (p.block_cols_off[b + 1] - p.block_cols_off[b]) : p.n;
        return std::max(0, std::min(bs, ncols - p.bcol[g] * bs));
    };
    auto block_entries = [&](int g) -> int64_t {
        const int o = p.bia_origin[g];
        const int t = p.type[g] & kTypeMask;
        return t == kCOO ? p.coo_count[o] : t == kCSR ? p.raw_nnz[o] : (int64_t)bs * block_cols_used(g);
    };
    std::vector<int64_t> strip_ptr(p.gb_m + 1, 0);
    for (int b = 0; b < p.gb_m; ++b) {
        int64_t c = 0;
        for (int q = sb_ptr[b]; q < sb_ptr[b + 1]; ++q) c += block_entries(sb_list[q]);
        strip_ptr[b + 1] = strip_ptr[b] + c;
    }
    const int64_t total = strip_ptr[p.gb_m];
    Entries E;
    E.row.resize(total);
    E.col.resize(total);
    E.val.resize(total);
    E.zero.assign(total, 0);

    const uint32_t rmask = (1u << bcb) - 1;
#pragma omp parallel
    {
        std::vector<int64_t> row_cnt(bs + 1);
        std::vector<uint32_t> tk;
        std::vector<double> tv;
        std::vector<int> tr;
#pragma omp for schedule(dynamic, 64)
        for (int b = 0; b < p.gb_m; ++b) {
            tk.clear(); tv.clear(); tr.clear();
            const int base_col_off = p.gather ? p.block_cols_off[b] : 0;
            for (int q = sb_ptr[b]; q < sb_ptr[b + 1]; ++q) {
                const int g = sb_list[q];
                const int o = p.bia_origin[g];
                const int t = p.type[g] & kTypeMask;
                auto orig_col = [&](int lc) -> int {
                    const int wc = p.bcol[g] * bs + lc;
                    return p.gather ? p.block_cols[base_col_off + wc] : wc;
                };
                if (t == kCOO) {
                    const uint32_t* idx = p.coo_idx.data() + p.pl_start[o];
                    const double* v = p.ent_val.data() + p.pl_start[o];
                    for (int k = 0; k < p.coo_count[o]; ++k) {
                        int r, c;
                        if (p.compact) { r = (int)((idx[k] >> 12) & rmask); c = (int)(idx[k] & 0xfff); }
                        else if (p.direct) { r = (int)((idx[k] >> 16) & rmask); c = (int)(idx[k] & 0xffff); }
                        else { r = (int)((idx[k] >> bcb) & rmask); c = orig_col((int)(idx[k] & rmask)); }
                        tr.push_back(r); tk.push_back((uint32_t)c); tv.push_back(v[k]);
                    }
                } else if (t == kCSR) {
                    const int* rp = p.csr_rp.data() + p.csr_rp_start[o];
                    const int* cc = p.csr_col.data() + p.pl_start[o];
                    const double* v = p.ent_val.data() + p.pl_start[o];
                    for (int r = 0; r < bs; ++r)
                        for (int k = rp[r]; k < rp[r + 1]; ++k) { tr.push_back(r); tk.push_back((uint32_t)orig_col(cc[k])); tv.push_back(v[k]); }
                } else {
                    const double* d = p.dense_val.data() + (size_t)p.dense_slot[o] * bs * bs;
                    const int used = block_cols_used(g);
                    for (int r = 0; r < bs; ++r)
                        for (int c = 0; c < used; ++c) { tr.push_back(r); tk.push_back((uint32_t)orig_col(c)); tv.push_back(d[r * bs + c]); }
                }
            }
            std::fill(row_cnt.begin(), row_cnt.end(), 0);
            for (int r : tr) row_cnt[r + 1]++;
            for (int r = 0; r < bs; ++r) row_cnt[r + 1] += row_cnt[r];
            const int64_t base = strip_ptr[b];
            for (size_t k = 0; k < tr.size(); ++k) {
                const int64_t at = base + row_cnt[tr[k]]++;
                E.row[at] = (uint32_t)(b * bs + tr[k]);
                E.col[at] = tk[k];
                E.val[at] = tv[k];
            }
        }
    }

    // rows without entries (the kernel stores their zeros)
    p.empty_rows.clear();
    {
        std::vector<uint8_t> has(p.m, 0);
        for (int64_t e = 0; e < total; ++e) if (E.row[e] < (uint32_t)p.m) has[E.row[e]] = 1;
        for (int r = 0; r < p.m; ++r) if (!has[r]) p.empty_rows.push_back((uint32_t)r);
    }

    std::vector<uint32_t> p_col, p_off;
    std::vector<uint8_t> p_zero;
    std::vector<int64_t> src;
    layout_tiles(p, sh, E, p_col, p_off, p_zero, src);

    // ---- items of equal cost: an entry whose x line no other entry of its
    // tile reads (a random gather: a vector-L1 miss served from L2 or
    // beyond) costs several times one whose x is staged in LDS or read again
    // from the L1, so items cut by entry count leave the CUs whose tiles
    // stage everything idle for much of the launch (the 1,048,576^2
    // stencil / power-law matrix in FP64: the stencil tiles end at ~78 us,
    // the power-law tiles at 132-140 us).  The
    // automatic cut is redone with each entry weighted as the first layout
    // placed it (kGatheredWeight for an unstaged entry whose line is read
    // once in its tile, kStagedWeight for the others) and the tiles' staged
    // lines chosen again.
    if (p.o.item_nnz == 0 && stream_panels(p) <= 1) {
        const int64_t wg = p.o.dtype == CBSPMV_F64 ? kGatheredWeightF64 : kGatheredWeightF32;
        std::vector<int64_t> wp((size_t)total + 1, kStagedWeight);
        wp[0] = 0;
        const size_t NI = p.items.size(), TI = (size_t)p.o.tile_items;
        const size_t NT = (NI + TI - 1) / TI;
        int64_t n_single = 0;
#pragma omp parallel reduction(+ : n_single)
        {
            std::vector<std::pair<uint32_t, int64_t>> ln;   // (line, entry of E)
#pragma omp for schedule(dynamic, 4)
            for (long long t = 0; t < (long long)NT; ++t) {
                const int64_t b = p.items[t * TI].e_begin;
                const int64_t e = p.items[std::min(NI, (size_t)(t + 1) * TI) - 1].e_end;
                ln.clear();
                for (int64_t o = b; o < e; ++o)
                    if (src[o] >= 0 && (p.s_key[o] & kKeyCached) == 0) ln.push_back({p_col[o] / kLineElems, src[o]});
                std::sort(ln.begin(), ln.end());
                for (size_t q = 0; q < ln.size(); ++q) {
                    const bool alone = (q == 0 || ln[q - 1].first != ln[q].first) &&
                                       (q + 1 == ln.size() || ln[q + 1].first != ln[q].first);
                    if (alone) {
                        wp[(size_t)ln[q].second + 1] = wg;
                        ++n_single;
                    }
                }
            }
        }
        if (n_single > 0) {
            for (int64_t e = 0; e < total; ++e) wp[(size_t)e + 1] += wp[(size_t)e];
            layout_tiles(p, sh, E, p_col, p_off, p_zero, src, &wp);
        }
    }

    // ---- column panels: the entries whose x line is not staged in LDS are
    // moved to panel items, one row-ordered stream per panel of columns, so
    // that the waves of one XCD gather x from one panel only (held in that
    // XCD's L2).  Every workgroup runs a tile's items and, on its other
    // waves, items of its XCD's panel; the panel sums reach y in a second,
    // small launch.  Rows cut into pieces keep all their entries in the
    // tiles.  A row left without tile entries keeps a zero entry there, so
    // that its tile stores its y before the panel sums are added.
    p.panels = 1;
    p.tile_items_count = (uint32_t)p.items.size();
    p.panel_item_ptr.clear();
    p.panel_mask.clear();
    const int P = stream_panels(p);
    if (P > 1) {
        const int64_t padded = (int64_t)p_col.size();
        std::vector<uint8_t> uncached(total, 0), split_row(p.m, 0);
        for (const StreamSplit& sp : p.splits) split_row[sp.row] = 1;
        int64_t n_out = 0;
        for (int64_t o = 0; o < padded; ++o) {
            if (src[o] < 0 || (p.s_key[o] & kKeyCached)) continue;
            if (split_row[E.row[src[o]]]) continue;
            uncached[src[o]] = 1;
            ++n_out;
        }
        if (n_out > 0) {
            const uint32_t width = (uint32_t)(((int64_t)p.n + P - 1) / P + kLineElems - 1) / kLineElems * kLineElems;
            Entries T;                 // tile entries
            std::vector<Entries> U(P); // panel entries, row order
            T.reserve((size_t)(total - n_out + 64));
            for (int64_t r0 = 0; r0 < total;) {
                int64_t r1 = r0 + 1;
                while (r1 < total && E.row[r1] == E.row[r0]) ++r1;
                bool kept = false;
                for (int64_t q = r0; q < r1; ++q) {
                    if (uncached[q]) U[E.col[q] / width].push(E.row[q], E.col[q], E.val[q], 0);
                    else { T.push(E.row[q], E.col[q], E.val[q], E.zero[q]); kept = true; }
                }
                if (!kept) T.push(E.row[r0], 0, 0.0, 1);
                r0 = r1;
            }
            E = Entries();
            layout_tiles(p, sh, T, p_col, p_off, p_zero, src);
            p.panels = P;
            p.tile_items_count = (uint32_t)p.items.size();
            p.panel_mask.assign(p.m, 0);
            // panel waves per panel: the workgroups on its XCDs times the
            // panel waves of each (as many as the tile's)
            const int64_t NT = (int64_t)(p.tiles.size() / 2);
            const int64_t grid = std::max<int64_t>(kXcds, (NT + kXcds - 1) / kXcds * kXcds);
            const int64_t waves = grid / P * (int64_t)sh.tile_items;
            p.panel_item_ptr.assign(P + 1, (uint32_t)p.items.size());
            std::vector<Piece> pieces;
            std::vector<StreamSplit> no_splits;
            uint32_t no_slots = 0;
            for (int q = 0; q < P; ++q) {
                const Entries& Uq = U[q];
                for (int64_t k = 0; k < Uq.size(); ++k) p.panel_mask[Uq.row[k]] |= (uint8_t)(1u << q);
                if (Uq.size() > 0) {
                    const int64_t target = std::max<int64_t>(256, (Uq.size() + waves - 1) / waves);
                    pieces.clear();
                    cut_items(Uq, row_runs(Uq, 0, Uq.size()), target, false, sh.omask, 0, pieces, no_splits, no_slots);
                    append_items(p, Uq, pieces, epl, p_col, p_off, p_zero, nullptr);
                }
                p.panel_item_ptr[q + 1] = (uint32_t)p.items.size();
            }
            // panel keys: column (never staged), padding as in the tiles
            const size_t start = p.s_key.size();
            p.s_key.resize(p_col.size());
            for (size_t o = start; o < p_col.size(); ++o) {
                const uint32_t off = (p_off[o] & sh.omask) << sh.cb;
                p.s_key[o] = p_zero[o] ? (kKeyCached | off | sh.zero_slot) : (off | p_col[o]);
            }
        }
    }

    // ---- device slab: per chunk, key planes then value planes (16 B/lane)
    const size_t NI = p.items.size();
    const int64_t padded = (int64_t)p_col.size();
    const int vb = stream_val_bytes(p.o.dtype);
    const int nk = epl * 4 / 16;
    const int64_t eb = 4 + vb;
    p.s_slab.assign((size_t)(padded * eb), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long i = 0; i < (long long)NI; ++i) {
        const StreamItem& it = p.items[i];
        for (int64_t e0 = it.e_begin; e0 < it.e_end; e0 += 64 * epl) {
            const int G = (int)std::min<int64_t>(64, (it.e_end - e0) / epl);
            uint8_t* base = p.s_slab.data() + e0 * eb;
            for (int l = 0; l < G; ++l)
                for (int j = 0; j < epl; ++j) {
                    const int64_t e = e0 + (int64_t)l * epl + j;
                    std::memcpy(base + (size_t)(j / 4) * G * 16 + l * 16 + (j % 4) * 4, &p.s_key[e], 4);
                    uint8_t* vdst = base + (size_t)(nk + j * vb / 16) * G * 16 + l * 16 + (j * vb) % 16;
                    if (vb == 8) {
                        std::memcpy(vdst, &p.s_val[e], 8);
                    } else if (vb == 4) {
                        const float f = (float)p.s_val[e];
                        std::memcpy(vdst, &f, 4);
                    } else {   // BF16, round to nearest even
                        const uint16_t h = bf16_bits(p.s_val[e]);
                        std::memcpy(vdst, &h, 2);
                    }
                }
        }
    }
    p.stream_ready = true;
}

// ---- stream entry <-> input entry ---------------------------------------------
size_t stream_value_offset(const Plan& p, uint32_t item, uint32_t e) {
    const int epl = stream_lane_entries(p.o.dtype);
    const size_t vb = (size_t)stream_val_bytes(p.o.dtype);
    const size_t eb = 4 + vb;
    const int nk = epl * 4 / 16;
    const StreamItem& it = p.items[item];
    const uint32_t chunk = (uint32_t)(64 * epl);
    const uint32_t e0 = it.e_begin + (e - it.e_begin) / chunk * chunk;
    const uint32_t G = std::min<uint32_t>(64u, (it.e_end - e0) / (uint32_t)epl);
    const uint32_t l = (e - e0) / (uint32_t)epl, j = (e - e0) % (uint32_t)epl;
    return (size_t)e0 * eb + (size_t)(nk + j * vb / 16) * G * 16 + (size_t)l * 16 + (j * vb) % 16;
}

// Decodes (row, column) of every real stream entry from its key, then pairs
// input entries with stream entries row by row on the column: a column that
// occurs once in the stream row takes every input entry of that (row, col)
// (COO and DENSE blocks merged them); a column that occurs as often in the
// stream row as in the input row (CSR blocks keep duplicates, in input
// order) is paired in order.
bool stream_entry_slots(const Plan& p, std::vector<uint32_t>& slot) {
    const int64_t nnz = p.nnz;
    slot.assign((size_t)nnz, 0u);
    if (nnz == 0) return true;
    const int cb = p.col_bits;
    const uint32_t cmask = (1u << cb) - 1u, omask = (1u << (31 - cb)) - 1u;
    const uint32_t zero_slot = (uint32_t)tile_slots(p.o.tile_bytes, p.o.dtype) - 1u;
    const size_t NI = p.items.size();
    // real entries in stream order (rows are non-decreasing inside the tile
    // items and inside the panel items), then grouped by row
    std::vector<uint32_t> srow, scol, spos;
    srow.reserve((size_t)nnz);
    scol.reserve((size_t)nnz);
    spos.reserve((size_t)nnz);
    for (size_t i = 0; i < NI; ++i) {
        const StreamItem& it = p.items[i];
        const size_t t = i / (size_t)p.o.tile_items;
        for (uint32_t e = it.e_begin; e < it.e_end; ++e) {
            const uint32_t k = p.s_key[e];
            const uint32_t low = k & cmask;
            const bool cached = (k & kKeyCached) != 0;
            if (cached && low == zero_slot) continue;   // padding
            const uint32_t col = cached ? p.tile_lines[p.tiles[2 * t] + low / kLineElems] * kLineElems + low % kLineElems : low;
            srow.push_back(it.row_base + ((k >> cb) & omask));
            scol.push_back(col);
            spos.push_back(e);
        }
    }
    const int m = p.m;
    std::vector<int64_t> sptr((size_t)m + 1, 0), iptr((size_t)m + 1, 0);
    for (uint32_t r : srow) {
        if (r >= (uint32_t)m) return false;
        sptr[r + 1]++;
    }
    for (int r : p.in_row) iptr[r + 1]++;
    for (int r = 0; r < m; ++r) { sptr[r + 1] += sptr[r]; iptr[r + 1] += iptr[r]; }
    {   // stable counting sort by row
        std::vector<uint32_t> c2(srow.size()), p2(srow.size());
        std::vector<int64_t> at(sptr.begin(), sptr.end() - 1);
        for (size_t q = 0; q < srow.size(); ++q) {
            const int64_t d = at[srow[q]]++;
            c2[d] = scol[q];
            p2[d] = spos[q];
        }
        scol.swap(c2);
        spos.swap(p2);
    }
    bool ok = true;
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, uint32_t>> sc, ic;   // (col, position) / (col, input index)
#pragma omp for schedule(dynamic, 256) reduction(&& : ok)
        for (int r = 0; r < m; ++r) {
            sc.clear(); ic.clear();
            for (int64_t q = sptr[r]; q < sptr[r + 1]; ++q) sc.push_back({scol[q], spos[q]});
            for (int64_t q = iptr[r]; q < iptr[r + 1]; ++q) ic.push_back({(uint32_t)p.in_col[q], (uint32_t)q});
            std::sort(sc.begin(), sc.end());
            std::sort(ic.begin(), ic.end());
            size_t a = 0, b = 0;
            while (b < ic.size()) {
                const uint32_t c = ic[b].first;
                while (a < sc.size() && sc[a].first < c) ++a;   // stream-only (zero) entries
                size_t a2 = a, b2 = b;
                while (a2 < sc.size() && sc[a2].first == c) ++a2;
                while (b2 < ic.size() && ic[b2].first == c) ++b2;
                const size_t ns = a2 - a, ni = b2 - b;
                if (ns == 1) {
                    for (size_t q = b; q < b2; ++q) slot[ic[q].second] = sc[a].second;
                } else if (ns == ni) {
                    for (size_t q = 0; q < ni; ++q) slot[ic[b + q].second] = sc[a + q].second;
                } else {
                    ok = false;
                    break;
                }
                a = a2;
                b = b2;
            }
        }
    }
    return ok;
}

// ---- parity export -----------------------------------------------------------
namespace {
template <typename T>
long copy_out(const std::vector<T>& v, void* dst, size_t cap) {
    if (dst && cap >= v.size() && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
    return (long)v.size();
}
}  // namespace

long export_array(const Plan& p, const std::string& s, void* dst, size_t cap, int& status) {
    status = CBSPMV_OK;
    if (s == "gb_row_idx") return copy_out(p.brow, dst, cap);
    if (s == "gb_col_idx") return copy_out(p.bcol, dst, cap);
    if (s == "gb_row_ptr") return copy_out(p.row_ptr, dst, cap);
    if (s == "nnz_per_blk") return copy_out(p.nnzpb, dst, cap);
    if (s == "type_per_blk") return copy_out(p.type, dst, cap);
    if (s == "segment_rows") return copy_out(p.seg_rows, dst, cap);
    if (s == "segment_begin") return copy_out(p.seg_begin, dst, cap);
    if (s == "segment_end") return copy_out(p.seg_end, dst, cap);
    if (s == "segment_direct") return copy_out(p.seg_direct, dst, cap);
    if (s == "segment_uniform") return copy_out(p.seg_uniform, dst, cap);
    if (s == "block_row_uniform") return copy_out(p.row_uniform, dst, cap);
    if (s == "flat_row_ptr") return copy_out(p.flat_row_ptr, dst, cap);
    if (s == "flat_col_idx") return copy_out(p.flat_col, dst, cap);
    if (s == "pattern_row_ptr") return copy_out(p.pat_row_ptr, dst, cap);
    if (s == "pattern_col_idx") return copy_out(p.pat_col, dst, cap);
    if (s == "pattern_output_ptr") return copy_out(p.pat_out_ptr, dst, cap);
    if (s == "pattern_output_rows") return copy_out(p.pat_out_rows, dst, cap);
    if (s == "block_cols") return copy_out(p.block_cols, dst, cap);
    if (s == "block_cols_offset") return copy_out(p.block_cols_off, dst, cap);
    if (s == "bia_nnz") return copy_out(p.bia_nnz, dst, cap);
    if (s == "bia_origin") return copy_out(p.bia_origin, dst, cap);
    if (s == "bia_end") return copy_out(p.bia_end, dst, cap);
    if (s == "image") return copy_out(p.image, dst, cap);
    if (s == "stream_key") return copy_out(p.s_key, dst, cap);
    if (s == "stream_slab") return copy_out(p.s_slab, dst, cap);
    if (s == "stream_val") return copy_out(p.s_val, dst, cap);
    if (s == "stream_empty_rows") return copy_out(p.empty_rows, dst, cap);
    if (s == "stream_tiles") return copy_out(p.tiles, dst, cap);
    if (s == "stream_tile_lines") return copy_out(p.tile_lines, dst, cap);
    if (s == "stream_items") {
        std::vector<uint32_t> v;
        for (const StreamItem& it : p.items) {
            v.push_back(it.e_begin); v.push_back(it.e_end); v.push_back(it.row_base);
            v.push_back((uint32_t)it.first_partial | ((uint32_t)it.last_partial << 8));
        }
        return copy_out(v, dst, cap);
    }
    if (s == "stream_col_bits") return copy_out(std::vector<int>{p.col_bits}, dst, cap);
    if (s == "stream_panels") return copy_out(std::vector<int>{p.panels, (int)p.tile_items_count}, dst, cap);
    if (s == "stream_panel_items") return copy_out(p.panel_item_ptr, dst, cap);
    if (s == "stream_panel_mask") return copy_out(p.panel_mask, dst, cap);
    if (s == "plan_dtype") return copy_out(std::vector<int>{p.o.dtype}, dst, cap);
    if (s == "stream_value_pos") {   // byte offset in stream_slab of each entry's value
        std::vector<int64_t> v;
        v.reserve(p.s_key.size());
        for (size_t i = 0; i < p.items.size(); ++i)
            for (uint32_t e = p.items[i].e_begin; e < p.items[i].e_end; ++e)
                v.push_back((int64_t)stream_value_offset(p, (uint32_t)i, e));
        return copy_out(v, dst, cap);
    }
    if (s == "policy") {
        // the reference's dispatch decision (cb-spmv-dcu.hpp:1428-1450):
        // unit_values, use_pattern, use_unrolled, use_small_direct,
        // launch_waves, grid, aggregate_coo, pattern_count
        const long P = p.pat_row_ptr.empty() ? 0 : (long)p.pat_row_ptr.size() - 1;
        const bool pat = p.o.pattern_wave && p.unit_values && P > 0 && 2 * P < p.m;
        const bool unr = pat && (long)p.pat_col.size() >= P * 128;
        const bool small = p.o.row_owned && !pat && p.unit_values && p.gather && p.compact && p.nnz < 10000;
        const long waves = (pat && P < 1000) ? 9 : p.o.waves;
        const long grid = pat ? (P + waves - 1) / waves
                              : p.o.row_owned ? (long)p.seg_rows.size() : ((long)p.nnzb + p.o.waves - 1) / p.o.waves;
        return copy_out(std::vector<int>{(int)p.unit_values, (int)pat, (int)unr, (int)small, (int)waves, (int)grid,
                                         (int)(p.nnz >= 10000), (int)P}, dst, cap);
    }
    if (s == "stream_tile_shape") return copy_out(std::vector<int>{p.o.tile_items, p.o.tile_bytes}, dst, cap);
    if (s == "pl_idx" || s == "pl_val") {
        const int bs = p.o.bs;
        std::vector<int32_t> idx;
        std::vector<double> val;
        for (int g = 0; g < p.nnzb; ++g) {
            if (p.nnzpb[g] <= 0) continue;
            const int o = p.bia_origin[g];
            const int t = p.type[g] & kTypeMask;
            if (t == kCOO) {
                for (int q = 0; q < p.coo_count[o]; ++q) idx.push_back((int32_t)p.coo_idx[p.pl_start[o] + q]);
                val.insert(val.end(), p.ent_val.begin() + p.pl_start[o], p.ent_val.begin() + p.pl_start[o] + p.coo_count[o]);
            } else if (t == kCSR) {
                idx.insert(idx.end(), p.csr_rp.begin() + p.csr_rp_start[o], p.csr_rp.begin() + p.csr_rp_start[o] + bs + 1);
                idx.insert(idx.end(), p.csr_col.begin() + p.pl_start[o], p.csr_col.begin() + p.pl_start[o] + p.raw_nnz[o]);
                val.insert(val.end(), p.ent_val.begin() + p.pl_start[o], p.ent_val.begin() + p.pl_start[o] + p.raw_nnz[o]);
            } else {
                const double* d = p.dense_val.data() + (size_t)p.dense_slot[o] * bs * bs;
                val.insert(val.end(), d, d + bs * bs);
            }
        }
        return s == "pl_idx" ? copy_out(idx, dst, cap) : copy_out(val, dst, cap);
    }
    status = CBSPMV_ERR_ARG;
    return -1;
}

}  // namespace cbs
