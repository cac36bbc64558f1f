// Plan serialisation: a built plan (builder output, canonical slab image and
// stream slab) written to one file and read back without rebuilding.  The
// reference rebuilds its GLOBAL_BLOCK from the COO on every run
// (cb-spmv-dcu/src/main.hip.cpp:46-70); a saved plan skips the build.
//
// File layout (little endian): 16-byte magic "CBSPMV-PLAN\0" + u32 format
// version + u32 byte order mark; the build options; the plan's scalars; then
// every array as (u64 element count, u64 element size, bytes).  The reader
// checks magic, version, element sizes and lengths, and fails with
// CBSPMV_ERR_ARG on any mismatch or truncation.  It then re-checks what the
// kernels index without bounds checks, so that a damaged file is refused
// rather than read or written out of bounds on the device: the options,
// every stream item's entry range and base row against the slab and m,
// every key's column (or LDS slot) and row, the tiles' line ranges against
// the line table and the LDS cache, the cut rows' pieces against the items
// and the partial buffer, the empty rows and the panel tables.  The
// transposed plan is not saved (cbspmv_build_transpose rebuilds it), nor a
// scattered-entry split (a loaded plan runs its full stream).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"

namespace cbs {
namespace {

constexpr char kMagic[16] = "CBSPMV-PLAN";
constexpr uint32_t kVersion = 4;
constexpr uint32_t kBom = 0x01020304u;

struct Writer {
    FILE* f;
    bool ok = true;
    void raw(const void* p, size_t n) { ok = ok && (n == 0 || std::fwrite(p, 1, n, f) == n); }
    template <typename T>
    void pod(const T& v) { raw(&v, sizeof(T)); }
    template <typename T>
    void vec(const std::vector<T>& v) {
        pod<uint64_t>(v.size());
        pod<uint64_t>(sizeof(T));
        raw(v.data(), v.size() * sizeof(T));
    }
};

struct Reader {
    FILE* f;
    bool ok = true;
    void raw(void* p, size_t n) { ok = ok && (n == 0 || std::fread(p, 1, n, f) == n); }
    template <typename T>
    void pod(T& v) { raw(&v, sizeof(T)); }
    template <typename T>
    void vec(std::vector<T>& v) {
        uint64_t n = 0, sz = 0;
        pod(n);
        pod(sz);
        if (!ok || sz != sizeof(T) || n > (1ull << 40)) { ok = false; return; }
        v.resize((size_t)n);
        raw(v.data(), (size_t)n * sizeof(T));
    }
};

// The same member list drives writing and reading.
template <typename IO>
void members(IO& io, Plan& p, cbspmv_opts& o) {
    io.pod(o);
    io.pod(p.o);
    io.pod(p.m); io.pod(p.n); io.pod(p.nnz);
    io.pod(p.gb_m); io.pod(p.gb_n); io.pod(p.nnzb);
    io.pod(p.gather); io.pod(p.direct); io.pod(p.compact);
    io.pod(p.super_sparse); io.pod(p.analyzed); io.pod(p.super_ratio);
    io.pod(p.balanced); io.pod(p.sd_before); io.pod(p.sd_after); io.pod(p.build_seconds);
    io.vec(p.brow); io.vec(p.bcol); io.vec(p.nnzpb); io.vec(p.row_ptr); io.vec(p.type);
    io.vec(p.seg_rows); io.vec(p.seg_begin); io.vec(p.seg_end);
    io.vec(p.seg_direct); io.vec(p.seg_uniform); io.vec(p.row_uniform);
    io.vec(p.flat_row_ptr); io.vec(p.flat_col);
    io.vec(p.pat_row_ptr); io.vec(p.pat_col); io.vec(p.pat_out_ptr); io.vec(p.pat_out_rows);
    io.vec(p.block_cols); io.vec(p.block_cols_off);
    io.vec(p.bia_nnz); io.vec(p.bia_origin); io.vec(p.bia_end);
    io.vec(p.pl_start); io.vec(p.raw_nnz); io.vec(p.coo_count); io.vec(p.dense_slot);
    io.vec(p.coo_idx); io.vec(p.ent_val); io.vec(p.csr_col); io.vec(p.csr_rp); io.vec(p.csr_rp_start);
    io.vec(p.dense_val);
    io.vec(p.in_row); io.vec(p.in_col); io.vec(p.in_val);
    io.pod(p.unit_values); io.pod(p.image_ready);
    io.vec(p.image);
    io.pod(p.stream_ready); io.pod(p.col_bits);
    io.vec(p.s_key); io.vec(p.s_val); io.vec(p.s_slab); io.vec(p.items);
    io.vec(p.item_split); io.vec(p.item_split_first); io.vec(p.splits); io.vec(p.empty_rows);
    io.vec(p.tiles); io.vec(p.tile_lines); io.pod(p.partial_slots);
    io.pod(p.panels); io.pod(p.tile_items_count);
    io.vec(p.panel_item_ptr); io.vec(p.panel_mask);
}

void write_offsets(Writer& w, const Plan& p) {
    w.pod<uint64_t>(p.off.size());
    for (const auto& kv : p.off) {
        w.pod<uint64_t>(kv.first.size());
        w.raw(kv.first.data(), kv.first.size());
        w.pod<uint64_t>(kv.second);
    }
}

void read_offsets(Reader& r, Plan& p) {
    uint64_t n = 0;
    r.pod(n);
    if (n > 4096) { r.ok = false; return; }
    p.off.clear();
    for (uint64_t i = 0; i < n && r.ok; ++i) {
        uint64_t len = 0, v = 0;
        r.pod(len);
        if (len > 256) { r.ok = false; return; }
        std::string k(len, '\0');
        r.raw(&k[0], len);
        r.pod(v);
        p.off[k] = (size_t)v;
    }
}

// What the stream kernels index without bounds checks, re-checked on load.
bool validate(const Plan& h, std::string& why) {
    auto bad = [&why](const char* w) { why = w; return false; };
    const Options& o = h.o;
    if (o.dtype < CBSPMV_F64 || o.dtype > CBSPMV_BF16 || (o.accum != CBSPMV_ACC_NATIVE && o.accum != CBSPMV_ACC_COMPENSATED))
        return bad("value type or accumulation");
    if (o.bs != 2 && o.bs != 4 && o.bs != 8 && o.bs != 16 && o.bs != 32 && o.bs != 64) return bad("block size");
    if (o.tile_items < 1 || o.tile_items > kMaxTileItems || o.tile_bytes < 4096 || o.tile_bytes > kMaxTileBytes)
        return bad("stream tile shape");
    if (h.panels != 1 && h.panels != 2 && h.panels != 4 && h.panels != 8) return bad("panel count");
    if (h.panels > 1 && o.tile_items > kMaxTileItems / 2) return bad("panels with more than 8 tile waves");
    if (h.col_bits < 1 || h.col_bits > 30) return bad("column bits");
    if (h.m < 0 || h.n < 0 || (int64_t)h.n > (1ll << h.col_bits)) return bad("matrix shape");
    const uint32_t m = (uint32_t)h.m, n = (uint32_t)h.n;
    const int epl = stream_lane_entries(o.dtype);
    const size_t eb = 4 + (size_t)stream_val_bytes(o.dtype);
    const uint32_t slots = (uint32_t)tile_slots(o.tile_bytes, o.dtype);
    const uint32_t cmask = (1u << h.col_bits) - 1u, omask = (1u << (31 - h.col_bits)) - 1u;
    const size_t NI = h.items.size();
    if (h.tile_items_count > NI || h.item_split.size() != NI || h.item_split_first.size() != NI)
        return bad("item tables");
    if (h.tiles.size() % 2 || (uint64_t)(h.tiles.size() / 2) * (uint64_t)o.tile_items < h.tile_items_count)
        return bad("tile table");
    for (size_t t = 0; t < h.tiles.size() / 2; ++t) {
        const uint64_t first = h.tiles[2 * t], cnt = h.tiles[2 * t + 1];
        if (first + cnt > h.tile_lines.size() || cnt * kLineElems > (uint64_t)slots - 1u) return bad("tile lines");
    }
    for (size_t i = 0; i < NI; ++i) {
        const StreamItem& it = h.items[i];
        if (it.e_begin > it.e_end || it.e_begin % epl || it.e_end % epl || (size_t)it.e_end * eb > h.s_slab.size())
            return bad("item entry range");
        if (it.e_end > it.e_begin && it.row_base >= m) return bad("item base row");
        if (it.first_partial > 1 || it.last_partial > 1) return bad("item flags");
        if ((it.first_partial && h.item_split_first[i] >= h.splits.size()) ||
            (it.last_partial && h.item_split[i] >= h.splits.size()))
            return bad("item split ids");
        const uint32_t chunk = (uint32_t)(64 * epl);
        for (uint32_t e = it.e_begin; e < it.e_end; ++e) {
            const uint32_t e0 = it.e_begin + (e - it.e_begin) / chunk * chunk;
            const uint32_t G = std::min<uint32_t>(64u, (it.e_end - e0) / (uint32_t)epl);
            const uint32_t l = (e - e0) / (uint32_t)epl, j = (e - e0) % (uint32_t)epl;
            const size_t at = (size_t)e0 * eb + (size_t)(j * 4 / 16) * G * 16 + (size_t)l * 16 + (j * 4) % 16;
            uint32_t k;
            std::memcpy(&k, h.s_slab.data() + at, 4);
            const uint32_t low = k & cmask;
            if (k & kKeyCached) {
                // panel items read no LDS (such keys are padding there)
                if (i < h.tile_items_count && low >= slots) return bad("key LDS slot");
            } else if (low >= n) {
                return bad("key column");
            }
            if ((uint64_t)it.row_base + ((k >> h.col_bits) & omask) >= m) return bad("key row");
        }
    }
    for (const StreamSplit& sp : h.splits) {
        if (sp.row >= m || sp.pieces < 1 || (uint64_t)sp.first_item + sp.pieces > NI ||
            (uint64_t)sp.partial_base + sp.pieces > h.partial_slots)
            return bad("cut rows");
    }
    for (uint32_t r : h.empty_rows)
        if (r >= m) return bad("empty rows");
    if (h.panels > 1) {
        if (h.panel_item_ptr.size() != (size_t)h.panels + 1 || h.panel_item_ptr[0] != h.tile_items_count ||
            h.panel_item_ptr.back() > NI || h.panel_mask.size() != (size_t)m)
            return bad("panel tables");
        for (size_t q = 0; q + 1 < h.panel_item_ptr.size(); ++q)
            if (h.panel_item_ptr[q] > h.panel_item_ptr[q + 1]) return bad("panel tables");
    } else if (h.tile_items_count != NI) {
        return bad("items outside the tiles");
    }
    return true;
}

}  // namespace
}  // namespace cbs

using cbs::fail;

extern "C" {

int cbspmv_save(const cbspmv_plan* plan, const char* path) {
    if (!plan || !path) return fail(CBSPMV_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CBSPMV_ERR_ARG, std::string("cannot open for writing: ") + path);
    cbs::Writer w{f};
    w.raw(cbs::kMagic, sizeof(cbs::kMagic));
    w.pod(cbs::kVersion);
    w.pod(cbs::kBom);
    cbspmv_opts o = plan->opts;
    cbs::members(w, const_cast<cbs::Plan&>(plan->host), o);
    cbs::write_offsets(w, plan->host);
    const bool ok = w.ok && std::fclose(f) == 0;
    if (!ok) return fail(CBSPMV_ERR_ARG, std::string("write failed: ") + path);
    return CBSPMV_OK;
}

int cbspmv_load(const char* path, cbspmv_plan** plan) {
    if (!path || !plan) return fail(CBSPMV_ERR_ARG, "null argument");
    *plan = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CBSPMV_ERR_ARG, std::string("cannot open: ") + path);
    std::unique_ptr<cbspmv_plan> p(new (std::nothrow) cbspmv_plan());
    if (!p) { std::fclose(f); return fail(CBSPMV_ERR_NOMEM, "out of host memory"); }
    cbs::Reader r{f};
    char magic[16];
    uint32_t version = 0, bom = 0;
    r.raw(magic, sizeof(magic));
    r.pod(version);
    r.pod(bom);
    if (!r.ok || std::memcmp(magic, cbs::kMagic, sizeof(magic)) != 0 || version != cbs::kVersion || bom != cbs::kBom) {
        std::fclose(f);
        return fail(CBSPMV_ERR_ARG, std::string("not a plan file of this version: ") + path);
    }
    try {
        cbs::members(r, p->host, p->opts);
        cbs::read_offsets(r, p->host);
    } catch (const std::bad_alloc&) {
        std::fclose(f);
        return fail(CBSPMV_ERR_NOMEM, "out of host memory while reading the plan");
    }
    char extra;
    const bool at_end = std::fread(&extra, 1, 1, f) == 0;
    std::fclose(f);
    if (!r.ok || !at_end) return fail(CBSPMV_ERR_ARG, std::string("truncated or malformed plan file: ") + path);
    const cbs::Plan& h = p->host;
    if (h.m < 0 || h.n < 0 || h.nnz < 0 || (int64_t)h.in_row.size() != h.nnz || !h.stream_ready)
        return fail(CBSPMV_ERR_ARG, std::string("inconsistent plan file: ") + path);
    std::string why;
    if (!cbs::validate(h, why)) return fail(CBSPMV_ERR_ARG, std::string("invalid plan file (") + why + "): " + path);
    p->schedule = cbs::choose_schedule(*p);
    *plan = p.release();
    return CBSPMV_OK;
}

}  // extern "C"
