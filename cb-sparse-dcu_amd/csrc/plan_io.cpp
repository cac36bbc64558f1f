// Plan serialisation: a built plan (builder output, canonical slab image and
// stream slab) written to one file and read back without rebuilding.  The
// reference rebuilds its GLOBAL_BLOCK from the COO on every run
// (cb-spmv-dcu/src/main.hip.cpp:46-70); a saved plan skips the build.
//
// File layout (little endian): 16-byte magic "CBSPMV-PLAN\0" + u32 format
// version + u32 byte order mark; the build options; the plan's scalars; then
// every array as (u64 element count, u64 element size, bytes).  The reader
// checks magic, version, element sizes and lengths, and fails with
// CBSPMV_ERR_ARG on any mismatch or truncation.  The transposed plan is not
// saved (cbspmv_build_transpose rebuilds it).
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
constexpr uint32_t kVersion = 2;
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
    p->schedule = cbs::choose_schedule(*p);
    *plan = p.release();
    return CBSPMV_OK;
}

}  // extern "C"
