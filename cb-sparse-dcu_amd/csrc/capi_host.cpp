// C ABI, host half: options, plan build, statistics and parity export.  The
// device half is runtime.hip (upload, SpMV, bench) and adjoint.hip
// (transposed plan, backward, value replacement).
#include <cstring>
#include <new>
#include <string>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"

namespace cbs {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace cbs

using cbs::fail;

namespace cbs {
int build_plan_object(const cbspmv_opts& o, int m, int n, int nnz, const int* row, const int* col,
                      const double* val, bool canonical, cbspmv_plan** out) {
    *out = nullptr;
    const int bs = o.block_size;
    if (bs != 2 && bs != 4 && bs != 8 && bs != 16 && bs != 32 && bs != 64)
        return fail(CBSPMV_ERR_ARG, "block_size must be 2, 4, 8, 16, 32 or 64");
    if (o.waves_per_block < 1 || o.waves_per_block > 16)
        return fail(CBSPMV_ERR_ARG, "waves_per_block must be in 1..16");
    if (o.dense_sixteenths < 1 || o.dense_sixteenths > 16)
        return fail(CBSPMV_ERR_ARG, "dense_sixteenths must be in 1..16");
    if (o.builder != CBSPMV_BUILDER_GATHER && o.builder != CBSPMV_BUILDER_COMP)
        return fail(CBSPMV_ERR_ARG, "unknown builder");
    if (o.value_dtype < CBSPMV_F64 || o.value_dtype > CBSPMV_BF16)
        return fail(CBSPMV_ERR_ARG, "unknown value dtype");
    if (o.schedule < CBSPMV_SCHED_AUTO || o.schedule > CBSPMV_SCHED_STREAM)
        return fail(CBSPMV_ERR_ARG, "unknown schedule");
    if (o.stream_panels != 0 && o.stream_panels != 1 && o.stream_panels != 2 && o.stream_panels != 4 &&
        o.stream_panels != 8)
        return fail(CBSPMV_ERR_ARG, "stream_panels must be 0, 1, 2, 4 or 8");
    if (o.stream_tiles < 0 || o.stream_tiles > 65536) return fail(CBSPMV_ERR_ARG, "stream_tiles must be in 0..65536");
    if (o.stream_split < CBSPMV_SPLIT_AUTO || o.stream_split > CBSPMV_SPLIT_ON)
        return fail(CBSPMV_ERR_ARG, "stream_split must be 0 (auto), 1 (off) or 2 (on)");
    // stream keys hold column + row-in-strip + 2 flag bits in 32 bits
    if ((long long)n > (1ll << (30 - coord_bits(bs))))
        return fail(CBSPMV_ERR_UNSUPPORTED, "n above 2^(30 - log2 block_size) is not supported");

    cbspmv_plan* p = new (std::nothrow) cbspmv_plan();
    if (!p) return fail(CBSPMV_ERR_NOMEM, "out of host memory");
    p->opts = o;
    Plan& h = p->host;
    h.o.bs = bs;
    h.o.waves = o.waves_per_block;
    h.o.row_owned = o.row_owned ? 1 : 0;
    h.o.pattern_wave = o.pattern_wave ? 1 : 0;
    h.o.force_coo64 = o.force_coo64 ? 1 : 0;
    h.o.dense_sixteenths = o.dense_sixteenths;
    h.o.builder = o.builder;
    h.o.dtype = o.value_dtype;
    h.o.accum = o.accumulate;
    h.o.schedule = canonical ? o.schedule : CBSPMV_SCHED_STREAM;
    h.o.item_nnz = o.stream_item_nnz;
    h.o.panels = o.stream_panels;
    h.o.tiles = o.stream_tiles > 0 ? o.stream_tiles : device_shape().cus;
    if (o.stream_tile_waves) h.o.tile_items = o.stream_tile_waves;
    else if (h.o.panels <= 1) h.o.tile_items = kWideTileItemsOpt;
    if (o.stream_tile_kib) h.o.tile_bytes = o.stream_tile_kib * 1024;
    if (h.o.panels > 1 && h.o.tile_items > kMaxTileItems / 2) {
        delete p;
        return fail(CBSPMV_ERR_ARG, "stream_panels > 1 needs stream_tile_waves <= 8 (a workgroup adds as many panel waves)");
    }
    if (h.o.tile_items < 1 || h.o.tile_items > kMaxTileItems || h.o.tile_bytes < 4096 || h.o.tile_bytes > kMaxTileBytes) {
        delete p;
        return fail(CBSPMV_ERR_ARG, "stream tile shape out of range (1..16 waves, 4..156 KiB)");
    }
    std::string err;
    int rc;
    try {
        rc = build_plan(h, m, n, nnz, row, col, val, err);
        if (rc == CBSPMV_OK) {
            if (canonical) pack_canonical(h);
            build_stream(h);
        }
    } catch (const std::bad_alloc&) {
        delete p;
        return fail(CBSPMV_ERR_NOMEM, "out of host memory while building the plan");
    }
    if (rc != CBSPMV_OK) {
        delete p;
        return fail(rc, err);
    }
    p->schedule = canonical ? choose_schedule(*p) : CBSPMV_SCHED_STREAM;
    if (canonical) {
        try {
            rc = build_split(*p, nnz, row, col, val);
        } catch (const std::bad_alloc&) {
            rc = fail(CBSPMV_ERR_NOMEM, "out of host memory while splitting the plan");
        }
        if (rc != CBSPMV_OK) {
            cbspmv_destroy(p);
            return rc;
        }
    }
    *out = p;
    return CBSPMV_OK;
}
}  // namespace cbs

extern "C" {

void cbspmv_default_opts(cbspmv_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->block_size = 16;        // cb-spmv-dcu/src/Makefile:4
    o->waves_per_block = 4;    // Makefile:5
    o->row_owned = 1;          // Makefile:6
    o->pattern_wave = 1;       // Makefile:7
    o->force_coo64 = 0;        // macros.h:89-91
    o->dense_sixteenths = 15;  // macros.h:92-94
    o->builder = CBSPMV_BUILDER_GATHER;
    o->value_dtype = CBSPMV_F64;
    o->accumulate = CBSPMV_ACC_NATIVE;
    o->schedule = CBSPMV_SCHED_AUTO;
    o->stream_item_nnz = 0;
}

int cbspmv_build(int m, int n, int nnz, const int* row, const int* col, const double* val,
                 const cbspmv_opts* opts, cbspmv_plan** plan) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "plan output pointer is null");
    *plan = nullptr;
    if (m < 0 || n < 0 || nnz < 0) return fail(CBSPMV_ERR_ARG, "negative matrix size");
    if (nnz > 0 && (!row || !col || !val)) return fail(CBSPMV_ERR_ARG, "null COO array");
    cbspmv_opts d;
    cbspmv_default_opts(&d);
    return cbs::build_plan_object(opts ? *opts : d, m, n, nnz, row, col, val, true, plan);
}

int cbspmv_plan_stats(const cbspmv_plan* plan, cbspmv_stats* s) {
    if (!plan || !s) return fail(CBSPMV_ERR_ARG, "null argument");
    const cbs::Plan& h = plan->host;
    std::memset(s, 0, sizeof(*s));
    s->m = h.m; s->n = h.n; s->nnz = h.nnz;
    s->gb_m = h.gb_m; s->gb_n = h.gb_n; s->gb_nnzb = h.nnzb;
    s->gather_flag = h.gather; s->direct_coo_columns = h.direct; s->compact_coo_indices = h.compact;
    s->super_sparse_blocks = h.super_sparse; s->analyzed_blocks = h.analyzed;
    s->super_sparse_ratio = h.super_ratio;
    s->load_balanced = h.balanced; s->load_stddev_before = h.sd_before; s->load_stddev_after = h.sd_after;
    for (int g = 0; g < h.nnzb; ++g) {
        if (h.nnzpb[g] <= 0) continue;
        const int t = h.type[g] & cbs::kTypeMask;
        s->coo_blocks += t == cbs::kCOO;
        s->csr_blocks += t == cbs::kCSR;
        s->dense_blocks += t == cbs::kDENSE;
        s->uniform_blocks += (h.type[g] & cbs::kUniform) != 0;
    }
    s->segments = (int)h.seg_rows.size();
    s->patterns = h.pat_row_ptr.empty() ? 0 : (int)h.pat_row_ptr.size() - 1;
    s->unit_values = h.unit_values;
    auto it = h.off.find("total");
    s->slab_bytes = it == h.off.end() ? 0 : (long long)it->second;
    s->stream_bytes = cbs::stream_bytes_per_spmv(*plan);
    s->restore_entries = (long long)h.block_cols.size();
    s->build_seconds = h.build_seconds;
    s->split_entries = (long long)plan->sweep.key.size();
    s->split_blocks = (int)plan->sweep.blocks();
    return CBSPMV_OK;
}

int cbspmv_export(const cbspmv_plan* plan, const char* name, void* dst, size_t* count) {
    if (!plan || !name || !count) return fail(CBSPMV_ERR_ARG, "null argument");
    if (std::strncmp(name, "transpose.", 10) == 0) {
        if (!plan->transpose) return fail(CBSPMV_ERR_STATE, "no transposed plan");
        return cbspmv_export(plan->transpose, name + 10, dst, count);
    }
    if (std::strncmp(name, "near.", 5) == 0) {
        if (!plan->near) return fail(CBSPMV_ERR_STATE, "the plan has no scattered-entry split");
        return cbspmv_export(plan->near, name + 5, dst, count);
    }
    if (std::strncmp(name, "split_", 6) == 0) {
        const cbs::SweepPart& sw = plan->sweep;
        const uint32_t shape[3] = {(uint32_t)sw.col_bits, (uint32_t)cbs::kSweepWaves, (uint32_t)sw.max_rows};
        const void* src = nullptr;
        size_t n = 0, esz = 4;
        const char* k = name + 6;
        if (!std::strcmp(k, "key")) { src = sw.key.data(); n = sw.key.size(); }
        else if (!std::strcmp(k, "val")) { src = sw.val.data(); n = sw.val.size(); esz = 8; }
        else if (!std::strcmp(k, "src")) { src = sw.src.data(); n = sw.src.size(); }
        else if (!std::strcmp(k, "wave_off")) { src = sw.wave_off.data(); n = sw.wave_off.size(); }
        else if (!std::strcmp(k, "blk_row")) { src = sw.blk_row.data(); n = sw.blk_row.size(); }
        else if (!std::strcmp(k, "blk_span")) { src = sw.blk_span.data(); n = sw.blk_span.size(); }
        else if (!std::strcmp(k, "shape")) { src = shape; n = plan->near ? 3 : 0; }
        else return fail(CBSPMV_ERR_ARG, std::string("unknown export name: ") + name);
        if (dst && *count >= n && n) std::memcpy(dst, src, n * esz);
        *count = n;
        return CBSPMV_OK;
    }
    if (std::strcmp(name, "grad_slot") == 0 || std::strcmp(name, "slot_of") == 0) {
        const std::vector<uint32_t>& v = name[0] == 'g' ? plan->grad_slot : plan->slot_of;
        if (dst && *count >= v.size() && !v.empty()) std::memcpy(dst, v.data(), v.size() * 4);
        *count = v.size();
        return CBSPMV_OK;
    }
    int status = CBSPMV_OK;
    const long c = cbs::export_array(plan->host, name, dst, *count, status);
    if (status != CBSPMV_OK) return fail(status, std::string("unknown export name: ") + name);
    *count = (size_t)c;
    return CBSPMV_OK;
}

long long cbspmv_layout_offset(const cbspmv_plan* plan, const char* name) {
    if (!plan || !name) return -1;
    auto it = plan->host.off.find(name);
    return it == plan->host.off.end() ? -1 : (long long)it->second;
}

const char* cbspmv_last_error(void) { return cbs::g_last_error.c_str(); }
const char* cbspmv_version(void) { return "cbspmv-mi355x 0.1 (gfx950)"; }

}  // extern "C"
