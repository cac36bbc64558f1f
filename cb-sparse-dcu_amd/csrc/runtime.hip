// C ABI, device half: upload, SpMV dispatch and the cb_spmv() bench protocol.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "blocked.hpp"
#include "cb_runtime.hpp"
#include "kernels.hpp"
#include "stream.hpp"

namespace cbs {

int hip_fail(hipError_t e, const char* what) {
    return fail(CBSPMV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void release(DeviceBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

int upload_bytes(DeviceBuf& b, const void* src, size_t bytes, hipStream_t s) {
    release(b);
    if (bytes == 0) return CBSPMV_OK;
    const hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) return fail(CBSPMV_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    b.bytes = bytes;
    if (src) CBS_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
    else CBS_HIP(hipMemsetAsync(b.ptr, 0, bytes, s));
    return CBSPMV_OK;
}

size_t vec_elem_bytes(const Plan& h) { return h.o.dtype == CBSPMV_F64 ? 8 : 4; }

DeviceShape device_shape() {
    static std::mutex mu;
    static std::vector<DeviceShape> known;
    DeviceShape def{kStreamCUs, kXcds};
    int count = 0, dev = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return def;
    }
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= known.size()) known.resize((size_t)dev + 1, DeviceShape{0, 0});
    if (known[dev].cus == 0) {
        int cus = 0, xcds = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = def.cus;
        if (hipDeviceGetAttribute(&xcds, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || xcds <= 0) xcds = 1;
        known[dev] = DeviceShape{cus, std::min(xcds, kMaxXcds)};
    }
    return known[dev];
}

void release_device(DevicePlan& d) {
    for (DeviceBuf* b : {&d.slab, &d.zero_strips, &d.key, &d.val, &d.items, &d.item_split,
                         &d.item_split_first, &d.splits, &d.partial, &d.counters, &d.empty, &d.tiles, &d.tile_lines,
                         &d.sv_pos, &d.sv_ptr, &d.sv_src, &d.ss_pos, &d.ss_ptr, &d.ss_src, &d.grad_slot,
                         &d.dvals_stored, &d.in_row, &d.in_col,
                         &d.panel_items, &d.panel_mask, &d.panel_out,
                         &d.sw_key, &d.sw_val, &d.sw_wave_off, &d.sw_blk_row, &d.sw_blk_span})
        release(*b);
    d.uploaded = false;
    d.sv_groups = 0;
    d.ss_groups = 0;
}

int choose_schedule(const cbspmv_plan& p) {
    const Plan& h = p.host;
    if (h.o.dtype != CBSPMV_F64) return CBSPMV_SCHED_STREAM;
    if (h.o.schedule != CBSPMV_SCHED_AUTO) return h.o.schedule;
    // unit-valued plans with repeated row patterns: the pattern-reuse schedule
    const long P = h.pat_row_ptr.empty() ? 0 : (long)h.pat_row_ptr.size() - 1;
    if (h.o.pattern_wave && h.unit_values && P > 0 && 2 * P < h.m) return CBSPMV_SCHED_BLOCKED;
    return CBSPMV_SCHED_STREAM;
}

long long stream_bytes_per_spmv(const cbspmv_plan& p) {
    const Plan& h = p.host;
    return (long long)h.s_slab.size() + (long long)h.items.size() * (long long)sizeof(StreamItem) +
           (long long)h.tile_lines.size() * 4 + (long long)h.tiles.size() * 4 +
           (long long)h.m * (long long)vec_elem_bytes(h) + (long long)h.n * (long long)vec_elem_bytes(h);
}

int upload_plan(cbspmv_plan& p, hipStream_t s) {
    Plan& h = p.host;
    DevicePlan& d = p.dev;
    int dev = 0;
    CBS_HIP(hipGetDevice(&dev));
    if (d.uploaded && d.device == dev) return CBSPMV_OK;
    release_device(d);
    d.device = dev;
    int rc;
    // canonical slab + blocked-schedule zero list (F64 plans only)
    if (h.o.dtype == CBSPMV_F64 && h.nnz > 0) {
        if ((rc = upload_vec(d.slab, h.image, s))) return rc;
        std::vector<uint32_t> zs;
        d.zero_all = !h.o.row_owned;
        if (h.o.row_owned) {
            std::vector<uint8_t> covered(h.gb_m, 0), direct(h.gb_m, 0);
            for (size_t q = 0; q < h.seg_rows.size(); ++q) {
                covered[h.seg_rows[q]] = 1;
                if (h.seg_direct[q]) direct[h.seg_rows[q]] = 1;
            }
            for (int b = 0; b < h.gb_m; ++b)
                if (!covered[b] || !direct[b]) zs.push_back((uint32_t)b);
        }
        d.n_zero_strips = (uint32_t)zs.size();
        if ((rc = upload_vec(d.zero_strips, zs, s))) return rc;
    }
    // stream slab
    if ((rc = upload_vec(d.key, h.s_slab, s))) return rc;
    if ((rc = upload_vec(d.items, h.items, s))) return rc;
    if ((rc = upload_vec(d.item_split, h.item_split, s))) return rc;
    if ((rc = upload_vec(d.item_split_first, h.item_split_first, s))) return rc;
    if ((rc = upload_vec(d.splits, h.splits, s))) return rc;
    // compensated plans keep (hi, lo) per row piece
    const size_t parts = h.o.accum == CBSPMV_ACC_COMPENSATED && h.o.dtype != CBSPMV_F64 ? 2 : 1;
    if ((rc = upload_bytes(d.partial, nullptr, (size_t)h.partial_slots * parts * vec_elem_bytes(h), s))) return rc;
    if ((rc = upload_bytes(d.counters, nullptr, h.splits.size() * 4, s))) return rc;
    if ((rc = upload_vec(d.empty, h.empty_rows, s))) return rc;
    if ((rc = upload_vec(d.tiles, h.tiles, s))) return rc;
    if ((rc = upload_vec(d.tile_lines, h.tile_lines, s))) return rc;
    if (h.panels > 1) {
        if ((rc = upload_vec(d.panel_items, h.panel_item_ptr, s))) return rc;
        if ((rc = upload_vec(d.panel_mask, h.panel_mask, s))) return rc;
        if ((rc = upload_bytes(d.panel_out, nullptr, (size_t)h.panels * h.m * vec_elem_bytes(h), s))) return rc;
    }
    if (p.near) {   // scattered-entry split: the other entries' plan and the sweep arrays
        if ((rc = upload_plan(*p.near, s))) return rc;
        const SweepPart& sw = p.sweep;
        if ((rc = upload_vec(d.sw_key, sw.key, s))) return rc;
        if (h.o.dtype == CBSPMV_F64) {
            if ((rc = upload_vec(d.sw_val, sw.val, s))) return rc;
        } else if (h.o.dtype == CBSPMV_F32) {
            const std::vector<float> v32(sw.val.begin(), sw.val.end());
            if ((rc = upload_vec(d.sw_val, v32, s))) return rc;
            CBS_HIP(hipStreamSynchronize(s));   // v32 is freed on return
        } else {   // BF16: rounded as the stream stores them
            std::vector<uint16_t> v16(sw.val.size());
            for (size_t i = 0; i < v16.size(); ++i) v16[i] = bf16_bits(sw.val[i]);
            if ((rc = upload_vec(d.sw_val, v16, s))) return rc;
            CBS_HIP(hipStreamSynchronize(s));   // v16 is freed on return
        }
        if ((rc = upload_vec(d.sw_wave_off, sw.wave_off, s))) return rc;
        if ((rc = upload_vec(d.sw_blk_row, sw.blk_row, s))) return rc;
        if ((rc = upload_vec(d.sw_blk_span, sw.blk_span, s))) return rc;
    }
    CBS_HIP(hipStreamSynchronize(s));
    d.uploaded = true;
    return CBSPMV_OK;
}

int stream_grid(const Plan& h) {
    const long long tiles = (long long)h.tiles.size() / 2;
    return (int)std::max<long long>(8, (tiles + 7) / 8 * 8);
}

// The stream kernels address x and y through buffer resources of at most
// 2 GiB each.
bool stream_vectors_fit(const Plan& h) {
    const size_t elem = h.o.dtype == CBSPMV_F64 ? 8 : 4;
    return (size_t)std::max(h.m, h.n) * elem < (size_t(1) << 31);
}

int run_stream(cbspmv_plan& p, const void* x, void* y, hipStream_t s, const void* xrow, void* dvals) {
    const Plan& h = p.host;
    const DevicePlan& d = p.dev;
    if (!stream_vectors_fit(h)) return fail(CBSPMV_ERR_UNSUPPORTED, "the stream schedule needs x and y below 2 GiB each");
    const int grid = stream_grid(h);
    hipError_t e;
    auto fill = [&](auto& a) {
        a.slab = static_cast<const uint8_t*>(d.key.ptr);
        a.items = static_cast<const StreamItem*>(d.items.ptr);
        a.item_split = static_cast<const uint32_t*>(d.item_split.ptr);
        a.item_split_first = static_cast<const uint32_t*>(d.item_split_first.ptr);
        a.splits = static_cast<const StreamSplit*>(d.splits.ptr);
        a.counters = static_cast<uint32_t*>(d.counters.ptr);
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
        a.panels = (uint32_t)h.panels;
        a.panel_item_ptr = static_cast<const uint32_t*>(d.panel_items.ptr);
        a.paced = p.paced ? 1u : 0u;
    };
    const bool panels = h.panels > 1;
    const uint8_t* mask = static_cast<const uint8_t*>(d.panel_mask.ptr);
    if (h.o.dtype == CBSPMV_F64) {
        StreamArgs<double, double> a{};
        fill(a);
        a.partial = static_cast<double*>(d.partial.ptr);
        a.x = static_cast<const double*>(x);
        a.y = static_cast<double*>(y);
        a.xrow = static_cast<const double*>(xrow);
        a.dvals = static_cast<double*>(dvals);
        a.panel_out = static_cast<double*>(d.panel_out.ptr);
        e = launch_stream(a, grid, dvals != nullptr, s);
        if (e == hipSuccess && panels) e = launch_panel_sums(a.y, a.panel_out, mask, (uint32_t)h.m, a.panels, s);
    } else if (h.o.dtype == CBSPMV_F32) {
        StreamArgs<float, float> a{};
        fill(a);
        a.partial = static_cast<float*>(d.partial.ptr);
        a.x = static_cast<const float*>(x);
        a.y = static_cast<float*>(y);
        a.xrow = static_cast<const float*>(xrow);
        a.dvals = static_cast<float*>(dvals);
        a.panel_out = static_cast<float*>(d.panel_out.ptr);
        e = launch_stream(a, grid, h.o.accum == CBSPMV_ACC_COMPENSATED, dvals != nullptr, s);
        if (e == hipSuccess && panels) e = launch_panel_sums(a.y, a.panel_out, mask, (uint32_t)h.m, a.panels, s);
    } else {
        StreamArgs<dev::bf16_t, float> a{};
        fill(a);
        a.partial = static_cast<float*>(d.partial.ptr);
        a.x = static_cast<const float*>(x);
        a.y = static_cast<float*>(y);
        a.xrow = static_cast<const float*>(xrow);
        a.dvals = static_cast<float*>(dvals);
        a.panel_out = static_cast<float*>(d.panel_out.ptr);
        e = launch_stream(a, grid, h.o.accum == CBSPMV_ACC_COMPENSATED, dvals != nullptr, s);
        if (e == hipSuccess && panels) e = launch_panel_sums(a.y, a.panel_out, mask, (uint32_t)h.m, a.panels, s);
    }
    if (e != hipSuccess) return hip_fail(e, "stream schedule launch");
    return CBSPMV_OK;
}

int run_blocked(cbspmv_plan& p, const void* x, void* y, hipStream_t s) {
    const Plan& h = p.host;
    const DevicePlan& d = p.dev;
    if (h.o.dtype != CBSPMV_F64) return fail(CBSPMV_ERR_UNSUPPORTED, "the blocked schedule is FP64 only");
    const uint8_t* base = static_cast<const uint8_t*>(d.slab.ptr);
    auto at = [&](const char* name) -> const uint8_t* {
        auto it = h.off.find(name);
        return it == h.off.end() ? nullptr : base + it->second;
    };
    double* yd = static_cast<double*>(y);
    const double* xd = static_cast<const double*>(x);
    // policy (cb-spmv-dcu.hpp:1428-1450)
    const long P = h.pat_row_ptr.empty() ? 0 : (long)h.pat_row_ptr.size() - 1;
    const bool use_pattern = h.o.pattern_wave && h.unit_values && P > 0 && 2 * P < h.m;
    hipError_t e;
    if (use_pattern) {
        PatternArgs a{};
        a.row_ptr = reinterpret_cast<const int*>(at("pattern_row_ptr"));
        a.col = reinterpret_cast<const int*>(at("pattern_col_idx"));
        a.out_ptr = reinterpret_cast<const int*>(at("pattern_output_ptr"));
        a.out_rows = reinterpret_cast<const int*>(at("pattern_output_rows"));
        a.x = xd;
        a.y = yd;
        a.patterns = (int)P;
        const bool unrolled = (long)h.pat_col.size() >= P * 128;
        const int waves = P < 1000 ? 9 : h.o.waves;
        e = launch_pattern(a, unrolled, waves, s);
        if (e != hipSuccess) return hip_fail(e, "pattern launch");
        return CBSPMV_OK;
    }
    if (d.zero_all) e = launch_zero_rows(yd, h.m, nullptr, 0, h.o.bs, s);
    else e = launch_zero_rows(yd, h.m, static_cast<const uint32_t*>(d.zero_strips.ptr), d.n_zero_strips, h.o.bs, s);
    if (e != hipSuccess) return hip_fail(e, "zero launch");
    BlockedArgs a{};
    a.slab = base;
    a.brow = reinterpret_cast<const int*>(at("gb_row_idx"));
    a.bcol = reinterpret_cast<const int*>(at("gb_col_idx"));
    a.nnzpb = reinterpret_cast<const int*>(at("nnz_per_blk"));
    a.type = at("type_per_blk");
    a.addr = reinterpret_cast<const uint64_t*>(at("block_addr"));
    a.seg_rows = reinterpret_cast<const int*>(at("segment_rows"));
    a.seg_begin = reinterpret_cast<const int*>(at("segment_begin"));
    a.seg_end = reinterpret_cast<const int*>(at("segment_end"));
    a.seg_direct = at("segment_direct");
    a.seg_uniform = at("segment_uniform");
    a.block_cols = reinterpret_cast<const int*>(at("block_cols"));
    a.block_cols_off = reinterpret_cast<const int*>(at("block_cols_offset"));
    a.x = xd;
    a.y = yd;
    a.m = h.m; a.n = h.n; a.nnzb = h.nnzb; a.waves = h.o.waves;
    a.gather = h.gather; a.direct = h.direct; a.compact = h.compact; a.unit = h.unit_values;
    a.row_owned = h.o.row_owned; a.aggregate = h.nnz >= 10000;
    const int grid = h.o.row_owned ? (int)h.seg_rows.size() : (h.nnzb + h.o.waves - 1) / h.o.waves;
    if (grid == 0) return CBSPMV_OK;
    e = launch_blocked(h.o.bs, a, grid, s);
    if (e != hipSuccess) return hip_fail(e, "blocked launch");
    return CBSPMV_OK;
}

int run_spmv(cbspmv_plan& p, const void* x, void* y, hipStream_t s) {
    const Plan& h = p.host;
    if (h.m == 0) return CBSPMV_OK;
    if (h.nnz == 0) {
        const hipError_t e = launch_zero_words(y, (size_t)h.m * vec_elem_bytes(h) / 4, s);
        return e == hipSuccess ? CBSPMV_OK : hip_fail(e, "zero launch");
    }
    int sched = p.forced != CBSPMV_SCHED_AUTO ? p.forced : p.schedule;
    if (p.values_replaced) sched = CBSPMV_SCHED_STREAM;   // only the stream slab takes new values
    if (sched == CBSPMV_SCHED_BLOCKED) return run_blocked(p, x, y, s);
    if (p.near) {   // scattered-entry split: y = the other entries' product, then y += the sweep
        const int rc = run_spmv(*p.near, x, y, s);
        return rc != CBSPMV_OK ? rc : run_sweep(p, x, y, s);
    }
    return run_stream(p, x, y, s);
}

}  // namespace cbs

using cbs::fail;

extern "C" {

void cbspmv_destroy(cbspmv_plan* plan) {
    if (!plan) return;
    if (plan->transpose) cbspmv_destroy(plan->transpose);
    if (plan->near) cbspmv_destroy(plan->near);
    cbs::release_device(plan->dev);
    delete plan;
}

int cbspmv_upload(cbspmv_plan* plan, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int rc = cbs::upload_plan(*plan, s);
    if (rc || !plan->transpose) return rc;
    return cbs::upload_plan(*plan->transpose, s);
}

int cbspmv_spmv(cbspmv_plan* plan, const void* d_x, void* d_y, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (!plan->dev.uploaded) return fail(CBSPMV_ERR_STATE, "plan not uploaded (call cbspmv_upload first)");
    if ((plan->host.m > 0 && !d_y) || (plan->host.nnz > 0 && !d_x)) return fail(CBSPMV_ERR_ARG, "null vector");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != plan->dev.device)
        return fail(CBSPMV_ERR_STATE, "current device is not the one the plan was uploaded to");
    return cbs::run_spmv(*plan, d_x, d_y, static_cast<hipStream_t>(stream));
}

int cbspmv_set_schedule(cbspmv_plan* plan, int schedule) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (schedule < CBSPMV_SCHED_AUTO || schedule > CBSPMV_SCHED_STREAM) return fail(CBSPMV_ERR_ARG, "unknown schedule");
    if (schedule == CBSPMV_SCHED_BLOCKED && plan->host.o.dtype != CBSPMV_F64)
        return fail(CBSPMV_ERR_UNSUPPORTED, "the blocked schedule is FP64 only");
    if (schedule == CBSPMV_SCHED_BLOCKED && plan->values_replaced)
        return fail(CBSPMV_ERR_UNSUPPORTED, "values were replaced: only the stream schedule holds them");
    plan->forced = schedule;
    return CBSPMV_OK;
}

int cbspmv_get_schedule(const cbspmv_plan* plan) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (plan->values_replaced) return CBSPMV_SCHED_STREAM;
    return plan->forced != CBSPMV_SCHED_AUTO ? plan->forced : plan->schedule;
}

double cbspmv_last_kernel_ms(const cbspmv_plan* plan) { return plan ? plan->last_kernel_ms : -1.0; }

// cb_spmv() protocol (cb-spmv-dcu.hpp:1088-1667): nnz == 0 -> y := 0; else
// upload, one launch whose y is returned, 100 warm-up launches, 1000 timed
// launches between events; runtime = ms / 1000, GFLOPS = 2 nnz 1e-6 / ms.
int cbspmv_bench(cbspmv_plan* plan, const double* x, double* y, cbspmv_metrics* metrics) {
    if (!plan || !metrics || (plan->host.m > 0 && !y) || (plan->host.n > 0 && !x))
        return fail(CBSPMV_ERR_ARG, "null argument");
    cbs::Plan& h = plan->host;
    metrics->runtime = 0.0;
    metrics->gflops = 0.0;
    if (h.nnz == 0) {
        std::fill(y, y + h.m, 0.0);
        return CBSPMV_OK;
    }
    if (h.o.dtype != CBSPMV_F64) return fail(CBSPMV_ERR_UNSUPPORTED, "cbspmv_bench takes FP64 vectors: use an F64 plan");
    hipStream_t s = nullptr;
    int rc = cbs::upload_plan(*plan, s);
    if (rc) return rc;
    void *dx = nullptr, *dy = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if (dx) (void)hipFree(dx);
        if (dy) (void)hipFree(dy);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    hipError_t e;
    if ((e = hipMalloc(&dx, (size_t)std::max(h.n, 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&dy, (size_t)std::max(h.m, 1) * 8)) != hipSuccess) {
        cleanup();
        return fail(CBSPMV_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    if ((e = hipMemcpy(dx, x, (size_t)h.n * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipEventCreate(&e0)) != hipSuccess || (e = hipEventCreate(&e1)) != hipSuccess) {
        cleanup();
        return cbs::fail(CBSPMV_ERR_HIP, std::string("setup: ") + hipGetErrorString(e));
    }
    rc = cbs::run_spmv(*plan, dx, dy, s);
    if (!rc && (e = hipStreamSynchronize(s)) != hipSuccess) rc = cbs::fail(CBSPMV_ERR_HIP, hipGetErrorString(e));
    if (!rc && (e = hipMemcpy(y, dy, (size_t)h.m * 8, hipMemcpyDeviceToHost)) != hipSuccess)
        rc = cbs::fail(CBSPMV_ERR_HIP, hipGetErrorString(e));
    for (int i = 0; !rc && i < 100; ++i) rc = cbs::run_spmv(*plan, dx, dy, s);
    if (!rc && (e = hipStreamSynchronize(s)) != hipSuccess) rc = cbs::fail(CBSPMV_ERR_HIP, hipGetErrorString(e));
    if (!rc) (void)hipEventRecord(e0, s);
    for (int i = 0; !rc && i < 1000; ++i) rc = cbs::run_spmv(*plan, dx, dy, s);
    if (!rc) {
        (void)hipEventRecord(e1, s);
        if ((e = hipEventSynchronize(e1)) != hipSuccess) rc = cbs::fail(CBSPMV_ERR_HIP, hipGetErrorString(e));
    }
    if (!rc) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        metrics->runtime = (double)ms / 1000.0;
        if (metrics->runtime > 0.0) metrics->gflops = 2.0 * (double)h.nnz * 1.0e-6 / metrics->runtime;
        plan->last_kernel_ms = metrics->runtime;
    }
    cleanup();
    return rc;
}

}  // extern "C"
