// Device-side state of a plan and the helpers shared by the C ABI files.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "cb_plan.hpp"

namespace cbs {

int fail(int code, const std::string& msg);

// One device allocation per purpose; all owned by the plan.
struct DeviceBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct DevicePlan {
    int device = -1;
    bool uploaded = false;
    // canonical slab (blocked schedule): bytes up to the y region
    DeviceBuf slab;
    DeviceBuf zero_strips;   // u32 block rows the blocked schedule zeroes first
    uint32_t n_zero_strips = 0;
    bool zero_all = false;   // LPT: every row is accumulated atomically
    // stream slab
    DeviceBuf key, val, items, item_split, item_split_first, splits,
        partial, counters, empty, tiles, tile_lines;
    // column panels
    DeviceBuf panel_items, panel_mask, panel_out;
    // value replacement: per stream entry holding input entries, its value's
    // byte offset in the stream slab and its input entries (CSR, input order)
    DeviceBuf sv_pos, sv_ptr, sv_src;
    uint32_t sv_groups = 0;
    // the same for values given in stored order (the transposed plan's
    // stream order): each group's sources are stored entries
    DeviceBuf ss_pos, ss_ptr, ss_src;
    uint32_t ss_groups = 0;
    // backward (forward plan only): input entry -> transposed stream index,
    // the stored-order dvals buffer, and the input coordinates
    DeviceBuf grad_slot, dvals_stored, in_row, in_col;
    // scattered-entry split: the sweep kernel's arrays
    DeviceBuf sw_key, sw_val, sw_wave_off, sw_blk_row, sw_blk_span;
};


}  // namespace cbs

struct cbspmv_plan {
    cbs::Plan host;
    cbs::DevicePlan dev;
    cbspmv_opts opts{};                      // options the plan was built with
    int schedule = CBSPMV_SCHED_STREAM;      // chosen schedule
    int forced = CBSPMV_SCHED_AUTO;          // cbspmv_set_schedule override
    bool values_replaced = false;            // cbspmv_set_values ran: stream schedule only
    double last_kernel_ms = 0.0;
    cbspmv_plan* transpose = nullptr;        // A^T plan, built on request
    std::vector<uint32_t> grad_slot;         // input entry -> transposed stream index
    std::vector<uint32_t> slot_of;           // input entry -> own stream index
    // scattered-entry split: the stream plan of the other entries, and the
    // scattered ones (cbspmv_spmv runs `near`, then the sweep kernel)
    cbspmv_plan* near = nullptr;
    cbs::SweepPart sweep;
    bool paced = false;   // this plan is another's `near` (StreamArgs::paced)
};

namespace cbs {
// Compute units and XCDs of the current device (cached per device); MI355X's
// 256 and 8 when no device is visible (a plan built on a host without one).
struct DeviceShape {
    int cus;
    int xcds;
};
DeviceShape device_shape();
int choose_schedule(const cbspmv_plan& p);
long long stream_bytes_per_spmv(const cbspmv_plan& p);
// Builds a plan object (builder, optional canonical slab, stream slab).
int build_plan_object(const cbspmv_opts& o, int m, int n, int nnz, const int* row, const int* col,
                      const double* val, bool canonical, cbspmv_plan** out);

// ---- device helpers (runtime.hip)
int hip_fail(hipError_t e, const char* what);
#define CBS_HIP(call)                                     \
    do {                                                  \
        const hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return cbs::hip_fail(e_, #call); \
    } while (0)
void release(DeviceBuf& b);
// (Re)allocates b and fills it from src, or with zeros when src is null.
int upload_bytes(DeviceBuf& b, const void* src, size_t bytes, hipStream_t s);
template <typename T>
int upload_vec(DeviceBuf& b, const std::vector<T>& v, hipStream_t s) {
    return upload_bytes(b, v.empty() ? nullptr : v.data(), v.size() * sizeof(T), s);
}
size_t vec_elem_bytes(const Plan& h);
void release_device(DevicePlan& d);
int upload_plan(cbspmv_plan& p, hipStream_t s);
// Stream schedule launch; with dvals != null also the SDDMM outputs.
bool stream_vectors_fit(const Plan& h);
int run_stream(cbspmv_plan& p, const void* x, void* y, hipStream_t s, const void* xrow = nullptr,
               void* dvals = nullptr);
int run_spmv(cbspmv_plan& p, const void* x, void* y, hipStream_t s);
// Scattered-entry split (split.cpp; its layout in split_layout.cpp): built
// after the plan when its options and the matrix call for it; dropped when
// values are replaced.
int build_split(cbspmv_plan& p, int nnz, const int* row, const int* col, const double* val);
void drop_split(cbspmv_plan& p);
// The sweep kernel: y += the scattered entries' products (runtime.hip).
int run_sweep(cbspmv_plan& p, const void* x, void* y, hipStream_t s);
}  // namespace cbs
