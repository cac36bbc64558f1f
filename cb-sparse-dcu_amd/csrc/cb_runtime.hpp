// Device-side state of a plan and the host helpers shared by the C ABI files.
#pragma once

#include <string>

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
    // transposed plan (backward)
    bool has_transpose = false;
    // SDDMM input coordinates
    DeviceBuf in_row, in_col;
};

}  // namespace cbs

struct cbspmv_plan {
    cbs::Plan host;
    cbs::DevicePlan dev;
    int schedule = CBSPMV_SCHED_STREAM;      // chosen schedule
    int forced = CBSPMV_SCHED_AUTO;          // cbspmv_set_schedule override
    double last_kernel_ms = 0.0;
    cbspmv_plan* transpose = nullptr;        // A^T plan, built on request
};

namespace cbs {
int choose_schedule(const cbspmv_plan& p);
long long stream_bytes_per_spmv(const cbspmv_plan& p);
}  // namespace cbs
