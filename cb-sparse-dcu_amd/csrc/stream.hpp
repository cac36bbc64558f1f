// Launch interface of the stream schedule (stream.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "cb_plan.hpp"

namespace cbs {

namespace dev { struct bf16_t; }

constexpr int kStreamThreads = 256;   // 4 waves per workgroup

template <typename V, typename A>
struct StreamArgs {
    const uint32_t* key;
    const V* val;
    const StreamItem* items;
    const uint32_t* strip_ids;
    const uint32_t* item_split;
    const uint32_t* item_split_first;
    const StreamSplit* splits;
    A* partial;
    uint32_t* counters;
    const uint32_t* empty;
    uint32_t n_empty;
    uint32_t n_items;
    int col_bits;
    int m;
    const A* x;
    A* y;
};

hipError_t launch_stream(int bs, const StreamArgs<double, double>& a, int grid, hipStream_t s);
hipError_t launch_stream(int bs, const StreamArgs<float, float>& a, int grid, bool comp, hipStream_t s);
hipError_t launch_stream(int bs, const StreamArgs<dev::bf16_t, float>& a, int grid, bool comp, hipStream_t s);

}  // namespace cbs
