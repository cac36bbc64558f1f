// Launch interface of the blocked schedule (blocked.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cbs {

constexpr int kMaxWaves = 16;

struct BlockedArgs {
    const uint8_t* slab;       // canonical slab base
    const int* brow;
    const int* bcol;
    const int* nnzpb;
    const uint8_t* type;
    const uint64_t* addr;      // payload byte offsets from slab
    const int* seg_rows;
    const int* seg_begin;
    const int* seg_end;
    const uint8_t* seg_direct;
    const uint8_t* seg_uniform;
    const int* block_cols;
    const int* block_cols_off;
    const double* x;
    double* y;
    int m, n, nnzb, waves;
    int gather, direct, compact, unit, row_owned, aggregate;
};

struct PatternArgs {
    const int* row_ptr;
    const int* col;
    const int* out_ptr;
    const int* out_rows;
    const double* x;
    double* y;
    int patterns;
};

hipError_t launch_blocked(int bs, const BlockedArgs& a, int grid, hipStream_t s);
hipError_t launch_zero_rows(double* y, int m, const uint32_t* strips, uint32_t n, int bs, hipStream_t s);
hipError_t launch_zero_words(void* p, size_t words, hipStream_t s);
hipError_t launch_pattern(const PatternArgs& a, bool unrolled, int waves, hipStream_t s);

}  // namespace cbs
