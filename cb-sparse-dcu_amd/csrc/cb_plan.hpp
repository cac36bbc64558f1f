// Host-side CB plan: the builder's output (the reference's GLOBAL_BLOCK,
// cb-spmv-dcu/src/format.h:62-183), the canonical packed slab and the
// transcoded stream slab, plus the device-side state of an uploaded plan.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "cbspmv.h"

namespace cbs {

// Block type codes and flags (cb-spmv-dcu/src/macros.h:99-105).
enum : uint8_t { kCOO = 0, kCSR = 1, kDENSE = 3, kTypeMask = 0x7f, kUniform = 0x80 };

struct Options {
    int bs = 16;
    int waves = 4;
    int row_owned = 1;
    int pattern_wave = 1;
    int force_coo64 = 0;
    int dense_sixteenths = 15;
    int builder = CBSPMV_BUILDER_GATHER;
    int dtype = CBSPMV_F64;
    int accum = CBSPMV_ACC_NATIVE;
    int schedule = CBSPMV_SCHED_AUTO;
    int item_nnz = 0;
};

inline int coord_bits(int bs) { int b = 0; while ((1 << b) < bs) ++b; return b; }
inline size_t orig_index_bytes(int bs) { return bs <= 16 ? 1 : 2; }   // IndexTypeOrig
inline size_t dense_mask_bytes(int bs) { return bs <= 8 ? 1 : bs == 16 ? 2 : bs == 32 ? 4 : 8; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

// Stream slab (device layout read by the stream schedule).  Within each block
// row ("strip") the entries of all its blocks are ordered by in-block row,
// then block order, then in-block order; each entry is one 32-bit key and one
// value.  key = original column | (row-in-strip << col_bits) | strip-start
// flag at bit 31.  Work is cut into wave "items" of about item_nnz entries
// covering at most kMaxItemStrips strips; a strip longer than an item is cut
// into pieces whose partial rows meet in a small partial buffer.
struct StreamItem {        // 16 bytes, read once per wave
    uint32_t e_begin;      // first entry
    uint32_t e_end;        // one past the last entry
    uint32_t s_begin;      // first strip (index into strip_ids)
    uint16_t n_strips;     // strips touched by the item
    uint8_t first_partial; // first strip continues from the previous item
    uint8_t last_partial;  // last strip continues into the next item
};
constexpr int kMaxItemStrips = 16;

struct StreamSplit {       // one per strip cut into pieces
    uint32_t strip;        // block-row id
    uint32_t first_item;   // item holding piece 0
    uint32_t pieces;       // number of pieces (items) of this strip
    uint32_t partial_base; // first slot in the partial buffer (slots of bs rows)
};

struct Plan {
    Options o;
    int m = 0, n = 0, nnz = 0;
    int gb_m = 0, gb_n = 0, nnzb = 0;
    bool gather = false, direct = false, compact = false;
    int super_sparse = 0, analyzed = 0;
    double super_ratio = 0.0;
    bool balanced = false;
    double sd_before = 0.0, sd_after = 0.0;
    double build_seconds = 0.0;

    // GLOBAL_BLOCK arrays (final block order)
    std::vector<int> brow, bcol, nnzpb, row_ptr;
    std::vector<uint8_t> type;
    std::vector<int> seg_rows, seg_begin, seg_end;
    std::vector<uint8_t> seg_direct, seg_uniform, row_uniform;
    std::vector<int> flat_row_ptr, flat_col;
    std::vector<int> pat_row_ptr, pat_col, pat_out_ptr, pat_out_rows;
    std::vector<int> block_cols, block_cols_off;
    std::vector<int> bia_nnz, bia_origin, bia_end;

    // Payloads indexed by origin (pre-balance) block id.  COO/CSR entries use
    // slots [pl_start[b], pl_start[b] + raw_nnz[b]) of the flat arrays;
    // coo_count[b] is the merged COO count.  DENSE blocks use dense_slot[b].
    std::vector<int64_t> pl_start;
    std::vector<int> raw_nnz, coo_count, dense_slot;
    std::vector<uint32_t> coo_idx;   // merged/encoded COO index words
    std::vector<double> ent_val;     // COO merged values | CSR values
    std::vector<int> csr_col;        // CSR local columns (IndexTypeOrig range)
    std::vector<int> csr_rp;         // (bs+1) per origin block (CSR only)
    std::vector<int64_t> csr_rp_start;
    std::vector<double> dense_val;   // bs*bs per DENSE block

    // Input COO kept for the SDDMM / value updates (row-major, input order).
    std::vector<int> in_row, in_col;

    // Canonical packed slab (cb-spmv-dcu.hpp:1130-1348), block addresses as
    // byte offsets.  Built on demand.
    bool unit_values = false;
    bool image_ready = false;
    std::vector<uint8_t> image;
    std::map<std::string, size_t> off;

    // Stream slab
    bool stream_ready = false;
    int col_bits = 0;
    std::vector<uint32_t> s_key;
    std::vector<double> s_val;       // FP64 copy; narrowed at upload
    std::vector<int64_t> s_src;      // value source: input COO position of each entry (-1: summed/zero)
    std::vector<StreamItem> items;
    std::vector<uint32_t> strip_ids; // non-empty strips in order
    std::vector<uint32_t> item_split;// per item: split id of its last strip or 0xffffffff
    std::vector<uint32_t> item_split_first; // per item: split id of its first strip or 0xffffffff
    std::vector<StreamSplit> splits;
    std::vector<uint32_t> empty_strips;
    uint32_t partial_slots = 0;
};

// Builder (builder.cpp): bit-exact with coo2block_gather / coo2block_comp.
int build_plan(Plan& p, int m, int n, int nnz, const int* row, const int* col,
               const double* val, std::string& err);
// Packer (packer.cpp)
void pack_canonical(Plan& p);
void build_stream(Plan& p);
long export_array(const Plan& p, const std::string& name, void* dst, size_t cap, int& status);

}  // namespace cbs
