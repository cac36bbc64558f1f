// Host-side CB plan: the builder's output (the reference's GLOBAL_BLOCK,
// cb-spmv-dcu/src/format.h:62-183), the canonical packed slab and the
// transcoded stream slab, plus the device-side state of an uploaded plan.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cbspmv.h"

namespace cbs {

// Block type codes and flags (cb-spmv-dcu/src/macros.h:99-105).
enum : uint8_t { kCOO = 0, kCSR = 1, kDENSE = 3, kTypeMask = 0x7f, kUniform = 0x80 };

// Waves per stream tile when the options leave it open: 16 without column
// panels (the forward launch hides more gather latency with 16 waves per
// CU), else 8 (a workgroup adds as many panel waves).
constexpr int kDefaultTileItemsOpt = 8;
constexpr int kWideTileItemsOpt = 16;
constexpr int kDefaultTileBytesOpt = 96 * 1024;
constexpr int kStreamCUs = 256;         // MI355X compute units: the tile count without a device

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
    int tile_items = kDefaultTileItemsOpt;
    int tile_bytes = kDefaultTileBytesOpt;
    int panels = 0;          // column panels of the stream schedule (0 = by x size)
    int tiles = kStreamCUs;  // stream tiles (one round of workgroups)
};

inline int coord_bits(int bs) { int b = 0; while ((1 << b) < bs) ++b; return b; }
inline size_t orig_index_bytes(int bs) { return bs <= 16 ? 1 : 2; }   // IndexTypeOrig
inline size_t dense_mask_bytes(int bs) { return bs <= 8 ? 1 : bs == 16 ? 2 : bs == 32 ? 4 : 8; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

// Stream slab (device layout read by the stream schedule).
//
// Entries: inside each strip (block row) the entries of all its blocks are
// ordered by in-block row, then block order, then in-block order, so every
// matrix row is one contiguous run.  Each entry is a 32-bit key and a value:
//   bit 31       column is staged in the workgroup's LDS x cache
//   [cb, 31)     row offset from the item's base row
//   [0, cb)      original column, or LDS slot when bit 31 is set
// Work: the stream is cut into wave items of about item_nnz entries at row
// boundaries (an item's rows span less than 2^(31-cb)); only a row longer
// than 4 items is cut, and its pieces meet in a small partial buffer through
// an atomic ticket.  Every item is padded with zero entries (value 0 times
// the always-zero LDS slot) to a multiple of the entries each lane owns, so
// lanes are either full or idle.  tile_items consecutive items form a tile
// run by one workgroup, which first stages the tile's hot x cache lines (16
// elements each, lines hit at least twice) into LDS.
struct StreamItem {        // 16 bytes
    uint32_t e_begin;      // first entry (a multiple of the lane width)
    uint32_t e_end;        // one past the last entry (padded)
    uint32_t row_base;     // row of the first entry
    uint8_t first_partial; // first run continues a row from the previous item
    uint8_t last_partial;  // last run continues into the next item
    uint8_t pad[2];
};
constexpr int kLineElems = 16;          // x elements per cached line
constexpr uint32_t kKeyCached = 0x80000000u;
// Tile shape: waves (items) per workgroup and LDS bytes of x cache; the
// defaults give one 8-wave workgroup per CU with 96 KB of its 160 KB of LDS.
constexpr int kMaxTileItems = 16;
// LDS x cache per workgroup at most: the 160 KB of a CU less room for the
// kernels' static LDS
constexpr int kMaxTileBytes = 156 * 1024;
// LDS x-cache elements per workgroup; the last slot always holds 0
inline int tile_slots(int bytes, int dtype) { return bytes / (dtype == CBSPMV_F64 ? 8 : 4); }
inline int tile_lines(int bytes, int dtype) { return tile_slots(bytes, dtype) / kLineElems - 1; }
// Entries one lane owns per chunk (a chunk is 64 lanes' worth).  The slab
// stores each chunk as 16-byte lane planes: the keys (1 or 2 planes), then
// the values (1 or 2 planes); plane p holds 16 bytes per lane, lane-major, so
// every load instruction reads one contiguous run of 64 x 16 bytes.  A chunk
// with G < 64 lanes (an item's last) stores planes of G x 16 bytes.
inline int stream_lane_entries(int dtype) { return dtype == CBSPMV_F64 ? 4 : 8; }
inline int stream_val_bytes(int dtype) { return dtype == CBSPMV_F64 ? 8 : dtype == CBSPMV_F32 ? 4 : 2; }

// Column panels.  When x is larger than one XCD's L2 can hold, the entries a
// tile cannot serve from its LDS x cache (scattered columns) are moved to
// panel items: panel q holds columns [q w, (q + 1) w) in row order, and only
// waves on the XCDs q, q + P, ... run panel q's items, so each XCD gathers
// from a w-column slice of x that stays in its L2.  A panel item writes its
// rows' partial sums to row r of panel q's output vector; a second launch
// adds them to y.
// MI355X XCDs.  Column panels deal their items by workgroup index with this
// period (workgroup b on XCD b % 8 is what makes a panel's x slice stay in
// one L2; another placement only costs locality); the batched launch reads
// the XCD each workgroup runs on from the hardware instead.
constexpr int kXcds = 8;
constexpr int kMaxXcds = 16;                    // per-XCD queues of the batched launch at most
constexpr int64_t kPanelL2Bytes = 4ll << 20;    // x up to this size is gathered whole
constexpr int64_t kPanelBytes = 2ll << 20;      // panel width target (bytes of x)
// stream_panels = 0 picks panels by the size of x only when this is set
constexpr bool kPanelsByDefault = false;

// Automatic item size: the smallest that cuts the stream into one round of
// tiles, but at least this many entries per item.
constexpr int kMinItemEntries = 1024;

struct StreamSplit {       // one per row cut into pieces
    uint32_t row;          // matrix row
    uint32_t first_item;   // item holding piece 0
    uint32_t pieces;       // number of pieces (items) of this row
    uint32_t partial_base; // first slot in the partial buffer
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

    // Input COO kept for the transposed build, the SDDMM and value updates
    // (row-major, input order).
    std::vector<int> in_row, in_col;
    std::vector<double> in_val;

    // Canonical packed slab (cb-spmv-dcu.hpp:1130-1348), block addresses as
    // byte offsets.  Built on demand.
    bool unit_values = false;
    bool image_ready = false;
    std::vector<uint8_t> image;
    std::map<std::string, size_t> off;

    // Stream slab
    bool stream_ready = false;
    int col_bits = 0;
    std::vector<uint32_t> s_key;     // logical entry order (export, tests)
    std::vector<double> s_val;
    std::vector<uint8_t> s_slab;     // device layout: chunk lane planes
    std::vector<StreamItem> items;
    std::vector<uint32_t> item_split;       // per item: split id of its last run or ~0
    std::vector<uint32_t> item_split_first; // per item: split id of its first run or ~0
    std::vector<StreamSplit> splits;
    std::vector<uint32_t> empty_rows;       // rows without entries (zeroed by the kernel)
    std::vector<uint32_t> tiles;            // per tile: first line, line count
    std::vector<uint32_t> tile_lines;       // cached x line ids (element / 16)
    uint32_t partial_slots = 0;
    // column panels (panels > 1): items [0, tile_items_count) belong to the
    // tiles, the rest to the panels
    int panels = 1;
    uint32_t tile_items_count = 0;
    std::vector<uint32_t> panel_item_ptr;  // panel q's items: [ptr[q], ptr[q + 1])
    std::vector<uint8_t> panel_mask;       // per row: panels holding entries of the row
};

// What the stream kernel reads of a plan, in brief: two plans of one matrix
// with the same shape and no DENSE blocks (whose zero fill depends on the
// block size) stream the same entries in the same items and tiles.
struct StreamPlanShape {
    int tile_items = 0;
    size_t items = 0, slab = 0, lines = 0;
    bool dense = false;
    bool operator==(const StreamPlanShape& b) const {
        return tile_items == b.tile_items && items == b.items && slab == b.slab && lines == b.lines &&
               dense == b.dense;
    }
};
inline StreamPlanShape stream_plan_shape(const Plan& h) {
    StreamPlanShape sh;
    sh.tile_items = h.o.tile_items;
    sh.items = h.items.size();
    sh.slab = h.s_slab.size();
    sh.lines = h.tile_lines.size();
    for (uint8_t t : h.type) sh.dense = sh.dense || (t & kTypeMask) == kDENSE;
    return sh;
}

// Builder (builder.cpp): bit-exact with coo2block_gather / coo2block_comp.
int build_plan(Plan& p, int m, int n, int nnz, const int* row, const int* col,
               const double* val, std::string& err);
// Packer (packer.cpp)
void pack_canonical(Plan& p);
void build_stream(Plan& p);
// Column panels the stream schedule uses for this plan (1 = none).
int stream_panels(const Plan& p);
long export_array(const Plan& p, const std::string& name, void* dst, size_t cap, int& status);
// For every input entry k (in_row/in_col order), its index in the padded
// entry stream (duplicates merged into one stream entry share it).  Returns
// false if the stream does not hold the input's entries.
bool stream_entry_slots(const Plan& p, std::vector<uint32_t>& slot);
// Byte offset in s_slab of the value of padded stream entry e.
size_t stream_value_offset(const Plan& p, uint32_t item, uint32_t e);

// Scattered entries of a split plan, as the sweep kernel reads them: row
// blocks of at most kSweepMaxRows rows (the LDS row tile), each run by one
// workgroup of kSweepWaves waves; a block's rows are cut into kSweepWaves
// runs of about equal entry counts, one per wave; a wave's entries are in
// column order (then row, then input order), and every 64 of them from the
// wave's first are reordered by row (stable), so that one segmented scan
// sums each 64-entry chunk per row.
constexpr int kSweepWaves = 4;
constexpr int kSweepMaxRows = 2048;
struct SweepPart {
    int col_bits = 0;                 // key: column below, row in block above
    int max_rows = 0;                 // rows per block at most
    std::vector<uint32_t> key;
    std::vector<double> val;          // input values (the device copy is in the plan dtype)
    std::vector<uint32_t> src;        // input entry of each
    std::vector<uint32_t> wave_off;   // blocks * kSweepWaves + 1
    std::vector<int32_t> blk_row, blk_span;
    size_t blocks() const { return blk_row.size(); }
};

// The BF16 bits of v (through FP32, round to nearest even; NaNs stay NaN).
inline uint16_t bf16_bits(double v) {
    const float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Scattered-entry split, host layout (split_layout.cpp).  1 for the entries
// of a row-major COO whose x line at most kScatteredReads entries of their
// group (about nnz / tiles entries, cut at row boundaries) read; line_elems
// x elements per 128-byte line.
constexpr size_t kScatteredReads = 3;
std::vector<uint8_t> scattered_entries(int nnz, const int* row, const int* col, int tiles, uint32_t line_elems);
// The sweep layout of the entries fe (input indices, row-major order) of
// a matrix whose columns fit col_bits, blocks of at most max_rows rows.
SweepPart sweep_layout(const std::vector<uint32_t>& fe, const int* row, const int* col, const double* val,
                       int col_bits, int max_rows);

}  // namespace cbs
