/*
 * cbspmv.h - C ABI of the MI355X (gfx950) CB-SpMV library, libcbspmv.so.
 *
 * Drop-in for the host/device boundary of the reference CB-SpMV-DCU, which has
 * no library ABI of its own: its callers use three C++ entry points that this
 * ABI replaces one for one (reference paths relative to the reference root):
 *
 *   GLOBAL_BLOCK(m, n, nnz) + coo2block_gather(row, col, val, gb)
 *       cb-spmv-dcu/src/format.h:106, cb-spmv-dcu/src/coo2block.h:628
 *       -> cbspmv_build()
 *   coo2block_comp(row, col, val, gb)   (the never-gathering builder)
 *       cb-spmv-dcu/src/coo2block.h:605
 *       -> cbspmv_build() with opts.builder = CBSPMV_BUILDER_COMP
 *   PROGRAM_METRICS cb_spmv(const GLOBAL_BLOCK&, const double* x, double* y)
 *       cb-spmv-dcu/src/cb-spmv-dcu.hpp:1088
 *       -> cbspmv_bench() (same protocol) or cbspmv_upload() + cbspmv_spmv()
 *
 * Conventions: every function returns 0 (CBSPMV_OK) or a negative status and
 * never exits the process (the reference's CHECK_HIP calls exit(1),
 * macros.h:17-31).  cbspmv_last_error() returns thread-local text for the last
 * failure.  Host pointers are borrowed for the duration of a call only.
 * Device pointers (d_*) are caller-owned device memory.  `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  A plan is not
 * thread-safe; distinct plans are independent.
 */
#ifndef CBSPMV_H
#define CBSPMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define CBSPMV_OK 0
#define CBSPMV_ERR_ARG -1          /* null pointer, bad size or option value      */
#define CBSPMV_ERR_RANGE -2        /* a row/col index outside [0,m) / [0,n)       */
#define CBSPMV_ERR_ORDER -3        /* COO rows not in row-major (non-decreasing)  */
#define CBSPMV_ERR_HIP -4          /* a HIP runtime call failed                   */
#define CBSPMV_ERR_NOMEM -5        /* host or device allocation failed            */
#define CBSPMV_ERR_STATE -6        /* call out of order (e.g. spmv before upload) */
#define CBSPMV_ERR_UNSUPPORTED -7  /* option combination not supported            */

/* ---- options ---------------------------------------------------------------
 * The reference fixes these at compile time (cb-spmv-dcu/src/macros.h:38-97,
 * Makefile:4-7); here they are per plan.  cbspmv_default_opts() returns the
 * reference Makefile's build: 16x16 blocks, 4 waves, row-owned segments,
 * pattern reuse enabled, FP64 values. */
enum cbspmv_builder { CBSPMV_BUILDER_GATHER = 0, CBSPMV_BUILDER_COMP = 1 };
enum cbspmv_dtype { CBSPMV_F64 = 0, CBSPMV_F32 = 1, CBSPMV_BF16 = 2 };
enum cbspmv_accum { CBSPMV_ACC_NATIVE = 0, CBSPMV_ACC_COMPENSATED = 1 };
enum cbspmv_backward_mode {
    CBSPMV_BACKWARD_DIRECT = 0, /* dx over A^T; dvals by a pass over the input COO  */
    CBSPMV_BACKWARD_FUSED = 1   /* one pass over A^T writes dx and stored-order dvals,
                                   then a permutation pass to input order          */
};
enum cbspmv_schedule {
    CBSPMV_SCHED_AUTO = 0,    /* pick by plan size / autotuner                       */
    CBSPMV_SCHED_BLOCKED = 1, /* one wave per packed block over the canonical slab   */
    CBSPMV_SCHED_STREAM = 2   /* nnz-balanced strip stream over the transcoded slab  */
};

typedef struct cbspmv_opts {
    int block_size;       /* 2, 4, 8, 16, 32 or 64            (CB_BLOCK_SIZE)       */
    int waves_per_block;  /* 1..16                            (CB_WAVES_PER_BLOCK)  */
    int row_owned;        /* 1: row-owned segments, 0: LPT    (CB_ROW_OWNED)        */
    int pattern_wave;     /* 1: pattern-reuse schedule usable (CB_PATTERN_WAVE)     */
    int force_coo64;      /* 1: COO up to 64 nnz              (CB_FORCE_COO64)      */
    int dense_sixteenths; /* DENSE threshold, 1..16           (CB_DENSE_SIXTEENTHS) */
    int builder;          /* enum cbspmv_builder                                     */
    int value_dtype;      /* enum cbspmv_dtype: storage type of A's values           */
    int accumulate;       /* enum cbspmv_accum (F32/BF16 only)                       */
    int schedule;         /* enum cbspmv_schedule                                    */
    int stream_item_nnz;  /* stream schedule: target entries per wave item (0=auto)  */
    int stream_tile_waves;/* stream schedule: waves per workgroup, 1..16 (0 = auto:
                             16 without column panels, else 8; 8 for the
                             transposed plan of the backward pass)                    */
    int stream_tile_kib;  /* stream schedule: LDS x cache per workgroup, 4..156 KiB (0=96) */
    int backward_mode;    /* enum cbspmv_backward_mode (cbspmv_backward)              */
    int stream_panels;    /* stream schedule: x column panels 1, 2, 4 or 8 (0 = 1)    */
    int stream_tiles;     /* stream schedule: tiles (workgroups of one round), 0 = the
                             current device's compute units (256, MI355X's count, when
                             no device is visible at build time); 1..65536            */
    int stream_split;     /* enum cbspmv_split: scattered entries in a sweep kernel  */
} cbspmv_opts;

/* Scattered-entry split (plans with native accumulation on the stream
 * schedule, no column panels).  An entry whose x line at most three
 * entries among the ~nnz/tiles entries around it read (itself included)
 * costs the stream schedule a vector-L1 miss, and an L2 miss once x outgrows
 * one XCD's L2.  The split
 * takes those entries out: the others form a stream plan of their own, and
 * the scattered ones run after it in a second kernel that adds them to y,
 * each wave walking its rows' entries in column order so that the waves of
 * an XCD sweep x together and mostly hit L2.  cbspmv_spmv and cbspmv_bench
 * run the split; batches, the transposed plan, the backward pass and
 * cbspmv_save use the plan's full stream.  Replacing values
 * (cbspmv_set_values, cbspmv_set_values_stored) drops the split: the full
 * stream takes the new values.  A split plan keeps its full stream besides
 * the other entries' plan and the sweep arrays (about twice the device
 * memory of an unsplit plan's stream). */
enum cbspmv_split {
    CBSPMV_SPLIT_AUTO = 0, /* split when x holds more than 4 MiB (one XCD's L2)
                              and at most 16 MiB, and at least 1M entries, 15%
                              of all, are scattered                              */
    CBSPMV_SPLIT_OFF = 1,
    CBSPMV_SPLIT_ON = 2    /* split whenever an entry is scattered (tests)       */
};

/* Builder statistics: the GLOBAL_BLOCK scalars (format.h:63-77) plus counts. */
typedef struct cbspmv_stats {
    int m, n, nnz;
    int gb_m, gb_n, gb_nnzb;
    int gather_flag, direct_coo_columns, compact_coo_indices;
    int super_sparse_blocks, analyzed_blocks;
    double super_sparse_ratio;
    int load_balanced;
    double load_stddev_before, load_stddev_after;
    int coo_blocks, csr_blocks, dense_blocks, uniform_blocks;
    int segments, patterns;
    int unit_values;
    long long slab_bytes;      /* canonical packed slab incl. x, y, epochs       */
    long long stream_bytes;    /* transcoded stream slab read per SpMV           */
    long long restore_entries; /* |block_cols|                                   */
    double build_seconds;
    long long split_entries;   /* entries in the sweep kernel (0: no split)     */
    int split_blocks;          /* its row blocks (workgroups)                   */
} cbspmv_stats;

/* PROGRAM_METRICS (format.h:185-188): runtime in ms per launch, GFLOPS. */
typedef struct cbspmv_metrics {
    double runtime;
    double gflops;
} cbspmv_metrics;

typedef struct cbspmv_plan cbspmv_plan;

/* ---- plan lifecycle ---------------------------------------------------------- */
void cbspmv_default_opts(cbspmv_opts* opts);

/* Builds the CB structure from a row-major COO stream (rows non-decreasing;
 * within-row order as given; duplicates allowed).  Bit-exact with
 * coo2block_gather (or coo2block_comp) for the same options.  opts may be NULL
 * (defaults).  nnz == 0 is allowed (y := 0, as cb_spmv, cb-spmv-dcu.hpp:1091). */
int cbspmv_build(int m, int n, int nnz, const int* row, const int* col,
                 const double* val, const cbspmv_opts* opts, cbspmv_plan** plan);
void cbspmv_destroy(cbspmv_plan* plan);
int cbspmv_plan_stats(const cbspmv_plan* plan, cbspmv_stats* stats);

/* Parity export.  Copies the named host array into dst when dst != NULL and
 * *count is large enough; always sets *count to the element count.  Names:
 *   gb_row_idx gb_col_idx gb_row_ptr nnz_per_blk type_per_blk(u8)
 *   segment_rows segment_begin segment_end segment_direct(u8) segment_uniform(u8)
 *   block_row_uniform(u8) flat_row_ptr flat_col_idx pattern_row_ptr
 *   pattern_col_idx pattern_output_ptr pattern_output_rows block_cols
 *   block_cols_offset bia_nnz bia_origin bia_end       (int32 unless marked)
 *   pl_idx (int32) pl_val (f64): payload streams in final block order
 *   image (u8): canonical packed slab bytes up to the y region, block
 *               addresses stored as byte offsets (0 = none), x region zero
 * Stream schedule (u32 unless marked): stream_key stream_val(f64)
 *   stream_slab(u8) stream_items (4 per item) stream_tiles stream_tile_lines
 *   stream_empty_rows stream_value_pos(i64) stream_col_bits(i32)
 *   stream_panels (i32: panels, tile item count) stream_panel_items (panel q's
 *   items: [v[q], v[q + 1])) stream_panel_mask(u8)
 * Scattered-entry split (empty without one): split_key (row in block above
 *   col_bits | column) split_val(f64, the input values) split_src (input
 *   entry of each) split_wave_off (block b's wave w: [v[b W + w], v[b W + w +
 *   1]), W waves) split_blk_row(i32: first row) split_blk_span(i32: rows)
 *   split_shape (col_bits, W, rows per block at most); "near.<name>" exports
 *   <name> of the stream plan of the other entries.
 * Returns CBSPMV_ERR_ARG for an unknown name. */
int cbspmv_export(const cbspmv_plan* plan, const char* name, void* dst, size_t* count);
/* Byte offset of a slab region in the canonical image ("total" = size). */
long long cbspmv_layout_offset(const cbspmv_plan* plan, const char* name);

/* ---- persistence (no reference counterpart: the reference rebuilds its
 * GLOBAL_BLOCK on every run, cb-spmv-dcu/src/main.hip.cpp:46-70) ------------ */
/* Writes a built plan (builder output, canonical slab image, stream slab) to
 * one file; cbspmv_load() reads it back without rebuilding.  The format is
 * versioned and checked on load (magic, version, element sizes, lengths);
 * it trusts files written by cbspmv_save.  A transposed plan is not saved. */
int cbspmv_save(const cbspmv_plan* plan, const char* path);
int cbspmv_load(const char* path, cbspmv_plan** plan);

/* ---- device execution (stream-ordered on the caller's stream) --------------- */
/* Copies the plan's device slabs to the current device (allocates); also
 * the transposed plan's, when cbspmv_build_transpose has built it. */
int cbspmv_upload(cbspmv_plan* plan, void* stream);
/* y = A x.  d_x has n elements, d_y has m elements, both of the plan's
 * vector type: double for F64 plans, float for F32/BF16 plans.  y is
 * overwritten (no accumulation).  No allocation, no host synchronisation.
 * The current device must be the one the plan was uploaded to
 * (CBSPMV_ERR_STATE otherwise). */
int cbspmv_spmv(cbspmv_plan* plan, const void* d_x, void* d_y, void* stream);
/* Forces one schedule for the next calls (CBSPMV_SCHED_*); AUTO restores the
 * plan's choice.  Returns CBSPMV_ERR_UNSUPPORTED if the plan lacks it. */
int cbspmv_set_schedule(cbspmv_plan* plan, int schedule);
int cbspmv_get_schedule(const cbspmv_plan* plan);

/* cb_spmv() semantics and protocol (cb-spmv-dcu.hpp:1088-1667): upload,
 * one validated launch whose y is copied to host y, 100 warm-up and 1000
 * timed launches between hipEvents; runtime = ms / 1000, GFLOPS = 2 nnz / t.
 * x, y are host arrays of n and m doubles.  Timings go to *metrics. */
int cbspmv_bench(cbspmv_plan* plan, const double* x, double* y, cbspmv_metrics* metrics);

/* ---- differentiation (no reference counterpart; extension) ------------------ */
/* Builds the transposed CB structure once (A^T from the input COO reordered
 * column-major, stable within a column; same options) and the maps between
 * input entries and stream entries.  Host only; the device copy is made on
 * the first transposed call. */
int cbspmv_build_transpose(cbspmv_plan* plan);
/* Number of entries of the "stored order" (the transposed plan's padded
 * stream), the length of cbspmv_backward_stored's dvals; -1 before
 * cbspmv_build_transpose.  cbspmv_export(plan, "grad_slot") gives, for every
 * input entry k, its stored-order index (uint32), "slot_of" its index in the
 * plan's own stream, and "transpose.<name>" exports array <name> of A^T. */
long long cbspmv_stored_entries(const cbspmv_plan* plan);
/* d_dx = A^T d_g  (d_g: m elements, d_dx: n, plan's vector type). */
int cbspmv_spmv_transpose(cbspmv_plan* plan, const void* d_g, void* d_dx, void* stream);
/* Gradients of y = A x for an upstream gradient g = dL/dy:
 *   d_dx    = A^T g                       (n elements; skipped when NULL)
 *   d_dvals = g[row[k]] * x[col[k]]       (nnz elements, input COO order;
 *                                          skipped when NULL)
 * duplicates of one (row, col) each get the full product.  Mode from
 * opts.backward_mode. */
int cbspmv_backward(cbspmv_plan* plan, const void* d_g, const void* d_x,
                    void* d_dx, void* d_dvals, void* stream);
/* The fused pass alone: d_dx (may be NULL) and dvals in stored order
 * (cbspmv_stored_entries() elements, 16-byte aligned); stored entries that
 * hold no input entry (zero fill of dense blocks, padding) get g_i x_j or 0. */
int cbspmv_backward_stored(cbspmv_plan* plan, const void* d_g, const void* d_x,
                           void* d_dx, void* d_dvals_stored, void* stream);
/* cbspmv_backward_stored that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory during the call
 * (HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same outputs as cbspmv_backward_stored (bit for bit); differences:
 *   - the plan and its transpose must already be on the current device
 *     (cbspmv_upload uploads the transpose too once it is built), else
 *     CBSPMV_ERR_STATE;
 *   - when d_dx is NULL the pass writes dx into a library-owned device
 *     array of 16 MiB per GPU (held until the process ends), so then
 *     n * element size must be at most 16 MiB (2M FP64 or 4M FP32
 *     columns), else CBSPMV_ERR_ARG.  Calls that overlap in time on one GPU
 *     share that array; its contents are discarded. */
int cbspmv_backward_stored_noalloc(cbspmv_plan* plan, const void* d_g, const void* d_x,
                                   void* d_dx, void* d_dvals_stored, void* stream);
/* Replaces the values of A on the device: d_vals holds nnz values in input
 * COO order, of the plan's vector type (double for F64 plans, float for F32
 * and BF16 plans, rounded to the storage type).  Duplicates are summed in
 * input order in FP64.  Updates A^T too when it exists.  After this call
 * the plan runs the stream schedule only, and cbspmv_export still reports
 * the build-time values. */
int cbspmv_set_values(cbspmv_plan* plan, const void* d_vals, void* stream);

/* cbspmv_set_values with the values given in stored order: d_vals_stored
 * holds cbspmv_stored_entries() values (the order of cbspmv_backward_stored's
 * dvals), so a trainer can keep its parameters in the order its gradients
 * come in.  Stored entries that hold no input entry are ignored; a stored
 * entry that merges duplicate input entries carries their sum.  Both plans
 * are updated: A^T from its own entries, A by summing, per stream entry,
 * the stored entries whose first input entry it holds.  Needs
 * cbspmv_build_transpose and cbspmv_upload. */
int cbspmv_set_values_stored(cbspmv_plan* plan, const void* d_vals_stored, void* stream);

/* ---- batching (no reference counterpart; extension) -------------------------- */
typedef struct cbspmv_batch cbspmv_batch;
/* Groups `count` built plans (borrowed; they must outlive the batch) of one
 * value type and accumulation into a batch: uploads every plan and
 * allocates partial-sum buffers for up to 16 right-hand sides per plan. */
int cbspmv_batch_create(cbspmv_plan* const* plans, int count, cbspmv_batch** batch, void* stream);
/* Y_i = A_i X_i for every matrix i, with nrhs[i] (0..16) right-hand sides
 * each: d_x[i] is n_i x nrhs[i] (column-major, leading dim n_i), d_y[i] is
 * m_i x nrhs[i], of the plans' vector type.  Two launches on `stream`: a
 * prologue (x of the plans with several right-hand sides interleaved in
 * groups of eight into a batch-owned buffer, work queues reset) and the
 * persistent batched kernel.  The pointer table is uploaded (with a stream
 * synchronisation) only when the pointers or counts differ from the
 * previous call.  Calls, captures and replays of one batch share its
 * scratch buffers: they must not overlap in time (one stream, or
 * synchronise between streams). */
int cbspmv_batch_spmv(cbspmv_batch* batch, const void* const* d_x, void* const* d_y,
                      const int* nrhs, void* stream);
/* Captures that launch for these pointers into a hipGraph (stream must be a
 * created stream, not NULL), instantiates it and launches it once;
 * cbspmv_batch_replay() launches the instantiated graph again.  The graph
 * holds no memset nodes. */
int cbspmv_batch_capture(cbspmv_batch* batch, const void* const* d_x, void* const* d_y,
                         const int* nrhs, void* stream);
int cbspmv_batch_replay(cbspmv_batch* batch, void* stream);
void cbspmv_batch_destroy(cbspmv_batch* batch);

/* Autotuner: builds the candidate plans for one COO matrix, times each on
 * the device in a batch of copies of itself with nrhs right-hand sides
 * each (enough copies to fill the GPU; 3 warm-up, mean of 20; the time is
 * per copy, i.e. what the matrix costs inside a full batch; a plan a batch
 * does not take is timed with nrhs single launches) and returns the
 * fastest, destroying the others; a later candidate replaces an earlier
 * one only when it is at least 5% faster, so list the preferred one first.
 * *best_ms is that per-copy time.  `candidates` lists (block_size, waves)
 * pairs; waves sets the blocked schedule's waves per workgroup (the
 * reference's CB_WAVES_PER_BLOCK), 0 keeps `base`'s.
 * NULL = {16,0},{32,0},{16,4},{16,9},{32,4},{32,9}.
 * The stream tile shape is the batch's, not a candidate's: a batched launch
 * runs every unit with its widest tile and largest LDS cache.  Options
 * `base` leaves at 0 get the batch shape: 8 tile waves, items of 2048
 * entries, an LDS x cache of 32 KiB (4 KiB with nrhs > 1, whose x is
 * gathered rather than staged).  A stream-schedule candidate without DENSE
 * blocks whose stream plan has the shape of one already timed (same tile
 * waves, items, slab bytes and staged lines: the block size and the blocked
 * schedule's waves then change nothing the stream kernel reads) is not
 * timed; the earlier candidate stands for it.
 * Allocates device memory for its test vectors. */
int cbspmv_autotune(int m, int n, int nnz, const int* row, const int* col, const double* val,
                    const cbspmv_opts* base, const int* candidates, int ncandidates,
                    int nrhs, cbspmv_plan** best, double* best_ms, void* stream);

/* ---- input -------------------------------------------------------------------- */
/* Reads a Matrix Market coordinate file the way the reference's reader does
 * (cb-spmv-dcu/src/io/mmio_highlevel.h:62-403, read_matrix_coo_data): every
 * value is forced to 1.0, symmetric/hermitian/skew files are expanded, and
 * the result is row-major COO (rows ascending; inside a row, file order with
 * mirrored entries placed when their source entry is read).  Arrays are
 * malloc'ed; release them with cbspmv_free().  Negative returns -1..-7 are the
 * reference reader's codes (open, banner, type, size, entry, expanded nnz
 * overflow, allocation). */
int cbspmv_read_mtx(const char* path, int* m, int* n, int* nnz, int** row, int** col, double** val);
void cbspmv_free(void* p);

/* ---- diagnostics ---------------------------------------------------------------- */
const char* cbspmv_last_error(void);
const char* cbspmv_version(void);
/* Average duration (ms) of the dominant kernel over the last cbspmv_bench
 * timed loop, from hipEvents on the plan's stream. */
double cbspmv_last_kernel_ms(const cbspmv_plan* plan);

#ifdef __cplusplus
}
#endif
#endif /* CBSPMV_H */
