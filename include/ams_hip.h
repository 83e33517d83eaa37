/*
 * ams_hip.h -- device C ABI of the MI355X hybrid-accumulation SpGEMM.
 *
 * C = A * B for CSR matrices resident in HBM: int64 row pointers, int32
 * column indices, FP64 or FP32 values.  Every output row has ascending
 * columns.  All work of a call is ordered on the context's stream; calls on
 * one context must not overlap (one host thread per context).
 *
 * The reference has no device API.  These entry points replace the phase
 * functions that its driver calls (src/ams_hybaccspgemm.c:36-220): the
 * reference-shaped host entry points in ams_dropin.h are built on them.
 * Each declaration cites the reference function whose job it takes over.
 *
 * Return codes: 0 = ok, negative = error; ams_last_error() describes it.
 */
#ifndef AMS_HIP_H
#define AMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { AMS_OK = 0, AMS_ERR_ARG = -1, AMS_ERR_HIP = -2, AMS_ERR_ALLOC = -3,
       AMS_ERR_STATE = -4, AMS_ERR_LIMIT = -5 };

enum { AMS_F64 = 0, AMS_F32 = 1 };

/* Numeric accumulator policy: the hybrid router, or one accumulator for all
 * rows as the reference's ams_hashspgemm / ams_hserscspgemm binaries do
 * (src/ams_hashspgemm.c:191, src/ams_hserscspgemm.c:198).  "hash" keeps the
 * dense column-block bin for rows too wide for an LDS table. */
enum { AMS_MODE_HYBRID = 0, AMS_MODE_HASH = 1, AMS_MODE_ESC = 2 };

/* A CSR matrix in device memory (csr_t of src/inc/store.h:49-56, with
 * device index widths).  Like the reference, the library trusts the
 * entries: row_ptr[0] = 0, row_ptr non-decreasing, row_ptr[num_row] =
 * num_nnz and every column in [0, num_col) are the caller's to ensure (the
 * calls check sizes and pointers only).  Columns of a row may come in any
 * order and may repeat; each copy adds its products. */
typedef struct {
	int64_t num_row;
	int64_t num_col;
	int64_t num_nnz;
	const int64_t *row_ptr;  /* num_row + 1 */
	const int32_t *col;      /* num_nnz */
	const void *val;         /* num_nnz, double or float per dtype */
} ams_dcsr_t;

/* Tunables.  They take the place of the reference's compile-time AMS_*
 * knobs (cmake/Options.cmake:104-191); ams_autotune() picks them. */
typedef struct {
	int32_t sym_hash_max;   /* rows with flops <= this use the LDS-hash symbolic      */
	int32_t num_hash_max;   /* bitmap-symbolic rows with nnz(C row) <= this may use an
	                           LDS accumulator, larger ones take the dense chunks;
	                           LDS-hash-symbolic rows always fit one                 */
	int32_t esc_max;        /* rows with flops <= this may use ESC                   */
	double  acc_thres;      /* flops / nnz(C row) below this -> ESC (AMS_ACC_THRES)  */
	int32_t dense_width;    /* columns per dense LDS chunk (multiple of 64) for FP64;
	                           FP32 chunks are 7/4 as wide, rounded down to 64     */
	int32_t bitmap_bits;    /* LDS bitmap width of the heavy symbolic, bits           */
	double  bitmap_ratio;   /* bitmap symbolic when flops >= ratio * words(n)         */
	double  dense_density;  /* dense bin also when nnz(C row) >= density * n          */
	int32_t mode;           /* AMS_MODE_*                                             */
	int32_t wgs_per_cu;     /* persistent workgroups per CU for the queue kernels     */
	int32_t esc_sort;       /* ESC row sort: 0 register bitonic, 1 bucket (64-bit keys),
	                           2 LSD radix, 3 merge of the row's sorted B rows (rows of
	                           one window of A nonzeros whose B rows ascend; others
	                           bitonic), 4 counting sort over the row's column range with
	                           rank fix-up per bucket (rows with a bucket of more than 16
	                           keys: bitonic) */
	int32_t sym_blocked;    /* wide-row symbolic per (row, column chunk) on blocked B */
	double  ghash_ratio;    /* rows with flops > sym_hash_max and flops < ratio * words(n)
	                           count and accumulate in global-memory hash tables
	                           (synchronised calls; 0 = off)                      */
	int32_t light_max;      /* dense-bin (row, column chunk) morsels of at most this many
	                           C entries (and at most 128 A nonzeros) are accumulated by
	                           one wave each: at their columns' ranks when the symbolic
	                           phase kept the row bitmaps (up to 512 entries), else in an
	                           LDS hash table (up to 341); 0 = off, at most 512 */
	int32_t row_bitmaps;    /* 1: synchronised symbolic calls keep the wide rows' column
	                           bitmaps for the numeric phase (at most 1/12 of the GPU's
	                           memory; ams_spgemm releases them when C would not fit
	                           beside them); 0: off */
} ams_params_t;

/* Counters of the last symbolic / numeric call (rows per bin, tasks). */
typedef struct {
	int64_t rows_total;
	int64_t rows_empty;
	int64_t rows_sym_hash;
	int64_t rows_sym_bitmap;
	int64_t rows_num_hash;
	int64_t rows_num_esc;
	int64_t rows_num_dense;
	int64_t dense_chunks;      /* column chunks of the dense bin                  */
	int64_t flops;             /* P = sum of row flops                            */
	int64_t nnz_c;
	/* per numeric route (hash, esc, dense, global hash): products, nnz(C), nnz(A) */
	int64_t flops_route[4];
	int64_t nnzc_route[4];
	int64_t nnza_route[4];
	/* per symbolic route (hash, bitmap, global hash): products, nnz(A) */
	int64_t flops_sym[3];
	int64_t nnza_sym[3];
	int64_t rows_sym_ghash;
	int64_t rows_num_ghash;
} ams_stats_t;

/* Phases timed with HIP events on the context stream when timing is on. */
enum { AMS_PH_FLOPS = 0, AMS_PH_CLASSIFY = 1, AMS_PH_SYMB_HASH = 2, AMS_PH_SYMB_BITMAP = 3,
       AMS_PH_SCAN = 4, AMS_PH_BBLK = 5, AMS_PH_NUMC_DENSE = 6, AMS_PH_NUMC_HASH = 7,
       AMS_PH_NUMC_ESC = 8, AMS_PH_SYMB_GHASH = 9, AMS_PH_NUMC_GHASH = 10, AMS_PH_NUMC_LIGHT = 11,
       AMS_PH_COUNT = 12 };

typedef struct ams_ctx ams_ctx_t;

/* Context: owns scratch and the one-shot output pool; stream may be NULL
 * (the null stream).  No counterpart in the reference (its state is global,
 * src/ams/ams_utils.c:9,20,31-32). */
int ams_ctx_create(ams_ctx_t **ctx, void *hip_stream);
int ams_ctx_destroy(ams_ctx_t *ctx);
int ams_set_stream(ams_ctx_t *ctx, void *hip_stream);
const char *ams_last_error(const ams_ctx_t *ctx);
void ams_default_params(ams_params_t *p);
int ams_set_params(ams_ctx_t *ctx, const ams_params_t *p);
int ams_get_params(const ams_ctx_t *ctx, ams_params_t *p);
int ams_get_stats(const ams_ctx_t *ctx, ams_stats_t *s);

/* Event timing of the phases above (no counterpart in the reference, whose
 * drivers time INIT/SYMB/MGMT/NUMC with clock_gettime,
 * src/ams_hybaccspgemm.c:41-217).  ams_get_timings synchronises and returns
 * the milliseconds and launch counts accumulated since timing was enabled. */
int ams_enable_timing(ams_ctx_t *ctx, int on);
int ams_get_timings(ams_ctx_t *ctx, double *ms, int64_t *launches);

/* Per-row flop counts: row_flp[i] = sum over A(i,:) of |B(k,:)|, int64, and
 * the total P.  Replaces flp_populate (src/mm/mm_utils.c:1063-1125) and
 * get_intprod (:981-1055).  row_flp may be NULL; total may be NULL. */
int ams_row_flops(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b,
	int64_t *row_flp_dev, int64_t *total_host);

/* Exclusive scan of n int64 values: out[0] = 0, out[i+1] = out[i] + in[i],
 * n outputs (in and out may alias).  Replaces scan (src/mm/mm_utils.c:866-957). */
int ams_scan_i64(ams_ctx_t *ctx, const int64_t *in_dev, int64_t *out_dev, int64_t n);

/* Symbolic phase: row_ptr_c (num_row + 1, device) receives C's row pointers
 * and *nnz_c the total.  Also records the row routing that the following
 * ams_spgemm_numeric on the same A and B uses.  Replaces flp_populate + scan
 * + ams_hash_symb_dnr + scan (src/ams_hybaccspgemm.c:41-129,
 * src/ams/ams_utils.c:1600-1752). */
int ams_spgemm_symbolic(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b,
	int dtype, int64_t *row_ptr_c_dev, int64_t *nnz_c_host);

/* Numeric phase: fills col_c / val_c (nnz_c entries, device) for the row
 * pointers of the preceding symbolic call, which must have been made on the
 * same A and B (same sizes and the same row-pointer and column arrays; the
 * values may differ), else AMS_ERR_STATE.  Synchronises the stream at the
 * end and returns AMS_ERR_STATE when a kernel raised an error flag (see
 * ams_get_error_flags).  Replaces pack_ams_pack + ams_hybacc_numc
 * (src/ams/ams_utils.c:1410-1596,1080-1338). */
int ams_spgemm_numeric(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b, int dtype,
	const int64_t *row_ptr_c_dev, int32_t *col_c_dev, void *val_c_dev);

/* One shot: symbolic, C allocated from the context pool (valid until the
 * next call on ctx), numeric.  Replaces the whole hybaccspgemm driver
 * (src/ams_hybaccspgemm.c:36-220) minus host I/O. */
int ams_spgemm(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b, int dtype,
	ams_dcsr_t *c_out);

/* Batched: independent products C_i = A_i * B_i stacked into one
 * block-diagonal product, so every phase is one launch and the heavy-row
 * morsels of all problems share one work queue.  The stacked inputs hold
 * problem i's rows of A at [a_row_off[i], a_row_off[i+1]), its rows of B
 * (= columns of A) at [b_row_off[i], b_row_off[i+1]) and its output columns
 * at [c_col_off[i], c_col_off[i+1]) (host arrays of nprob + 1 entries).
 * ams_batched_plan uploads the problem table, sizes every scratch buffer and
 * returns the exact nnz of the stacked C in *c_cap_host.  After it,
 * ams_batched_run neither allocates nor synchronises: it can be captured
 * into a hipGraph (ams_batched_capture does that) and replayed.
 * c_row_ptr (sum m_i + 1) receives stacked row pointers, c_col / c_val are
 * caller buffers of c_cap entries, *nnz_dev (device int64) the total; a run
 * whose nnz exceeds c_cap writes no entries.  No counterpart in the
 * reference. */
int ams_batched_plan(ams_ctx_t *ctx, const ams_dcsr_t *a_stack, const ams_dcsr_t *b_stack,
	int dtype, int32_t nprob, const int64_t *a_row_off, const int64_t *b_row_off,
	const int64_t *c_col_off, int64_t *c_cap_host);
int ams_batched_run(ams_ctx_t *ctx, const ams_dcsr_t *a_stack, const ams_dcsr_t *b_stack,
	int dtype, int64_t *c_row_ptr_dev, int32_t *c_col_dev, void *c_val_dev, int64_t c_cap,
	int64_t *nnz_dev);
/* Captures one ams_batched_run into a graph on ctx's stream (which must not
 * be the null stream) and returns an opaque graph handle.  The graph's
 * kernels point into ctx's scratch: ams_graph_launch replays it on ctx only,
 * and only while that scratch has neither moved (a later call on ctx needing
 * a larger buffer reallocates it) nor been re-planned (another
 * ams_batched_plan); otherwise it returns AMS_ERR_STATE and the graph must
 * be captured again.  Replays and other calls on ctx are ordered on its
 * stream.  ams_graph_destroy releases the handle. */
int ams_batched_capture(ams_ctx_t *ctx, const ams_dcsr_t *a_stack, const ams_dcsr_t *b_stack,
	int dtype, int64_t *c_row_ptr_dev, int32_t *c_col_dev, void *c_val_dev, int64_t c_cap,
	int64_t *nnz_dev, void **graph_exec);
int ams_graph_launch(ams_ctx_t *ctx, void *graph_exec);
int ams_graph_destroy(void *graph_exec);

/* One-shot product that allocates nothing, for callers that cannot allocate
 * device memory (code under hipGraph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.
 * It runs the unsynchronised pipeline of ams_batched_run for one product
 * with every scratch buffer carved from a fixed array of 64 MiB per GPU
 * (reserved when the library loads on that GPU, held until the process
 * ends; G500 R-MAT scale 16 edge factor 16 fits), so a product whose
 * scratch would exceed it fails with AMS_ERR_LIMIT and does nothing.  It
 * neither allocates nor synchronises: c_row_ptr (num_row(A) + 1) receives
 * C's row pointers and *nnz_dev (device int64) nnz(C); c_col / c_val are
 * caller buffers of c_cap entries, and a C of more than c_cap entries
 * writes none (read *nnz_dev, then call again with a larger buffer).
 * ams_get_error_flags(ctx) reads the call's kernel error flags.  Heavy
 * rows are counted whole (no split into pieces).  All calls on one GPU
 * share its fixed array and leave their contents in it: calls that overlap
 * on one GPU (other streams or contexts) race, so order them on one stream.
 * The call does not change ctx's own scratch; it invalidates ctx's
 * symbolic plan (an ams_spgemm_numeric must follow a new symbolic call). */
int ams_spgemm_fixed(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b, int dtype,
	int64_t *c_row_ptr_dev, int32_t *c_col_dev, void *c_val_dev, int64_t c_cap, int64_t *nnz_dev);

/* Error flags the kernels of the last run on ctx raised, read after the
 * stream drains (so it also serves after a graph replay or a batched run,
 * which do not synchronise): 0x1 LDS-hash row count differs from its
 * symbolic count, 0x2 ESC row count differs, 0x4 dense chunk count
 * differs, 0x8 symbolic hash table full, 0x10 numeric hash table full,
 * 0x20 global-hash row count differs, 0x40 / 0x80 symbolic / numeric
 * global hash table full.
 * Any non-zero value means C is not to be trusted.  No counterpart in the
 * reference (its kernels cannot detect these). */
int ams_get_error_flags(ams_ctx_t *ctx, int32_t *flags);

/* Masked-SpGEMM backward of C = A * B with upstream gradient G on pattern(C)
 * (same row pointers and columns as C): dA = (G * B^T) masked to pattern(A),
 * dB = (A^T * G) masked to pattern(B).  One pass over the products with G's
 * row as the pre-seeded key set; dA_val / dB_val (nnz(A), nnz(B)) are
 * overwritten.  No counterpart in the reference. */
int ams_spgemm_masked_bwd(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b,
	const ams_dcsr_t *g, int dtype, void *da_val_dev, void *db_val_dev);

/* Autotuner: times candidate parameter sets on (A, B) and installs the
 * fastest in ctx; *best (may be NULL) receives it.  No counterpart in the
 * reference (its thresholds are compile-time constants). */
int ams_autotune(ams_ctx_t *ctx, const ams_dcsr_t *a, const ams_dcsr_t *b, int dtype,
	int reps, ams_params_t *best, double *best_ms);

/* The reference's IN_VERIFY checksum line (src/ams_hybaccspgemm.c:320-329):
 * nnz(C), the sum of all column indices and the sum of all values, plus the
 * number of rows whose columns are not strictly ascending.  Σ val is
 * accumulated in FP64 with atomics (exact for integer-valued C). */
int ams_verify_result(ams_ctx_t *ctx, const ams_dcsr_t *c, int dtype, int64_t *nnz,
	int64_t *sum_col, double *sum_val, int64_t *unsorted_rows);

/* Device memory helpers for host code written in plain C (the drop-in). */
int ams_dev_alloc(void **ptr, int64_t bytes);
int ams_dev_free(void *ptr);
int ams_memcpy_h2d(ams_ctx_t *ctx, void *dst, const void *src, int64_t bytes);
int ams_memcpy_d2h(ams_ctx_t *ctx, void *dst, const void *src, int64_t bytes);
int ams_sync(ams_ctx_t *ctx);
/* Conversions between the reference's int64 host layout and device int32. */
int ams_i64_to_i32(ams_ctx_t *ctx, const int64_t *in_dev, int32_t *out_dev, int64_t n);
int ams_i32_to_i64(ams_ctx_t *ctx, const int32_t *in_dev, int64_t *out_dev, int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* AMS_HIP_H */
