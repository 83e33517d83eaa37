/*
 * ams_dropin.h -- the reference's own C entry points for the hybrid
 * accumulation SpGEMM, served by the MI355X library (libams_hip.so).
 *
 * A program written against the reference's headers (src/inc/store.h,
 * src/mm/mm_utils.h, src/ams/ams_utils.h) links against libams_hip.so
 * instead of libAms/libMm/libInc and keeps its calls.  Types keep the
 * reference's byte layout (int64 indices, double values).  Inputs are host
 * arrays: a call uploads what it needs (once per driver sequence, see
 * ams_dropin_counters), runs the device pipeline on one GPU and downloads
 * its outputs.  Errors print a message and exit, as the
 * reference does (src/inc/utils.c:205-216).  Not reentrant: one process-wide
 * device context, like the reference's globals (src/ams/ams_utils.c:9,20,31-32).
 *
 * Column order: every output row has ascending columns.  The reference emits
 * hash rows in hash-slot order (src/ams/ams_utils.c:1008-1016) and ESC rows
 * sorted; a caller that needs slot order must not rely on this library.
 */
#ifndef AMS_DROPIN_H
#define AMS_DROPIN_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/inc/store.h:19-20 */
typedef int64_t my_cnt_t;
typedef double my_rat_t;

/* src/inc/store.h:49-56 */
typedef struct {
	my_cnt_t num_row;
	my_cnt_t num_col;
	my_cnt_t num_nnz;
	my_cnt_t *row_ptrs;
	my_cnt_t *cols;
	my_rat_t *rats;
} csr_t;

/* src/inc/store.h:66-69 */
typedef struct {
	my_cnt_t key;
	my_rat_t val;
} keyval_t;

/* src/mm/mm_utils.h:48-54 */
typedef struct {
	double init_time;
	double symb_time;
	double mgmt_time;
	double numc_time;
	double total_time;
} runtime_t;

/* src/inc/utils.h:100-107: NUMA placement has no meaning on one GPU; the
 * pointer is accepted and ignored. */
typedef struct ams_numa_cfg numa_cfg_t;

/* src/ams/ams_utils.h:11-37 (three-region form). */
typedef struct {
	keyval_t *b_roff_maxc;  /* key: first row of the morsel, val: max row flops */
	keyval_t *b_arr;        /* key: morsel id, val: compression factor          */
	my_cnt_t total_cnt;
	my_cnt_t high_head_idx;
	my_cnt_t high_tail_idx;
	my_cnt_t low_head_idx;
	my_cnt_t low_tail_idx;
} ams_pack_t;

/* src/ams/ams_utils.h:40-46 */
typedef struct {
	csr_t *csr;
	ams_pack_t *ams_pack;
	my_cnt_t *row_flp;
	my_cnt_t *ps_row_flp;
	my_cnt_t *row_nz;
} ams_bin_t;

/* Whole products.  The reference's drivers are static functions of their
 * mains: hybaccspgemm src/ams_hybaccspgemm.c:36, hashspgemm
 * src/ams_hashspgemm.c:36, hserscspgemm src/ams_hserscspgemm.c:36.  C's
 * arrays are allocated here with mmap (free with free_csr). */
runtime_t hybaccspgemm(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, numa_cfg_t *numa_cfg);
runtime_t hashspgemm(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, numa_cfg_t *numa_cfg);
runtime_t hserscspgemm(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, numa_cfg_t *numa_cfg);

/* src/mm/mm_utils.h:72 (src/mm/mm_utils.c:26-33) */
void free_csr(csr_t *csr);
/* src/mm/mm_utils.h:184 (src/mm/mm_utils.c:866-957) */
void scan(my_cnt_t *out, my_cnt_t *in, my_cnt_t num);
/* src/mm/mm_utils.h:195 (src/mm/mm_utils.c:981-1055) */
my_cnt_t get_intprod(csr_t *csr_a, csr_t *csr_b, my_cnt_t *row_nz);
/* src/mm/mm_utils.h:196 (src/mm/mm_utils.c:1063-1125) */
void flp_populate(csr_t *csr_a, csr_t *csr_b, my_cnt_t *row_flp);
/* src/mm/mm_utils.h:111 (src/mm/mm_utils.c:963-974) */
my_cnt_t *lower_bound(my_cnt_t *start, my_cnt_t *end, my_cnt_t val);

/* src/inc/store.h:22-26 */
typedef struct {
	my_cnt_t user_id;
	my_cnt_t item_id;
	my_rat_t rating;
} triple_t;

/* MatrixMarket / PaRMAT text I/O (ams_mtx.c), src/mm/mm_utils.h:75-101 with
 * the file semantics of src/mm/mm_utils.c:451-862.  Triple arrays are
 * mmap'ed; release them with dealloc_memory(ptr, bytes) (src/inc/utils.c:254-260). */
void dealloc_memory(void *ptr, size_t size);
int compare_triples(const void *a, const void *b);
bool mtx2trpl(triple_t **trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t *num_nnz, const char *filename,
	const int numa_node_idx);
void trpl2mtx(const char *filename, triple_t *trpl, my_cnt_t num_row, my_cnt_t num_col, my_cnt_t num_nnz,
	bool isSymmetric);
void trpl2csr(csr_t *csr, triple_t *trpl, const my_cnt_t num_row, const my_cnt_t num_col, const my_cnt_t num_nnz,
	numa_cfg_t *numa_cfg);
void mtx2csr(csr_t *csr, const char *filename, numa_cfg_t *numa_cfg);
/* src/mm/mm_utils.c:730-732: what the drivers call (src/ams_hybaccspgemm.c:294-299) */
void raw2csr(csr_t *csr, const char *filename, numa_cfg_t *numa_cfg);
void trpl_ones(triple_t *trpl, my_cnt_t num_nnz);
void trpl_transpose(triple_t *trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t num_nnz);
void txt2trpl(triple_t **trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t *num_nnz, const char *filename,
	const int numa_node_idx);
/* src/rmat_ones.c:6-37: PaRMAT text -> symmetric all-ones .org.mtx and its transpose .trs.mtx */
void txt_transpose(const char *filename_dest_trs, const char *filename_dest_org, const char *filename_src,
	const int numa_node_idx);

/* src/ams/ams_utils.h:60-94 */
void free_ams_pack(ams_pack_t *ams_pack);
void init_ams_bin(ams_bin_t *ams_bin, ams_pack_t *ams_pack, csr_t *csr, numa_cfg_t *numa_cfg);
void free_ams_bin(ams_bin_t *ams_bin);
my_cnt_t pack_ams_pack(my_cnt_t *ps_row_nz, ams_bin_t *ams_bin, numa_cfg_t *numa_cfg);
extern my_rat_t ams_ht_load;
my_cnt_t ams_ht_slots_for(my_cnt_t nz);
void ams_hash_symb(csr_t *csr_a, csr_t *csr_b, my_cnt_t *ht_check, ams_bin_t *ams_bin);
void ams_hash_symb_dnr(csr_t *csr_a, csr_t *csr_b, my_cnt_t *ht_check, ams_bin_t *ams_bin);
/* src/ams/ams_utils.h:88 (src/ams/ams_utils.c:1754-1846): the OpenMP form of
 * the dynamic-range symbolic; the same row counts. */
void ams_hash_symb_omp_dnr(csr_t *csr_a, csr_t *csr_b, my_cnt_t *ht_check, ams_bin_t *ams_bin);
/* src/ams/ams_utils.h:92-94 (src/ams/ams_utils.c:1851-2046): the morsel table
 * over nonzeros of csr (ps_row_coef: scaled prefix of nonzeros per row). */
my_cnt_t pack_ams_pack_mmmv(csr_t *csr, ams_pack_t *ams_pack, my_cnt_t *ps_row_coef, my_cnt_t *nnz_per_row,
	my_cnt_t coef, numa_cfg_t *numa_cfg);
void ams_hash_numc(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, my_cnt_t *ht_check, my_rat_t *ht_value,
	ams_bin_t *ams_bin);
void ams_hsersc_numc(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, keyval_t *keyval_org, keyval_t *aux,
	my_cnt_t max_row_flp, ams_bin_t *ams_bin);
void ams_hybacc_numc(csr_t *csr_a, csr_t *csr_b, csr_t *csr_c, keyval_t *keyval_org, keyval_t *aux,
	my_cnt_t max_row_flp, my_cnt_t *ht_check, my_rat_t *ht_value, ams_bin_t *ams_bin);

/* Device copies of a driver sequence.  Between init_ams_bin and free_ams_bin
 * of one ams_bin_t (the reference's driver sequence,
 * src/ams_hybaccspgemm.c:53-199) the library keeps A and B on the GPU and
 * reuses the symbolic phase's plan in the numeric phase, so each input is
 * uploaded once and the symbolic phase runs once; A and B must not change
 * during the sequence.  Outside a sequence every call uploads and frees.
 * No counterpart in the reference: ams_dropin_counters reports the matrices
 * uploaded and the device symbolic phases run since the process started,
 * ams_dropin_release drops the kept copies (for a caller that changes A or
 * B in place inside a sequence). */
void ams_dropin_counters(my_cnt_t *uploads, my_cnt_t *symbolic_runs);
void ams_dropin_release(void);

/* Morsel parameters of pack_ams_pack, compile-time in the reference
 * (cmake/Options.cmake:104-128; values of scripts/ams_run_common.sh:26 and
 * scripts/ams_rmat_run.sh:33).  No counterpart function in the reference. */
void ams_dropin_set_pack_params(my_cnt_t ams_size, my_cnt_t ams_min_num, double ams_prt);

#ifdef __cplusplus
}
#endif

#endif /* AMS_DROPIN_H */
