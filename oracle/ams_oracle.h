/*
 * ams_oracle.h -- CPU restatement of the reference's AMS hybrid-accumulation
 * SpGEMM (ams_hybaccspgemm) and its two single-accumulator siblings.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the correctness oracle and the
 * reported CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * Types mirror src/inc/store.h:19-69 (int64 indices, double values) and
 * src/mm/mm_utils.h:48-54 (runtime_t).  Every function cites the reference
 * file:line it restates.
 */
#ifndef AMS_ORACLE_H
#define AMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* csr_t of src/inc/store.h:49-56, same field order. */
typedef struct {
	int64_t num_row;
	int64_t num_col;
	int64_t num_nnz;
	int64_t *row_ptrs;
	int64_t *cols;
	double *rats;
} orc_csr_t;

/* keyval_t of src/inc/store.h:66-69. */
typedef struct {
	int64_t key;
	double val;
} orc_keyval_t;

/* runtime_t of src/mm/mm_utils.h:48-54. */
typedef struct {
	double init_time;
	double symb_time;
	double mgmt_time;
	double numc_time;
	double total_time;
} orc_runtime_t;

/* ams_pack_t of src/ams/ams_utils.h:11-37 (three-region variant). */
typedef struct {
	orc_keyval_t *b_roff_maxc; /* key: first row of morsel, val: max row flops */
	orc_keyval_t *b_arr;       /* key: morsel id, val: compression factor      */
	int64_t total_cnt;
	int64_t high_head_idx;
	int64_t high_tail_idx;
	int64_t low_head_idx;
	int64_t low_tail_idx;
} orc_pack_t;

/* Compile-time knobs of cmake/Options.cmake, made run-time parameters. */
typedef struct {
	int thread_num;          /* THREAD_NUM                                   */
	int smt_core_num;        /* SOCKET_CORE_NUM: threads kept under SMTACT   */
	int high_thread_num;     /* AMS_HIGH_THREAD_NUM                          */
	int64_t ams_size;        /* AMS_SIZE (bytes)                             */
	int64_t ams_min_num;     /* AMS_MIN_NUM                                  */
	double ams_prt;          /* AMS_PRT                                      */
	double acc_thres;        /* AMS_ACC_THRES                                */
	int64_t smt_symb_thres;  /* SMTACT_SYMB_THRES (0 disables SMTACT)        */
	int64_t smt_numc_thres;  /* SMTACT_NUMC_THRES                            */
	int64_t dn_min;          /* DN_MIN                                       */
	int64_t hash_constant;   /* HASH_CONSTANT                                */
} orc_params_t;

/* Numeric accumulator selection: the three reference binaries. */
enum { ORC_HYBACC = 0, ORC_HASH = 1, ORC_ESC = 2 };

void orc_default_params(orc_params_t *p, int thread_num);

int64_t orc_next_pow2(int64_t x);
int64_t *orc_lower_bound(int64_t *start, int64_t *end, int64_t val);
void orc_scan(int64_t *out, const int64_t *in, int64_t num);
void orc_flp_populate(const orc_csr_t *a, const orc_csr_t *b, int64_t *row_flp, int thread_num);
int64_t orc_get_intprod(const orc_csr_t *a, const orc_csr_t *b, int thread_num);
void orc_hash_symb_dnr(const orc_csr_t *a, const orc_csr_t *b, const int64_t *row_flp,
	int64_t *row_nz, const orc_params_t *p, int thread_num);
int64_t orc_pack(const int64_t *ps_row_nz, const int64_t *row_flp, const int64_t *ps_row_flp,
	int64_t num_row, const orc_params_t *p, orc_pack_t *pack);
int64_t orc_pack_mmmv(const int64_t *row_ptrs, int64_t num_row, int64_t num_nnz, const int64_t *ps_row_coef,
	const int64_t *nnz_per_row, int64_t coef, const orc_params_t *p, orc_pack_t *pack);
void orc_free_pack(orc_pack_t *pack);
orc_keyval_t *orc_radix_sort_keyval(orc_keyval_t *arr, orc_keyval_t *aux, int64_t n);
void orc_numc(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c, const int64_t *row_flp,
	const orc_pack_t *pack, int64_t max_row_flp, const orc_params_t *p, int mode, int thread_num);

/* Whole driver (src/ams_hybaccspgemm.c:36-220).  Allocates c->row_ptrs,
 * c->cols, c->rats with malloc; release with orc_free_csr. */
orc_runtime_t orc_spgemm(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c,
	const orc_params_t *p, int mode);
void orc_free_csr(orc_csr_t *c);

/* Per-row ascending column sort (stable); values follow their columns. */
void orc_canonicalize(orc_csr_t *c, int thread_num);

/* Symbolic-only driver: row_nz[i] = nnz of C row i (no morsels needed). */
void orc_symbolic(const orc_csr_t *a, const orc_csr_t *b, int64_t *row_nz, const orc_params_t *p);

#ifdef __cplusplus
}
#endif

#endif /* AMS_ORACLE_H */
