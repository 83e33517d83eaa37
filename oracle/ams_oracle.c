/*
 * ams_oracle.c -- CPU restatement of the reference's AMS hybrid-accumulation
 * SpGEMM path.  TEST INFRASTRUCTURE ONLY (correctness oracle and the reported
 * CPU baseline): only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product library never links it.
 *
 * What is restated, and from where (paths relative to the reference root):
 *   orc_next_pow2          src/inc/utils.h:92                (NEXT_POW_2)
 *   orc_lower_bound        src/mm/mm_utils.c:963-974         (lower_bound)
 *   orc_scan               src/mm/mm_utils.c:866-957         (scan)
 *   orc_flp_populate       src/mm/mm_utils.c:1063-1125       (flp_populate)
 *   orc_get_intprod        src/mm/mm_utils.c:981-1055        (get_intprod)
 *   orc_hash_symb_dnr      src/ams/ams_utils.c:1600-1752     (ams_hash_symb_dnr)
 *   orc_pack               src/ams/ams_utils.c:1410-1596     (pack_ams_pack)
 *                          + src/mm/mm_utils.c:57-65         (compare_keyvals_3)
 *   orc_pack_mmmv          src/ams/ams_utils.c:1851-2046     (pack_ams_pack_mmmv)
 *   orc_radix_sort_keyval  src/mm/mm_utils.c:1273-1375       (radix_sort_keyval_toggling_unrolling)
 *   hash_row               src/ams/ams_utils.c:956-1019      (hybacc_hash_krnl)
 *   esc_row                src/ams/ams_utils.c:1023-1075     (hybacc_hsersc_krnl)
 *   orc_numc               src/ams/ams_utils.c:1080-1338     (ams_hybacc_numc + _krnl),
 *                          and its hash-only / ESC-only forms :168-328, :335-475
 *   orc_spgemm             src/ams_hybaccspgemm.c:36-220     (hybaccspgemm driver)
 *
 * Deviations, all of them guards that change no result:
 *   - pack: total_cnt = nnz/T is raised to 1 when it would be 0 (the
 *     reference then calls mmap(0) and exits, src/inc/utils.c:204-210).
 *   - numc: AMS_HIGH_THREAD_NUM is clamped so that at least one thread is not
 *     in region a; with every thread in region a the morsel at high_tail_idx
 *     would never be claimed.
 *   - NUMA/CXL placement, core pinning and libconfig are dropped (malloc).
 */
#define _GNU_SOURCE
#include "ams_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static void *xmalloc(size_t bytes) {
	void *p = malloc(bytes ? bytes : 1);
	if (!p) {
		fprintf(stderr, "ams_oracle: out of memory (%zu bytes)\n", bytes);
		exit(EXIT_FAILURE);
	}
	return p;
}

static double elapsed(struct timespec a, struct timespec b) {
	return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

void orc_default_params(orc_params_t *p, int thread_num) {
	/* Values of the reference's hybacc R-MAT profile:
	 * scripts/ams_run_common.sh:20-35 and scripts/ams_rmat_run.sh:32-33. */
	if (thread_num < 1) thread_num = 1;
	p->thread_num = thread_num;
	p->smt_core_num = thread_num;          /* physical cores only: SMTACT keeps them all */
	p->high_thread_num = (int)lround(22.0 / 64.0 * thread_num);
	p->ams_size = 524288;
	p->ams_min_num = 32768;
	p->ams_prt = 0.85;
	p->acc_thres = 2.00;
	p->smt_symb_thres = 7000000;
	p->smt_numc_thres = 2100000;
	p->dn_min = 256;
	p->hash_constant = 2654435761LL;
}

/* src/inc/utils.h:92 -- the same libm expression, so the quirks above 2^29
 * are reproduced too. */
int64_t orc_next_pow2(int64_t x) {
	return (int64_t)pow(2, ceil(log((double)x) / log(2)));
}

/* src/mm/mm_utils.c:963-974: first element not less than val. */
int64_t *orc_lower_bound(int64_t *start, int64_t *end, int64_t val) {
	while (start < end) {
		int64_t *mid = start + (end - start) / 2;
		if (*mid < val) start = mid + 1;
		else end = mid;
	}
	return start;
}

/* src/mm/mm_utils.c:866-957: num outputs, out[0] = 0, out[i+1] = out[i] + in[i].
 * The reference splits num >= 2^17 over threads in two passes; integer sums
 * make that identical to this serial form. */
void orc_scan(int64_t *out, const int64_t *in, int64_t num) {
	if (num <= 0) return;
	out[0] = 0;
	for (int64_t i = 0; i < num - 1; i++) out[i + 1] = out[i] + in[i];
}

/* src/mm/mm_utils.c:1063-1125: row_flp[i] = sum over A(i,:) of |B(k,:)|. */
void orc_flp_populate(const orc_csr_t *a, const orc_csr_t *b, int64_t *row_flp, int thread_num) {
	#pragma omp parallel for num_threads(thread_num) schedule(static)
	for (int64_t i = 0; i < a->num_row; i++) {
		int64_t f = 0;
		for (int64_t p = a->row_ptrs[i]; p < a->row_ptrs[i + 1]; p++) {
			int64_t k = a->cols[p];
			f += b->row_ptrs[k + 1] - b->row_ptrs[k];
		}
		row_flp[i] = f;
	}
}

/* src/mm/mm_utils.c:981-1055: total intermediate products P. */
int64_t orc_get_intprod(const orc_csr_t *a, const orc_csr_t *b, int thread_num) {
	int64_t total = 0;
	#pragma omp parallel for num_threads(thread_num) schedule(static) reduction(+:total)
	for (int64_t i = 0; i < a->num_row; i++) {
		for (int64_t p = a->row_ptrs[i]; p < a->row_ptrs[i + 1]; p++) {
			int64_t k = a->cols[p];
			total += b->row_ptrs[k + 1] - b->row_ptrs[k];
		}
	}
	return total;
}

/* Table size used by both hash phases: NEXT_POW_2(min(row_flp, num_col))
 * (ams_utils.c:1647-1651 symbolic, :961-962 numeric; ams_ht_load is 1.0). */
static int64_t table_size(int64_t row_flp, int64_t num_col) {
	int64_t s = row_flp > num_col ? num_col : row_flp;
	return orc_next_pow2(s);
}

/* Symbolic count of one row with a linear-probing table of int64 keys,
 * empty = -1 (ams_utils.c:1647-1686). */
static int64_t symb_row(const orc_csr_t *a, const orc_csr_t *b, int64_t row, int64_t row_flp,
	int64_t *ht, int64_t hc) {
	if (row_flp <= 0) return 0;
	int64_t size = table_size(row_flp, b->num_col);
	memset(ht, -1, sizeof(int64_t) * (size_t)size);
	int64_t nz = 0;
	for (int64_t p = a->row_ptrs[row]; p < a->row_ptrs[row + 1]; p++) {
		int64_t k = a->cols[p];
		for (int64_t q = b->row_ptrs[k]; q < b->row_ptrs[k + 1]; q++) {
			int64_t key = b->cols[q];
			int64_t h = (key * hc) & (size - 1);
			for (;;) {
				if (ht[h] == key) break;
				if (ht[h] == -1) { ht[h] = key; nz++; break; }
				h = (h + 1) & (size - 1);
			}
		}
	}
	return nz;
}

/* src/ams/ams_utils.c:1600-1752: dynamic-range symbolic.  Threads claim
 * DN_MIN rows at a time from a shared cursor with CAS (get_next_rng :1615). */
void orc_hash_symb_dnr(const orc_csr_t *a, const orc_csr_t *b, const int64_t *row_flp,
	int64_t *row_nz, const orc_params_t *p, int thread_num) {
	volatile int64_t cursor = 0;
	int64_t tsize = orc_next_pow2(b->num_col > 0 ? b->num_col : 1);
	int64_t m = a->num_row;
	int64_t dn = p->dn_min > 0 ? p->dn_min : 1;
	#pragma omp parallel num_threads(thread_num)
	{
		int64_t *ht = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)tsize);
		for (;;) {
			int64_t start, stop;
			do {
				start = cursor;
				stop = start + dn;
				if (stop > m) stop = m;
			} while (__sync_val_compare_and_swap(&cursor, start, stop) != start);
			if (start >= m) break;
			for (int64_t r = start; r < stop; r++)
				row_nz[r] = symb_row(a, b, r, row_flp[r], ht, p->hash_constant);
		}
		free(ht);
	}
}

/* compare_keyvals_3 (src/mm/mm_utils.c:57-65): val ascending, then key. */
static int cmp_val_key(const void *x, const void *y) {
	const orc_keyval_t *u = (const orc_keyval_t *)x, *v = (const orc_keyval_t *)y;
	int c = (u->val > v->val) - (u->val < v->val);
	if (c) return c;
	return (u->key > v->key) - (u->key < v->key);
}

/* src/ams/ams_utils.c:1410-1596 (AMS_MAXFLP off, SMTACT_REAR on).
 * Morsel k spans rows [lower_bound(ps, T*k), lower_bound(ps, T*(k+1))),
 * T = MIN(AMS_MIN_NUM, AMS_SIZE/8), the last ends at num_row.  Each valid
 * morsel gets cf = flops/nnz and the max row flops; empty morsels are
 * counted into low_head_idx.  Morsels are sorted by (cf, id) and split into
 * the low region and the top (1 - AMS_PRT) high region.  Returns the max row
 * flops over all morsels. */
int64_t orc_pack(const int64_t *ps_row_nz, const int64_t *row_flp, const int64_t *ps_row_flp,
	int64_t num_row, const orc_params_t *p, orc_pack_t *pack) {
	int64_t t = p->ams_min_num;
	if (p->ams_size / 8 < t) t = p->ams_size / 8;
	if (t < 1) t = 1;
	int64_t total = ps_row_nz[num_row] / t;
	if (total < 1) total = 1;  /* guard: the reference exits here */
	pack->total_cnt = total;
	pack->b_arr = (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)total);
	pack->b_roff_maxc = (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)(total + 1));
	pack->b_roff_maxc[0].key = 0;
	pack->b_roff_maxc[0].val = 0.0;
	for (int64_t k = 0; k < total; k++) {
		int64_t *lb = orc_lower_bound((int64_t *)ps_row_nz, (int64_t *)ps_row_nz + num_row + 1,
			t * (k + 1));
		pack->b_roff_maxc[k + 1].key = lb - ps_row_nz;
		pack->b_roff_maxc[k + 1].val = 0.0;
		pack->b_arr[k].key = k;
	}
	pack->b_roff_maxc[total].key = num_row;

	int64_t invalid = 0;
	double max_flp = 0.0;
	for (int64_t k = 0; k < total; k++) {
		int64_t s = pack->b_roff_maxc[k].key, e = pack->b_roff_maxc[k + 1].key;
		if (e - s > 0) {
			pack->b_arr[k].val = (double)(ps_row_flp[e] - ps_row_flp[s]) /
				(double)(ps_row_nz[e] - ps_row_nz[s]);
			int64_t mx = row_flp[s];
			for (int64_t r = s + 1; r < e; r++) if (row_flp[r] > mx) mx = row_flp[r];
			pack->b_roff_maxc[k].val = (double)mx;
			if ((double)mx > max_flp) max_flp = (double)mx;
		} else {
			pack->b_roff_maxc[k].val = 0.0;
			pack->b_arr[k].val = 0.0;
			invalid++;
		}
	}
	qsort(pack->b_arr, (size_t)total, sizeof(orc_keyval_t), cmp_val_key);
	pack->low_head_idx = invalid;
	pack->high_head_idx = total - 1;
	pack->high_tail_idx = (int64_t)((double)(total - invalid) * p->ams_prt + (double)invalid);
	pack->low_tail_idx = pack->high_tail_idx - 1;
	return (int64_t)max_flp;
}

/* src/ams/ams_utils.c:1851-2046 (AMS_MAXFLP off): morsels over the scaled
 * nonzero prefix ps_row_coef, total = nnz * coef / T; a valid morsel's sort
 * value is its mean nonzeros per row, (row_ptrs[e] - row_ptrs[s]) / (e - s),
 * and its b_roff_maxc value the largest nnz_per_row of its rows.  The rest
 * (empties, sort, regions) as orc_pack; the same total = 0 guard. */
int64_t orc_pack_mmmv(const int64_t *row_ptrs, int64_t num_row, int64_t num_nnz, const int64_t *ps_row_coef,
	const int64_t *nnz_per_row, int64_t coef, const orc_params_t *p, orc_pack_t *pack) {
	int64_t t = p->ams_min_num;
	if (p->ams_size / 8 < t) t = p->ams_size / 8;
	if (t < 1) t = 1;
	int64_t total = num_nnz * coef / t;
	if (total < 1) total = 1;  /* guard: the reference exits here */
	pack->total_cnt = total;
	pack->b_arr = (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)total);
	pack->b_roff_maxc = (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)(total + 1));
	pack->b_roff_maxc[0].key = 0;
	pack->b_roff_maxc[0].val = 0.0;
	for (int64_t k = 0; k < total; k++) {
		int64_t *lb = orc_lower_bound((int64_t *)ps_row_coef, (int64_t *)ps_row_coef + num_row + 1,
			t * (k + 1));
		pack->b_roff_maxc[k + 1].key = lb - ps_row_coef;
		pack->b_roff_maxc[k + 1].val = 0.0;
		pack->b_arr[k].key = k;
	}
	pack->b_roff_maxc[total].key = num_row;
	int64_t invalid = 0;
	double max_nz = 0.0;
	for (int64_t k = 0; k < total; k++) {
		int64_t s = pack->b_roff_maxc[k].key, e = pack->b_roff_maxc[k + 1].key;
		if (e - s > 0) {
			pack->b_arr[k].val = (double)(row_ptrs[e] - row_ptrs[s]) / (double)(e - s);
			int64_t mx = nnz_per_row[s];
			for (int64_t r = s + 1; r < e; r++) if (nnz_per_row[r] > mx) mx = nnz_per_row[r];
			pack->b_roff_maxc[k].val = (double)mx;
			if ((double)mx > max_nz) max_nz = (double)mx;
		} else {
			pack->b_roff_maxc[k].val = 0.0;
			pack->b_arr[k].val = 0.0;
			invalid++;
		}
	}
	qsort(pack->b_arr, (size_t)total, sizeof(orc_keyval_t), cmp_val_key);
	pack->low_head_idx = invalid;
	pack->high_head_idx = total - 1;
	pack->high_tail_idx = (int64_t)((double)(total - invalid) * p->ams_prt + (double)invalid);
	pack->low_tail_idx = pack->high_tail_idx - 1;
	return (int64_t)max_nz;
}

void orc_free_pack(orc_pack_t *pack) {
	free(pack->b_arr);
	free(pack->b_roff_maxc);
	memset(pack, 0, sizeof(*pack));
}

/* src/mm/mm_utils.c:1273-1375: LSD radix sort on the key bytes that the
 * maximum key needs, stable, toggling between the two buffers.  Returns the
 * buffer that holds the sorted data. */
orc_keyval_t *orc_radix_sort_keyval(orc_keyval_t *arr, orc_keyval_t *aux, int64_t n) {
	if (n < 2) return arr;
	bool same = true;
	for (int64_t i = 1; i < n; i++) if (arr[i].key != arr[0].key) { same = false; break; }
	if (same) return arr;
	int64_t mx = 0;
	for (int64_t i = 0; i < n; i++) if (arr[i].key > mx) mx = arr[i].key;
	int passes = 0;
	while (mx > 0) { mx >>= 8; passes++; }
	if (passes == 0) passes = 1;
	if (passes > 8) passes = 8;
	orc_keyval_t *in = arr, *out = aux;
	for (int pass = 0; pass < passes; pass++) {
		size_t count[256];
		memset(count, 0, sizeof(count));
		int sh = 8 * pass;
		for (int64_t i = 0; i < n; i++) count[((uint64_t)in[i].key >> sh) & 0xff]++;
		for (int d = 1; d < 256; d++) count[d] += count[d - 1];
		for (int64_t j = n - 1; j >= 0; j--) out[--count[((uint64_t)in[j].key >> sh) & 0xff]] = in[j];
		orc_keyval_t *tmp = in; in = out; out = tmp;
	}
	return in;
}

/* Per-thread scratch of the numeric phase (src/ams_hybaccspgemm.c:187-196). */
typedef struct {
	int64_t *ht_check;
	double *ht_value;
	orc_keyval_t *kv_org;
	orc_keyval_t *kv_aux;
} scratch_t;

/* hybacc_hash_krnl (src/ams/ams_utils.c:956-1019): linear-probing table;
 * the first product of a key is assigned, later ones added; output in slot
 * order. */
static void hash_row(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c, int64_t row,
	int64_t row_flp, scratch_t *s, int64_t hc) {
	int64_t size = table_size(row_flp, b->num_col);
	int64_t *ht = s->ht_check;
	double *hv = s->ht_value;
	memset(ht, -1, sizeof(int64_t) * (size_t)size);
	for (int64_t p = a->row_ptrs[row]; p < a->row_ptrs[row + 1]; p++) {
		int64_t k = a->cols[p];
		double av = a->rats[p];
		for (int64_t q = b->row_ptrs[k]; q < b->row_ptrs[k + 1]; q++) {
			int64_t key = b->cols[q];
			int64_t h = (key * hc) & (size - 1);
			volatile double prod = av * b->rats[q]; /* func_mul: never fused */
			for (;;) {
				if (ht[h] == key) { hv[h] += prod; break; }
				if (ht[h] == -1) { ht[h] = key; hv[h] = prod; break; }
				h = (h + 1) & (size - 1);
			}
		}
	}
	int64_t off = c->row_ptrs[row], idx = 0;
	for (int64_t h = 0; h < size; h++) {
		if (ht[h] != -1) {
			c->cols[off + idx] = ht[h];
			c->rats[off + idx] = hv[h];
			idx++;
		}
	}
}

/* hybacc_hsersc_krnl (src/ams/ams_utils.c:1023-1075): expand the row's
 * products in A-row order, stable radix sort by column, then compress runs of
 * equal columns left to right (first product assigned, the rest added). */
static void esc_row(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c, int64_t row,
	scratch_t *s) {
	int64_t n0 = 0;
	orc_keyval_t *kv = s->kv_org;
	for (int64_t p = a->row_ptrs[row]; p < a->row_ptrs[row + 1]; p++) {
		int64_t k = a->cols[p];
		double av = a->rats[p];
		for (int64_t q = b->row_ptrs[k]; q < b->row_ptrs[k + 1]; q++) {
			volatile double prod = av * b->rats[q];
			kv[n0].key = b->cols[q];
			kv[n0].val = prod;
			n0++;
		}
	}
	orc_keyval_t *sorted = orc_radix_sort_keyval(s->kv_org, s->kv_aux, n0);
	orc_keyval_t *dst = sorted == s->kv_org ? s->kv_aux : s->kv_org;
	int64_t n1 = 0;
	dst[0] = sorted[0];
	for (int64_t i = 1; i < n0; i++) {
		if (sorted[i].key == sorted[i - 1].key) {
			dst[n1].val = dst[n1].val + sorted[i].val;
		} else {
			n1++;
			dst[n1] = sorted[i];
		}
	}
	int64_t off = c->row_ptrs[row];
	for (int64_t i = 0; i <= n1; i++) {
		c->cols[off + i] = dst[i].key;
		c->rats[off + i] = dst[i].val;
	}
}

/* One claimant of the morsel queue (ams_hybacc_numc_krnl,
 * src/ams/ams_utils.c:1080-1154).  `ascend` selects the direction: +1 with
 * stop when head > tail, -1 with stop when head <= tail. */
static void claim_loop(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c, const int64_t *row_flp,
	const orc_pack_t *pack, volatile int64_t *head, volatile int64_t *tail, int ascend,
	const orc_params_t *p, int mode, scratch_t *s) {
	for (;;) {
		int64_t cur, nxt;
		do {
			cur = *head;
			int64_t t = *tail;
			if (ascend ? (cur > t) : (cur <= t)) return;
			nxt = cur + (ascend ? 1 : -1);
		} while (__sync_val_compare_and_swap(head, cur, nxt) != cur);
		int64_t id = pack->b_arr[cur].key;
		int64_t r0 = pack->b_roff_maxc[id].key, r1 = pack->b_roff_maxc[id + 1].key;
		for (int64_t r = r0; r < r1; r++) {
			int64_t f = row_flp[r];
			if (f <= 0) continue;
			int use_esc;
			if (mode == ORC_HASH) use_esc = 0;
			else if (mode == ORC_ESC) use_esc = 1;
			else {
				double cf = (double)f / (double)(c->row_ptrs[r + 1] - c->row_ptrs[r]);
				use_esc = cf < p->acc_thres;  /* ams_utils.c:1141-1149 */
			}
			if (use_esc) esc_row(a, b, c, r, s);
			else hash_row(a, b, c, r, f, s, p->hash_constant);
		}
	}
}

/* ams_hybacc_numc (src/ams/ams_utils.c:1159-1338), three-region form:
 * the last high/2 threads descend from high_head (region a), the next high/2
 * ascend from high_tail (region b), the rest ascend from low_head (region d);
 * then a/b threads help the low region and d threads help from high_tail. */
void orc_numc(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c, const int64_t *row_flp,
	const orc_pack_t *pack, int64_t max_row_flp, const orc_params_t *p, int mode, int thread_num) {
	volatile int64_t high_head = pack->high_head_idx;
	volatile int64_t high_tail = pack->high_tail_idx;
	volatile int64_t low_head = pack->low_head_idx;
	volatile int64_t low_tail = pack->low_tail_idx;
	int high = p->high_thread_num;
	if (high / 2 >= thread_num) high = 2 * (thread_num - 1);  /* guard, see header */
	if (high < 0) high = 0;
	int64_t tsize = orc_next_pow2(b->num_col > 0 ? b->num_col : 1);
	int64_t kvn = max_row_flp > 0 ? max_row_flp : 1;
	#pragma omp parallel num_threads(thread_num)
	{
		int tid = omp_get_thread_num();
		scratch_t s;
		s.ht_check = mode == ORC_ESC ? NULL : (int64_t *)xmalloc(sizeof(int64_t) * (size_t)tsize);
		s.ht_value = mode == ORC_ESC ? NULL : (double *)xmalloc(sizeof(double) * (size_t)tsize);
		s.kv_org = mode == ORC_HASH ? NULL : (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)kvn);
		s.kv_aux = mode == ORC_HASH ? NULL : (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)kvn);
		if (tid >= thread_num - high / 2) {
			claim_loop(a, b, c, row_flp, pack, &high_head, &high_tail, 0, p, mode, &s);
		} else if (tid >= thread_num - high) {
			claim_loop(a, b, c, row_flp, pack, &high_tail, &high_head, 1, p, mode, &s);
		} else {
			claim_loop(a, b, c, row_flp, pack, &low_head, &low_tail, 1, p, mode, &s);
		}
		if (tid >= thread_num - high) {
			claim_loop(a, b, c, row_flp, pack, &low_head, &low_tail, 1, p, mode, &s);
		} else {
			claim_loop(a, b, c, row_flp, pack, &high_tail, &high_head, 1, p, mode, &s);
		}
		free(s.ht_check); free(s.ht_value); free(s.kv_org); free(s.kv_aux);
	}
}

/* The driver, src/ams_hybaccspgemm.c:36-220 (AMS_DNR and SMTACT_REAR on). */
orc_runtime_t orc_spgemm(const orc_csr_t *a, const orc_csr_t *b, orc_csr_t *c,
	const orc_params_t *p, int mode) {
	struct timespec t0, t1;
	orc_runtime_t rt;
	int64_t m = a->num_row;
	int threads = p->thread_num;

	/* INIT :41-82 */
	clock_gettime(CLOCK_MONOTONIC, &t0);
	c->num_row = m;
	c->num_col = b->num_col;
	c->row_ptrs = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m + 1));
	memset(c->row_ptrs, 0, sizeof(int64_t) * (size_t)(m + 1));
	int64_t *row_flp = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
	int64_t *ps_row_flp = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m + 1));
	int64_t *row_nz = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
	orc_flp_populate(a, b, row_flp, threads);
	orc_scan(ps_row_flp, row_flp, m + 1);
	int symb_threads = threads;
	orc_pack_t pack;
	memset(&pack, 0, sizeof(pack));
	if (p->smt_symb_thres > 0) {
		/* SMTACT: pack by A's row pointers only to decide the thread count. */
		int64_t mx = orc_pack(a->row_ptrs, row_flp, ps_row_flp, m, p, &pack);
		if (mx > p->smt_symb_thres && p->smt_core_num < threads) symb_threads = p->smt_core_num;
	}
	clock_gettime(CLOCK_MONOTONIC, &t1);
	rt.init_time = elapsed(t0, t1);

	/* SYMB :85-129 */
	clock_gettime(CLOCK_MONOTONIC, &t0);
	orc_hash_symb_dnr(a, b, row_flp, row_nz, p, symb_threads);
	orc_scan(c->row_ptrs, row_nz, m + 1);
	c->num_nnz = c->row_ptrs[m];
	clock_gettime(CLOCK_MONOTONIC, &t1);
	rt.symb_time = elapsed(t0, t1);

	/* MGMT :141-182 */
	clock_gettime(CLOCK_MONOTONIC, &t0);
	c->cols = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)c->num_nnz);
	c->rats = (double *)xmalloc(sizeof(double) * (size_t)c->num_nnz);
	orc_free_pack(&pack);
	int64_t max_row_flp = orc_pack(c->row_ptrs, row_flp, ps_row_flp, m, p, &pack);
	int numc_threads = threads;
	if (p->smt_symb_thres > 0 && max_row_flp > p->smt_numc_thres && p->smt_core_num < threads)
		numc_threads = p->smt_core_num;
	clock_gettime(CLOCK_MONOTONIC, &t1);
	rt.mgmt_time = elapsed(t0, t1);

	/* NUMC :185-215 */
	clock_gettime(CLOCK_MONOTONIC, &t0);
	orc_numc(a, b, c, row_flp, &pack, max_row_flp, p, mode, numc_threads);
	orc_free_pack(&pack);
	free(row_flp); free(ps_row_flp); free(row_nz);
	clock_gettime(CLOCK_MONOTONIC, &t1);
	rt.numc_time = elapsed(t0, t1);
	rt.total_time = rt.init_time + rt.symb_time + rt.mgmt_time + rt.numc_time;
	return rt;
}

void orc_free_csr(orc_csr_t *c) {
	free(c->row_ptrs); free(c->cols); free(c->rats);
	c->row_ptrs = NULL; c->cols = NULL; c->rats = NULL;
	c->num_row = c->num_col = c->num_nnz = 0;
}

static int cmp_key(const void *x, const void *y) {
	const orc_keyval_t *u = (const orc_keyval_t *)x, *v = (const orc_keyval_t *)y;
	return (u->key > v->key) - (u->key < v->key);
}

void orc_canonicalize(orc_csr_t *c, int thread_num) {
	#pragma omp parallel num_threads(thread_num)
	{
		int64_t cap = 0;
		orc_keyval_t *buf = NULL;
		#pragma omp for schedule(dynamic, 256)
		for (int64_t r = 0; r < c->num_row; r++) {
			int64_t s = c->row_ptrs[r], n = c->row_ptrs[r + 1] - s;
			if (n < 2) continue;
			if (n > cap) { free(buf); cap = n; buf = (orc_keyval_t *)xmalloc(sizeof(orc_keyval_t) * (size_t)cap); }
			for (int64_t i = 0; i < n; i++) { buf[i].key = c->cols[s + i]; buf[i].val = c->rats[s + i]; }
			qsort(buf, (size_t)n, sizeof(orc_keyval_t), cmp_key);
			for (int64_t i = 0; i < n; i++) { c->cols[s + i] = buf[i].key; c->rats[s + i] = buf[i].val; }
		}
		free(buf);
	}
}

void orc_symbolic(const orc_csr_t *a, const orc_csr_t *b, int64_t *row_nz, const orc_params_t *p) {
	int64_t m = a->num_row;
	int64_t *row_flp = (int64_t *)xmalloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
	orc_flp_populate(a, b, row_flp, p->thread_num);
	orc_hash_symb_dnr(a, b, row_flp, row_nz, p, p->thread_num);
	free(row_flp);
}
