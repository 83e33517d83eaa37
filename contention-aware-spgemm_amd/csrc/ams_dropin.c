/*
 * ams_dropin.c -- the reference's C entry points (include/ams_dropin.h) on
 * top of the device C ABI (include/ams_hip.h).  Plain C host code: it moves
 * the caller's int64/double host arrays to the GPU, calls the device
 * pipeline and brings the results back.
 *
 * Mapping of the reference's phases (src/ams_hybaccspgemm.c:36-220):
 *   INIT  upload A and B, device row flops                  (init_time)
 *   SYMB  device symbolic, row pointers of C downloaded     (symb_time)
 *   MGMT  C's host arrays mapped                            (mgmt_time)
 *   NUMC  device numeric, columns and values downloaded     (numc_time)
 * The morsel table of pack_ams_pack is host bookkeeping in the reference
 * and stays host bookkeeping here; the device pipeline schedules its own
 * work (row bins and the persistent morsel queue) and does not read it.
 */
#define _GNU_SOURCE
#include "ams_dropin.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <time.h>

#include "ams_hip.h"

my_rat_t ams_ht_load = 1.0;

static ams_ctx_t *g_ctx = NULL;
static my_cnt_t g_ams_size = 524288;  /* AMS_SIZE (scripts/ams_run_common.sh:26) */
static my_cnt_t g_ams_min_num = 32768; /* AMS_MIN_NUM */
static double g_ams_prt = 0.85;        /* AMS_PRT (scripts/ams_rmat_run.sh:33) */

static void die(const char *what, int rc) {
	fprintf(stderr, "ams_hip: %s failed (%d): %s\n", what, rc, g_ctx ? ams_last_error(g_ctx) : "");
	exit(EXIT_FAILURE);
}

static ams_ctx_t *ctx(void) {
	if (!g_ctx) {
		int rc = ams_ctx_create(&g_ctx, NULL);
		if (rc != AMS_OK) die("ams_ctx_create", rc);
	}
	return g_ctx;
}

static double now(void) {
	struct timespec t;
	clock_gettime(CLOCK_MONOTONIC, &t);
	return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Host memory with the reference's ownership rule: mmap'ed, released by
 * munmap(size) in free_csr / dealloc_memory (src/inc/utils.c:254-260). */
static void *map_bytes(size_t bytes) {
	if (bytes == 0) bytes = 8;
	void *p = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
	if (p == MAP_FAILED) {
		perror("mmap");
		exit(EXIT_FAILURE);
	}
	return p;
}

static void unmap_bytes(void *p, size_t bytes) {
	if (p) munmap(p, bytes ? bytes : 8);
}

/* ------------------------------------------------------------ device copies */

typedef struct {
	ams_dcsr_t d;
	void *rp, *col, *val;
} dev_csr_t;

static void *dalloc(int64_t bytes) {
	void *p = NULL;
	int rc = ams_dev_alloc(&p, bytes);
	if (rc != AMS_OK) die("ams_dev_alloc", rc);
	return p;
}

static void upload(const csr_t *h, dev_csr_t *d) {
	ams_ctx_t *c = ctx();
	int64_t m = h->num_row, nnz = h->row_ptrs[m];
	d->rp = dalloc(8 * (m + 1));
	d->col = dalloc(4 * (nnz > 0 ? nnz : 1));
	d->val = dalloc(8 * (nnz > 0 ? nnz : 1));
	int rc = ams_memcpy_h2d(c, d->rp, h->row_ptrs, 8 * (m + 1));
	if (rc == AMS_OK && nnz > 0) {
		/* int64 columns pass through the value buffer, then narrow on device */
		void *tmp = dalloc(8 * nnz);
		rc = ams_memcpy_h2d(c, tmp, h->cols, 8 * nnz);
		if (rc == AMS_OK) rc = ams_i64_to_i32(c, (const int64_t *)tmp, (int32_t *)d->col, nnz);
		if (rc == AMS_OK) rc = ams_sync(c);
		ams_dev_free(tmp);
		if (rc == AMS_OK) rc = ams_memcpy_h2d(c, d->val, h->rats, 8 * nnz);
	}
	if (rc != AMS_OK) die("upload", rc);
	d->d.num_row = m;
	d->d.num_col = h->num_col;
	d->d.num_nnz = nnz;
	d->d.row_ptr = (const int64_t *)d->rp;
	d->d.col = (const int32_t *)d->col;
	d->d.val = d->val;
}

static void release(dev_csr_t *d) {
	ams_dev_free(d->rp);
	ams_dev_free(d->col);
	ams_dev_free(d->val);
	memset(d, 0, sizeof(*d));
}

static void check_shapes(const csr_t *a, const csr_t *b) {
	if (a->num_col != b->num_row) {
		fprintf(stderr, "ams_hip: A has %ld columns but B has %ld rows\n", (long)a->num_col, (long)b->num_row);
		exit(EXIT_FAILURE);
	}
}

/* ------------------------------------------------- session device copies
 *
 * The reference's driver calls its phases on the same A and B between
 * init_ams_bin and free_ams_bin (src/ams_hybaccspgemm.c:53-199): row flops,
 * symbolic, numeric.  For that session the drop-in keeps the device copies
 * of A and B (two slots; C = A*A shares one) and the device row pointers and
 * routing of the last symbolic phase, so that every matrix goes up once and
 * the numeric phase reuses the symbolic phase's plan.  A and B must not
 * change during the session (the reference reads them only, too).  Outside a
 * session every call uploads what it needs and frees it before returning. */

typedef struct {
	const my_cnt_t *rp, *cols;
	const my_rat_t *rats;
	my_cnt_t m, n, nnz;
} host_key_t;

typedef struct {
	host_key_t key;
	dev_csr_t dev;
	int used;
} slot_t;

static struct {
	const ams_bin_t *bin;  /* the session's bin, NULL outside a session */
	slot_t slot[2];
	/* symbolic phase of the session: operands, device row pointers of C */
	const slot_t *sym_a, *sym_b;
	int64_t *sym_rp;
	int64_t sym_nnz;
} g_sess;  /* zero: no session, empty slots, no symbolic plan (sym_rp NULL) */

static int64_t g_uploads = 0, g_symbolics = 0;

static host_key_t key_of(const csr_t *h) {
	host_key_t k = {h->row_ptrs, h->cols, h->rats, h->num_row, h->num_col, h->row_ptrs[h->num_row]};
	return k;
}

static int key_eq(const host_key_t *x, const host_key_t *y) {
	return x->rp == y->rp && x->cols == y->cols && x->rats == y->rats && x->m == y->m && x->n == y->n &&
	       x->nnz == y->nnz;
}

static void drop_symbolic(void) {
	if (g_sess.sym_rp) ams_dev_free(g_sess.sym_rp);
	g_sess.sym_rp = NULL;
	g_sess.sym_a = g_sess.sym_b = NULL;
	g_sess.sym_nnz = 0;
}

static void drop_slot(slot_t *s) {
	if (!s->used) return;
	if (g_sess.sym_a == s || g_sess.sym_b == s) drop_symbolic();
	release(&s->dev);
	memset(s, 0, sizeof(*s));
}

static void drop_all(void) {
	drop_symbolic();
	drop_slot(&g_sess.slot[0]);
	drop_slot(&g_sess.slot[1]);
}

static slot_t *find_slot(const host_key_t *k) {
	for (int i = 0; i < 2; i++)
		if (g_sess.slot[i].used && key_eq(&g_sess.slot[i].key, k)) return &g_sess.slot[i];
	return NULL;
}

/* Device copies of A and B (b may be the same matrix), uploading what the
 * slots do not hold yet. */
static void operands(const csr_t *a, const csr_t *b, slot_t **sa, slot_t **sb) {
	host_key_t ka = key_of(a), kb = key_of(b);
	slot_t *xa = find_slot(&ka), *xb = find_slot(&kb);
	if (!xa) {
		xa = (xb == &g_sess.slot[0]) ? &g_sess.slot[1] : &g_sess.slot[0];
		drop_slot(xa);
		upload(a, &xa->dev);
		xa->key = ka;
		xa->used = 1;
		g_uploads++;
	}
	if (!xb) {
		if (key_eq(&ka, &kb)) {
			xb = xa;
		} else {
			xb = (xa == &g_sess.slot[0]) ? &g_sess.slot[1] : &g_sess.slot[0];
			drop_slot(xb);
			upload(b, &xb->dev);
			xb->key = kb;
			xb->used = 1;
			g_uploads++;
		}
	}
	*sa = xa;
	*sb = xb;
}

/* End of a public call: outside a session nothing stays on the GPU. */
static void call_done(void) {
	if (!g_sess.bin) drop_all();
}

/* Symbolic phase of A*B on the slots' device copies, kept as the session's
 * plan; returns nnz(C). */
static int64_t run_symbolic_on(slot_t *sa, slot_t *sb) {
	ams_ctx_t *c = ctx();
	int64_t m = sa->dev.d.num_row, nnz = 0;
	drop_symbolic();
	g_sess.sym_rp = (int64_t *)dalloc(8 * (m + 1));
	int rc = ams_spgemm_symbolic(c, &sa->dev.d, &sb->dev.d, AMS_F64, g_sess.sym_rp, &nnz);
	if (rc != AMS_OK) die("ams_spgemm_symbolic", rc);
	g_symbolics++;
	g_sess.sym_a = sa;
	g_sess.sym_b = sb;
	g_sess.sym_nnz = nnz;
	return nnz;
}

void ams_dropin_counters(my_cnt_t *uploads, my_cnt_t *symbolic_runs) {
	if (uploads) *uploads = g_uploads;
	if (symbolic_runs) *symbolic_runs = g_symbolics;
}

void ams_dropin_release(void) {
	drop_all();
	g_sess.bin = NULL;
}

static void set_mode(int mode) {
	ams_params_t p;
	ams_get_params(ctx(), &p);
	p.mode = mode;
	int rc = ams_set_params(ctx(), &p);
	if (rc != AMS_OK) die("ams_set_params", rc);
}

/* Downloads columns (int32 on device) into int64 host memory. */
static void download_cols(int32_t *dcol, my_cnt_t *hcol, int64_t nnz) {
	if (nnz <= 0) return;
	ams_ctx_t *c = ctx();
	void *tmp = dalloc(8 * nnz);
	int rc = ams_i32_to_i64(c, dcol, (int64_t *)tmp, nnz);
	if (rc == AMS_OK) rc = ams_memcpy_d2h(c, hcol, tmp, 8 * nnz);
	ams_dev_free(tmp);
	if (rc != AMS_OK) die("download", rc);
}

/* ------------------------------------------------------------- products */

static runtime_t product(csr_t *a, csr_t *b, csr_t *c, int mode) {
	runtime_t rt;
	check_shapes(a, b);
	set_mode(mode);
	ams_ctx_t *cx = ctx();
	double t0 = now();
	slot_t *sa, *sb;
	operands(a, b, &sa, &sb);
	int64_t m = a->num_row;
	c->num_row = m;
	c->num_col = b->num_col;
	c->row_ptrs = (my_cnt_t *)map_bytes(sizeof(my_cnt_t) * (size_t)(m + 1));
	double t1 = now();
	rt.init_time = t1 - t0;

	int64_t nnz = run_symbolic_on(sa, sb);
	int rc = ams_memcpy_d2h(cx, c->row_ptrs, g_sess.sym_rp, 8 * (m + 1));
	if (rc != AMS_OK) die("download row_ptrs", rc);
	c->num_nnz = nnz;
	double t2 = now();
	rt.symb_time = t2 - t1;

	c->cols = (my_cnt_t *)map_bytes(sizeof(my_cnt_t) * (size_t)nnz);
	c->rats = (my_rat_t *)map_bytes(sizeof(my_rat_t) * (size_t)nnz);
	int32_t *dcol = (int32_t *)dalloc(4 * (nnz > 0 ? nnz : 1));
	double *dval = (double *)dalloc(8 * (nnz > 0 ? nnz : 1));
	double t3 = now();
	rt.mgmt_time = t3 - t2;

	rc = ams_spgemm_numeric(cx, &sa->dev.d, &sb->dev.d, AMS_F64, g_sess.sym_rp, dcol, dval);
	if (rc != AMS_OK) die("ams_spgemm_numeric", rc);
	download_cols(dcol, c->cols, nnz);
	rc = ams_memcpy_d2h(cx, c->rats, dval, 8 * nnz);
	if (rc != AMS_OK) die("download values", rc);
	ams_dev_free(dcol);
	ams_dev_free(dval);
	call_done();
	rt.numc_time = now() - t3;
	rt.total_time = rt.init_time + rt.symb_time + rt.mgmt_time + rt.numc_time;
	return rt;
}

runtime_t hybaccspgemm(csr_t *a, csr_t *b, csr_t *c, numa_cfg_t *cfg) {
	(void)cfg;
	return product(a, b, c, AMS_MODE_HYBRID);
}

runtime_t hashspgemm(csr_t *a, csr_t *b, csr_t *c, numa_cfg_t *cfg) {
	(void)cfg;
	return product(a, b, c, AMS_MODE_HASH);
}

runtime_t hserscspgemm(csr_t *a, csr_t *b, csr_t *c, numa_cfg_t *cfg) {
	(void)cfg;
	return product(a, b, c, AMS_MODE_ESC);
}

void free_csr(csr_t *csr) {
	if (!csr) return;
	unmap_bytes(csr->row_ptrs, sizeof(my_cnt_t) * (size_t)(csr->num_row + 1));
	unmap_bytes(csr->cols, sizeof(my_cnt_t) * (size_t)csr->num_nnz);
	unmap_bytes(csr->rats, sizeof(my_rat_t) * (size_t)csr->num_nnz);
	csr->row_ptrs = NULL;
	csr->cols = NULL;
	csr->rats = NULL;
	csr->num_row = csr->num_col = csr->num_nnz = 0;
}

/* ------------------------------------------------------------ mm_utils */

void scan(my_cnt_t *out, my_cnt_t *in, my_cnt_t num) {
	if (num <= 0) return;
	ams_ctx_t *c = ctx();
	int64_t *d = (int64_t *)dalloc(8 * num);
	int rc = num > 1 ? ams_memcpy_h2d(c, d, in, 8 * (num - 1)) : AMS_OK;
	if (rc == AMS_OK) rc = ams_scan_i64(c, d, d, num);
	if (rc == AMS_OK) rc = ams_memcpy_d2h(c, out, d, 8 * num);
	ams_dev_free(d);
	if (rc != AMS_OK) die("scan", rc);
}

static my_cnt_t flops_to_host(csr_t *a, csr_t *b, my_cnt_t *row_flp) {
	check_shapes(a, b);
	ams_ctx_t *c = ctx();
	slot_t *sa, *sb;
	operands(a, b, &sa, &sb);
	int64_t *d = (int64_t *)dalloc(8 * (a->num_row > 0 ? a->num_row : 1));
	int64_t total = 0;
	int rc = ams_row_flops(c, &sa->dev.d, &sb->dev.d, d, &total);
	if (rc == AMS_OK && row_flp) rc = ams_memcpy_d2h(c, row_flp, d, 8 * a->num_row);
	ams_dev_free(d);
	call_done();
	if (rc != AMS_OK) die("ams_row_flops", rc);
	return total;
}

void flp_populate(csr_t *a, csr_t *b, my_cnt_t *row_flp) {
	(void)flops_to_host(a, b, row_flp);
}

my_cnt_t get_intprod(csr_t *a, csr_t *b, my_cnt_t *row_nz) {
	return flops_to_host(a, b, row_nz);
}

my_cnt_t *lower_bound(my_cnt_t *start, my_cnt_t *end, my_cnt_t val) {
	while (start < end) {
		my_cnt_t *mid = start + (end - start) / 2;
		if (*mid < val) start = mid + 1;
		else end = mid;
	}
	return start;
}

/* ----------------------------------------------------------- ams_utils */

my_cnt_t ams_ht_slots_for(my_cnt_t nz) {
	if (ams_ht_load >= 1.0 || nz <= 0) return nz;
	return (my_cnt_t)((double)nz / ams_ht_load);
}

void ams_dropin_set_pack_params(my_cnt_t ams_size, my_cnt_t ams_min_num, double ams_prt) {
	g_ams_size = ams_size;
	g_ams_min_num = ams_min_num;
	g_ams_prt = ams_prt;
}

void free_ams_pack(ams_pack_t *pk) {
	if (!pk) return;
	unmap_bytes(pk->b_roff_maxc, sizeof(keyval_t) * (size_t)(pk->total_cnt + 1));
	unmap_bytes(pk->b_arr, sizeof(keyval_t) * (size_t)pk->total_cnt);
	memset(pk, 0, sizeof(*pk));
}

void init_ams_bin(ams_bin_t *bin, ams_pack_t *pk, csr_t *csr, numa_cfg_t *cfg) {
	(void)cfg;
	/* a new driver sequence: device copies of an earlier one are dropped */
	drop_all();
	g_sess.bin = bin;
	bin->ams_pack = pk;
	bin->csr = csr;
	bin->row_flp = (my_cnt_t *)map_bytes(sizeof(my_cnt_t) * (size_t)csr->num_row);
	bin->ps_row_flp = (my_cnt_t *)map_bytes(sizeof(my_cnt_t) * (size_t)(csr->num_row + 1));
	bin->row_nz = (my_cnt_t *)map_bytes(sizeof(my_cnt_t) * (size_t)csr->num_row);
}

void free_ams_bin(ams_bin_t *bin) {
	if (bin && bin == g_sess.bin) {  /* end of the driver sequence */
		drop_all();
		g_sess.bin = NULL;
	}
	if (!bin || !bin->csr) return;
	unmap_bytes(bin->row_flp, sizeof(my_cnt_t) * (size_t)bin->csr->num_row);
	unmap_bytes(bin->ps_row_flp, sizeof(my_cnt_t) * (size_t)(bin->csr->num_row + 1));
	unmap_bytes(bin->row_nz, sizeof(my_cnt_t) * (size_t)bin->csr->num_row);
	bin->row_flp = bin->ps_row_flp = bin->row_nz = NULL;
	bin->csr = NULL;
	bin->ams_pack = NULL;
}

static int cmp_cf_id(const void *x, const void *y) {
	const keyval_t *u = (const keyval_t *)x, *v = (const keyval_t *)y;
	int c = (u->val > v->val) - (u->val < v->val);
	return c ? c : (u->key > v->key) - (u->key < v->key);
}

/* Morsel table of src/ams/ams_utils.c:1410-1596: morsel k covers rows
 * [lower_bound(ps, T*k), lower_bound(ps, T*(k+1))), T = MIN(AMS_MIN_NUM,
 * AMS_SIZE/8), the last ending at num_row; cf and max row flops per valid
 * morsel, empties counted into low_head_idx, sorted by (cf, id), regions
 * split at AMS_PRT.  A count of 0 morsels is raised to 1 (the reference
 * exits on mmap(0) there). */
my_cnt_t pack_ams_pack(my_cnt_t *ps, ams_bin_t *bin, numa_cfg_t *cfg) {
	(void)cfg;
	ams_pack_t *pk = bin->ams_pack;
	my_cnt_t m = bin->csr->num_row;
	my_cnt_t t = g_ams_min_num < g_ams_size / 8 ? g_ams_min_num : g_ams_size / 8;
	if (t < 1) t = 1;
	my_cnt_t total = ps[m] / t;
	if (total < 1) total = 1;
	pk->total_cnt = total;
	pk->b_arr = (keyval_t *)map_bytes(sizeof(keyval_t) * (size_t)total);
	pk->b_roff_maxc = (keyval_t *)map_bytes(sizeof(keyval_t) * (size_t)(total + 1));
	pk->b_roff_maxc[0].key = 0;
	for (my_cnt_t k = 0; k < total; k++) {
		pk->b_roff_maxc[k + 1].key = lower_bound(ps, ps + m + 1, t * (k + 1)) - ps;
		pk->b_arr[k].key = k;
	}
	pk->b_roff_maxc[total].key = m;
	my_cnt_t invalid = 0;
	double mx_all = 0.0;
	for (my_cnt_t k = 0; k < total; k++) {
		my_cnt_t s = pk->b_roff_maxc[k].key, e = pk->b_roff_maxc[k + 1].key;
		if (e > s) {
			pk->b_arr[k].val = (double)(bin->ps_row_flp[e] - bin->ps_row_flp[s]) / (double)(ps[e] - ps[s]);
			my_cnt_t mx = bin->row_flp[s];
			for (my_cnt_t r = s + 1; r < e; r++) mx = bin->row_flp[r] > mx ? bin->row_flp[r] : mx;
			pk->b_roff_maxc[k].val = (double)mx;
			mx_all = (double)mx > mx_all ? (double)mx : mx_all;
		} else {
			pk->b_roff_maxc[k].val = 0.0;
			pk->b_arr[k].val = 0.0;
			invalid++;
		}
	}
	qsort(pk->b_arr, (size_t)total, sizeof(keyval_t), cmp_cf_id);
	pk->low_head_idx = invalid;
	pk->high_head_idx = total - 1;
	pk->high_tail_idx = (my_cnt_t)((double)(total - invalid) * g_ams_prt + (double)invalid);
	pk->low_tail_idx = pk->high_tail_idx - 1;
	return (my_cnt_t)mx_all;
}

/* Morsel table over nonzeros of src/ams/ams_utils.c:1851-2046
 * (pack_ams_pack_mmmv): T = MIN(AMS_MIN_NUM, AMS_SIZE/8) and
 * total_cnt = nnz(csr) * coef / T morsels, morsel k ending at
 * lower_bound(ps_row_coef, T*(k+1)) and the last at num_row; a valid morsel
 * gets the mean nonzeros per row of csr as its sort value and the largest
 * nnz_per_row of its rows in b_roff_maxc; empty morsels count into
 * low_head_idx; sorted by (value, id); regions split at AMS_PRT.  The count
 * is raised to 1 where the reference would map 0 bytes and exit. */
my_cnt_t pack_ams_pack_mmmv(csr_t *csr, ams_pack_t *pk, my_cnt_t *ps_row_coef, my_cnt_t *nnz_per_row, my_cnt_t coef,
	numa_cfg_t *cfg) {
	(void)cfg;
	my_cnt_t m = csr->num_row;
	my_cnt_t t = g_ams_min_num < g_ams_size / 8 ? g_ams_min_num : g_ams_size / 8;
	if (t < 1) t = 1;
	my_cnt_t total = csr->num_nnz * coef / t;
	if (total < 1) total = 1;
	pk->total_cnt = total;
	pk->b_arr = (keyval_t *)map_bytes(sizeof(keyval_t) * (size_t)total);
	pk->b_roff_maxc = (keyval_t *)map_bytes(sizeof(keyval_t) * (size_t)(total + 1));
	pk->b_roff_maxc[0].key = 0;
	for (my_cnt_t k = 0; k < total; k++) {
		pk->b_roff_maxc[k + 1].key = lower_bound(ps_row_coef, ps_row_coef + m + 1, t * (k + 1)) - ps_row_coef;
		pk->b_arr[k].key = k;
	}
	pk->b_roff_maxc[total].key = m;
	my_cnt_t invalid = 0;
	double mx_all = 0.0;
	for (my_cnt_t k = 0; k < total; k++) {
		my_cnt_t s = pk->b_roff_maxc[k].key, e = pk->b_roff_maxc[k + 1].key;
		if (e > s) {
			pk->b_arr[k].val = (double)(csr->row_ptrs[e] - csr->row_ptrs[s]) / (double)(e - s);
			my_cnt_t mx = nnz_per_row[s];
			for (my_cnt_t r = s + 1; r < e; r++) mx = nnz_per_row[r] > mx ? nnz_per_row[r] : mx;
			pk->b_roff_maxc[k].val = (double)mx;
			mx_all = (double)mx > mx_all ? (double)mx : mx_all;
		} else {
			pk->b_roff_maxc[k].val = 0.0;
			pk->b_arr[k].val = 0.0;
			invalid++;
		}
	}
	qsort(pk->b_arr, (size_t)total, sizeof(keyval_t), cmp_cf_id);
	pk->low_head_idx = invalid;
	pk->high_head_idx = total - 1;
	pk->high_tail_idx = (my_cnt_t)((double)(total - invalid) * g_ams_prt + (double)invalid);
	pk->low_tail_idx = pk->high_tail_idx - 1;
	return (my_cnt_t)mx_all;
}

/* Symbolic: ams_bin->row_nz[i] = nnz of C row i.  Both reference variants
 * (morsel-claimed, src/ams/ams_utils.c:483-602, and dynamic-range,
 * :1600-1752, with its OpenMP form :1754-1846) compute the same counts; one
 * device symbolic serves all three, and in a session it is kept for the
 * numeric phase. */
static void symbolic_rows(csr_t *a, csr_t *b, ams_bin_t *bin) {
	check_shapes(a, b);
	ams_ctx_t *c = ctx();
	slot_t *sa, *sb;
	operands(a, b, &sa, &sb);
	int64_t m = a->num_row;
	run_symbolic_on(sa, sb);
	my_cnt_t *rp = (my_cnt_t *)malloc(sizeof(my_cnt_t) * (size_t)(m + 1));
	if (!rp) die("malloc", AMS_ERR_ALLOC);
	int rc = ams_memcpy_d2h(c, rp, g_sess.sym_rp, 8 * (m + 1));
	if (rc != AMS_OK) die("download row_ptrs", rc);
	for (int64_t i = 0; i < m; i++) bin->row_nz[i] = rp[i + 1] - rp[i];
	free(rp);
	call_done();
}

void ams_hash_symb(csr_t *a, csr_t *b, my_cnt_t *ht_check, ams_bin_t *bin) {
	(void)ht_check;
	symbolic_rows(a, b, bin);
}

void ams_hash_symb_dnr(csr_t *a, csr_t *b, my_cnt_t *ht_check, ams_bin_t *bin) {
	(void)ht_check;
	symbolic_rows(a, b, bin);
}

void ams_hash_symb_omp_dnr(csr_t *a, csr_t *b, my_cnt_t *ht_check, ams_bin_t *bin) {
	(void)ht_check;
	symbolic_rows(a, b, bin);
}

/* Numeric for a C whose row_ptrs the caller already holds (from the
 * symbolic phase) and whose cols/rats arrays it allocated.  The session's
 * symbolic plan is reused when it was made for these operands (the device
 * numeric phase re-routes the rows itself when the mode differs); otherwise,
 * or if the device plan no longer matches, the symbolic phase runs first. */
static void numeric_rows(csr_t *a, csr_t *b, csr_t *c, int mode) {
	check_shapes(a, b);
	set_mode(mode);
	ams_ctx_t *cx = ctx();
	slot_t *sa, *sb;
	operands(a, b, &sa, &sb);
	int64_t m = a->num_row;
	if (!(g_sess.sym_rp && g_sess.sym_a == sa && g_sess.sym_b == sb)) run_symbolic_on(sa, sb);
	int64_t nnz = g_sess.sym_nnz;
	if (nnz != c->row_ptrs[m]) {
		fprintf(stderr, "ams_hip: C has %ld entries but A*B has %ld\n", (long)c->row_ptrs[m], (long)nnz);
		exit(EXIT_FAILURE);
	}
	int32_t *dcol = (int32_t *)dalloc(4 * (nnz > 0 ? nnz : 1));
	double *dval = (double *)dalloc(8 * (nnz > 0 ? nnz : 1));
	int rc = ams_spgemm_numeric(cx, &sa->dev.d, &sb->dev.d, AMS_F64, g_sess.sym_rp, dcol, dval);
	if (rc == AMS_ERR_STATE) {
		/* refused before any work (no kernel flags): another call replaced
		 * the device plan, so plan again; kernel flags stay fatal */
		int32_t flags = -1;
		if (ams_get_error_flags(cx, &flags) == AMS_OK && flags == 0) {
			run_symbolic_on(sa, sb);
			rc = ams_spgemm_numeric(cx, &sa->dev.d, &sb->dev.d, AMS_F64, g_sess.sym_rp, dcol, dval);
		}
	}
	if (rc != AMS_OK) die("ams_spgemm_numeric", rc);
	download_cols(dcol, c->cols, nnz);
	rc = ams_memcpy_d2h(cx, c->rats, dval, 8 * nnz);
	if (rc != AMS_OK) die("download values", rc);
	c->num_nnz = nnz;
	ams_dev_free(dcol);
	ams_dev_free(dval);
	call_done();
}

void ams_hash_numc(csr_t *a, csr_t *b, csr_t *c, my_cnt_t *ht_check, my_rat_t *ht_value, ams_bin_t *bin) {
	(void)ht_check;
	(void)ht_value;
	(void)bin;
	numeric_rows(a, b, c, AMS_MODE_HASH);
}

void ams_hsersc_numc(csr_t *a, csr_t *b, csr_t *c, keyval_t *kv, keyval_t *aux, my_cnt_t max_row_flp,
	ams_bin_t *bin) {
	(void)kv;
	(void)aux;
	(void)max_row_flp;
	(void)bin;
	numeric_rows(a, b, c, AMS_MODE_ESC);
}

void ams_hybacc_numc(csr_t *a, csr_t *b, csr_t *c, keyval_t *kv, keyval_t *aux, my_cnt_t max_row_flp,
	my_cnt_t *ht_check, my_rat_t *ht_value, ams_bin_t *bin) {
	(void)kv;
	(void)aux;
	(void)max_row_flp;
	(void)ht_check;
	(void)ht_value;
	(void)bin;
	numeric_rows(a, b, c, AMS_MODE_HYBRID);
}
