/* ams_mtx.c -- MatrixMarket and PaRMAT text I/O of the drop-in layer.
 *
 * The reference reads its matrices with mtx2trpl / trpl2csr and writes the
 * symmetric "ones" pairs of its R-MAT runs with txt_transpose
 * (src/mm/mm_utils.c:451-862, src/rmat_ones.c:6-37).  These functions keep
 * their names, signatures and file semantics so that a caller of those can
 * relink against this library:
 *
 *   reading (mtx2trpl, src/mm/mm_utils.c:451-604)
 *     - lines starting with '%' are comments; any comment that contains
 *       "symmetric" (the banner included) marks the file symmetric
 *     - the first other line is "rows cols entries"
 *     - entry lines are "i j [value]", 1-based; a missing value is 1.0
 *     - entries whose value is exactly 0.0 are dropped, every kept value
 *       becomes 1.0
 *     - a symmetric file gets (j, i) added for every off-diagonal (i, j)
 *     - entries are sorted by (row, column); duplicates are kept
 *   writing (trpl2mtx, :607-642)
 *     - symmetric: only entries with row <= column, count in the header
 *     - every value is written as 1.0 with "%f"
 *   PaRMAT text (txt2trpl, :757-835; rmat_ones)
 *     - "src dst [value]" 0-based, no self loops, scale and edge factor
 *       taken from the file name "<name>_<scale>_<ef>.txt"; every edge is
 *       mirrored
 *
 * Lines are read with getline, so any line length and file size work (the
 * reference reads with O_DIRECT into one file-sized buffer).  Triple arrays
 * are mmap'ed like the reference's (release with dealloc_memory).
 */
#define _GNU_SOURCE
#include "ams_dropin.h"

#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>

static void *mtx_map(size_t bytes) {
	if (bytes == 0) bytes = 8;
	void *p = mmap(NULL, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
	if (p == MAP_FAILED) {
		perror("mmap");
		exit(EXIT_FAILURE);
	}
	return p;
}

void dealloc_memory(void *ptr, size_t size) {
	if (ptr) munmap(ptr, size ? size : 8);
}

/* (row, column) order.  The reference returns the int64 difference
 * truncated to int (src/mm/mm_utils.c:47-55); this compares, which agrees
 * wherever the difference fits an int. */
int compare_triples(const void *a, const void *b) {
	const triple_t *x = (const triple_t *)a, *y = (const triple_t *)b;
	if (x->user_id != y->user_id) return x->user_id < y->user_id ? -1 : 1;
	if (x->item_id != y->item_id) return x->item_id < y->item_id ? -1 : 1;
	return 0;
}

/* Growable triple buffer (malloc while reading, mmap'ed at the end). */
typedef struct {
	triple_t *t;
	my_cnt_t n, cap;
} tbuf_t;

static void tbuf_push(tbuf_t *b, my_cnt_t r, my_cnt_t c) {
	if (b->n == b->cap) {
		b->cap = b->cap ? 2 * b->cap : 1024;
		b->t = (triple_t *)realloc(b->t, sizeof(triple_t) * (size_t)b->cap);
		if (!b->t) {
			perror("realloc");
			exit(EXIT_FAILURE);
		}
	}
	b->t[b->n].user_id = r;
	b->t[b->n].item_id = c;
	b->t[b->n].rating = 1.0;
	++b->n;
}

static int cmp_cnt(const void *a, const void *b) {
	const my_cnt_t x = *(const my_cnt_t *)a, y = *(const my_cnt_t *)b;
	return (x > y) - (x < y);
}

/* Sorts by (row, column) into an mmap'ed array.  Every kept value is 1.0, so
 * entries with equal keys are identical and any sort gives the same bytes:
 * rows in [0, num_row) are bucketed by a counting pass and each row's
 * columns sorted on their own; anything else falls back to qsort. */
static triple_t *tbuf_finish(tbuf_t *b, my_cnt_t *num_nnz, my_cnt_t num_row) {
	const my_cnt_t n = b->n;
	triple_t *out = (triple_t *)mtx_map(sizeof(triple_t) * (size_t)n);
	bool bucketed = num_row > 0;
	for (my_cnt_t k = 0; bucketed && k < n; ++k)
		bucketed = b->t[k].user_id >= 0 && b->t[k].user_id < num_row;
	my_cnt_t *start = bucketed ? (my_cnt_t *)calloc((size_t)num_row + 1, sizeof(my_cnt_t)) : NULL;
	my_cnt_t *cols = bucketed && start ? (my_cnt_t *)malloc(sizeof(my_cnt_t) * (size_t)(n ? n : 1)) : NULL;
	if (cols) {
		for (my_cnt_t k = 0; k < n; ++k) ++start[b->t[k].user_id + 1];
		for (my_cnt_t r = 0; r < num_row; ++r) start[r + 1] += start[r];
		for (my_cnt_t k = 0; k < n; ++k) cols[start[b->t[k].user_id]++] = b->t[k].item_id;
		my_cnt_t lo = 0;
		for (my_cnt_t r = 0; r < num_row; ++r) {
			const my_cnt_t hi = start[r];
			if (hi - lo > 1) qsort(cols + lo, (size_t)(hi - lo), sizeof(my_cnt_t), cmp_cnt);
			for (my_cnt_t k = lo; k < hi; ++k) {
				out[k].user_id = r;
				out[k].item_id = cols[k];
				out[k].rating = 1.0;
			}
			lo = hi;
		}
	} else {
		qsort(b->t, (size_t)n, sizeof(triple_t), compare_triples);
		if (n) memcpy(out, b->t, sizeof(triple_t) * (size_t)n);
	}
	free(cols);
	free(start);
	*num_nnz = n;
	free(b->t);
	b->t = NULL;
	return out;
}

/* Parses "i j [v]" as the reference's sscanf("%ld %ld %lf") does and
 * returns the number of fields read (leading blanks skipped, parsing stops at
 * the first field that does not convert). */
static int parse_entry(const char *line, my_cnt_t *i, my_cnt_t *j, double *v) {
	char *end;
	const char *p = line;
	const long a = strtol(p, &end, 10);
	if (end == p) return 0;
	*i = a;
	p = end;
	const long b = strtol(p, &end, 10);
	if (end == p) return 1;
	*j = b;
	p = end;
	const double x = strtod(p, &end);
	if (end == p) return 2;
	*v = x;
	return 3;
}

bool mtx2trpl(triple_t **trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t *num_nnz, const char *filename,
              const int numa_node_idx) {
	(void)numa_node_idx;
	FILE *f = fopen(filename, "r");
	if (!f) {
		fprintf(stderr, "Error opening file: %s\n", filename);
		exit(EXIT_FAILURE);
	}
	bool symmetric = false, header = false;
	tbuf_t b = {0};
	char *line = NULL;
	size_t cap = 0;
	ssize_t len;
	while ((len = getline(&line, &cap, f)) != -1) {
		if (len > 0 && line[len - 1] == '\n') line[--len] = '\0';
		if (line[0] == '%') {
			if (strstr(line, "symmetric")) symmetric = true;
			continue;
		}
		if (!header) {
			long r = 0, c = 0, z = 0;
			if (sscanf(line, "%ld %ld %ld", &r, &c, &z) != 3) {
				fprintf(stderr, "Error reading header line: %s\n", line);
				exit(EXIT_FAILURE);
			}
			*num_row = r;
			*num_col = c;
			header = true;
			continue;
		}
		my_cnt_t i, j;
		double v;
		const int n = parse_entry(line, &i, &j, &v);
		if (n == 2) {
			v = 1.0;
		} else if (n != 3) {
			fprintf(stderr, "Error reading line: %s\n", line);
			exit(EXIT_FAILURE);
		}
		if (v == 0.0) continue;
		tbuf_push(&b, i - 1, j - 1);
		if (symmetric && i != j) tbuf_push(&b, j - 1, i - 1);
	}
	free(line);
	fclose(f);
	if (!header) {
		fprintf(stderr, "Error reading header line: %s\n", filename);
		exit(EXIT_FAILURE);
	}
	*trpl = tbuf_finish(&b, num_nnz, *num_row);
	return symmetric;
}

void trpl2mtx(const char *filename, triple_t *trpl, my_cnt_t num_row, my_cnt_t num_col, my_cnt_t num_nnz,
              bool isSymmetric) {
	FILE *f = fopen(filename, "w");
	if (!f) {
		perror("Error opening file");
		exit(EXIT_FAILURE);
	}
	if (isSymmetric) {
		my_cnt_t kept = 0;
		for (my_cnt_t i = 0; i < num_nnz; ++i) kept += trpl[i].user_id <= trpl[i].item_id;
		fprintf(f, "%%%%MatrixMarket matrix coordinate real symmetric\n");
		fprintf(f, "%ld %ld %ld\n", (long)num_row, (long)num_col, (long)kept);
		for (my_cnt_t i = 0; i < num_nnz; ++i)
			if (trpl[i].user_id <= trpl[i].item_id)
				fprintf(f, "%ld %ld %f\n", (long)trpl[i].user_id + 1, (long)trpl[i].item_id + 1, 1.0);
	} else {
		fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
		fprintf(f, "%ld %ld %ld\n", (long)num_row, (long)num_col, (long)num_nnz);
		for (my_cnt_t i = 0; i < num_nnz; ++i)
			fprintf(f, "%ld %ld %f\n", (long)trpl[i].user_id + 1, (long)trpl[i].item_id + 1, 1.0);
	}
	fclose(f);
}

void trpl2csr(csr_t *csr, triple_t *trpl, const my_cnt_t num_row, const my_cnt_t num_col, const my_cnt_t num_nnz,
              numa_cfg_t *numa_cfg) {
	(void)numa_cfg;
	memset(csr, 0, sizeof(csr_t));
	csr->num_row = num_row;
	csr->num_col = num_col;
	csr->num_nnz = num_nnz;
	csr->row_ptrs = (my_cnt_t *)mtx_map(sizeof(my_cnt_t) * (size_t)(num_row + 1));
	csr->cols = (my_cnt_t *)mtx_map(sizeof(my_cnt_t) * (size_t)num_nnz);
	csr->rats = (my_rat_t *)mtx_map(sizeof(my_rat_t) * (size_t)num_nnz);
	for (my_cnt_t i = 0; i < num_nnz; ++i) {
		csr->row_ptrs[trpl[i].user_id + 1]++;
		csr->cols[i] = trpl[i].item_id;
		csr->rats[i] = trpl[i].rating;
	}
	for (my_cnt_t r = 1; r <= num_row; ++r) csr->row_ptrs[r] += csr->row_ptrs[r - 1];
}

void mtx2csr(csr_t *csr, const char *filename, numa_cfg_t *numa_cfg) {
	triple_t *trpl = NULL;
	my_cnt_t num_row = 0, num_col = 0, num_nnz = 0;
	mtx2trpl(&trpl, &num_row, &num_col, &num_nnz, filename, 0);
	trpl2csr(csr, trpl, num_row, num_col, num_nnz, numa_cfg);
	dealloc_memory(trpl, sizeof(triple_t) * (size_t)num_nnz);
}

/* src/mm/mm_utils.c:730-732: the drivers' name for mtx2csr. */
void raw2csr(csr_t *csr, const char *filename, numa_cfg_t *numa_cfg) { mtx2csr(csr, filename, numa_cfg); }

void trpl_ones(triple_t *trpl, my_cnt_t num_nnz) {
	for (my_cnt_t i = 0; i < num_nnz; ++i) trpl[i].rating = 1.0;
}

void trpl_transpose(triple_t *trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t num_nnz) {
	for (my_cnt_t i = 0; i < num_nnz; ++i) {
		const my_cnt_t t = trpl[i].user_id;
		trpl[i].user_id = trpl[i].item_id;
		trpl[i].item_id = t;
	}
	qsort(trpl, (size_t)num_nnz, sizeof(triple_t), compare_triples);
	const my_cnt_t t = *num_row;
	*num_row = *num_col;
	*num_col = t;
}

void txt2trpl(triple_t **trpl, my_cnt_t *num_row, my_cnt_t *num_col, my_cnt_t *num_nnz, const char *filename,
              const int numa_node_idx) {
	(void)numa_node_idx;
	/* "<name>_<scale>_<ef>.txt": the last two numbers of the base name */
	const char *base = strrchr(filename, '/');
	base = base ? base + 1 : filename;
	int scale = 0, ef = 0;
	const char *u = strchr(base, '_');
	if (!u || sscanf(u + 1, "%d_%d.txt", &scale, &ef) != 2) {
		fprintf(stderr, "Failed to extract scale and edge factor from %s\n", filename);
		exit(EXIT_FAILURE);
	}
	*num_row = *num_col = (my_cnt_t)1 << scale;
	FILE *f = fopen(filename, "r");
	if (!f) {
		fprintf(stderr, "Error opening file, %s\n", filename);
		exit(EXIT_FAILURE);
	}
	tbuf_t b = {0};
	char *line = NULL;
	size_t cap = 0;
	my_cnt_t lines = 0;
	while (getline(&line, &cap, f) != -1) {
		++lines;
		my_cnt_t i, j;
		double v;
		const int n = parse_entry(line, &i, &j, &v);
		if (n == 2) {
			v = 1.0;
		} else if (n != 3) {
			fprintf(stderr, "Error reading line: %s\n", line);
			exit(EXIT_FAILURE);
		}
		if (v == 0.0) continue;
		if (i == j) {
			fprintf(stderr, "self loop %ld in %s\n", (long)i, filename);
			exit(EXIT_FAILURE);
		}
		tbuf_push(&b, i, j);
		tbuf_push(&b, j, i);
	}
	free(line);
	fclose(f);
	if (2 * lines != *num_row * (my_cnt_t)ef * 2) {  /* the reference asserts this */
		fprintf(stderr, "%s: %ld edges, scale %d edge factor %d expects %ld\n", filename, (long)lines, scale, ef,
		        (long)(*num_row * (my_cnt_t)ef));
		exit(EXIT_FAILURE);
	}
	*trpl = tbuf_finish(&b, num_nnz, *num_row);
}

void txt_transpose(const char *filename_dest_trs, const char *filename_dest_org, const char *filename_src,
                   const int numa_node_idx) {
	triple_t *trpl = NULL;
	my_cnt_t num_row, num_col, num_nnz;
	txt2trpl(&trpl, &num_row, &num_col, &num_nnz, filename_src, numa_node_idx);
	trpl_ones(trpl, num_nnz);
	trpl2mtx(filename_dest_org, trpl, num_row, num_col, num_nnz, true);
	trpl_transpose(trpl, &num_row, &num_col, num_nnz);
	trpl2mtx(filename_dest_trs, trpl, num_row, num_col, num_nnz, true);
	dealloc_memory(trpl, sizeof(triple_t) * (size_t)num_nnz);
}
