/*
 * xtn_oracle.c -- CPU restatement of the XTNeighbor-streaming hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.
 * The product (xt-neighbor_amd/) never links or calls this code.
 *
 * Citations are to the reference repository (heartnetkung/XT-neighbor),
 * relative to its root. The restatement follows the reference's algorithm
 * (variants of the 18-residue prefix -> stable key sort -> groups -> all
 * in-group index pairs -> sort/unique -> full-string distance -> filter) but is
 * written as plain sequential C; where the reference uses CUB device
 * primitives this file uses a stable merge sort and linear scans.
 */
#include "xtn_oracle.h"

#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define XO_MAX_PREFIX 18 /* xtn.h:8 MAX_COMPRESSED_LENGTH */
#define XO_HAMMING_SENTINEL 77 /* kernel.cu:292 */

/* ---------------------------------------------------------------- codec -- */

/* codec.cu:25-38: 'A'..'Y' minus B J O U X map to c-'@' (1..25), else -1. */
int xo_char_encode(char c) {
	if (c < 'A' || c > 'Y')
		return -1;
	if (c == 'B' || c == 'J' || c == 'O' || c == 'U' || c == 'X')
		return -1;
	return c - '@';
}

static int slot_get(const uint32_t w[3], int i) {
	return (int)((w[i / 6] >> (27 - 5 * (i % 6))) & 0x1F);
}

static void slot_or(uint32_t w[3], int i, int code) {
	w[i / 6] |= (uint32_t)code << (27 - 5 * (i % 6));
}

/* codec.cu:43-59: pack up to 18 residues, six per 32-bit word, residue i in
 * bits [31-5(i%6) .. 27-5(i%6)] of word i/6. Stops at NUL, LF, ',' or TAB; an
 * invalid residue clears word 0 and stops. */
void xo_str_encode(const char* s, uint32_t out[3]) {
	out[0] = out[1] = out[2] = 0;
	for (int i = 0; i < XO_MAX_PREFIX; i++) {
		char c = s[i];
		if (c == '\0' || c == '\n' || c == ',' || c == '\t')
			break;
		int code = xo_char_encode(c);
		if (code < 0) {
			out[0] = 0;
			break;
		}
		slot_or(out, i, code);
	}
}

/* codec.cu:64-80: inverse of the packing; stops at the first empty slot. */
void xo_str_decode(const uint32_t in[3], char out[19]) {
	int i = 0;
	for (; i < XO_MAX_PREFIX; i++) {
		int code = slot_get(in, i);
		if (code == 0)
			break;
		out[i] = (char)('@' + code);
	}
	out[i] = '\0';
}

/* codec.cu:86-105: packed length = index of the last non-empty slot + 1. */
int xo_len_decode(const uint32_t in[3]) {
	int len = XO_MAX_PREFIX;
	while (len > 0 && slot_get(in, len - 1) == 0)
		len--;
	return len;
}

/* codec.cu:110-149: delete slot `position`, move later slots up by one and
 * leave the last slot empty. */
void xo_remove_char(const uint32_t in[3], int position, uint32_t out[3]) {
	int codes[XO_MAX_PREFIX];
	for (int i = 0; i < XO_MAX_PREFIX; i++)
		codes[i] = slot_get(in, i);
	out[0] = out[1] = out[2] = 0;
	int j = 0;
	for (int i = 0; i < XO_MAX_PREFIX; i++) {
		if (i == position)
			continue;
		slot_or(out, j++, codes[i]);
	}
}

/* ------------------------------------------------------------- variants -- */

/* kernel.cu:33-57: number of deletion variants, sum_{k<=min(d,L)} C(L,k). */
int xo_combination_len(int len, int distance) {
	int kmax = distance < len ? distance : len;
	long long c = 1, total = 1;
	for (int k = 1; k <= kmax; k++) {
		c = c * (len - k + 1) / k;
		total += c;
	}
	return (int)total;
}

typedef struct {
	uint32_t* out;
	int n, pos, len, maxdel;
} expand_ctx;

/* Post-order walk of the tree of deletion sets (children in increasing
 * position order, a node's own variant after all of its children, the
 * untouched prefix last). This is the order produced by the stack machine of
 * generate_combination.cu:42-93 (pinned by test_generate_combination.cu). */
static void expand_rec(expand_ctx* cx, const uint32_t cur[3], int removed, int next) {
	if (removed < cx->maxdel) {
		for (int b = next; b < cx->len; b++) {
			uint32_t child[3];
			/* original position b sits at b - removed in the shortened key */
			xo_remove_char(cur, b - removed, child);
			expand_rec(cx, child, removed + 1, b + 1);
		}
	}
	if (cx->pos < cx->n) {
		memcpy(cx->out + 3 * (size_t)cx->pos, cur, 3 * sizeof(uint32_t));
	}
	cx->pos++;
}

int xo_expand_keys(const uint32_t seq[3], int distance, uint32_t* out, int n) {
	expand_ctx cx;
	cx.out = out;
	cx.n = n;
	cx.pos = 0;
	cx.len = xo_len_decode(seq);
	cx.maxdel = distance < cx.len ? distance : cx.len;
	expand_rec(&cx, seq, 0, 0);
	return cx.pos;
}

/* --------------------------------------------------------- stable sorts -- */

typedef struct {
	uint32_t k[3];
	int32_t v;
} xo_rec;

/* cub.cu:14-27: word 1 first, then word 0, then word 2. */
static int rec_less(const xo_rec* a, const xo_rec* b) {
	if (a->k[1] != b->k[1])
		return a->k[1] < b->k[1];
	if (a->k[0] != b->k[0])
		return a->k[0] < b->k[0];
	return a->k[2] < b->k[2];
}

static int rec_key_eq(const xo_rec* a, const xo_rec* b) {
	return a->k[0] == b->k[0] && a->k[1] == b->k[1] && a->k[2] == b->k[2];
}

/* bottom-up merge sort; stable, as the reference's device merge sort is
 * (cub.cu:95-104, relied on by test_stream1.cu's value order). */
static void rec_sort(xo_rec* a, int64_t n) {
	if (n < 2)
		return;
	xo_rec* tmp = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)n);
	xo_rec* src = a;
	xo_rec* dst = tmp;
	for (int64_t width = 1; width < n; width *= 2) {
		for (int64_t lo = 0; lo < n; lo += 2 * width) {
			int64_t mid = lo + width < n ? lo + width : n;
			int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
			int64_t i = lo, j = mid, o = lo;
			while (i < mid && j < hi) {
				if (rec_less(&src[j], &src[i]))
					dst[o++] = src[j++];
				else
					dst[o++] = src[i++];
			}
			while (i < mid)
				dst[o++] = src[i++];
			while (j < hi)
				dst[o++] = src[j++];
		}
		xo_rec* t = src;
		src = dst;
		dst = t;
	}
	if (src != a)
		memcpy(a, src, sizeof(xo_rec) * (size_t)n);
	free(tmp);
}

typedef struct {
	int32_t x, y;
} xo_pair;

static int pair_cmp(const void* pa, const void* pb) {
	const xo_pair* a = (const xo_pair*)pa;
	const xo_pair* b = (const xo_pair*)pb;
	if (a->x != b->x)
		return a->x < b->x ? -1 : 1;
	if (a->y != b->y)
		return a->y < b->y ? -1 : 1;
	return 0;
}

/* ------------------------------------------------------------- stages --- */

static xo_rec* make_variants(const uint32_t* seqs, int n, int distance, int64_t* count) {
	int64_t total = 0;
	for (int i = 0; i < n; i++)
		total += xo_combination_len(xo_len_decode(seqs + 3 * (size_t)i), distance);
	xo_rec* recs = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)(total > 0 ? total : 1));
	uint32_t* tmp = NULL;
	int tmp_cap = 0;
	int64_t pos = 0;
	for (int i = 0; i < n; i++) {
		const uint32_t* s = seqs + 3 * (size_t)i;
		int c = xo_combination_len(xo_len_decode(s), distance);
		if (c > tmp_cap) {
			tmp_cap = c;
			tmp = (uint32_t*)realloc(tmp, sizeof(uint32_t) * 3 * (size_t)c);
		}
		xo_expand_keys(s, distance, tmp, c);
		for (int k = 0; k < c; k++) {
			memcpy(recs[pos].k, tmp + 3 * (size_t)k, 3 * sizeof(uint32_t));
			recs[pos].v = i; /* generate_combination.cu:25-28 value = sequence id */
			pos++;
		}
	}
	free(tmp);
	*count = total;
	return recs;
}

/* xtn_inner.cu:329-358 (stage 1): variants, then sort_key_values. */
int64_t xo_stage1(const uint32_t* seqs, int n, int distance, uint32_t** keys, int32_t** values) {
	int64_t v = 0;
	xo_rec* recs = make_variants(seqs, n, distance, &v);
	rec_sort(recs, v);
	*keys = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)(v > 0 ? v : 1));
	*values = (int32_t*)malloc(sizeof(int32_t) * (size_t)(v > 0 ? v : 1));
	for (int64_t i = 0; i < v; i++) {
		memcpy(*keys + 3 * (size_t)i, recs[i].k, 3 * sizeof(uint32_t));
		(*values)[i] = recs[i].v;
	}
	free(recs);
	return v;
}

/* xtn_inner.cu:379 (stage 2 re-sort). */
void xo_sort_stage(uint32_t* keys, int32_t* values, int64_t n) {
	xo_rec* recs = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)(n > 0 ? n : 1));
	for (int64_t i = 0; i < n; i++) {
		memcpy(recs[i].k, keys + 3 * (size_t)i, 3 * sizeof(uint32_t));
		recs[i].v = values[i];
	}
	rec_sort(recs, n);
	for (int64_t i = 0; i < n; i++) {
		memcpy(keys + 3 * (size_t)i, recs[i].k, 3 * sizeof(uint32_t));
		values[i] = recs[i].v;
	}
	free(recs);
}

/* One lower-bound round over sorted records.
 * Pairs: kernel.cu:157-191 (every position pair i<j of a group, as
 * (min id, max id), kept when min id <= lowerbound, self pairs included),
 * then sort_int2 (xtn_inner.cu:108). Kept records: kernel.cu:501-521 and
 * xtn_inner.cu:181-201 (ids > lowerbound, in groups keeping at least two). */
static int64_t round_pairs(xo_rec* recs, int64_t* n_records, int lowerbound, xo_pair** pairs_out) {
	int64_t n = *n_records;
	int64_t count = 0;
	for (int64_t s = 0; s < n;) {
		int64_t e = s + 1;
		while (e < n && rec_key_eq(&recs[s], &recs[e]))
			e++;
		for (int64_t i = s; i < e; i++)
			for (int64_t j = i + 1; j < e; j++) {
				int32_t a = recs[i].v, b = recs[j].v;
				int32_t mn = a < b ? a : b;
				if (mn <= lowerbound)
					count++;
			}
		s = e;
	}
	xo_pair* pairs = (xo_pair*)malloc(sizeof(xo_pair) * (size_t)(count > 0 ? count : 1));
	int64_t pos = 0, keep = 0;
	for (int64_t s = 0; s < n;) {
		int64_t e = s + 1;
		while (e < n && rec_key_eq(&recs[s], &recs[e]))
			e++;
		int valid = 0;
		for (int64_t i = s; i < e; i++) {
			if (recs[i].v > lowerbound)
				valid++;
			for (int64_t j = i + 1; j < e; j++) {
				int32_t a = recs[i].v, b = recs[j].v;
				xo_pair p;
				p.x = a < b ? a : b;
				p.y = a < b ? b : a;
				if (p.x <= lowerbound)
					pairs[pos++] = p;
			}
		}
		if (valid >= 2)
			for (int64_t i = s; i < e; i++)
				if (recs[i].v > lowerbound)
					recs[keep++] = recs[i];
		s = e;
	}
	qsort(pairs, (size_t)count, sizeof(xo_pair), pair_cmp);
	*n_records = keep;
	*pairs_out = pairs;
	return count;
}

int64_t xo_gen_pairs(uint32_t* keys, int32_t* values, int64_t* n_records, int lowerbound,
                     int32_t** pairs) {
	int64_t n = *n_records;
	xo_rec* recs = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)(n > 0 ? n : 1));
	for (int64_t i = 0; i < n; i++) {
		memcpy(recs[i].k, keys + 3 * (size_t)i, 3 * sizeof(uint32_t));
		recs[i].v = values[i];
	}
	xo_pair* p = NULL;
	int64_t count = round_pairs(recs, n_records, lowerbound, &p);
	for (int64_t i = 0; i < *n_records; i++) {
		memcpy(keys + 3 * (size_t)i, recs[i].k, 3 * sizeof(uint32_t));
		values[i] = recs[i].v;
	}
	free(recs);
	*pairs = (int32_t*)p;
	return count;
}

/* ------------------------------------------------------------- distance -- */

/* kernel.cu:235-252 / 262-280: Wagner-Fischer edit distance, one column. */
static int levenshtein(const char* a, int la, const char* b, int lb) {
	int stack_col[64] = {0};
	int* col = la + 1 <= 64 ? stack_col : (int*)malloc(sizeof(int) * (size_t)(la + 1));
	for (int y = 0; y <= la; y++)
		col[y] = y;
	for (int x = 1; x <= lb; x++) {
		int diag = col[0];
		col[0] = x;
		for (int y = 1; y <= la; y++) {
			int up = col[y];
			int best = col[y] + 1;
			if (col[y - 1] + 1 < best)
				best = col[y - 1] + 1;
			int sub = diag + (a[y - 1] == b[x - 1] ? 0 : 1);
			if (sub < best)
				best = sub;
			col[y] = best;
			diag = up;
		}
	}
	int d = col[la];
	if (col != stack_col)
		free(col);
	return d;
}

/* kernel.cu:290-300: Hamming distance; strings of unequal length give 77. */
static int hamming(const char* a, int la, const char* b, int lb) {
	if (la != lb)
		return XO_HAMMING_SENTINEL;
	int d = 0;
	for (int i = 0; i < la; i++)
		d += a[i] != b[i];
	return d;
}

int xo_distance(const char* a, int la, const char* b, int lb, int measure) {
	return measure == 1 ? hamming(a, la, b, lb) : levenshtein(a, la, b, lb);
}

/* ------------------------------------------------------------ NN search -- */

static void encode_all(const char* chars, const uint32_t* offsets, int n, uint32_t* seqs) {
	/* kernel.cu:594-610 toInt3: first min(len, 18) residues, no validation */
	for (int i = 0; i < n; i++) {
		uint32_t* w = seqs + 3 * (size_t)i;
		w[0] = w[1] = w[2] = 0;
		int len = (int)(offsets[i + 1] - offsets[i]);
		if (len > XO_MAX_PREFIX)
			len = XO_MAX_PREFIX;
		for (int k = 0; k < len; k++)
			slot_or(w, k, chars[offsets[i] + k] - '@');
	}
}

/* xtn.cu:267-509 (NN mode) with the id range split into `rounds` lower-bound
 * rounds (xtn.cu:385-470); stage 4 per round is xtn_inner.cu:487-500:
 * sort+unique (xtn_inner.cu:138-142), cal_distance (kernel.cu:377-392: x == y
 * dropped, keep dist <= d) and output in ascending (x, y). */
int64_t xo_nn_port(const char* chars, const uint32_t* offsets, int n, int distance,
                   int measure, int rounds, int32_t** out, xo_stats* stats) {
	xo_stats st;
	memset(&st, 0, sizeof(st));
	uint32_t* seqs = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)(n > 0 ? n : 1));
	encode_all(chars, offsets, n, seqs);
	int64_t nrec = 0;
	xo_rec* recs = make_variants(seqs, n, distance, &nrec);
	free(seqs);
	st.n_variants = nrec;
	rec_sort(recs, nrec);

	int64_t cap = 1024, len = 0;
	int32_t* res = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)cap);
	if (rounds < 1)
		rounds = 1;
	for (int r = 0; r < rounds; r++) {
		int lowerbound = (int)((int64_t)n * (r + 1) / rounds) - 1;
		xo_pair* pairs = NULL;
		int64_t np = round_pairs(recs, &nrec, lowerbound, &pairs);
		st.n_candidates += np;
		for (int64_t k = 0; k < np; k++) {
			if (k > 0 && pairs[k].x == pairs[k - 1].x && pairs[k].y == pairs[k - 1].y)
				continue;
			st.n_unique++;
			int x = pairs[k].x, y = pairs[k].y;
			if (x == y)
				continue;
			int d = xo_distance(chars + offsets[x], (int)(offsets[x + 1] - offsets[x]),
			                    chars + offsets[y], (int)(offsets[y + 1] - offsets[y]), measure);
			d = (int)(signed char)d; /* stored as char (kernel.cu:386-391) */
			if (d > distance)
				continue;
			if (len == cap) {
				cap *= 2;
				res = (int32_t*)realloc(res, sizeof(int32_t) * 3 * (size_t)cap);
			}
			res[3 * len] = x;
			res[3 * len + 1] = y;
			res[3 * len + 2] = d;
			len++;
		}
		free(pairs);
	}
	free(recs);
	st.n_output = len;
	if (stats)
		*stats = st;
	*out = res;
	return len;
}

/* ------------------------------------------------- multi-threaded port -- */

static uint32_t rec_part(const xo_rec* r, int parts) {
	uint64_t h = ((uint64_t)r->k[0] << 32 | r->k[1]) * 0x9E3779B97F4A7C15ull;
	h ^= (uint64_t)r->k[2] * 0xC2B2AE3D27D4EB4Full;
	h ^= h >> 29;
	return (uint32_t)(h % (uint64_t)parts);
}

/* every i<j position pair of the equal-key groups of sorted recs[0, n), as
 * (min id, max id), self pairs included (kernel.cu:157-191, one round) */
static xo_pair* group_pairs(const xo_rec* recs, int64_t n, int64_t* count) {
	int64_t c = 0;
	for (int64_t s = 0; s < n;) {
		int64_t e = s + 1;
		while (e < n && rec_key_eq(&recs[s], &recs[e]))
			e++;
		c += (e - s) * (e - s - 1) / 2;
		s = e;
	}
	xo_pair* p = (xo_pair*)malloc(sizeof(xo_pair) * (size_t)(c > 0 ? c : 1));
	int64_t pos = 0;
	for (int64_t s = 0; s < n;) {
		int64_t e = s + 1;
		while (e < n && rec_key_eq(&recs[s], &recs[e]))
			e++;
		for (int64_t i = s; i < e; i++)
			for (int64_t j = i + 1; j < e; j++) {
				int32_t a = recs[i].v, b = recs[j].v;
				p[pos].x = a < b ? a : b;
				p[pos].y = a < b ? b : a;
				pos++;
			}
		s = e;
	}
	*count = c;
	return p;
}

int64_t xo_nn_port_mt(const char* chars, const uint32_t* offsets, int n, int distance,
                      int measure, int threads, int32_t** out, xo_stats* stats) {
	if (threads < 1)
		threads = 1;
	const int T = threads;
	xo_stats st;
	memset(&st, 0, sizeof(st));
	uint32_t* seqs = (uint32_t*)malloc(sizeof(uint32_t) * 3 * (size_t)(n > 0 ? n : 1));
	encode_all(chars, offsets, n, seqs);
	/* variant counts per sequence range, then generation in place */
	int64_t* vstart = (int64_t*)calloc((size_t)T + 1, sizeof(int64_t));
#pragma omp parallel num_threads(T)
	{
		const int t = omp_get_thread_num();
		const int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
		int64_t c = 0;
		for (int i = lo; i < hi; i++)
			c += xo_combination_len(xo_len_decode(seqs + 3 * (size_t)i), distance);
		vstart[t + 1] = c;
	}
	for (int t = 0; t < T; t++)
		vstart[t + 1] += vstart[t];
	const int64_t V = vstart[T];
	st.n_variants = V;
	xo_rec* recs = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)(V > 0 ? V : 1));
	xo_rec* part = (xo_rec*)malloc(sizeof(xo_rec) * (size_t)(V > 0 ? V : 1));
	int64_t* pcnt = (int64_t*)calloc((size_t)T * T, sizeof(int64_t)); /* [thread][partition] */
	int64_t* pbase = (int64_t*)calloc((size_t)T * T + 1, sizeof(int64_t));
	int64_t* pstart = (int64_t*)calloc((size_t)T + 1, sizeof(int64_t));
	xo_pair** cand = (xo_pair**)calloc((size_t)T, sizeof(xo_pair*));
	int64_t* ncand = (int64_t*)calloc((size_t)T, sizeof(int64_t));
	int64_t* xcnt = (int64_t*)calloc((size_t)T * T, sizeof(int64_t)); /* [thread][x range] */
	int64_t* xbase = (int64_t*)calloc((size_t)T * T + 1, sizeof(int64_t));
	int64_t* xstart = (int64_t*)calloc((size_t)T + 1, sizeof(int64_t));
	xo_pair* byx = NULL;
	int32_t** res = (int32_t**)calloc((size_t)T, sizeof(int32_t*));
	int64_t* nres = (int64_t*)calloc((size_t)T, sizeof(int64_t));
	int64_t* nuniq = (int64_t*)calloc((size_t)T, sizeof(int64_t));
#pragma omp parallel num_threads(T)
	{
		const int t = omp_get_thread_num();
		const int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
		/* 1. variants of sequences [lo, hi) (generate_combination.cu) */
		{
			int64_t pos = vstart[t];
			uint32_t* tmp = NULL;
			int cap = 0;
			for (int i = lo; i < hi; i++) {
				const uint32_t* sq = seqs + 3 * (size_t)i;
				const int c = xo_combination_len(xo_len_decode(sq), distance);
				if (c > cap) {
					cap = c;
					tmp = (uint32_t*)realloc(tmp, sizeof(uint32_t) * 3 * (size_t)c);
				}
				xo_expand_keys(sq, distance, tmp, c);
				for (int k = 0; k < c; k++) {
					memcpy(recs[pos].k, tmp + 3 * (size_t)k, 3 * sizeof(uint32_t));
					recs[pos].v = i;
					pos++;
				}
			}
			free(tmp);
			for (int64_t r = vstart[t]; r < vstart[t + 1]; r++)
				pcnt[(size_t)t * T + rec_part(&recs[r], T)]++;
		}
#pragma omp barrier
#pragma omp single
		{
			/* partition-major bases: partition p holds thread 0's, then 1's ... records */
			int64_t acc = 0;
			for (int p = 0; p < T; p++) {
				pstart[p] = acc;
				for (int q = 0; q < T; q++) {
					pbase[(size_t)q * T + p] = acc;
					acc += pcnt[(size_t)q * T + p];
				}
			}
			pstart[T] = acc;
		}
		/* 2. records into their key partitions */
		{
			int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)T);
			for (int p = 0; p < T; p++)
				cur[p] = pbase[(size_t)t * T + p];
			for (int64_t r = vstart[t]; r < vstart[t + 1]; r++)
				part[cur[rec_part(&recs[r], T)]++] = recs[r];
			free(cur);
		}
#pragma omp barrier
		/* 3. partition t: stable key sort (cub.cu:95-104), groups, pairs */
		{
			xo_rec* pr = part + pstart[t];
			const int64_t m = pstart[t + 1] - pstart[t];
			rec_sort(pr, m);
			cand[t] = group_pairs(pr, m, &ncand[t]);
			const xo_pair* cp = cand[t];
			for (int64_t k = 0; k < ncand[t]; k++)
				xcnt[(size_t)t * T + (int)((int64_t)cp[k].x * T / (n > 0 ? n : 1))]++;
		}
#pragma omp barrier
#pragma omp single
		{
			int64_t acc = 0;
			for (int r = 0; r < T; r++) {
				xstart[r] = acc;
				for (int q = 0; q < T; q++) {
					xbase[(size_t)q * T + r] = acc;
					acc += xcnt[(size_t)q * T + r];
				}
			}
			xstart[T] = acc;
			byx = (xo_pair*)malloc(sizeof(xo_pair) * (size_t)(acc > 0 ? acc : 1));
		}
		/* 4. pairs into x ranges */
		{
			int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)T);
			for (int r = 0; r < T; r++)
				cur[r] = xbase[(size_t)t * T + r];
			const xo_pair* cp = cand[t];
			for (int64_t k = 0; k < ncand[t]; k++)
				byx[cur[(int)((int64_t)cp[k].x * T / (n > 0 ? n : 1))]++] = cp[k];
			free(cur);
			free(cand[t]);
		}
#pragma omp barrier
		/* 5. range t: sort + unique (xtn_inner.cu:138-142), distance and
		 * filter (kernel.cu:377-392) */
		{
			xo_pair* xp = byx + xstart[t];
			const int64_t m = xstart[t + 1] - xstart[t];
			qsort(xp, (size_t)m, sizeof(xo_pair), pair_cmp);
			int64_t cap = 1024, len = 0, u = 0;
			int32_t* o = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)cap);
			for (int64_t k = 0; k < m; k++) {
				if (k > 0 && xp[k].x == xp[k - 1].x && xp[k].y == xp[k - 1].y)
					continue;
				u++;
				const int x = xp[k].x, y = xp[k].y;
				if (x == y)
					continue;
				int d = xo_distance(chars + offsets[x], (int)(offsets[x + 1] - offsets[x]),
				                    chars + offsets[y], (int)(offsets[y + 1] - offsets[y]), measure);
				d = (int)(signed char)d;
				if (d > distance)
					continue;
				if (len == cap) {
					cap *= 2;
					o = (int32_t*)realloc(o, sizeof(int32_t) * 3 * (size_t)cap);
				}
				o[3 * len] = x;
				o[3 * len + 1] = y;
				o[3 * len + 2] = d;
				len++;
			}
			res[t] = o;
			nres[t] = len;
			nuniq[t] = u;
		}
	}
	int64_t total = 0;
	for (int t = 0; t < T; t++) {
		st.n_candidates += ncand[t];
		st.n_unique += nuniq[t];
		total += nres[t];
	}
	int32_t* all = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(total > 0 ? total : 1));
	int64_t pos = 0;
	for (int t = 0; t < T; t++) {
		memcpy(all + 3 * pos, res[t], sizeof(int32_t) * 3 * (size_t)nres[t]);
		pos += nres[t];
		free(res[t]);
	}
	st.n_output = total;
	free(seqs);
	free(vstart);
	free(recs);
	free(part);
	free(pcnt);
	free(pbase);
	free(pstart);
	free(cand);
	free(ncand);
	free(xcnt);
	free(xbase);
	free(xstart);
	free(byx);
	free(res);
	free(nres);
	free(nuniq);
	if (stats)
		*stats = st;
	*out = all;
	return total;
}

int64_t xo_nn_brute(const char* chars, const uint32_t* offsets, int n, int distance,
                    int measure, int32_t** out) {
	int64_t cap = 1024, len = 0;
	int32_t* res = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)cap);
	for (int i = 0; i < n; i++) {
		int li = (int)(offsets[i + 1] - offsets[i]);
		for (int j = i + 1; j < n; j++) {
			int lj = (int)(offsets[j + 1] - offsets[j]);
			/* a length gap above d already rules out an edit distance <= d */
			if (measure == 0 && (li - lj > distance || lj - li > distance))
				continue;
			int d = xo_distance(chars + offsets[i], li, chars + offsets[j], lj, measure);
			if (d > distance)
				continue;
			if (len == cap) {
				cap *= 2;
				res = (int32_t*)realloc(res, sizeof(int32_t) * 3 * (size_t)cap);
			}
			res[3 * len] = i;
			res[3 * len + 1] = j;
			res[3 * len + 2] = d;
			len++;
		}
	}
	*out = res;
	return len;
}

/* every j != row with dist(row, j) <= distance, ascending j (the brute-force
 * definition restricted to one row). *out receives 2 int32 per hit (j, dist). */
int64_t xo_neighbors_of(const char* chars, const uint32_t* offsets, int n, int row, int distance,
                        int measure, int32_t** out) {
	int64_t cap = 64, len = 0;
	int32_t* res = (int32_t*)malloc(sizeof(int32_t) * 2 * (size_t)cap);
	const int lr = (int)(offsets[row + 1] - offsets[row]);
	for (int j = 0; j < n; j++) {
		if (j == row)
			continue;
		int lj = (int)(offsets[j + 1] - offsets[j]);
		if (measure == 0 && (lr - lj > distance || lj - lr > distance))
			continue;
		int d = xo_distance(chars + offsets[row], lr, chars + offsets[j], lj, measure);
		if (d > distance)
			continue;
		if (len == cap) {
			cap *= 2;
			res = (int32_t*)realloc(res, sizeof(int32_t) * 2 * (size_t)cap);
		}
		res[2 * len] = j;
		res[2 * len + 1] = d;
		len++;
	}
	*out = res;
	return len;
}

/* -------------------------------------------------------------- overlap -- */

typedef struct {
	int32_t frequency, repertoire, original;
} xo_info;

typedef struct {
	const char* chars;
	const uint32_t* offsets;
} str_ctx;

/* cub.cu:55-78: lexicographic on the full strings, shorter first on a tie. */
static int info_less(const str_ctx* c, const xo_info* a, const xo_info* b) {
	uint32_t sa = c->offsets[a->original], sb = c->offsets[b->original];
	int la = (int)(c->offsets[a->original + 1] - sa);
	int lb = (int)(c->offsets[b->original + 1] - sb);
	int m = la < lb ? la : lb;
	for (int i = 0; i < m; i++) {
		char x = c->chars[sa + i], y = c->chars[sb + i];
		if (x != y)
			return x < y;
	}
	return la < lb;
}

static int info_equal(const str_ctx* c, const xo_info* a, const xo_info* b) {
	return !info_less(c, a, b) && !info_less(c, b, a);
}

static void info_sort(const str_ctx* c, xo_info* a, int n) {
	if (n < 2)
		return;
	xo_info* tmp = (xo_info*)malloc(sizeof(xo_info) * (size_t)n);
	xo_info* src = a;
	xo_info* dst = tmp;
	for (int width = 1; width < n; width *= 2) {
		for (int lo = 0; lo < n; lo += 2 * width) {
			int mid = lo + width < n ? lo + width : n;
			int hi = lo + 2 * width < n ? lo + 2 * width : n;
			int i = lo, j = mid, o = lo;
			while (i < mid && j < hi) {
				if (info_less(c, &src[j], &src[i]))
					dst[o++] = src[j++];
				else
					dst[o++] = src[i++];
			}
			while (i < mid)
				dst[o++] = src[i++];
			while (j < hi)
				dst[o++] = src[j++];
		}
		xo_info* t = src;
		src = dst;
		dst = t;
	}
	if (src != a)
		memcpy(a, src, sizeof(xo_info) * (size_t)n);
	free(tmp);
}

typedef struct {
	int32_t a, b;
	uint64_t v;
} xo_rep_entry;

static int rep_entry_cmp(const void* pa, const void* pb) {
	const xo_rep_entry* a = (const xo_rep_entry*)pa;
	const xo_rep_entry* b = (const xo_rep_entry*)pb;
	if (a->a != b->a)
		return a->a < b->a ? -1 : 1;
	if (a->b != b->b)
		return a->b < b->b ? -1 : 1;
	return 0;
}

typedef struct {
	xo_rep_entry* e;
	int64_t len, cap;
} rep_list;

static void rep_push(rep_list* l, int32_t a, int32_t b, uint64_t v) {
	if (l->len == l->cap) {
		l->cap = l->cap ? l->cap * 2 : 1024;
		l->e = (xo_rep_entry*)realloc(l->e, sizeof(xo_rep_entry) * (size_t)l->cap);
	}
	l->e[l->len].a = a;
	l->e[l->len].b = b;
	l->e[l->len].v = v;
	l->len++;
}

/* which: bit 0 = exact-match term (init_overlap_output, kernel.cu:565-584),
 *        bit 1 = neighbour term (pair2rep, kernel.cu:405-435).
 * Flow: xtn_overlap_inner.cu:57-124 (stable string sort, run-length groups,
 * init term), NN on the unique strings, xtn_overlap_inner.cu:141-203 (verify
 * through the last member of each group, expand to repertoire pairs) and
 * mergeOutput (xtn_overlap_inner.cu:212-245: sort by (A, B), sum by key). */
static int64_t overlap_terms(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                             const int32_t* rep, int distance, int measure, int which, int threads,
                             uint64_t** out) {
	str_ctx sc = {chars, offsets};
	xo_info* info = (xo_info*)malloc(sizeof(xo_info) * (size_t)(n > 0 ? n : 1));
	for (int i = 0; i < n; i++) {
		info[i].frequency = freq[i];
		info[i].repertoire = rep[i];
		info[i].original = i;
	}
	info_sort(&sc, info, n);

	/* group starts (run-length encode by string equality) */
	int* gstart = (int*)malloc(sizeof(int) * (size_t)(n + 1));
	int ng = 0;
	for (int i = 0; i < n; i++)
		if (i == 0 || !info_equal(&sc, &info[i - 1], &info[i]))
			gstart[ng++] = i;
	gstart[ng] = n;

	rep_list L = {NULL, 0, 0};
	if (which & 1) {
		for (int g = 0; g < ng; g++)
			for (int i = gstart[g]; i < gstart[g + 1]; i++)
				for (int j = i; j < gstart[g + 1]; j++)
					rep_push(&L, info[i].repertoire, info[j].repertoire,
					         (uint64_t)(int64_t)info[i].frequency * (uint64_t)(int64_t)info[j].frequency);
	}

	if ((which & 2) && ng > 1) {
		/* unique strings, one per group, in sorted-group order */
		char* uchars;
		uint32_t* uoff = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ng + 1));
		size_t total = 0;
		for (int g = 0; g < ng; g++) {
			int o = info[gstart[g + 1] - 1].original; /* seqOffset[x]-1, kernel.cu:337-338 */
			total += offsets[o + 1] - offsets[o];
		}
		uchars = (char*)malloc(total > 0 ? total : 1);
		uoff[0] = 0;
		for (int g = 0; g < ng; g++) {
			int o = info[gstart[g + 1] - 1].original;
			uint32_t l = offsets[o + 1] - offsets[o];
			memcpy(uchars + uoff[g], chars + offsets[o], l);
			uoff[g + 1] = uoff[g] + l;
		}
		int32_t* nn = NULL;
		/* threads > 1: the same NN output from the OpenMP port */
		int64_t nnn = threads > 1 ? xo_nn_port_mt(uchars, uoff, ng, distance, measure, threads, &nn, NULL)
		                          : xo_nn_port(uchars, uoff, ng, distance, measure, 1, &nn, NULL);
		for (int64_t k = 0; k < nnn; k++) {
			int x = nn[3 * k], y = nn[3 * k + 1];
			for (int i = gstart[x]; i < gstart[x + 1]; i++)
				for (int j = gstart[y]; j < gstart[y + 1]; j++) {
					int ra = info[i].repertoire, rb = info[j].repertoire;
					uint64_t v = (uint64_t)(int64_t)info[i].frequency * (uint64_t)(int64_t)info[j].frequency;
					if (ra == rb)
						v *= 2;
					rep_push(&L, ra < rb ? ra : rb, ra < rb ? rb : ra, v);
				}
		}
		free(nn);
		free(uchars);
		free(uoff);
	}

	qsort(L.e, (size_t)L.len, sizeof(xo_rep_entry), rep_entry_cmp);
	uint64_t* res = (uint64_t*)malloc(sizeof(uint64_t) * 3 * (size_t)(L.len > 0 ? L.len : 1));
	int64_t m = 0;
	for (int64_t k = 0; k < L.len; k++) {
		if (m > 0 && (int64_t)res[3 * (m - 1)] == L.e[k].a && (int64_t)res[3 * (m - 1) + 1] == L.e[k].b) {
			res[3 * (m - 1) + 2] += L.e[k].v;
			continue;
		}
		res[3 * m] = (uint64_t)L.e[k].a;
		res[3 * m + 1] = (uint64_t)L.e[k].b;
		res[3 * m + 2] = L.e[k].v;
		m++;
	}
	free(L.e);
	free(gstart);
	free(info);
	*out = res;
	return m;
}

int64_t xo_overlap_port(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                        const int32_t* rep, int distance, int measure, uint64_t** out) {
	return overlap_terms(chars, offsets, n, freq, rep, distance, measure, 3, 1, out);
}

int64_t xo_overlap_port_mt(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                           const int32_t* rep, int distance, int measure, int threads, uint64_t** out) {
	return overlap_terms(chars, offsets, n, freq, rep, distance, measure, 3, threads, out);
}

int64_t xo_overlap_terms(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                         const int32_t* rep, int distance, int measure, int which, uint64_t** out) {
	return overlap_terms(chars, offsets, n, freq, rep, distance, measure, which, 1, out);
}

void xo_free(void* p) {
	free(p);
}
