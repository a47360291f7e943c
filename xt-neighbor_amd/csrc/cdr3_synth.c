/*
 * cdr3_synth.c -- seeded synthetic CDR3 workloads (libxtn_synth.so).
 *
 * Not part of the reference: the benchmark inputs it generates stand in for
 * the Emerson repertoires the reference's notebooks used, which are not
 * available offline. PRNG: SplitMix64 (Steele, Lea, Flood 2014); normals by
 * Box-Muller.
 *
 * Sequence rule: length = clip(round(N(mu, sigma)), lmin, lmax), first residue
 * 'C', last 'F' or 'W' (50/50), interior residues uniform over
 * ACDEFGHIKLMNPQRSTVWY. A planted sequence copies a uniformly chosen earlier
 * sequence and applies k random edits (substitution, insertion or deletion,
 * uniform type, interior position), k uniform in [1, max_edits]. Exact
 * duplicates are rejected and regenerated.
 */
#include "cdr3_synth.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char AA20[] = "ACDEFGHIKLMNPQRSTVWY";

typedef struct {
	uint64_t s;
} rng_t;

static uint64_t rng_next(rng_t* r) {
	uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static double rng_unif(rng_t* r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static uint64_t rng_below(rng_t* r, uint64_t n) { return n ? rng_next(r) % n : 0; }

static double rng_normal(rng_t* r) {
	double u1 = rng_unif(r), u2 = rng_unif(r);
	if (u1 < 1e-300)
		u1 = 1e-300;
	return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ---------------------------------------------------------- string set -- */

typedef struct {
	uint64_t* keys; /* hash, 0 = empty */
	int32_t* idx;
	uint64_t mask;
	const char* chars;
	const uint32_t* offsets;
} strset;

static uint64_t str_hash(const char* s, int len) {
	uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)len;
	for (int i = 0; i < len; i++)
		h = (h ^ (unsigned char)s[i]) * 0x100000001b3ull;
	h ^= h >> 29;
	return h | 1ull;
}

static int strset_init(strset* ss, int64_t n) {
	uint64_t cap = 16;
	while (cap < (uint64_t)(2 * n + 16))
		cap <<= 1;
	ss->keys = (uint64_t*)calloc(cap, sizeof(uint64_t));
	ss->idx = (int32_t*)malloc(cap * sizeof(int32_t));
	ss->mask = cap - 1;
	return ss->keys && ss->idx ? 0 : -1;
}

static void strset_free(strset* ss) {
	free(ss->keys);
	free(ss->idx);
}

static void strset_clear(strset* ss) { memset(ss->keys, 0, (ss->mask + 1) * sizeof(uint64_t)); }

/* 1 if s is present; otherwise inserts (with index i) unless i < 0 */
static int strset_find_or_add(strset* ss, const char* s, int len, int32_t i) {
	uint64_t h = str_hash(s, len);
	uint64_t p = h & ss->mask;
	while (ss->keys[p]) {
		if (ss->keys[p] == h) {
			int32_t j = ss->idx[p];
			int lj = (int)(ss->offsets[j + 1] - ss->offsets[j]);
			if (lj == len && memcmp(ss->chars + ss->offsets[j], s, (size_t)len) == 0)
				return 1;
		}
		p = (p + 1) & ss->mask;
	}
	if (i >= 0) {
		ss->keys[p] = h;
		ss->idx[p] = i;
	}
	return 0;
}

/* ------------------------------------------------------------- editing -- */

static int fresh_seq(rng_t* r, char* buf, double mu, double sigma, int lmin, int lmax) {
	int len = (int)lround(mu + sigma * rng_normal(r));
	if (len < lmin)
		len = lmin;
	if (len > lmax)
		len = lmax;
	if (len < 2)
		len = 2;
	buf[0] = 'C';
	for (int k = 1; k < len - 1; k++)
		buf[k] = AA20[rng_below(r, 20)];
	buf[len - 1] = rng_unif(r) < 0.5 ? 'F' : 'W';
	return len;
}

static int edit_seq(rng_t* r, char* buf, int len, int cap) {
	int type = (int)rng_below(r, 3);
	if (len < 3 && type != 1)
		type = 1;
	if (type == 1 && len + 1 > cap)
		type = 0;
	if (type == 0) { /* substitution */
		int p = 1 + (int)rng_below(r, (uint64_t)(len - 2));
		char c;
		do
			c = AA20[rng_below(r, 20)];
		while (c == buf[p]);
		buf[p] = c;
		return len;
	}
	if (type == 1) { /* insertion before position p, 1 <= p <= len-1 */
		int p = 1 + (int)rng_below(r, (uint64_t)(len - 1));
		memmove(buf + p + 1, buf + p, (size_t)(len - p));
		buf[p] = AA20[rng_below(r, 20)];
		return len + 1;
	}
	{ /* deletion */
		int p = 1 + (int)rng_below(r, (uint64_t)(len - 2));
		memmove(buf + p, buf + p + 1, (size_t)(len - p - 1));
		return len - 1;
	}
}

/* --------------------------------------------------------------- NN set -- */

int64_t xs_generate_nn(uint64_t seed, int32_t n, double mu, double sigma, int32_t lmin, int32_t lmax,
                       double planted_frac, int32_t max_edits, char* chars, int64_t cap_chars,
                       uint32_t* offsets) {
	if (n < 0 || lmin < 2 || lmax < lmin || lmax > 200 || max_edits < 1)
		return -1;
	rng_t r = {seed * 0x2545F4914F6CDD1Dull + 0x1234567ull};
	strset ss;
	if (strset_init(&ss, n))
		return -2;
	ss.chars = chars;
	ss.offsets = offsets;
	offsets[0] = 0;
	char buf[256];
	int64_t total = 0;
	for (int32_t i = 0; i < n; i++) {
		int len = 0;
		for (int attempt = 0;; attempt++) {
			if (i > 0 && rng_unif(&r) < planted_frac) {
				int32_t par = (int32_t)rng_below(&r, (uint64_t)i);
				len = (int)(offsets[par + 1] - offsets[par]);
				memcpy(buf, chars + offsets[par], (size_t)len);
				int k = 1 + (int)rng_below(&r, (uint64_t)max_edits);
				for (int e = 0; e < k; e++)
					len = edit_seq(&r, buf, len, (int)sizeof(buf) - 1);
				if (len < lmin || len > lmax)
					continue;
			} else {
				len = fresh_seq(&r, buf, mu, sigma, lmin, lmax);
			}
			if (total + len > cap_chars) {
				strset_free(&ss);
				return -3;
			}
			if (!strset_find_or_add(&ss, buf, len, -1))
				break;
			if (attempt > 1000000) {
				strset_free(&ss);
				return -4;
			}
		}
		memcpy(chars + total, buf, (size_t)len);
		total += len;
		offsets[i + 1] = (uint32_t)total;
		strset_find_or_add(&ss, buf, len, i);
	}
	strset_free(&ss);
	return total;
}

/* ------------------------------------------------------ repertoire sets -- */

int64_t xs_overlap_sizes(uint64_t seed, int32_t n_reps, int32_t size_min, int32_t size_max,
                         int32_t* sizes) {
	rng_t r = {seed * 0x9E3779B97F4A7C15ull + 77ull};
	int64_t total = 0;
	double a = log((double)size_min), b = log((double)size_max);
	for (int i = 0; i < n_reps; i++) {
		sizes[i] = (int32_t)lround(exp(a + (b - a) * rng_unif(&r)));
		total += sizes[i];
	}
	return total;
}

int64_t xs_generate_overlap(uint64_t seed, int32_t n_reps, const int32_t* sizes, double shared_frac,
                            double mu, double sigma, int32_t lmin, int32_t lmax, double zipf_s,
                            int32_t zipf_max, char* chars, int64_t cap_chars, uint32_t* offsets,
                            int32_t* freq) {
	if (n_reps < 1 || zipf_max < 1 || lmin < 2 || lmax < lmin || lmax > 200)
		return -1;
	rng_t r = {seed * 0xD1B54A32D192ED03ull + 99ull};
	/* Zipf CDF over 1..zipf_max */
	double* cdf = (double*)malloc(sizeof(double) * (size_t)zipf_max);
	double acc = 0.0;
	for (int k = 1; k <= zipf_max; k++) {
		acc += pow((double)k, -zipf_s);
		cdf[k - 1] = acc;
	}
	int32_t maxsize = 0;
	int64_t ntot = 0;
	for (int i = 0; i < n_reps; i++) {
		if (sizes[i] > maxsize)
			maxsize = sizes[i];
		ntot += sizes[i];
	}
	strset ss;
	if (strset_init(&ss, maxsize)) {
		free(cdf);
		return -2;
	}
	ss.chars = chars;
	ss.offsets = offsets;
	offsets[0] = 0;
	char buf[256];
	int64_t total = 0, row = 0;
	for (int rep = 0; rep < n_reps; rep++) {
		strset_clear(&ss);
		int64_t rep_start = row;
		for (int32_t i = 0; i < sizes[rep]; i++) {
			int len = 0;
			for (int attempt = 0;; attempt++) {
				if (rep_start > 0 && rng_unif(&r) < shared_frac) {
					int64_t src = (int64_t)rng_below(&r, (uint64_t)rep_start);
					len = (int)(offsets[src + 1] - offsets[src]);
					memcpy(buf, chars + offsets[src], (size_t)len);
					if (rng_unif(&r) < 0.5)
						len = edit_seq(&r, buf, len, (int)sizeof(buf) - 1);
					if (len < lmin || len > lmax)
						continue;
				} else {
					len = fresh_seq(&r, buf, mu, sigma, lmin, lmax);
				}
				if (total + len > cap_chars) {
					strset_free(&ss);
					free(cdf);
					return -3;
				}
				if (!strset_find_or_add(&ss, buf, len, -1))
					break;
				if (attempt > 1000000) {
					strset_free(&ss);
					free(cdf);
					return -4;
				}
			}
			memcpy(chars + total, buf, (size_t)len);
			total += len;
			offsets[row + 1] = (uint32_t)total;
			strset_find_or_add(&ss, buf, len, (int32_t)row);
			double u = rng_unif(&r) * acc;
			int lo = 0, hi = zipf_max - 1;
			while (lo < hi) {
				int mid = (lo + hi) / 2;
				if (cdf[mid] < u)
					lo = mid + 1;
				else
					hi = mid;
			}
			freq[row] = lo + 1;
			row++;
		}
	}
	strset_free(&ss);
	free(cdf);
	(void)ntot;
	return total;
}

/* ---------------------------------------------------------- query sets -- */

int64_t xs_generate_queries(uint64_t seed, const char* idx_chars, const uint32_t* idx_offsets,
                            int32_t n_index, int32_t n, double mu, double sigma, int32_t lmin,
                            int32_t lmax, double planted_frac, char* chars, int64_t cap_chars,
                            uint32_t* offsets) {
	if (n < 0 || lmin < 2 || lmax < lmin || lmax > 200)
		return -1;
	rng_t r = {seed * 0xA0761D6478BD642Full + 5ull};
	strset ss;
	if (strset_init(&ss, n))
		return -2;
	ss.chars = chars;
	ss.offsets = offsets;
	offsets[0] = 0;
	char buf[256];
	int64_t total = 0;
	for (int32_t i = 0; i < n; i++) {
		int len = 0;
		for (int attempt = 0;; attempt++) {
			if (n_index > 0 && rng_unif(&r) < planted_frac) {
				int32_t par = (int32_t)rng_below(&r, (uint64_t)n_index);
				len = (int)(idx_offsets[par + 1] - idx_offsets[par]);
				memcpy(buf, idx_chars + idx_offsets[par], (size_t)len);
				len = edit_seq(&r, buf, len, (int)sizeof(buf) - 1);
				if (len < lmin || len > lmax)
					continue;
			} else {
				len = fresh_seq(&r, buf, mu, sigma, lmin, lmax);
			}
			if (total + len > cap_chars) {
				strset_free(&ss);
				return -3;
			}
			if (!strset_find_or_add(&ss, buf, len, -1))
				break;
			if (attempt > 1000000) {
				strset_free(&ss);
				return -4;
			}
		}
		memcpy(chars + total, buf, (size_t)len);
		total += len;
		offsets[i + 1] = (uint32_t)total;
		strset_find_or_add(&ss, buf, len, i);
	}
	strset_free(&ss);
	return total;
}

/* ------------------------------------------------------------ CSV out -- */

int32_t xs_write_csv(const char* path, const char* chars, const uint32_t* offsets, int32_t n,
                     const int32_t* freq) {
	FILE* f = fopen(path, "wb");
	if (!f)
		return -1;
	static char buf[1 << 20];
	setvbuf(f, buf, _IOFBF, sizeof(buf));
	fputs(freq ? "cdr3,count\n" : "cdr3\n", f);
	for (int32_t i = 0; i < n; i++) {
		fwrite(chars + offsets[i], 1, offsets[i + 1] - offsets[i], f);
		if (freq)
			fprintf(f, ",%d\n", freq[i]);
		else
			fputc('\n', f);
	}
	const int bad = ferror(f);
	return (fclose(f) != 0 || bad) ? -1 : 0;
}
