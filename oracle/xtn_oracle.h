/*
 * xtn_oracle.h -- CPU restatement of the XTNeighbor-streaming hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker for the HIP
 * implementation in xt-neighbor_amd/. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker (or as
 * the timed CPU baseline). The product path never calls into it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the XT-neighbor repository root).
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   xo_char_encode/xo_str_encode/xo_str_decode/xo_len_decode/xo_remove_char
 *       -> known-answer vectors of xtneighbor_streaming/test/test_codec.cu
 *   xo_expand_keys      -> test_generate_combination.cu vectors (DFS order) and
 *                          the closed-form count sum_k C(L,k)
 *   xo_combination_len  -> closed-form binomial sums
 *   xo_stage1           -> test_stream1.cu vector (sorted keys, stable values)
 *   xo_sort_stage       -> test_stream2.cu vector (stable re-sort)
 *   xo_gen_pairs        -> test_stream3.cu vectors (pairs with self pairs, the
 *                          records kept for the next lower-bound round)
 *   xo_distance         -> test_stream4.cu vector (incl. 22-char strings),
 *                          textbook identities (d(s,s)=0, symmetry, |la-lb|
 *                          lower bound, brute-force DP over all alignments on
 *                          tiny strings)
 *   xo_nn_port          -> xo_nn_brute (independent O(n^2) definition) plus
 *                          the reference fixture data/dummy_input.txt
 *   xo_nn_port_mt       -> xo_nn_port (triples and counters) at several
 *                          thread counts
 *   xo_overlap_port     -> test_overlap_init.cu, test_stream4_overlap.cu,
 *                          data/dummy_input_airr.tsv expected matrix, and the
 *                          closed form R^T (A + I) R on duplicate-free input
 *   xo_overlap_port_mt  -> xo_overlap_port at several thread counts
 */
#ifndef XTN_ORACLE_H
#define XTN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
	int64_t n_variants;      /* deletion records generated (V) */
	int64_t n_candidates;    /* pairs emitted from groups, self pairs included (C) */
	int64_t n_unique;        /* unique candidate pairs incl. (i,i) ("4 afterdup") */
	int64_t n_output;        /* pairs written (O) */
} xo_stats;

int  xo_char_encode(char c);
void xo_str_encode(const char* s, uint32_t out[3]);
void xo_str_decode(const uint32_t in[3], char out[19]);
int  xo_len_decode(const uint32_t in[3]);
void xo_remove_char(const uint32_t in[3], int position, uint32_t out[3]);
int  xo_combination_len(int len, int distance);
/* writes n keys (3 words each) in the reference's DFS order; returns the count
 * that the sequence produces (callers pass n = xo_combination_len). */
int  xo_expand_keys(const uint32_t seq[3], int distance, uint32_t* out, int n);

/* stage 1: variants of every sequence, then the reference's stable key sort.
 * keys: 3*V words, values: V ints. Returns V. Buffers are malloc'ed. */
int64_t xo_stage1(const uint32_t* seqs, int n, int distance, uint32_t** keys, int32_t** values);
/* stable re-sort of (key, value) records by the reference comparator, in place */
void xo_sort_stage(uint32_t* keys, int32_t* values, int64_t n);
/* pairs of one lower-bound round over sorted records (min id <= lowerbound),
 * sorted; then the records kept for the next round are compacted in place.
 * Returns the pair count; *n_records receives the kept record count. */
int64_t xo_gen_pairs(uint32_t* keys, int32_t* values, int64_t* n_records, int lowerbound,
                     int32_t** pairs);

int xo_distance(const char* a, int la, const char* b, int lb, int measure);

/* NN search, the reference algorithm restated (single lower-bound round unless
 * rounds > 1, in which case the id range is split into that many rounds).
 * *out receives 3 int32 per output pair (i, j, dist), sorted by (i, j). */
int64_t xo_nn_port(const char* chars, const uint32_t* offsets, int n, int distance,
                   int measure, int rounds, int32_t** out, xo_stats* stats);
/* xo_nn_port on `threads` OpenMP threads (the timed multi-core CPU baseline):
 * variants generated per sequence range, records hash-partitioned by key so
 * that a key group stays in one partition, each partition sorted and paired
 * by one thread, pairs partitioned by x range, sorted, deduplicated and
 * verified per range, ranges concatenated in order. Same output and counters
 * as xo_nn_port with one round. */
int64_t xo_nn_port_mt(const char* chars, const uint32_t* offsets, int n, int distance,
                      int measure, int threads, int32_t** out, xo_stats* stats);
/* NN search by definition: every i<j with dist <= distance. */
int64_t xo_nn_brute(const char* chars, const uint32_t* offsets, int n, int distance,
                    int measure, int32_t** out);

/* the brute force for one row: every j != row within the distance,
 * ascending; *out receives (j, dist) int32 pairs. */
int64_t xo_neighbors_of(const char* chars, const uint32_t* offsets, int n, int row, int distance,
                        int measure, int32_t** out);

/* repertoire overlap, the reference algorithm restated. freq/rep are per input
 * row. *out receives 3 uint64 per entry (A, B, sum), sorted by (A, B). */
int64_t xo_overlap_port(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                        const int32_t* rep, int distance, int measure, uint64_t** out);
/* xo_overlap_port with the neighbour search on `threads` OpenMP threads
 * (xo_nn_port_mt); same output. */
int64_t xo_overlap_port_mt(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                           const int32_t* rep, int distance, int measure, int threads, uint64_t** out);
/* the same with a term selector: bit 0 = exact-match term, bit 1 = neighbour
 * term (both = xo_overlap_port). */
int64_t xo_overlap_terms(const char* chars, const uint32_t* offsets, int n, const int32_t* freq,
                         const int32_t* rep, int distance, int measure, int which, uint64_t** out);

void xo_free(void* p);

#ifdef __cplusplus
}
#endif
#endif
