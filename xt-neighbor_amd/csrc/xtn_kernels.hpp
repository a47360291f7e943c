// xtn_kernels.hpp -- the XTNeighbor hot-path kernels for gfx950.
//
// Data layout in HBM (one key-space chunk):
//   pk[2*i], pk[2*i+1]   18-residue prefix of sequence i packed as a 90-bit
//                        integer (residue 0 most significant, 5 bits each,
//                        code = c - '@'), the query set in bits 26..55 and
//                        the prefix length in bits 56..63 of the high word.
//   voff[i]              first deletion record of sequence i (exclusive scan
//                        of the per-sequence variant counts), voff[n] = V.
//   rec[r]               one deletion record: 64-bit word, the high 64-idb
//                        bits are a hash of the variant's packed prefix, the
//                        low idb bits the sequence id.
//   pairs[q]             candidate pair (i << idb) | j, i < j.
//
// A hashed key only widens the candidate set (two variants that collide
// produce extra candidates that verification removes); every pair of strings
// within distance d still shares an exact variant of their 18-prefixes, so
// the verified output is unchanged. The reference keys the same variants
// exactly as a 12-byte Int3 (xtn.h:19-25, kernel.cu:594-610).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "xtn_prims.hpp"

namespace xtn {

constexpr int MAX_PREFIX = 18;      // xtn.h:8 MAX_COMPRESSED_LENGTH
constexpr int PK_SET_SHIFT = 26;    // packed high word: 26 residue bits, then the set id
constexpr uint32_t PK_SET_MASK = (1u << 30) - 1u;
constexpr uint64_t PK_HI_MASK = (1ull << 56) - 1ull;  // residues + set, without the length byte
constexpr int HAMMING_SENTINEL = 77;  // kernel.cu:292
constexpr int SMALL_RUN_MAX = 32;   // runs longer than this use the flattened pair kernel

typedef unsigned __int128 u128;

struct RunCtrs {
	unsigned long long n_small;      // runs of 2..SMALL_RUN_MAX records
	unsigned long long n_big;        // longer runs
	unsigned long long pairs_small;  // sum len*(len-1)/2 over small runs
	unsigned long long pairs_big;    // same over big runs
	unsigned long long n_cand;       // candidate pairs written (append cursor)
	unsigned long long n_unique;     // unique candidate pairs verified
	unsigned long long append;       // partition append cursor
	unsigned long long coarse_ovf;   // the generator's coarse pass lost records (sub-range full)
	unsigned long long n_records;    // records the generator's coarse pass wrote (k_seg_tiles)
};

struct Run {
	uint32_t start, len;
};

// ------------------------------------------------------------ packing --

__device__ __forceinline__ uint64_t variant_count(int L, int d) {
	const int m = d < L ? d : L;
	uint64_t c = 1, total = 1;
	for (int j = 1; j <= m; ++j) {
		c = c * (uint64_t)(L - j + 1) / (uint64_t)j;
		total += c;
	}
	return total;
}

// one thread per sequence: pack the 18-prefix, count its deletion variants
// (kernel.cu:33-57 counts the same sum of binomials)
// pk[2i], pk[2i+1]: the 90-bit prefix (residue 0 most significant, 5 bits
// each), then in the high word bits 26..55 the query set of the id (batched
// form, set_of non-null; those bits lie above every residue slot, so they
// ride through the deletions into the variant hash) and bits 56..63 the
// prefix length.
// Returns sequence i's variant count.
__device__ __forceinline__ uint32_t pack_one(const uint8_t* __restrict__ chars, const uint32_t* __restrict__ offs,
                                             const uint32_t* __restrict__ seq_map, uint32_t row_base,
                                             const uint32_t* __restrict__ set_of, int32_t i, int distance,
                                             uint64_t* __restrict__ pk) {
	const uint32_t o = seq_map ? seq_map[i] : row_base + (uint32_t)i;
	const uint32_t s = offs[o];
	const uint32_t len = offs[o + 1] - s;
	const int L = len < (uint32_t)MAX_PREFIX ? (int)len : MAX_PREFIX;
	// the (at most 18) prefix residues with aligned dword loads: six words
	// cover 18 bytes at any alignment, and the residue buffer carries 16
	// bytes of slack past its end
	const uintptr_t addr = (uintptr_t)(chars + s);
	const uint32_t* base = (const uint32_t*)(addr & ~(uintptr_t)3);
	const uint32_t ash = (uint32_t)(addr & 3);
	const int nw = L > 0 ? (int)((ash + (uint32_t)L - 1) >> 2) + 1 : 0;
	uint32_t raw[6];
#pragma unroll
	for (int k = 0; k < 6; ++k)
		raw[k] = k < nw ? base[k] : 0u;
	uint32_t wd[5];
#pragma unroll
	for (int k = 0; k < 5; ++k)
		wd[k] = __builtin_amdgcn_alignbyte(raw[k + 1], raw[k], ash);
	uint64_t lo = 0, hi = 0;
#pragma unroll
	for (int k = 0; k < MAX_PREFIX; ++k) {
		const uint32_t c = (wd[k >> 2] >> (8 * (k & 3))) & 0xFFu;
		const uint64_t code = k < L ? (uint64_t)((c - (uint32_t)'@') & 0x1Fu) : 0ull;
		const int sh = 5 * (MAX_PREFIX - 1 - k);
		if (sh >= 64)
			hi |= code << (sh - 64);
		else {
			lo |= code << sh;
			if (sh > 59)
				hi |= code >> (64 - sh);
		}
	}
	if (set_of)
		hi |= (uint64_t)(set_of[i] & PK_SET_MASK) << PK_SET_SHIFT;
	pk[2 * (int64_t)i] = lo;
	pk[2 * (int64_t)i + 1] = hi | ((uint64_t)L << 56);
	return (uint32_t)variant_count(L, distance);
}

__global__ __launch_bounds__(256) void k_pack(const uint8_t* __restrict__ chars,
                                              const uint32_t* __restrict__ offs,
                                              const uint32_t* __restrict__ seq_map, uint32_t row_base,
                                              const uint32_t* __restrict__ set_of,
                                              int32_t n, int distance, uint64_t* __restrict__ pk,
                                              uint64_t* __restrict__ vcount) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	vcount[i] = pack_one(chars, offs, seq_map, row_base, set_of, i, distance, pk);
}

// ---------------------------------------------------------- generation --

// 90-bit packed variant -> 64-bit key, in two steps:
//   H(v) = sum over the 18 slots i of slot_i(v) * G^i  (mod 2^64, G odd)
//          + set * HASH_SET                              (batched form)
//   key  = xorshift-multiply-xorshift of H
// The polynomial makes deletions cheap: deleting slot q moves every later
// slot down one place, which multiplies its term by G^-1. With
// Pre[j] = sum_{i<j} slot_i(p) G^i, T = Pre[18] and
// A[q] = Pre[q] - G^-1 Pre[q+1] (A[18] = T - G^-1 T), the variant of p with
// slots q1 < q2 < ... < qm deleted (q = 18 stands for "no deletion") has
//   H = A[q1] + G^-1 A[q2] + ... + G^-(m-1) A[qm] + G^-m T,
// so a generator holding a sequence's A table and G^-m T pays two table
// reads and one multiply per variant at m = 2. The key is a function of the
// variant string alone (and the set), however it was reached.
constexpr uint64_t HASH_G = 0x9E3779B97F4A7C15ull;
constexpr uint64_t inv_mod64(uint64_t a) {  // a odd; Newton's iteration, 3 -> 96 correct bits
	uint64_t x = a;
	for (int i = 0; i < 5; ++i)
		x *= 2ull - a * x;
	return x;
}
constexpr uint64_t HASH_GINV = inv_mod64(HASH_G);
static_assert(HASH_G * HASH_GINV == 1ull, "G^-1");
constexpr uint64_t HASH_GINV2 = HASH_GINV * HASH_GINV;
constexpr uint64_t HASH_SET = 0x5851F42D4C957F2Dull;
constexpr uint32_t PK_SLOT_HI_MASK = (1u << 26) - 1u;  // residue bits of the packed high word

__device__ __forceinline__ uint64_t hash_mix(uint64_t x) {
	x ^= x >> 32;
	x *= 0xD6E8FEB86659FD93ull;
	x ^= x >> 29;
	return x;
}

__device__ __forceinline__ uint32_t slot_of(uint64_t lo, uint32_t hi26, int i) {
	const int b = 5 * (MAX_PREFIX - 1 - i);  // slot 17 at bits 0..4, slot 0 at 85..89
	const uint32_t x = b >= 64 ? (hi26 >> (b - 64)) : b > 59 ? (uint32_t)(lo >> b) | (hi26 << (64 - b))
	                                                          : (uint32_t)(lo >> b);
	return x & 31u;
}

constexpr uint64_t pow_mod64(uint64_t b, int e) {
	uint64_t x = 1;
	for (int i = 0; i < e; ++i)
		x *= b;
	return x;
}
constexpr uint64_t HASH_G17 = pow_mod64(HASH_G, MAX_PREFIX - 1);

// Distances 2 and 3 key records by H (their tiles hold few enough sequences
// for the tables); other distances, whose variants are always built
// explicitly, by hash_word: the variant's two words through an odd
// multiplier and the same finaliser. Both are functions of the variant (and
// the set) alone; one pipeline run, and the index and queries of the served
// form, use one distance and therefore one of them.
__device__ __forceinline__ bool poly_keys(int distance) {
	return distance == 2 || distance == 3;
}

template <bool wide>
__device__ __forceinline__ uint64_t hash_word(u128 v) {
	const uint64_t lo = (uint64_t)v;
	const uint64_t hi = wide ? (uint64_t)(v >> 64) : (uint64_t)(uint32_t)(v >> 64);  // 26 bits unless wide
	return hash_mix(lo ^ (hi * HASH_G) ^ HASH_SET);
}

// H of a packed variant (set id in bits 90..119 when `wide`), slot by slot:
// the form for variants built explicitly
template <bool wide>
__device__ __forceinline__ uint64_t hash_variant(u128 v) {
	uint64_t lo = (uint64_t)v;
	const uint64_t hi = (uint64_t)(v >> 64);
	uint32_t hi26 = (uint32_t)hi & PK_SLOT_HI_MASK;
	uint64_t h = 0, g = HASH_G17;
#pragma unroll 1
	for (int i = MAX_PREFIX - 1; i >= 0; --i) {  // slot i sits in the low 5 bits
		h += (lo & 31u) * g;
		lo = (lo >> 5) | ((uint64_t)hi26 << 59);
		hi26 >>= 5;
		g *= HASH_GINV;
	}
	if (wide)
		h += ((hi >> PK_SET_SHIFT) & PK_SET_MASK) * HASH_SET;
	return hash_mix(h);
}

// Deletes slot pos of a packed prefix (slot pos occupies bits
// [5*(17-pos), 5*(17-pos)+5)): the slots before it stay, the slots after it
// move up by one, the last slot becomes empty. hm[pos] keeps slots 0..pos-1,
// lm[pos] slots pos+1..17.
__device__ __forceinline__ u128 del_slot_t(u128 p, int pos, const u128* __restrict__ hm,
                                           const u128* __restrict__ lm) {
	return (p & hm[pos]) | ((p & lm[pos]) << 5);
}

// slot NO_DEL (= MAX_PREFIX) is a no-op: hm keeps every slot, lm none
constexpr int NO_DEL = MAX_PREFIX;

__device__ __forceinline__ void init_slot_masks(u128* hm, u128* lm, int t, int nt) {
	for (int pos = t; pos <= MAX_PREFIX; pos += nt) {
		const u128 one = 1;
		if (pos == NO_DEL) {
			hm[pos] = ~(u128)0;
			lm[pos] = 0;
			continue;
		}
		const int low_bits = 5 * (MAX_PREFIX - 1 - pos);
		lm[pos] = (one << low_bits) - 1;
		hm[pos] = ~((one << (low_bits + 5)) - 1);
	}
}

constexpr int PAIR_LUT = 153;    // C(18, 2) deletion pairs
constexpr int TRIPLE_LUT = 816;  // C(18, 3) deletion triples

// Tables of the generator, fixed at compile time: binomials C(n, k) for
// n, k <= 18, and the colex-ranked deletion pairs and triples. Colex ranks do
// not depend on L: the first C(L, j) j-subsets of {0..17} in colex order are
// exactly the j-subsets of {0..L-1}. A table entry holds the positions of its
// rank, largest first, one byte each.
struct GenTables {
	uint32_t binom[19 * 19];
	uint32_t triple[TRIPLE_LUT];
	uint16_t pair[PAIR_LUT];
	uint16_t pad;
};

constexpr GenTables make_gen_tables() {
	GenTables g{};
	for (int n = 0; n < 19; ++n)
		for (int k = 0; k < 19; ++k) {
			uint32_t c = 0;
			if (k <= n) {
				c = 1;
				for (int j = 1; j <= k; ++j)
					c = c * (uint32_t)(n - j + 1) / (uint32_t)j;
			}
			g.binom[n * 19 + k] = c;
		}
	int q = 0;
	for (int hi = 1; hi < MAX_PREFIX; ++hi)
		for (int lo = 0; lo < hi; ++lo)
			g.pair[q++] = (uint16_t)((hi << 8) | lo);
	q = 0;
	for (int c = 2; c < MAX_PREFIX; ++c)
		for (int b = 1; b < c; ++b)
			for (int a = 0; a < b; ++a)
				g.triple[q++] = (uint32_t)((c << 16) | (b << 8) | a);
	return g;
}

__constant__ GenTables c_gen_tables = make_gen_tables();

// Deletion sets of at most d <= 3 positions ordered by their largest
// position: the sets inside a prefix of length L (largest position < L) are
// then exactly the first sum_{j <= d} C(L, j) entries, so a variant's rank
// indexes the table directly whatever L is. An entry holds three positions,
// ascending in bytes 0, 1, 2, padded with NO_DEL (18) above the real ones
// where the set has fewer: bytes 0..m-1 are the q1 < ... < qm of the
// polynomial hash, and slot deletions applied from byte m-1 down to byte 0
// (largest first, so the smaller positions still point at their slots)
// build the variant.
constexpr int DEL_SETS1 = 1 + 18;
constexpr int DEL_SETS2 = DEL_SETS1 + 153;
constexpr int DEL_SETS3 = DEL_SETS2 + 816;

struct DelTables {
	uint32_t d1[DEL_SETS1];
	uint32_t d2[DEL_SETS2];
	uint32_t d3[DEL_SETS3];
};

constexpr uint32_t del_entry(int a, int b, int c) {
	return (uint32_t)a | ((uint32_t)b << 8) | ((uint32_t)c << 16);
}

constexpr DelTables make_del_tables() {
	DelTables t{};
	t.d1[0] = t.d2[0] = t.d3[0] = del_entry(NO_DEL, NO_DEL, NO_DEL);
	int q1 = 1, q2 = 1, q3 = 1;
	for (int m = 0; m < MAX_PREFIX; ++m) {
		t.d1[q1++] = del_entry(m, NO_DEL, NO_DEL);
		t.d2[q2++] = del_entry(m, NO_DEL, NO_DEL);
		t.d3[q3++] = del_entry(m, NO_DEL, NO_DEL);
		for (int j = 0; j < m; ++j) {
			t.d2[q2++] = del_entry(j, m, NO_DEL);
			t.d3[q3++] = del_entry(j, m, NO_DEL);
			for (int i = 0; i < j; ++i)
				t.d3[q3++] = del_entry(i, j, m);
		}
	}
	return t;
}

__constant__ DelTables c_del_tables = make_del_tables();

// LDS copy of the tables; every thread of the workgroup calls it
__device__ __forceinline__ void load_gen_tables(GenTables* s, int t, int nt) {
	const uint32_t* src = reinterpret_cast<const uint32_t*>(&c_gen_tables);
	uint32_t* dst = reinterpret_cast<uint32_t*>(s);
	for (int i = t; i < (int)(sizeof(GenTables) / 4); i += nt)
		dst[i] = src[i];
}

// the k-th deletion variant of a prefix of length L with at most m deletions.
// Ranks enumerate deletion sets by size (0, 1, 2, ...), each size in colex
// order. The order is internal: records are sorted by key afterwards.
__device__ __forceinline__ u128 unrank_variant(u128 p, int L, int m, uint32_t k,
                                               const uint32_t* __restrict__ binom /*19x19*/,
                                               const u128* __restrict__ hm,
                                               const u128* __restrict__ lm,
                                               const uint16_t* __restrict__ pair_lut,
                                               const uint32_t* __restrict__ triple_lut) {
	if (k == 0)
		return p;
	k -= 1;
	if (k < (uint32_t)L)
		return del_slot_t(p, (int)k, hm, lm);
	k -= (uint32_t)L;
	if (m >= 2) {
		const uint32_t c2 = binom[L * 19 + 2];
		if (k < c2) {
			const uint32_t ab = pair_lut[k];
			return del_slot_t(del_slot_t(p, (int)(ab >> 8), hm, lm), (int)(ab & 0xFF), hm, lm);
		}
		k -= c2;
	}
	if (m >= 3) {
		const uint32_t c3 = binom[L * 19 + 3];
		if (k < c3) {
			const uint32_t abc = triple_lut[k];
			u128 v = del_slot_t(p, (int)(abc >> 16), hm, lm);
			v = del_slot_t(v, (int)((abc >> 8) & 0xFF), hm, lm);
			return del_slot_t(v, (int)(abc & 0xFF), hm, lm);
		}
		k -= c3;
	}
	for (int j = 4; j <= m; ++j) {
		const uint32_t cj = binom[L * 19 + j];
		if (k < cj) {
			u128 v = p;
			uint32_t q = k;
			for (int t = j; t >= 1; --t) {
				int c = t - 1;
				while (c + 1 < L && binom[(c + 1) * 19 + t] <= q)
					++c;
				q -= binom[c * 19 + t];
				v = del_slot_t(v, c, hm, lm);  // positions come out in decreasing order
			}
			return v;
		}
		k -= cj;
	}
	return p;  // unreachable for k < variant_count
}

constexpr int GEN_NT = 256;
constexpr int GEN_IT = 16;  // GEN_TILE == RS_TILE: a generator tile is a radix tile
constexpr int GEN_TILE = GEN_NT * GEN_IT;

enum GenMode { GEN_DIRECT = 0, GEN_HIST = 1, GEN_APPEND = 2, GEN_COARSE = 3 };
// threads of the generator's coarse pass: 512 holds the tile's sequences
// in a workgroup for half as many records per lane, and the tile's set-up
// (sequence lookups and hash tables, a chain of dependent loads) then weighs
// twice as much; measured 4% slower at 1M CDR3s, d=2
constexpr int GEN_COARSE_NT = 256;

struct GenArgs {
	const uint64_t* pk;
	const uint64_t* voff;  // n+1 entries
	int32_t n;
	int32_t distance;
	int32_t idb;
	int32_t part_bits;  // partition = top part_bits of the hash
	uint32_t part_id;
	int32_t pad;
	int64_t V;
	uint64_t* out;
	int64_t cap;
	unsigned long long* part_hist;  // GEN_HIST: 2^part_bits counters
	unsigned long long* append;     // GEN_APPEND cursor
	const uint32_t* set_of;         // non-null in the batched form: the set ids k_pack put in pk
	const uint32_t* tile_seq;       // nullable: k_tile_seq's table (else a binary search)
	uint32_t* digit_hist;           // GEN_DIRECT, nullable: per-tile counts of the
	int64_t ntiles;                 // 8-bit digit at digit_shift, one row of
	int32_t digit_shift;            // RS_BINS per tile as k_radix_hist writes them
	                                // (GEN_COARSE: the coarse digit)
	int32_t sub_bits;               // GEN_COARSE: 2^sub_bits sub-ranges per digit
	uint32_t* cursor;               // GEN_COARSE: slots taken per (digit, sub-range)
	unsigned long long* coarse_ovf; // GEN_COARSE: set when a sub-range outgrew sub_cap
	int64_t sub_cap;                // GEN_COARSE: slots per sub-range
};

// First slot of sub-range j of coarse digit c in the generator's coarse
// pass (digit-major: a digit's sub-ranges are adjacent)
__device__ __forceinline__ int64_t seg_base(uint32_t c, uint32_t j, int sub_bits, int64_t sub_cap) {
	return (int64_t)((c << sub_bits) | j) * sub_cap;
}

__device__ __forceinline__ int64_t upper_bound_u64(const uint64_t* a, int64_t n, uint64_t x) {
	int64_t lo = 0, hi = n;
	while (lo < hi) {
		int64_t mid = (lo + hi) >> 1;
		if (a[mid] <= x)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

// tile_seq[b] = the sequence holding record b * GEN_TILE for the ntiles
// tiles of V records, and tile_seq[ntiles] = the one holding record V - 1:
// each sequence writes the entries of the tile starts inside its records
__global__ __launch_bounds__(256) void k_tile_seq(const uint64_t* __restrict__ voff, int32_t n, uint64_t V,
                                                  uint32_t ntiles, uint32_t* __restrict__ tile_seq) {
	const int i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	const uint64_t o = voff[i], e = voff[i + 1];
	for (uint64_t b = (o + GEN_TILE - 1) / GEN_TILE; b * GEN_TILE < e; ++b)
		tile_seq[b] = (uint32_t)i;
	if (o < e && o <= V - 1 && V - 1 < e)
		tile_seq[ntiles] = (uint32_t)i;
}

// Flattened generation: each workgroup owns GEN_TILE consecutive record
// indexes, so stores are contiguous whatever the sequence lengths
// (generate_combination.cu:212-229 instead runs one thread per sequence and
// writes a strided run per lane).
constexpr int GEN_TBL = 56;               // sequences per tile that get hash tables
constexpr int GEN_PKL = 448;              // sequences per tile whose prefixes are staged in LDS
constexpr int GEN_TBL_W = MAX_PREFIX + 2;  // A[0..18], then G^-d T (+ the set term)

// the hash table of one sequence (the A[q] and G^-d T of hash_variant's
// comment), from its packed prefix
__device__ __forceinline__ void seq_hash_table(const GenArgs& a, int64_t sid, uint64_t* __restrict__ tab) {
	const uint64_t lo = a.pk[2 * sid];
	const uint64_t hw = a.pk[2 * sid + 1];
	const uint32_t hi26 = (uint32_t)hw & PK_SLOT_HI_MASK;
	uint64_t pre = 0, g = 1;
#pragma unroll
	for (int q = 0; q < MAX_PREFIX; ++q) {
		const uint64_t nxt = pre + (uint64_t)slot_of(lo, hi26, q) * g;
		tab[q] = pre - HASH_GINV * nxt;
		pre = nxt;
		g *= HASH_G;
	}
	tab[MAX_PREFIX] = pre - HASH_GINV * pre;
	uint64_t c = pre;
	for (int j = 0; j < a.distance; ++j)
		c *= HASH_GINV;
	if (a.set_of)
		c += ((hw >> PK_SET_SHIFT) & PK_SET_MASK) * HASH_SET;
	tab[MAX_PREFIX + 1] = c;
}

// SMALL: only distances up to 3 (no unranking of larger deletion sets, whose
// registers would then set the whole kernel's budget)
// Positions q < L of a packed prefix whose residue equals the one before it.
// Deleting any position of a run of equal residues leaves the same string, so
// a deletion set D gives the same variant as a set that deletes, in every
// run, a prefix of the run; only those sets need a record:
// D is kept when no q in D has its bit here while q - 1 is not in D.
__device__ __forceinline__ uint32_t run_mask(uint64_t lo, uint64_t hw) {
	const uint32_t hi26 = (uint32_t)hw & PK_SLOT_HI_MASK;
	const int L = (int)(hw >> 56);
	uint32_t m = 0, prev = 0xFFFFFFFFu;
#pragma unroll
	for (int q = 0; q < MAX_PREFIX; ++q) {
		const uint32_t c = slot_of(lo, hi26, q);
		if (q < L && c == prev)
			m |= 1u << q;
		prev = c;
	}
	return m;
}

__device__ __forceinline__ bool dup_deletion_set(uint32_t e, uint32_t runs) {
	const uint32_t dm = ((1u << (e & 0xFFu)) | (1u << ((e >> 8) & 0xFFu)) | (1u << (e >> 16))) &
	                    ((1u << MAX_PREFIX) - 1u);  // NO_DEL (18) drops out
	return (dm & runs & ~(dm << 1)) != 0u;
}

template <int MODE, int NT = GEN_NT, bool SMALL = false>
__global__ __launch_bounds__(NT) void k_gen(GenArgs a) {
	constexpr int IT = GEN_TILE / NT;  // records per lane
	// variant offsets of the tile's sequences minus r0, clamped to
	// [0, GEN_TILE]; the first one (<= 0) is kept whole in s_v0. A tile of
	// at most GEN_TBL sequences uses only the head of the offsets and keeps
	// the sequences' hash tables behind it; one of at most GEN_PKL, their
	// packed prefixes.
	__shared__ union {
		uint16_t vrel[GEN_TILE + 1];
		struct {
			uint16_t head[GEN_TBL + 4];
			uint64_t tab[GEN_TBL][GEN_TBL_W];
		} t;
		struct {
			uint16_t head[GEN_PKL + 8];
			uint64_t pk[GEN_PKL][2];
		} p;
		uint64_t keys[MODE == GEN_COARSE ? GEN_TILE / 2 : 1];  // GEN_COARSE, after the loop: half the sorted tile
	} s_v;
	uint16_t* const s_vrel = s_v.vrel;
	__shared__ int32_t s_v0;
	union GenU {
		GenTables tab;              // d >= 4
		uint32_t del[DEL_SETS3];    // d <= 3
	};
	struct GenSmallU {
		uint32_t del[DEL_SETS3];
	};
	__shared__ std::conditional_t<SMALL, GenSmallU, GenU> s_u;
	__shared__ u128 s_hm[MAX_PREFIX + 1], s_lm[MAX_PREFIX + 1];
	__shared__ int64_t s_lo;
	__shared__ int s_cnt;
	__shared__ uint32_t s_hist[MODE == GEN_HIST ? 4096 : 1];
	__shared__ uint32_t s_dig[MODE == GEN_DIRECT ? RS_BINS : MODE == GEN_COARSE ? RS_BINS + 1 : 1];
	__shared__ uint32_t s_run[MODE == GEN_COARSE ? GEN_TBL : 1];  // run_mask of the tile's sequences (form 0)

	const int t = threadIdx.x;
	const int64_t r0 = (int64_t)blockIdx.x * GEN_TILE;
	if (r0 >= a.V)
		return;
	const int64_t r1 = r0 + GEN_TILE < a.V ? r0 + GEN_TILE : a.V;

	const bool small_d = SMALL || a.distance <= 3;
	if (small_d) {
		const uint32_t* src = a.distance == 1 ? c_del_tables.d1 : a.distance == 2 ? c_del_tables.d2 : c_del_tables.d3;
		const int ns = a.distance == 1 ? DEL_SETS1 : a.distance == 2 ? DEL_SETS2 : DEL_SETS3;
		for (int i = t; i < ns; i += NT)
			s_u.del[i] = src[i];
	} else {
		if constexpr (!SMALL)
			load_gen_tables(&s_u.tab, t, NT);
	}
	init_slot_masks(s_hm, s_lm, t, NT);
	if (MODE == GEN_DIRECT || MODE == GEN_COARSE)
		for (int i = t; i < (MODE == GEN_COARSE ? RS_BINS + 1 : RS_BINS); i += NT)
			s_dig[i] = 0;
	if (MODE == GEN_HIST) {
		const int nb = 1 << a.part_bits;
		for (int i = t; i < nb; i += NT)
			s_hist[i] = 0;
	}
	if (t == 0) {
		// voff[n] == V, so both searches land in [0, n-1]; the clamps only
		// keep a caller's inconsistent V from reading past voff. The table
		// gives the sequence of the next tile's first record as hi, one
		// sequence more than needed where that sequence starts the next
		// tile: its offset clamps to GEN_TILE and no record selects it.
		int64_t lo, hi;
		if (a.tile_seq) {
			lo = a.tile_seq[blockIdx.x];
			hi = a.tile_seq[blockIdx.x + 1];
		} else {
			lo = upper_bound_u64(a.voff, (int64_t)a.n + 1, (uint64_t)r0) - 1;
			hi = upper_bound_u64(a.voff, (int64_t)a.n + 1, (uint64_t)(r1 - 1)) - 1;
		}
		lo = lo < a.n - 1 ? lo : a.n - 1;
		hi = hi < a.n - 1 ? hi : a.n - 1;
		hi = hi > lo ? hi : lo;
		hi = hi - lo + 1 <= GEN_TILE ? hi : lo + GEN_TILE - 1;
		s_lo = lo;
		s_cnt = (int)(hi - lo + 1);
	}
	__syncthreads();
	const int64_t slo = s_lo;
	const int cnt = s_cnt;  // sequences touching this tile, <= GEN_TILE
	// a sequence has at most 2^18 variants: the first offset fits 32 bits
	for (int i = t; i <= cnt; i += NT) {
		const int64_t rel = (int64_t)a.voff[slo + i] - r0;
		s_vrel[i] = (uint16_t)(rel < 0 ? 0 : (rel > GEN_TILE ? GEN_TILE : rel));
		if (i == 0)
			s_v0 = (int32_t)rel;
	}
	// uniform: hash tables for the tile's sequences (form 0), or variants
	// built slot by slot from prefixes staged in LDS (form 1) or read from
	// global memory (form 2, tiles of more than GEN_PKL sequences)
	const bool poly = poly_keys(a.distance);
	const int form = poly && cnt <= GEN_TBL ? 0 : cnt <= GEN_PKL ? 1 : 2;
	if (form == 0)
		for (int i = t; i < cnt; i += NT) {
			seq_hash_table(a, slo + i, s_v.t.tab[i]);
			if constexpr (MODE == GEN_COARSE)
				s_run[i] = run_mask(a.pk[2 * (slo + i)], a.pk[2 * (slo + i) + 1]);
		}
	else if (form == 1)
		for (int i = t; i < cnt; i += NT) {
			s_v.p.pk[i][0] = a.pk[2 * (slo + i)];
			s_v.p.pk[i][1] = a.pk[2 * (slo + i) + 1];
		}
	__syncthreads();
	const int v0 = s_v0;

	const uint64_t idmask = (1ull << a.idb) - 1ull;
	uint64_t kept[MODE == GEN_APPEND || MODE == GEN_COARSE ? IT : 1];
	uint32_t nkept = 0;
	uint32_t dropm = 0;  // GEN_COARSE: bit it set when record it repeats a variant of its sequence
	// The lane's first record: the sequence holding it (largest idx with
	// s_vrel[idx] <= t). Every iteration then finds the next record's
	// sequence before generating its own. Forms 0 and 1 read nothing from
	// global memory inside the loop, so nothing waits on its stores; form 2
	// fetches the next record's packed prefix ahead, so that load is in
	// flight during the unranking.
	auto gen_loop = [&](auto form_c) {
		constexpr int FORM = decltype(form_c)::value;
		constexpr bool TBL = FORM == 0;
		int idx_n = 0, start_n = 0;  // start: the sequence's first record in the tile (v0 for idx 0)
		uint64_t lo_n = 0, hi_n = 0;
		if (r0 + t < r1) {
			int lo = 0, hi = cnt;
			while (lo < hi) {
				int mid = (lo + hi + 1) >> 1;
				if (s_vrel[mid] <= t)
					lo = mid;
				else
					hi = mid - 1;
			}
			idx_n = lo;
			start_n = lo == 0 ? v0 : (int)s_vrel[lo];
			if constexpr (FORM == 2) {
				lo_n = a.pk[2 * (slo + lo)];
				hi_n = a.pk[2 * (slo + lo) + 1];
			}
		}
		// kept rolled: the body holds the whole unranking, sixteen copies of
		// it would not fit the instruction cache
#pragma unroll 1
		for (int it = 0; it < IT; ++it) {
			const int rr = it * NT + t;  // record index within the tile
			const int64_t r = r0 + rr;
			const bool valid = r < r1;
			const int idx = idx_n, start = start_n;
			const uint64_t lo64 = lo_n, hiw = hi_n;
			if (r + NT < r1) {  // records of a lane only move forward
				// the next four sequence starts per look (the offsets do not
				// decrease; indexes past cnt count as "beyond")
				const int rn = rr + NT;
				int j = idx, st = start;
				for (;;) {
					int sv[4];
#pragma unroll
					for (int u = 0; u < 4; ++u) {
						const int q = j + 1 + u;
						sv[u] = q < cnt ? (int)s_vrel[q < GEN_TILE ? q : GEN_TILE] : GEN_TILE + 1;
					}
					int nadv = 0;
#pragma unroll
					for (int u = 0; u < 4; ++u)
						if (sv[u] <= rn) {
							nadv = u + 1;
							st = sv[u];
						}
					j += nadv;
					if (nadv < 4)
						break;
				}
				idx_n = j;
				start_n = st;
				if constexpr (FORM == 2) {
					lo_n = a.pk[2 * (slo + j)];
					hi_n = a.pk[2 * (slo + j) + 1];
				}
			}
			uint64_t rec = 0;
			bool dup = false;
			if (valid) {
				const int64_t sid = slo + idx;
				const uint32_t k = (uint32_t)(rr - start);
				uint64_t hv;
				if constexpr (TBL) {  // the deletion set's table terms
					const uint32_t e = s_u.del[k];
					if constexpr (MODE == GEN_COARSE)
						dup = dup_deletion_set(e, s_run[idx]);
					const uint64_t* tb = s_v.t.tab[idx];
					// every byte of an entry is <= NO_DEL: the three reads go out together
					const uint64_t t1 = tb[e & 0xFFu], t2 = tb[(e >> 8) & 0xFFu], t3 = tb[e >> 16];
					uint64_t h = tb[MAX_PREFIX + 1];
					if (a.distance >= 1)
						h += t1;
					if (a.distance >= 2)
						h += HASH_GINV * t2;
					if (a.distance >= 3)
						h += HASH_GINV2 * t3;
					hv = hash_mix(h);
				} else {
					uint64_t plo = lo64, phi = hiw;
					if constexpr (FORM == 1) {
						plo = s_v.p.pk[idx][0];
						phi = s_v.p.pk[idx][1];
					}
					const u128 p = ((u128)(phi & PK_HI_MASK) << 64) | (u128)plo;
					u128 v;
					if (small_d) {  // uniform: branch-free unranking, largest position first
						const uint32_t e = s_u.del[k];
						v = p;
						if (a.distance >= 3)
							v = del_slot_t(v, (int)(e >> 16), s_hm, s_lm);
						if (a.distance >= 2)
							v = del_slot_t(v, (int)((e >> 8) & 0xFFu), s_hm, s_lm);
						v = del_slot_t(v, (int)(e & 0xFFu), s_hm, s_lm);
					} else {
						if constexpr (!SMALL) {
							const int L = (int)(phi >> 56);
							const int m = a.distance < L ? a.distance : L;
							v = unrank_variant(p, L, m, k, s_u.tab.binom, s_hm, s_lm, s_u.tab.pair, s_u.tab.triple);
						}
					}
					// the set bits make variants of different query sets different keys
					if (poly)
						hv = a.set_of ? hash_variant<true>(v) : hash_variant<false>(v);
					else
						hv = a.set_of ? hash_word<true>(v) : hash_word<false>(v);
				}
				rec = (hv & ~idmask) | ((uint64_t)sid & idmask);
			}
			if (MODE == GEN_DIRECT) {
				if (valid) {
					__builtin_nontemporal_store(rec, &a.out[r]);  // whole lines, read only by the next pass
					if (a.digit_hist)
						atomicAdd(&s_dig[(uint32_t)(rec >> a.digit_shift) & 0xFFu], 1u);
				}
			} else if (MODE == GEN_HIST) {
				if (valid)
					atomicAdd(&s_hist[a.part_bits ? (rec >> (64 - a.part_bits)) : 0], 1u);
			} else if (MODE == GEN_COARSE) {
				// the record stays in a register; lanes past the end count
				// into the spare bin s_dig[RS_BINS]
				// a repeated variant counts into the spare bin too and is not written
				kept[it] = rec;
				dropm |= (uint32_t)dup << it;
				atomicAdd(&s_dig[valid && !dup ? (uint32_t)(rec >> a.digit_shift) & 0xFFu : (uint32_t)RS_BINS], 1u);
			} else {
				const bool mine = valid && (a.part_bits == 0 || (uint32_t)(rec >> (64 - a.part_bits)) == a.part_id);
				if (mine)
					kept[nkept++ & (IT - 1)] = rec;
			}
		}
	};
	if (form == 0)
		gen_loop(std::integral_constant<int, 0>{});
	else if (form == 1)
		gen_loop(std::integral_constant<int, 1>{});
	else
		gen_loop(std::integral_constant<int, 2>{});
	if (MODE == GEN_APPEND) {
		__shared__ uint32_t s_tmp[NT / WAVE + 1];
		__shared__ unsigned long long s_base;
		unsigned long long slot = block_append<NT>(a.append, nkept, s_tmp, &s_base);
		for (uint32_t i = 0; i < nkept; ++i, ++slot)
			if ((int64_t)slot < a.cap)
				a.out[slot] = kept[i];
	}
	if (MODE == GEN_DIRECT && a.digit_hist) {
		__syncthreads();
		for (int i = t; i < RS_BINS; i += NT)
			a.digit_hist[(int64_t)blockIdx.x * RS_BINS + i] = s_dig[i];
	}
	if (MODE == GEN_HIST) {
		__syncthreads();
		const int nb = 1 << a.part_bits;
		for (int i = t; i < nb; i += NT)
			if (s_hist[i])
				atomicAdd(&a.part_hist[i], (unsigned long long)s_hist[i]);
	}
	if constexpr (MODE == GEN_COARSE) {
		// The tile's records go to their coarse digit's sub-range (the tile
		// index mod 2^sub_bits; tiles of one sub-range run on one XCD, so the
		// runs of neighbouring tiles meet in its L2): one slot claim per
		// digit, a position in the tile sorted by digit from one LDS atomic
		// per record (slot order inside a sub-range is free), then the
		// records written out in digit runs. The generator's tables are dead
		// after the loop: the sorted tile takes their place, and the slot
		// claims are in flight while it is placed.
		__syncthreads();
		uint32_t* const s_dpre = reinterpret_cast<uint32_t*>(&s_u);
		uint32_t* const s_cur = s_dpre + RS_BINS;
		uint32_t* const s_gbase = s_cur + RS_BINS + 1;
		uint64_t* const s_keys = s_v.keys;
		__shared__ uint32_t s_tmp[NT / WAVE + 1];
		static_assert(sizeof(s_u) >= (3 * RS_BINS + 1) * sizeof(uint32_t), "offsets fit the generator's tables");
		const uint32_t c = t < RS_BINS ? s_dig[t] : 0u;
		uint32_t n_kept;  // records written: those of the 256 digits
		const uint32_t ex = block_excl_scan<NT>(c, s_tmp, &n_kept);
		const uint32_t sub = blockIdx.x & ((1u << a.sub_bits) - 1u);
		if (t < RS_BINS) {
			s_dpre[t] = ex;
			s_cur[t] = ex;
		}
		if (t == 0)
			s_cur[RS_BINS] = GEN_TILE;  // lanes past the end: positions at or beyond the tile
		__syncthreads();
		uint32_t b = 0;
		if (t < RS_BINS && c)
			b = atomicAdd(&a.cursor[((uint32_t)t << a.sub_bits) | sub], c);
		uint32_t rank[IT];  // the record's position in the tile sorted by digit
#pragma unroll
		for (int r = 0; r < IT; ++r) {
			const bool valid = r0 + r * NT + t < r1 && !((dropm >> r) & 1u);
			rank[r] = atomicAdd(&s_cur[valid ? (uint32_t)(kept[r] >> a.digit_shift) & 0xFFu : (uint32_t)RS_BINS], 1u);
		}
		// the sorted tile in two halves, so that LDS holds half of it (more
		// workgroups per CU for the generation): a record goes to the half
		// its position falls in; the digit runs keep their length
		const int tile_n = (int)n_kept;
		constexpr uint32_t HALF = GEN_TILE / 2;
		for (uint32_t lo_pos = 0; lo_pos < (uint32_t)GEN_TILE; lo_pos += HALF) {
#pragma unroll
			for (int r = 0; r < IT; ++r) {
				const uint32_t pos = rank[r] - lo_pos;
				if (pos < HALF)  // also false below lo_pos (unsigned) and for lanes past the end
					s_keys[pos] = kept[r];
			}
			if (lo_pos == 0 && t < RS_BINS) {
				if (c && (int64_t)b + c > a.sub_cap)
					atomicOr(a.coarse_ovf, 1ull);  // records lost: the host redoes the run
				s_gbase[t] = b;
			}
			__syncthreads();
			const int n_half = tile_n - (int)lo_pos < (int)HALF ? tile_n - (int)lo_pos : (int)HALF;
			for (int i = t; i < n_half; i += NT) {
				const uint64_t key = s_keys[i];
				const uint32_t dg = (uint32_t)(key >> a.digit_shift) & 0xFFu;
				const int64_t slot = (int64_t)s_gbase[dg] + ((int)lo_pos + i - (int)s_dpre[dg]);
				if (slot < a.sub_cap)
					a.out[seg_base(dg, sub, a.sub_bits, a.sub_cap) + slot] = key;
			}
			__syncthreads();  // the half is written before the next one is placed
		}
	}
}

// --------------------------------------------------------------- runs --

// A run is a maximal segment of sorted records with equal (rec >> rs); runs
// of >= 2 records are listed (xtn_inner.cu:34-46 does this with
// DeviceRunLengthEncode over all keys). Each workgroup stages a 4096-record
// tile plus one record of halo on each side in LDS, finds the run heads of
// the tile and appends them with one atomic per list per workgroup.
constexpr int RN_NT = 256;
constexpr int RN_IT = 16;
constexpr int RN_TILE = RN_NT * RN_IT;

__global__ __launch_bounds__(RN_NT) void k_runs(const uint64_t* __restrict__ rec, int64_t n, int rs,
                                                Run* __restrict__ small, int64_t small_cap,
                                                Run* __restrict__ big, int64_t big_cap,
                                                RunCtrs* __restrict__ ctr) {
	__shared__ uint64_t s_k[RN_TILE + 2];
	__shared__ uint32_t s_tmp32[RN_NT / WAVE + 1];
	__shared__ unsigned long long s_tmp64[RN_NT / WAVE + 1];
	__shared__ unsigned long long s_base;
	const int t = threadIdx.x;
	const int64_t base = (int64_t)blockIdx.x * RN_TILE;
	const int64_t rem = n - base;
	const int tile_n = rem < RN_TILE ? (int)rem : RN_TILE;
	const uint64_t NONE = ~0ull;  // shifted keys have at most 64 - rs bits, rs >= 1
	for (int j = t; j < tile_n; j += RN_NT)
		s_k[j + 1] = rec[base + j] >> rs;
	if (t == 0) {
		s_k[0] = base > 0 ? (rec[base - 1] >> rs) : NONE;
		s_k[tile_n + 1] = base + tile_n < n ? (rec[base + tile_n] >> rs) : NONE;
	}
	__syncthreads();

	uint32_t lens[RN_IT];
	uint32_t ns = 0, nb = 0;
	unsigned long long ps = 0, pb = 0;
#pragma unroll
	for (int it = 0; it < RN_IT; ++it) {
		const int j = it * RN_NT + t;
		uint32_t len = 0;
		if (j < tile_n) {
			const uint64_t cur = s_k[j + 1];
			if (cur != s_k[j] && s_k[j + 2] == cur) {
				int jj = j + 2;
				while (jj < tile_n && s_k[jj + 1] == cur)
					++jj;
				int64_t g = base + jj;
				if (jj >= tile_n)  // the run may continue past the tile
					while (g < n && (rec[g] >> rs) == cur)
						++g;
				len = (uint32_t)(g - (base + j));
			}
		}
		lens[it] = len;
		const unsigned long long np = (unsigned long long)len * (len ? len - 1 : 0) / 2;
		if (len && len <= (uint32_t)SMALL_RUN_MAX) {
			++ns;
			ps += np;
		} else if (len) {
			++nb;
			pb += np;
		}
	}
	unsigned long long sb = block_append<RN_NT>(&ctr->n_small, ns, s_tmp32, &s_base);
	unsigned long long bb = block_append<RN_NT>(&ctr->n_big, nb, s_tmp32, &s_base);
	block_add<RN_NT>(&ctr->pairs_small, ps, s_tmp64);
	block_add<RN_NT>(&ctr->pairs_big, pb, s_tmp64);
#pragma unroll
	for (int it = 0; it < RN_IT; ++it) {
		const uint32_t len = lens[it];
		if (!len)
			continue;
		const Run r{(uint32_t)(base + it * RN_NT + t), len};
		if (len <= (uint32_t)SMALL_RUN_MAX) {
			if ((int64_t)sb < small_cap)
				small[sb] = r;
			++sb;
		} else {
			if ((int64_t)bb < big_cap)
				big[bb] = r;
			++bb;
		}
	}
}

// the candidate pair of two records holding the same stored hash bits and two
// different ids, whose smaller id lies in [lo, hi) (a lower-bound round,
// xtn_inner.cu:245-267 / kernel.cu:501-521, restricts pairs the same way)
template <bool RANGED = true>
__device__ __forceinline__ bool pair_of(uint64_t ka, uint64_t kb, int idb, uint32_t lo, uint32_t hi,
                                        uint64_t* out) {
	if ((ka ^ kb) >> idb)
		return false;  // same sorted bits, different variant hash
	const uint64_t mask = (1ull << idb) - 1ull;
	const uint64_t ia = ka & mask, ib = kb & mask;
	if (ia == ib)
		return false;  // a sequence paired with itself (kernel.cu:377-380 drops these later)
	const uint64_t mi = ia < ib ? ia : ib;
	if (RANGED && (mi < lo || mi >= hi))
		return false;
	*out = ia < ib ? ((ia << idb) | ib) : ((ib << idb) | ia);
	return true;
}

// every unordered record pair of a small run (kernel.cu:157-191 does this for
// every group, one thread per group)
__global__ __launch_bounds__(256) void k_pairs_small(const uint64_t* __restrict__ rec,
                                                     const Run* __restrict__ runs, int64_t nruns,
                                                     int idb, uint32_t lo, uint32_t hi,
                                                     uint64_t* __restrict__ pairs,
                                                     int64_t cap, RunCtrs* __restrict__ ctr) {
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	Run run{0, 0};
	if (t < nruns)
		run = runs[t];
	uint32_t count = 0;
	for (uint32_t a = 0; a < run.len; ++a) {
		const uint64_t ka = rec[run.start + a];
		for (uint32_t b = a + 1; b < run.len; ++b) {
			uint64_t p;
			count += pair_of(ka, rec[run.start + b], idb, lo, hi, &p) ? 1u : 0u;
		}
	}
	__shared__ uint32_t s_tmp[256 / WAVE + 1];
	__shared__ unsigned long long s_base;
	unsigned long long slot = block_append<256>(&ctr->n_cand, count, s_tmp, &s_base);
	for (uint32_t a = 0; a < run.len; ++a) {
		const uint64_t ka = rec[run.start + a];
		for (uint32_t b = a + 1; b < run.len; ++b) {
			uint64_t p;
			if (pair_of(ka, rec[run.start + b], idb, lo, hi, &p)) {
				if ((int64_t)slot < cap)
					pairs[slot] = p;
				++slot;
			}
		}
	}
}

// ---------------------------------------------------------- buckets --
//
// After two global LSD passes the records of a key-space chunk are ordered by
// the B = 16 hash bits below the chunk bits: every bucket (equal B bits) is a
// contiguous segment of ~2k records. One workgroup per bucket finishes the
// grouping in LDS: the segment is loaded, sorted by the next hash bits with
// stable LDS radix passes, runs of equal sort bits are found and every record
// pair of a run that shares the full stored hash and two different ids is
// emitted (generate_pairs, kernel.cu:157-191, over all groups at once).
// Segments longer than the LDS tile go to the global sort/run/pair kernels.

// bstart[b] = first record of bucket b (records sorted by bucket bits)
__global__ __launch_bounds__(256) void k_bucket_bounds(const uint64_t* __restrict__ rec, int64_t n,
                                                       int bshift, uint32_t bmask, int nb,
                                                       uint32_t* __restrict__ bstart) {
	const int b = blockIdx.x * 256 + threadIdx.x;
	if (b > nb)
		return;
	int64_t lo = 0, hi = n;
	while (lo < hi) {
		const int64_t mid = (lo + hi) >> 1;
		if ((uint32_t)((rec[mid] >> bshift) & bmask) < (uint32_t)b)
			lo = mid + 1;
		else
			hi = mid;
	}
	bstart[b] = (uint32_t)lo;
}

// ------------------------------------------- region (second) pass --
//
// Buckets of the B > 8 bucket bits are laid out as fixed regions of
// `stride` slots: bucket b owns [b * stride, b * stride + count[b]). The
// records come from an exact (stable) pass over the top 8 bucket bits, so a
// coarse bucket c (those 8 bits) is a contiguous range; tiles are cut inside
// coarse buckets, every record of a tile shares c, and a tile ranks its
// records by the remaining bucket bits in LDS and claims each run with one
// atomic on the bucket's cursor (16 records per atomic on average). Slots past
// the stride are not written: the cursor still counts them, which marks the
// bucket as overflowing (more than the LDS bucket kernel holds), and the
// chunk is then redone on the exact two-pass sort.

// ctl[0..256] = first record of each coarse bucket (cstart, cstart[256] = n),
// ctl[257..513] = first tile of each coarse bucket (tstart[256] = tiles)
constexpr int CTL_WORDS = 2 * (RS_BINS + 1);

__global__ __launch_bounds__(RS_BINS) void k_coarse_tiles(const uint32_t* __restrict__ dtot,
                                                          uint32_t* __restrict__ ctl) {
	__shared__ uint32_t s_tmp[RS_BINS / WAVE + 1];
	const int c = threadIdx.x;
	const uint32_t cnt = dtot[c];
	uint32_t n;
	const uint32_t cs = block_excl_scan<RS_BINS>(cnt, s_tmp, &n);
	const uint32_t tiles = (cnt + RS_TILE - 1) / RS_TILE;
	uint32_t nt;
	const uint32_t ts = block_excl_scan<RS_BINS>(tiles, s_tmp, &nt);
	ctl[c] = cs;
	ctl[RS_BINS + 1 + c] = ts;
	if (c == 0) {
		ctl[RS_BINS] = n;
		ctl[2 * RS_BINS + 1] = nt;
	}
}

// The coarse pass fused into the generator (k_gen<GEN_COARSE>) leaves the
// records in sub-ranges: sub-range j of coarse digit c (cursor
// e = (c << sub_bits) | j) holds min(cursor[e], sub_cap) records from slot
// seg_base(c, j). The passes after it read a coarse digit's sub-ranges as one
// run of records (position p of digit c lies in the sub-range j with
// segpre[c][j] <= p < segpre[c][j + 1]), cut into tiles as the dense layout
// would be. k_seg_tiles writes, in `seg`: the first tile of every coarse
// digit (SEG_TSTART, 257 words, as k_coarse_tiles does), the digit's first
// record in a dense layout (SEG_CSTART, 257 words) and the prefix counts of
// its sub-ranges (SEG_PRE, S + 1 words per digit at a stride of S + 4, so
// that every digit's row starts on 16 bytes). It also clears the region
// layout's bucket cursors, one launch fewer.
constexpr int SEG_NT = 1024;
constexpr int SEG_MAX_SUB = 64;  // sub-ranges per coarse digit, at most
constexpr int SEG_TSTART = 0;
constexpr int SEG_CSTART = RS_BINS + 1;
constexpr int SEG_PRE = 2 * (RS_BINS + 1) + 2;  // 16-byte aligned
constexpr int seg_stride(int nsub) { return nsub + 4; }
constexpr int seg_words(int nsub) { return SEG_PRE + RS_BINS * seg_stride(nsub); }

template <int NSUB>
__global__ __launch_bounds__(SEG_NT) void k_seg_tiles(const uint32_t* __restrict__ cursor, int64_t sub_cap,
                                                      uint32_t* __restrict__ seg, uint32_t* __restrict__ bzero,
                                                      int64_t nzero, unsigned long long* __restrict__ n_rec) {
	constexpr int PER = RS_BINS * NSUB / SEG_NT;  // a thread's segments, all of one digit
	constexpr int GPT = NSUB / PER;               // threads per digit
	static_assert(PER % 4 == 0 && NSUB % PER == 0 && NSUB <= SEG_MAX_SUB, "whole digits, u32x4 rows");
	__shared__ uint32_t s_part[SEG_NT];
	__shared__ uint32_t s_tot[RS_BINS];
	__shared__ uint32_t s_tmp[SEG_NT / WAVE + 1];
	const int t = threadIdx.x;
	const int c = t / GPT, q = t % GPT;
	uint32_t cnt[PER], sum = 0;
	const uint4* const cur4 = reinterpret_cast<const uint4*>(cursor + (int64_t)t * PER);
#pragma unroll
	for (int k = 0; k < PER / 4; ++k) {
		const uint4 v = cur4[k];
		cnt[4 * k] = v.x;
		cnt[4 * k + 1] = v.y;
		cnt[4 * k + 2] = v.z;
		cnt[4 * k + 3] = v.w;
	}
#pragma unroll
	for (int k = 0; k < PER; ++k) {
		cnt[k] = (int64_t)cnt[k] < sub_cap ? cnt[k] : (uint32_t)sub_cap;
		sum += cnt[k];
	}
	uint4* const z4 = reinterpret_cast<uint4*>(bzero);
	for (int64_t i = t; i < nzero / 4; i += SEG_NT)
		z4[i] = make_uint4(0u, 0u, 0u, 0u);
	s_part[t] = sum;
	__syncthreads();
	uint32_t off = 0, tot = 0;
#pragma unroll
	for (int u = 0; u < GPT; ++u) {
		const uint32_t v = s_part[c * GPT + u];
		off += u < q ? v : 0u;
		tot += v;
	}
	uint32_t* const pre = seg + SEG_PRE + (int64_t)c * seg_stride(NSUB);
	uint4* const pre4 = reinterpret_cast<uint4*>(pre + q * PER);
#pragma unroll
	for (int k = 0; k < PER / 4; ++k) {
		uint4 v;
		v.x = off;
		off += cnt[4 * k];
		v.y = off;
		off += cnt[4 * k + 1];
		v.z = off;
		off += cnt[4 * k + 2];
		v.w = off;
		off += cnt[4 * k + 3];
		pre4[k] = v;
	}
	if (q == 0) {
		pre[NSUB] = tot;
		s_tot[c] = tot;
	}
	__syncthreads();
	const uint32_t ct = t < RS_BINS ? s_tot[t] : 0u;
	uint32_t all_c, all_t;
	const uint32_t cs = block_excl_scan<SEG_NT>(ct, s_tmp, &all_c);
	const uint32_t ts = block_excl_scan<SEG_NT>((ct + RS_TILE - 1) / RS_TILE, s_tmp, &all_t);
	if (t == 0 && n_rec)
		*n_rec = all_c;  // the records in the layout
	if (t < RS_BINS) {
		seg[SEG_TSTART + t] = ts;
		seg[SEG_CSTART + t] = cs;
	}
	if (t == 0) {
		seg[SEG_TSTART + RS_BINS] = all_t;
		seg[SEG_CSTART + RS_BINS] = all_c;
	}
}

// Position p (< sp[nsub]) of coarse digit c in the sub-range layout, given
// the digit's prefix counts in LDS (sp[0..nsub]): the sub-range is the last
// j with sp[j] <= p (empty ones are passed over). j carries the sub-range of
// the lane's previous position (positions only grow; -1 before the first,
// which is found by bisection), and is updated.
__device__ __forceinline__ int64_t seg_addr(const uint32_t* sp, int nsub, uint32_t c, int sub_bits,
                                            int64_t sub_cap, uint32_t p, int& j) {
	if (j < 0) {
		int lo = 0, hi = nsub - 1;
		while (lo < hi) {
			const int mid = (lo + hi + 1) >> 1;
			if (sp[mid] <= p)
				lo = mid;
			else
				hi = mid - 1;
		}
		j = lo;
	}
	while (j + 1 < nsub && sp[j + 1] <= p)
		++j;
	return seg_base(c, (uint32_t)j, sub_bits, sub_cap) + (p - sp[j]);
}

// The tile of workgroup i in the coarse-digit tiling (ts[0..256] in LDS):
// its digit and its first position inside the digit, or false past the end
__device__ __forceinline__ bool coarse_tile(const uint32_t* ts, int64_t i, int64_t grid, uint32_t* c_out,
                                            uint32_t* k_out) {
	const uint32_t ntiles = ts[RS_BINS];
	const uint32_t tile = (uint32_t)xcd_tile(i, grid);
	if (tile >= ntiles)
		return false;
	int lo = 0, hi = RS_BINS - 1;  // the coarse digit: largest c with ts[c] <= tile
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (ts[mid] <= tile)
			lo = mid;
		else
			hi = mid - 1;
	}
	*c_out = (uint32_t)lo;
	*k_out = tile - ts[lo];
	return true;
}

// The digit holding coarse tile `tile` (ts: first tile of each digit, in
// LDS) and the tile's index inside the digit
__device__ __forceinline__ uint32_t tile_digit(const uint32_t* ts, uint32_t tile, uint32_t* k_out) {
	int lo = 0, hi = RS_BINS - 1;  // largest c with ts[c] <= tile
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (ts[mid] <= tile)
			lo = mid;
		else
			hi = mid - 1;
	}
	*k_out = tile - ts[lo];
	return (uint32_t)lo;
}

// A wave's keys from the sub-range layout: positions [p0, p0 + IT * WAVE)
// of digit c (round r lane l at p0 + r * WAVE + l), the first w_valid of
// them present. They mostly lie in one or two sub-ranges: the first one and
// the next two boundaries, found once for the wave, place every position
// without further LDS reads; with more boundaries inside the range each lane
// walks them. The loads carry the non-temporal hint (read once).
template <int IT>
__device__ __forceinline__ void seg_wave_keys(const uint32_t* sp, int nsub, uint32_t c, int sub_bits,
                                              int64_t sub_cap, const uint64_t* __restrict__ in, int64_t p0,
                                              int w_valid, uint64_t (&k)[IT]) {
	const int lane = lane_id();
	int j0 = -1;
	(void)seg_addr(sp, nsub, c, sub_bits, sub_cap, (uint32_t)p0, j0);
	const uint32_t b0 = sp[j0];
	const uint32_t b1 = j0 + 1 < nsub ? sp[j0 + 1] : 0xFFFFFFFFu;
	const uint32_t b2 = j0 + 2 < nsub ? sp[j0 + 2] : 0xFFFFFFFFu;
	const uint32_t b3 = j0 + 3 < nsub ? sp[j0 + 3] : 0xFFFFFFFFu;
	const bool few = (int64_t)b3 >= p0 + (int64_t)IT * WAVE;
	int j = j0;
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		const bool valid = r * WAVE + lane < w_valid;
		const uint32_t p = (uint32_t)(p0 + (int64_t)r * WAVE + lane);
		int64_t at;
		if (few) {
			const int jj = j0 + (p >= b1 ? 1 : 0) + (p >= b2 ? 1 : 0);
			const uint32_t bb = p >= b2 ? b2 : p >= b1 ? b1 : b0;
			at = seg_base(c, (uint32_t)jj, sub_bits, sub_cap) + (p - bb);
		} else {
			at = valid ? seg_addr(sp, nsub, c, sub_bits, sub_cap, p, j) : 0;
		}
		k[r] = valid ? __builtin_nontemporal_load(&in[at]) : 0ull;
	}
}

// The sub-range layout's coarse tiles [t0, t0 + ntiles) (the digits of one
// part; ntiles < 0: every tile from t0 on) copied densely, digits in order,
// the range's first digit at out[0]: the exact regrouping after a bucket
// outgrew its region, or a chunk sorted in place, starts from it. The grid
// covers the tile count.
__global__ __launch_bounds__(256) void k_seg_compact(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                     const uint32_t* __restrict__ seg, int sub_bits,
                                                     int64_t sub_cap, int64_t t0, int64_t ntiles) {
	__shared__ uint32_t s_ts[RS_BINS + 1];
	__shared__ uint32_t s_sp[SEG_MAX_SUB + 1];
	const int nsub = 1 << sub_bits;
	for (int i = threadIdx.x; i <= RS_BINS; i += 256)
		s_ts[i] = seg[SEG_TSTART + i];
	__syncthreads();
	const int64_t row = xcd_tile(blockIdx.x, gridDim.x);
	if (row >= (ntiles < 0 ? (int64_t)s_ts[RS_BINS] - t0 : ntiles))  // uniform
		return;
	uint32_t k;
	const uint32_t c = tile_digit(s_ts, (uint32_t)(t0 + row), &k);
	uint32_t k0;
	const uint32_t c0 = tile_digit(s_ts, (uint32_t)t0, &k0);
	for (int i = threadIdx.x; i <= nsub; i += 256)
		s_sp[i] = seg[SEG_PRE + (int64_t)c * seg_stride(nsub) + i];
	__syncthreads();
	const uint32_t total = s_sp[nsub];
	const uint32_t p0 = k * RS_TILE;
	const uint32_t n = total - p0 < (uint32_t)RS_TILE ? total - p0 : (uint32_t)RS_TILE;
	const int64_t dst = (int64_t)seg[SEG_CSTART + c] - (int64_t)seg[SEG_CSTART + c0] + p0;
	int j = -1;
	for (uint32_t i = threadIdx.x; i < n; i += 256)
		out[dst + i] = in[seg_addr(s_sp, nsub, c, sub_bits, sub_cap, p0 + i, j)];
}

// The first grouping pass of a chunk whose records sit in the sub-range
// layout: the chunk is coarse tiles [t0, t0 + ntiles) (the digits of its
// part). k_seg_hist writes tile t0 + b's row b of counts of the digit at
// `shift`; k_seg_scatter places the records densely by that digit at the
// rows' scanned bases.
__global__ __launch_bounds__(RS_NT) void k_seg_hist(const uint64_t* __restrict__ in, const uint32_t* __restrict__ seg,
                                                    int64_t t0, int sub_bits, int64_t sub_cap, int shift,
                                                    uint32_t* __restrict__ hist) {
	constexpr int SEG = RS_IT * WAVE;
	__shared__ uint32_t s_ts[RS_BINS + 1];
	__shared__ uint32_t s_sp[SEG_MAX_SUB + 1];
	__shared__ uint32_t s_h[RS_WAVES][RS_BINS];
	const int t = threadIdx.x, w = t / WAVE;
	const int nsub = 1 << sub_bits;
	for (int i = t; i <= RS_BINS; i += RS_NT)
		s_ts[i] = seg[SEG_TSTART + i];
	for (int i = t; i < RS_WAVES * RS_BINS; i += RS_NT)
		(&s_h[0][0])[i] = 0;
	__syncthreads();
	uint32_t kt;
	const uint32_t c = tile_digit(s_ts, (uint32_t)(t0 + blockIdx.x), &kt);
	for (int i = t; i <= nsub; i += RS_NT)
		s_sp[i] = seg[SEG_PRE + (int64_t)c * seg_stride(nsub) + i];
	__syncthreads();
	const int64_t base = (int64_t)kt * RS_TILE;
	const int64_t total = s_sp[nsub];
	const int tile_n = (int)(total - base < RS_TILE ? total - base : RS_TILE);  // >= 1
	const int wrem = tile_n - w * SEG;
	const int w_valid = wrem <= 0 ? 0 : (wrem < SEG ? wrem : SEG);
	uint64_t k[RS_IT];
	seg_wave_keys<RS_IT>(s_sp, nsub, c, sub_bits, sub_cap, in, base + (int64_t)w * SEG, w_valid, k);
#pragma unroll
	for (int r = 0; r < RS_IT; ++r)
		if (r * WAVE + lane_id() < w_valid)
			atomicAdd(&s_h[w][(uint32_t)(k[r] >> shift) & 0xFFu], 1u);
	__syncthreads();
	uint32_t cnt = 0;
#pragma unroll
	for (int ww = 0; ww < RS_WAVES; ++ww)
		cnt += s_h[ww][t];
	hist[(int64_t)blockIdx.x * RS_BINS + t] = cnt;
}

__global__ __launch_bounds__(RS_NT) void k_seg_scatter(const uint64_t* __restrict__ in,
                                                       const uint32_t* __restrict__ seg, int64_t t0,
                                                       int64_t ntiles, int sub_bits, int64_t sub_cap, int shift,
                                                       const uint32_t* __restrict__ offs,
                                                       uint64_t* __restrict__ out, int64_t n) {
	constexpr int SEG = RS_IT * WAVE;
	__shared__ uint64_t s_keys[RS_TILE];
	__shared__ uint32_t s_cnt[RS_WAVES][RS_BINS];
	__shared__ uint32_t s_dpre[RS_BINS + 1];
	__shared__ uint32_t s_gbase[RS_BINS];
	__shared__ uint32_t s_scan[RS_WAVES + 1];
	__shared__ uint32_t s_ts[RS_BINS + 1];
	__shared__ uint32_t s_sp[SEG_MAX_SUB + 1];
	const int t = threadIdx.x, w = t / WAVE;
	const int nsub = 1 << sub_bits;
	for (int i = t; i <= RS_BINS; i += RS_NT)
		s_ts[i] = seg[SEG_TSTART + i];
	__syncthreads();
	const int64_t row = xcd_tile(blockIdx.x, ntiles);
	uint32_t kt;
	const uint32_t c = tile_digit(s_ts, (uint32_t)(t0 + row), &kt);
	for (int i = t; i <= nsub; i += RS_NT)
		s_sp[i] = seg[SEG_PRE + (int64_t)c * seg_stride(nsub) + i];
	s_gbase[t] = offs[row * RS_BINS + t];
	__syncthreads();
	const int64_t base = (int64_t)kt * RS_TILE;
	const int64_t total = s_sp[nsub];
	const int tile_n = (int)(total - base < RS_TILE ? total - base : RS_TILE);  // >= 1
	const int wrem = tile_n - w * SEG;
	const int w_valid = wrem <= 0 ? 0 : (wrem < SEG ? wrem : SEG);
	uint64_t k[RS_IT];
	seg_wave_keys<RS_IT>(s_sp, nsub, c, sub_bits, sub_cap, in, base + (int64_t)w * SEG, w_valid, k);
	uint32_t lp[RS_IT];  // order inside a digit run is free: the unstable rank
	rank_tile_u<RS_NT, RS_IT>(k, w_valid, shift, lp, s_cnt, s_dpre, s_scan);
#pragma unroll
	for (int r = 0; r < RS_IT; ++r)
		if (lp[r] != 0xFFFFFFFFu)
			s_keys[lp[r]] = k[r];
	__syncthreads();
	for (int i = t; i < tile_n; i += RS_NT) {
		const uint64_t key = s_keys[i];
		const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
		const uint32_t g = s_gbase[d] + ((uint32_t)i - s_dpre[d]);
		if ((int64_t)g < n)
			out[g] = key;
	}
}

struct RegionArgs {
	const uint64_t* in;     // sorted by the coarse digit, or the segment layout
	uint64_t* out;          // the regions
	const uint32_t* ctl;    // k_coarse_tiles
	uint32_t* count;        // 2^B cursors, zeroed
	int64_t stride;         // slots per region
	int32_t shift;          // lowest bucket bit: the bucket is bits [shift, shift + B)
	int32_t fine_bits;      // B - 8
	// sub-range layout (k_seg_tiles): ctl is then seg
	int64_t sub_cap;
	int32_t sub_bits;
	int32_t pad;
};

template <int NT = RS_NT, bool SEGS = false>
__global__ __launch_bounds__(NT) void k_region_scatter(RegionArgs a) {
	constexpr int IT = RS_TILE / NT;   // keys per thread
	constexpr int SEG = IT * WAVE;     // keys per wave
	__shared__ uint64_t s_keys[RS_TILE];
	__shared__ uint32_t s_cnt[NT / WAVE][RS_BINS];
	__shared__ uint32_t s_dpre[RS_BINS + 1];
	__shared__ uint32_t s_gbase[RS_BINS];
	__shared__ uint32_t s_scan[NT / WAVE + 1];
	__shared__ uint32_t s_ts[RS_BINS + 1];
	__shared__ uint32_t s_sp[SEGS ? SEG_MAX_SUB + 1 : 1];

	const int t = threadIdx.x;
	const int lane = lane_id();
	const int w = t / WAVE;
	const uint32_t* const ts = a.ctl + (SEGS ? SEG_TSTART : RS_BINS + 1);  // first tile of each digit
	for (int i = t; i <= RS_BINS; i += NT)
		s_ts[i] = ts[i];
	__syncthreads();
	uint32_t c, kt;
	if (!coarse_tile(s_ts, blockIdx.x, gridDim.x, &c, &kt))  // uniform
		return;
	int64_t base;  // the tile's first position: in the input, or inside the digit (SEGS)
	int tile_n;
	const int nsub = SEGS ? 1 << a.sub_bits : 1;
	if constexpr (SEGS) {
		for (int i = t; i <= nsub; i += NT)
			s_sp[i] = a.ctl[SEG_PRE + (int64_t)c * seg_stride(nsub) + i];
		__syncthreads();
		base = (int64_t)kt * RS_TILE;
		const int64_t total = s_sp[nsub];
		tile_n = (int)(total - base < RS_TILE ? total - base : RS_TILE);  // >= 1
	} else {
		const int64_t cbeg = a.ctl[c], cend = a.ctl[c + 1];
		base = cbeg + (int64_t)kt * RS_TILE;
		tile_n = (int)(cend - base < RS_TILE ? cend - base : RS_TILE);  // >= 1
	}

	const int64_t seg = base + (int64_t)w * SEG;
	const int wrem = tile_n - w * SEG;
	const int w_valid = wrem <= 0 ? 0 : (wrem < SEG ? wrem : SEG);
	uint64_t k[IT];
	if constexpr (SEGS) {
		seg_wave_keys<IT>(s_sp, nsub, c, a.sub_bits, a.sub_cap, a.in, seg, w_valid, k);
	} else {
#pragma unroll
		for (int r = 0; r < IT; ++r) {
			const bool valid = r * WAVE + lane < w_valid;
			k[r] = valid ? __builtin_nontemporal_load(&a.in[seg + (int64_t)r * WAVE + lane]) : 0ull;  // read once
		}
	}
	uint32_t lp[IT];  // slot order inside a bucket is free: the unstable rank
	rank_tile_u<NT, IT>(k, w_valid, a.shift, lp, s_cnt, s_dpre, s_scan);
#pragma unroll
	for (int r = 0; r < IT; ++r)
		if (lp[r] != 0xFFFFFFFFu)
			s_keys[lp[r]] = k[r];
	if (t == 0)
		s_dpre[RS_BINS] = (uint32_t)tile_n;
	__syncthreads();
	// one cursor claim per digit run of the tile
	const uint32_t fmask = (1u << a.fine_bits) - 1u;
	if (t < RS_BINS) {
		const uint32_t cnt = s_dpre[t + 1] - s_dpre[t];
		s_gbase[t] = cnt ? atomicAdd(&a.count[(c << a.fine_bits) | ((uint32_t)t & fmask)], cnt) : 0u;
	}
	__syncthreads();
	for (int i = t; i < tile_n; i += NT) {
		const uint64_t key = s_keys[i];
		const uint32_t d = (uint32_t)(key >> a.shift) & 0xFFu;
		const uint32_t slot = s_gbase[d] + ((uint32_t)i - s_dpre[d]);
		if ((int64_t)slot < a.stride)
			a.out[(int64_t)((c << a.fine_bits) | (d & fmask)) * a.stride + slot] = key;
	}
}

struct BucketArgs {
	const uint64_t* rec;
	const uint32_t* bstart;  // sorted layout: bucket b is [bstart[b], bstart[b + 1])
	const uint32_t* bcount;  // region layout (non-null): [b * bstride, + bcount[b])
	int64_t bstride;
	int32_t nb;
	int32_t bin_shift;  // the bucket is binned by bits [bin_shift, bin_shift + bin_bits)
	int32_t bin_bits;   // <= BK_BIN_BITS
	int32_t idb;
	int32_t cap;        // largest bucket handled in LDS (<= BK_CAP)
	int32_t ovf_cap;
	uint32_t id_lo, id_hi;  // only pairs whose smaller id lies in [id_lo, id_hi)
	uint64_t* pairs;
	int64_t pcap;
	RunCtrs* ctr;
	uint32_t* ovf;      // overflowing bucket ids
	unsigned long long* n_ovf;
};

// records of bucket b: [*s, *e); e - s > cap marks an overflowing bucket
__device__ __forceinline__ void bucket_range(const BucketArgs& a, int b, uint32_t* s, uint32_t* e) {
	if (a.bcount) {
		*s = (uint32_t)((int64_t)b * a.bstride);
		*e = *s + a.bcount[b];
	} else {
		*s = a.bstart[b];
		*e = a.bstart[b + 1];
	}
}

constexpr int BK_BIN_BITS = 12;
constexpr int BK_BINS = 1 << BK_BIN_BITS;
constexpr int BK_NT = 512;
constexpr int BK_CAP = 3072;             // largest bucket grouped in LDS (default)
constexpr int BK_CAP_MAX = 4096;         // largest instantiated: 62 KB of LDS, two workgroups per CU
constexpr uint16_t BK_NONE = 0xFFFF;

constexpr int BK_STAGE = 376;  // pairs staged in LDS between flushes
constexpr int BK_WAVES = BK_NT / WAVE;
constexpr int BK_GROUP = 3;     // record slots per walk pass: a wave lists at most 3 * 64 walkers

// Buckets of <= BK_CAP records, each handled by one workgroup in LDS: every
// record is pushed on the list of its bin (12 more hash bits, one atomicExch
// each), then each record walks the records pushed before it in its bin and
// emits the pairs that share the full stored hash and hold two different
// ids. Every unordered pair of a bin is visited exactly once; the emission
// order is irrelevant because the pairs are sorted afterwards.
// The workgroups are persistent: workgroup w takes buckets w, w + grid, ...,
// loading the records of its next bucket into registers while it works on
// the current one, and collects pairs in LDS (LDS atomics), appending them to
// the global buffer with one atomic per flush. Every loop is bounded by the
// bucket count; nothing waits on another workgroup. 49 KB of LDS: three
// workgroups (24 waves) per CU (at 53.8 KB only two fit on MI355X).
// records of bucket b into registers (nothing for an overflowing bucket)
template <int PER>
__device__ __forceinline__ void bucket_load(const BucketArgs& a, int b, uint64_t (&k)[PER],
                                            uint32_t* s, uint32_t* e) {
	*s = 0;
	*e = 0;
	if (b < a.nb)
		bucket_range(a, b, s, e);
	const int c = (int)(*e - *s);
#pragma unroll
	for (int i = 0; i < PER; ++i) {
		const int j = i * BK_NT + (int)threadIdx.x;
		k[i] = (c <= a.cap && j < c) ? __builtin_nontemporal_load(&a.rec[*s + j]) : 0ull;  // read once
	}
}

// RANGED: the round covers only part of the smaller ids (the id test is
// compiled out of the common single-round case). CAP: the largest bucket
// the workgroup holds (2048, 3072 or 4096 records; a.cap <= CAP).
template <bool RANGED, int CAP>
__global__ __launch_bounds__(BK_NT) void k_bucket(BucketArgs a) {
	constexpr int BK_PER = CAP / BK_NT;  // records per thread
	static_assert(CAP % (2 * BK_NT) == 0, "two links per register");
	__shared__ uint64_t s_keys[CAP];
	__shared__ uint16_t s_next[CAP];
	__shared__ uint32_t s_head[BK_BINS];  // (iteration << 16) | record of the bin's last push
	__shared__ uint64_t s_stage[BK_STAGE];
	__shared__ uint16_t s_walk[BK_WAVES][BK_GROUP * WAVE];  // each wave's records that have a predecessor
	__shared__ uint32_t s_nstage;
	__shared__ unsigned long long s_base;
	const int t = threadIdx.x;
	const int lane = t & (WAVE - 1);
	const int w = t / WAVE;
	const uint32_t bmask = (1u << a.bin_bits) - 1u;
	if (t == 0)
		s_nstage = 0;
	// heads carry the iteration that pushed them, so a head left by an
	// earlier bucket reads as an empty bin and the bins need no clearing;
	// a workgroup handles far fewer than 0xFFFF buckets
	for (int i = t; i < BK_BINS; i += BK_NT)
		s_head[i] = 0xFFFF0000u;
	const int G = (int)gridDim.x;

	// bins, walks and flushes bucket b (records in k, cnt of them)
	auto process = [&](int b, uint32_t iter, const uint64_t (&k)[BK_PER], int cnt) {
		if (cnt > a.cap) {
			if (t == 0) {
				const unsigned long long slot = atomicAdd(a.n_ovf, 1ull);
				if ((int64_t)slot < a.ovf_cap)
					a.ovf[slot] = (uint32_t)b;
			}
		} else if (cnt >= 2) {
			const uint32_t tag = iter << 16;
			// each record's link also stays in a register, two per register,
			// so the walk starts without reading it back
			uint32_t pv[BK_PER / 2];
#pragma unroll
			for (int i = 0; i < BK_PER; ++i) {
				const int j = i * BK_NT + t;
				uint32_t link = BK_NONE;
				if (j < cnt) {
					s_keys[j] = k[i];
					const uint32_t prev = atomicExch(&s_head[(uint32_t)(k[i] >> a.bin_shift) & bmask], tag | (uint32_t)j);
					link = (prev & 0xFFFF0000u) == tag ? (prev & 0xFFFFu) : BK_NONE;
					s_next[j] = (uint16_t)link;
				}
				if (i & 1)
					pv[i / 2] |= link << 16;
				else
					pv[i / 2] = link;
			}
			__syncthreads();
			// Only records with a predecessor in their bin walk (about one in
			// five): each wave lists them in its own LDS row and walks the
			// list densely, a lane per walker, instead of running one walk per
			// record slot with most lanes idle. A group of BK_GROUP slots joins
			// the list when it fits, else the list is walked first. The row is
			// the wave's own, so no barrier guards it.
			auto walk = [&](uint32_t nw) {
				for (uint32_t q0 = 0; q0 < nw; q0 += WAVE) {  // uniform
					uint32_t n = BK_NONE;
					uint64_t kj = 0;
					if (q0 + lane < nw) {
						const uint32_t j = s_walk[w][q0 + lane];
						kj = s_keys[j];
						n = s_next[j];
					}
					for (; n != BK_NONE;) {
						const uint64_t kn = s_keys[n];  // key and link read together
						const uint32_t nn = s_next[n];
						uint64_t p;
						const bool ok = pair_of<RANGED>(kj, kn, a.idb, a.id_lo, a.id_hi, &p);
						n = nn;
						if (!ok)
							continue;
						const uint32_t slot = atomicAdd(&s_nstage, 1u);
						const bool spill = slot >= (uint32_t)BK_STAGE;
						if (!spill) {
							s_stage[slot] = p;
						} else {
							// stage full (a large group): straight to the global
							// buffer, one atomic for the wave's spilling lanes
							const uint64_t m = __ballot(spill);
							const int leader = __ffsll((unsigned long long)m) - 1;
							unsigned long long gs = 0;
							if (lane == leader)
								gs = atomicAdd(&a.ctr->n_cand, (unsigned long long)__popcll(m));
							gs = __shfl(gs, leader, WAVE) + (unsigned long long)__popcll(m & lanemask_lt());
							if ((int64_t)gs < a.pcap)
								a.pairs[gs] = p;
						}
					}
				}
			};
			uint32_t nw = 0;  // walkers listed, uniform over the wave
#pragma unroll
			for (int g = 0; g < BK_PER; g += BK_GROUP) {
				if (g * BK_NT >= cnt)
					break;  // uniform
				uint64_t m[BK_GROUP];
				uint32_t c = 0;
#pragma unroll
				for (int u = 0; u < BK_GROUP; ++u) {
					const int i = g + u;
					const int j = i * BK_NT + t;
					m[u] = __ballot(i < BK_PER && j < cnt && ((pv[i / 2] >> (16 * (i & 1))) & 0xFFFFu) != BK_NONE);
					c += (uint32_t)__popcll(m[u]);
				}
				if (nw + c > (uint32_t)(BK_GROUP * WAVE)) {
					walk(nw);
					nw = 0;
				}
#pragma unroll
				for (int u = 0; u < BK_GROUP; ++u) {
					if ((m[u] >> lane) & 1u)
						s_walk[w][nw + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[u] >> 32),
						                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m[u], 0u))] =
						    (uint16_t)((g + u) * BK_NT + t);
					nw += (uint32_t)__popcll(m[u]);
				}
			}
			walk(nw);
			__syncthreads();
		}
		const uint32_t ns = s_nstage;
		if (ns > (uint32_t)BK_STAGE / 2 || (b + G >= a.nb && ns > 0)) {  // uniform
			const uint32_t nf = ns < (uint32_t)BK_STAGE ? ns : (uint32_t)BK_STAGE;
			__syncthreads();
			if (t == 0)
				s_base = atomicAdd(&a.ctr->n_cand, (unsigned long long)nf);
			__syncthreads();
			const unsigned long long base = s_base;
			for (uint32_t i = t; i < nf; i += BK_NT)
				if ((int64_t)(base + i) < a.pcap)
					a.pairs[base + i] = s_stage[i];
			__syncthreads();
			if (t == 0)
				s_nstage = 0;
			__syncthreads();
		}
	};

	// two register buffers in turn: the next bucket's records are in flight
	// while the current one is binned and walked (no copies between them)
	uint64_t kA[BK_PER], kB[BK_PER];
	uint32_t sA, eA, sB, eB;
	bucket_load(a, blockIdx.x, kA, &sA, &eA);
	__syncthreads();
	uint32_t iter = 0;
	for (int b = blockIdx.x; b < a.nb; b += 2 * G) {  // uniform over the workgroup
		bucket_load(a, b + G, kB, &sB, &eB);
		process(b, iter++, kA, (int)(eA - sA));
		if (b + G >= a.nb)
			break;
		bucket_load(a, b + 2 * G, kA, &sA, &eA);
		process(b + G, iter++, kB, (int)(eB - sB));
	}
}

__global__ __launch_bounds__(256) void k_big_counts(const Run* __restrict__ runs, int64_t nruns,
                                                    uint64_t* __restrict__ counts) {
	const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (t < nruns) {
		const uint64_t l = runs[t].len;
		counts[t] = l * (l - 1) / 2;
	}
}

// bounded fan-out: the pairs of all long runs form one flat index space that
// the whole grid strides over, so a single huge group is spread over every CU
__global__ __launch_bounds__(256) void k_pairs_big(const uint64_t* __restrict__ rec,
                                                   const Run* __restrict__ runs,
                                                   const uint64_t* __restrict__ pre /* nruns+1 */,
                                                   int64_t nruns, int idb, uint32_t lo, uint32_t hi,
                                                   uint64_t* __restrict__ pairs, int64_t cap,
                                                   RunCtrs* __restrict__ ctr) {
	const uint64_t total = pre[nruns];
	const uint64_t stride = (uint64_t)gridDim.x * 256;
	for (uint64_t base = (uint64_t)blockIdx.x * 256; base < total; base += stride) {
		const uint64_t q = base + threadIdx.x;
		bool ok = false;
		uint64_t p = 0;
		if (q < total) {
			// the run with pre[run] <= q < pre[run+1] (its own names: lo and
			// hi are the id range the pairs are kept for)
			int64_t r_lo = 0, r_hi = nruns - 1;
			while (r_lo < r_hi) {
				int64_t mid = (r_lo + r_hi + 1) >> 1;
				if (pre[mid] <= q)
					r_lo = mid;
				else
					r_hi = mid - 1;
			}
			const uint64_t u = q - pre[r_lo];
			uint64_t b = (uint64_t)((1.0 + sqrt(1.0 + 8.0 * (double)u)) * 0.5);
			while (b * (b - 1) / 2 > u)
				--b;
			while ((b + 1) * b / 2 <= u)
				++b;
			const uint64_t a = u - b * (b - 1) / 2;
			const uint32_t st = runs[r_lo].start;
			ok = pair_of(rec[st + a], rec[st + b], idb, lo, hi, &p);
		}
		const unsigned long long slot = wave_append(&ctr->n_cand, ok ? 1u : 0u);
		if (ok && (int64_t)slot < cap)
			pairs[slot] = p;
	}
}

// ------------------------------------------------------- verification --

// Levenshtein restricted to the diagonal band |x - y| <= D: exact when the
// distance is <= D, otherwise some value > D (kernel.cu:235-280 computes the
// full Wagner-Fischer table). The band lives in registers, one pair per lane.
template <int D>
__device__ __forceinline__ int lev_band(const uint8_t* __restrict__ a, int la,
                                        const uint8_t* __restrict__ b, int lb) {
	constexpr int W = 2 * D + 1;
	constexpr int INF = D + 1;
	if (la - lb > D || lb - la > D)
		return INF;
	int row[W];
#pragma unroll
	for (int k = 0; k < W; ++k) {
		const int x = k - D;  // row y = 0: cell (0, x)
		row[k] = (x >= 0 && x <= lb) ? (x < INF ? x : INF) : INF;
	}
	for (int y = 1; y <= la; ++y) {
		const uint8_t ca = a[y - 1];
		int left = INF;
#pragma unroll
		for (int k = 0; k < W; ++k) {
			const int x = y - D + k;
			int val = INF;
			if (x >= 0 && x <= lb) {
				if (x == 0) {
					val = y < INF ? y : INF;
				} else {
					const int diag = row[k] + (ca != b[x - 1] ? 1 : 0);
					const int up = (k + 1 < W ? row[k + 1] : INF) + 1;
					val = diag < up ? diag : up;
					val = left + 1 < val ? left + 1 : val;
					val = val < INF ? val : INF;
				}
			}
			row[k] = val;  // row[k] of the previous line is no longer needed
			left = val;
		}
	}
	return row[lb - la + D];
}

// the same band with a run-time half width D <= 127 (distances beyond the
// register templates); the band is 2D+1 cells of private memory
__device__ int lev_band_rt(const uint8_t* __restrict__ a, int la, const uint8_t* __restrict__ b,
                           int lb, int D) {
	const int W = 2 * D + 1;
	const int INF = D + 1;
	if (la - lb > D || lb - la > D)
		return INF;
	int row[255];
	for (int k = 0; k < W; ++k) {
		const int x = k - D;
		row[k] = (x >= 0 && x <= lb) ? (x < INF ? x : INF) : INF;
	}
	for (int y = 1; y <= la; ++y) {
		const uint8_t ca = a[y - 1];
		int left = INF;
		for (int k = 0; k < W; ++k) {
			const int x = y - D + k;
			int val = INF;
			if (x >= 0 && x <= lb) {
				if (x == 0) {
					val = y < INF ? y : INF;
				} else {
					const int diag = row[k] + (ca != b[x - 1] ? 1 : 0);
					const int up = (k + 1 < W ? row[k + 1] : INF) + 1;
					val = diag < up ? diag : up;
					val = left + 1 < val ? left + 1 : val;
					val = val < INF ? val : INF;
				}
			}
			row[k] = val;
			left = val;
		}
	}
	return row[lb - la + D];
}

__device__ __forceinline__ int hamming(const uint8_t* __restrict__ a, int la,
                                       const uint8_t* __restrict__ b, int lb) {
	if (la != lb)
		return HAMMING_SENTINEL;  // kernel.cu:291-292
	int d = 0;
	for (int i = 0; i < la; ++i)
		d += a[i] != b[i];
	return d;
}

__device__ __forceinline__ int seq_distance(const uint8_t* a, int la, const uint8_t* b, int lb,
                                            int distance, int measure) {
	if (measure == 1)
		return hamming(a, la, b, lb);
	switch (distance) {
	case 1: return lev_band<1>(a, la, b, lb);
	case 2: return lev_band<2>(a, la, b, lb);
	case 3: return lev_band<3>(a, la, b, lb);
	case 4: return lev_band<4>(a, la, b, lb);
	case 5: return lev_band<5>(a, la, b, lb);
	case 6: return lev_band<6>(a, la, b, lb);
	case 7: return lev_band<7>(a, la, b, lb);
	default: return lev_band<8>(a, la, b, lb);  // callers take the wide band above 8
	}
}

// ---- register path for strings of at most REG_LEN residues (CDR3s) ----
// Both strings are loaded once into registers, four residues per word; the
// band DP is unrolled over the rows so every residue access is a constant
// shift. Same results as lev_band / hamming on the same strings.
constexpr int REG_LEN = 32;
constexpr int REG_WORDS = REG_LEN / 4;

// the residues p[0..len) (len <= REG_LEN) into w, byte j of the string in
// byte j % 4 of w[j / 4]. Loads only the aligned words holding a byte of the
// string: an aligned word never straddles a page, so no load leaves the
// allocation. Bytes of w past len are unspecified.
__device__ __forceinline__ void load_residues(const uint8_t* p, int len, uint32_t (&w)[REG_WORDS]) {
	const uintptr_t addr = (uintptr_t)p;
	const uint32_t* base = (const uint32_t*)(addr & ~(uintptr_t)3);
	const uint32_t sh = (uint32_t)(addr & 3);
	const int nw = len > 0 ? (int)((sh + (uint32_t)len - 1) >> 2) + 1 : 0;  // <= REG_WORDS + 1
	uint32_t raw[REG_WORDS + 1];
#pragma unroll
	for (int k = 0; k <= REG_WORDS; ++k)
		raw[k] = k < nw ? base[k] : 0u;
#pragma unroll
	for (int k = 0; k < REG_WORDS; ++k)
		w[k] = __builtin_amdgcn_alignbyte(raw[k + 1], raw[k], sh);
}

__device__ __forceinline__ uint32_t residue(const uint32_t (&w)[REG_WORDS], int j) {
	return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;  // j is a compile-time constant at every call
}

template <int D>
__device__ __forceinline__ int lev_band_reg(const uint32_t (&A)[REG_WORDS], int la,
                                            const uint32_t (&B)[REG_WORDS], int lb) {
	constexpr int W = 2 * D + 1;
	constexpr int INF = D + 1;
	if (la - lb > D || lb - la > D)
		return INF;
	int row[W];
#pragma unroll
	for (int k = 0; k < W; ++k) {
		const int x = k - D;
		row[k] = (x >= 0 && x <= lb) ? (x < INF ? x : INF) : INF;
	}
#pragma unroll
	for (int y = 1; y <= REG_LEN; ++y) {
		if (y <= la) {  // rows past la: skipped by the whole wave once no lane needs them
			const uint32_t ca = residue(A, y - 1);
			int left = INF;
#pragma unroll
			for (int k = 0; k < W; ++k) {
				const int x = y - D + k;
				int val = INF;
				if (x >= 0 && x <= REG_LEN && x <= lb) {
					if (x == 0) {
						val = y < INF ? y : INF;
					} else {
						const int diag = row[k] + (ca != residue(B, x - 1) ? 1 : 0);
						const int up = (k + 1 < W ? row[k + 1] : INF) + 1;
						val = diag < up ? diag : up;
						val = left + 1 < val ? left + 1 : val;
						val = val < INF ? val : INF;
					}
				}
				row[k] = val;
				left = val;
			}
		}
	}
	int r = INF;
#pragma unroll
	for (int k = 0; k < W; ++k)
		if (k == lb - la + D)
			r = row[k];
	return r;
}

__device__ __forceinline__ int hamming_reg(const uint32_t (&A)[REG_WORDS], int la,
                                           const uint32_t (&B)[REG_WORDS], int lb) {
	if (la != lb)
		return HAMMING_SENTINEL;
	int d = 0;
#pragma unroll
	for (int j = 0; j < REG_LEN; ++j)
		d += (j < la && residue(A, j) != residue(B, j)) ? 1 : 0;
	return d;
}

// seq_distance with the register path when both strings fit it. WIDE: the
// Levenshtein band of a distance above 8, whose private band array
// (lev_band_rt) only the kernels instantiated for such distances contain, so
// that the others need no scratch memory (the runtime reserves scratch for
// every resident wave of a kernel that uses it).
template <bool WIDE>
__device__ __forceinline__ int pair_distance(const uint8_t* a, int la, const uint8_t* b, int lb,
                                             int distance, int measure) {
	if (WIDE)
		return lev_band_rt(a, la, b, lb, distance);
	if (la > REG_LEN || lb > REG_LEN)
		return seq_distance(a, la, b, lb, distance, measure);
	uint32_t A[REG_WORDS], B[REG_WORDS];
	load_residues(a, la, A);
	load_residues(b, lb, B);
	if (measure == 1)
		return hamming_reg(A, la, B, lb);
	switch (distance) {
	case 1: return lev_band_reg<1>(A, la, B, lb);
	case 2: return lev_band_reg<2>(A, la, B, lb);
	case 3: return lev_band_reg<3>(A, la, B, lb);
	case 4: return lev_band_reg<4>(A, la, B, lb);
	case 5: return lev_band_reg<5>(A, la, B, lb);
	case 6: return lev_band_reg<6>(A, la, B, lb);
	case 7: return lev_band_reg<7>(A, la, B, lb);
	default: return lev_band_reg<8>(A, la, B, lb);
	}
}

// pair_distance with the distance fixed at compile time, so that a kernel
// holds only the band it uses: D in 1..8 is that distance's band (Hamming
// needs none), D = 0 the wide band of a Levenshtein distance above 8
template <int D>
__device__ __forceinline__ int pair_distance_d(const uint8_t* a, int la, const uint8_t* b, int lb,
                                               int distance, int measure) {
	if constexpr (D == 0) {
		return lev_band_rt(a, la, b, lb, distance);
	} else {
		if (la > REG_LEN || lb > REG_LEN)
			return measure == 1 ? hamming(a, la, b, lb) : lev_band<D>(a, la, b, lb);
		uint32_t A[REG_WORDS], B[REG_WORDS];
		load_residues(a, la, A);
		load_residues(b, lb, B);
		if (measure == 1)
			return hamming_reg(A, la, B, lb);
		return lev_band_reg<D>(A, la, B, lb);
	}
}

// ------------------------------------------- unique + verify + select --

// Sorted candidate pairs -> unique pairs -> distances -> output, every step
// order-preserving, so the output comes out sorted. Tiles of SEL_TILE items,
// SEL_IT consecutive items per thread for the order-preserving writes; the
// counts of a step land in tile_cnt, whose in-place exclusive scan gives the
// tile bases and (at tile_cnt[ntiles]) the total the next step runs over.
// Verification runs over the compacted unique pairs only, so its lanes are
// dense (a candidate pair repeats ~10 times in the sorted array).
// Replaces the sort + unique (xtn_inner.cu:48-63) and the distance stage
// with its output compaction (kernel.cu:235-300 cal_distance, xtn.cu:460-490).
constexpr int SEL_NT = 256;
constexpr int SEL_IT = 8;
constexpr int SEL_TILE = SEL_NT * SEL_IT;

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* s_tmp) {
	uint32_t total;
	block_excl_scan<SEL_NT>(v, s_tmp, &total);
	return total;
}

__global__ __launch_bounds__(SEL_NT) void k_uniq_count(const uint64_t* __restrict__ keys, int64_t n,
                                                       uint32_t* __restrict__ tile_cnt) {
	__shared__ uint32_t s_tmp[SEL_NT / WAVE + 1];
	const int64_t base = (int64_t)blockIdx.x * SEL_TILE + (int64_t)threadIdx.x * SEL_IT;
	uint32_t c = 0;
#pragma unroll
	for (int i = 0; i < SEL_IT; ++i) {
		const int64_t q = base + i;
		if (q < n && (q == 0 || keys[q - 1] != keys[q]))
			++c;
	}
	const uint32_t tot = block_sum(c, s_tmp);
	if (threadIdx.x == 0)
		tile_cnt[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SEL_NT) void k_uniq_write(const uint64_t* __restrict__ keys, int64_t n,
                                                       const uint32_t* __restrict__ tile_base,
                                                       uint64_t* __restrict__ out) {
	__shared__ uint32_t s_tmp[SEL_NT / WAVE + 1];
	const int64_t base = (int64_t)blockIdx.x * SEL_TILE + (int64_t)threadIdx.x * SEL_IT;
	uint32_t flags = 0, c = 0;
#pragma unroll
	for (int i = 0; i < SEL_IT; ++i) {
		const int64_t q = base + i;
		if (q < n && (q == 0 || keys[q - 1] != keys[q])) {
			flags |= 1u << i;
			++c;
		}
	}
	uint64_t pos = tile_base[blockIdx.x] + block_excl_scan<SEL_NT>(c, s_tmp, (uint32_t*)nullptr);
#pragma unroll
	for (int i = 0; i < SEL_IT; ++i)
		if (flags & (1u << i))
			out[pos++] = keys[base + i];
}

// distances of the unique pairs, one per thread (a pair's gathers and DP are
// latency-bound: the more pairs in flight the better); *n_dev of them (the
// scan total of the previous step), the grid covers the largest possible
// count. Tiles of SEL_NT pairs.
template <bool WIDE>
__global__ __launch_bounds__(SEL_NT) void k_verify_count(const uint64_t* __restrict__ upairs,
                                                         const uint32_t* __restrict__ n_dev, int idb,
                                                         const uint8_t* __restrict__ chars,
                                                         const uint32_t* __restrict__ offs,
                                                         const uint32_t* __restrict__ seq_map,
                                                         uint32_t row_base,
                                                         const uint32_t* __restrict__ set_of,
                                                         int distance, int measure,
                                                         int8_t* __restrict__ dist,
                                                         uint32_t* __restrict__ tile_cnt) {
	__shared__ uint32_t s_tmp[SEL_NT / WAVE + 1];
	const int64_t n = *n_dev;
	const int64_t q = (int64_t)blockIdx.x * SEL_NT + threadIdx.x;
	if ((int64_t)blockIdx.x * SEL_NT >= n) {  // uniform
		if (threadIdx.x == 0)
			tile_cnt[blockIdx.x] = 0;
		return;
	}
	uint32_t c = 0;
	{
		const uint64_t mask = (1ull << idb) - 1ull;
		if (q < n) {
			const uint64_t p = upairs[q];
			const uint32_t a = (uint32_t)(p >> idb), b = (uint32_t)(p & mask);
			int8_t dv = -1;
			if (!set_of || set_of[a] == set_of[b]) {
				const uint32_t oa = seq_map ? seq_map[a] : row_base + a;
				const uint32_t ob = seq_map ? seq_map[b] : row_base + b;
				const uint32_t sa = offs[oa], sb = offs[ob];
				const int d = pair_distance<WIDE>(chars + sa, (int)(offs[oa + 1] - sa), chars + sb,
				                            (int)(offs[ob + 1] - sb), distance, measure);
				if (d <= distance) {
					dv = (int8_t)d;
					++c;
				}
			}
			dist[q] = dv;
		}
	}
	const uint32_t tot = block_sum(c, s_tmp);
	if (threadIdx.x == 0)
		tile_cnt[blockIdx.x] = tot;
}

// the pairs within the distance, in order, as (i, j) and dist; tiles of
// SEL_NT pairs as in k_verify_count
__global__ __launch_bounds__(SEL_NT) void k_keep_write(const uint64_t* __restrict__ upairs,
                                                       const uint32_t* __restrict__ n_dev, int idb,
                                                       const int8_t* __restrict__ dist,
                                                       const uint32_t* __restrict__ tile_base,
                                                       int2* __restrict__ out_pairs,
                                                       int8_t* __restrict__ out_dist) {
	__shared__ uint32_t s_tmp[SEL_NT / WAVE + 1];
	const int64_t n = *n_dev;
	if ((int64_t)blockIdx.x * SEL_NT >= n)  // uniform
		return;
	const int64_t q = (int64_t)blockIdx.x * SEL_NT + threadIdx.x;
	const int8_t dv = q < n ? dist[q] : (int8_t)-1;
	const uint32_t keep = dv >= 0 ? 1u : 0u;
	const uint64_t pos = tile_base[blockIdx.x] + block_excl_scan<SEL_NT>(keep, s_tmp, (uint32_t*)nullptr);
	if (keep) {
		const uint64_t p = upairs[q];
		const uint64_t mask = (1ull << idb) - 1ull;
		out_pairs[pos] = make_int2((int)(p >> idb), (int)(p & mask));
		out_dist[pos] = dv;
	}
}

// set_of[r] = the query set holding row r0 + r: the largest k with
// offs[k] <= r0 + r (empty sets share their offset with the next set)
__global__ __launch_bounds__(256) void k_set_of(const uint32_t* __restrict__ offs, int32_t n_sets,
                                                uint32_t r0, uint32_t n, uint32_t* __restrict__ set_of) {
	const uint32_t r = blockIdx.x * 256 + threadIdx.x;
	if (r >= n)
		return;
	const uint32_t row = r0 + r;
	int lo = 0, hi = n_sets - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		if (offs[mid] <= row)
			lo = mid;
		else
			hi = mid - 1;
	}
	set_of[r] = (uint32_t)lo;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
	const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (i < n)
		p[i] = v;
}

}  // namespace xtn
