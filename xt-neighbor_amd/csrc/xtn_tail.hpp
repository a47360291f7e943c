// xtn_tail.hpp -- the candidate tail of the NN pipeline: sort the candidate
// pairs, keep the first of each run of equal pairs, verify those and write the
// pairs within the distance in ascending (i, j) order.
//
// The reference does this as sort_int2 + unique (xtn_inner.cu:138-142,
// cub.cu:116-124, 156-163), cal_distance (kernel.cu:359-392) and the flag
// compaction of postprocessing (xtn_inner.cu:147-165), each a separate device
// call with host round trips in between. Here it is P + 2 launches with no
// host involvement, all sized by a capacity and reading the actual candidate
// count from device memory:
//
//   k_tail_hist    one read of the candidates: the global count of every
//                  8-bit digit of every LSD pass (P x 256 counters); also
//                  clears the look-back state of the launches below
//   k_onesweep     one per digit (P launches): workgroup b ranks tile b
//                  stably by the digit in LDS, publishes its digit counts,
//                  adds the counts of the tiles before it by decoupled
//                  look-back (status words carrying a flag and a count) and
//                  scatters the tile in digit runs
//   k_select       workgroup b takes tile b of the sorted candidates, marks
//                  the first of each run of equal pairs, compacts those in
//                  LDS, verifies them with dense lanes, and writes the kept
//                  pairs at the offset the look-back over the tiles before it
//                  gives (unique and kept counts travel together in 64-bit
//                  status words; the last tile writes the totals)
//
// A workgroup only ever waits on tiles with a smaller index. Each XCD
// dispatches its share of a grid in index order, so the lowest unfinished
// tile never waits on a workgroup that is not running, and every wait ends.
// Tiles are indexed by blockIdx rather than by a ticket taken from one
// counter: on this part atomics on one address serialise at tens of
// nanoseconds each, which for thousands of workgroups costs more than the
// pass itself. Each wait is also bounded by a spin limit that, when hit,
// raises an error flag the host reads back (the run is then redone on the
// reduce-then-scan path) and lets the kernel finish.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xtn_kernels.hpp"

namespace xtn {

constexpr int TAIL_MAX_PASSES = 8;  // 2 * idb <= 62 bits
constexpr uint32_t LB_AGG = 1u << 30;  // status: the tile's own count
constexpr uint32_t LB_INC = 2u << 30;  // status: count of this and every earlier tile
constexpr uint32_t LB_VAL = (1u << 30) - 1u;
constexpr uint32_t LB_SPIN_LIMIT = 1u << 22;  // s_sleep(1) rounds a wait may take (well under a second)

// control block of the tail, cleared before every run; the scalars come
// first so that one small copy reads them back
struct TailCtl {
	uint32_t error;                            // a look-back spin hit its limit
	uint32_t overflow;                         // a select tile's group range outgrew its LDS
	unsigned long long n_unique;               // unique candidates verified
	unsigned long long n_out;                  // pairs within the distance
	uint32_t pad1[12];
	// k_tail_hist: digit counts, one copy per group of workgroups (blockIdx
	// mod TH_COPIES) so that fewer atomics meet on one address; a pass sums
	// the copies
	uint32_t ghist[8][TAIL_MAX_PASSES][RS_BINS];
};
constexpr int TH_COPIES = 8;

__device__ __forceinline__ void lb_store(uint32_t* p, uint32_t v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t lb_load(const uint32_t* p) {
	return __hip_atomic_load(const_cast<uint32_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back: the sum of the counts of the tiles before `tile`,
// from status words state[p * stride] (p < tile). A word is 0 until its tile
// publishes LB_AGG | count, then LB_INC | inclusive sum; tile -1 counts as an
// inclusive 0. Each round reads LB_WIN predecessors at once (the loads are
// independent, and on this part every such load crosses to the coherence
// point of the eight XCDs), sums aggregates from the nearest down to the
// first inclusive, and waits where a word is still 0 (its tile has a smaller
// index, so it runs).
constexpr int LB_WIN = 8;

__device__ __forceinline__ uint32_t lookback(const uint32_t* state, int64_t stride, int64_t tile,
                                             uint32_t spin_limit, uint32_t* error) {
	uint32_t excl = 0, spins = 0;
	int64_t p = tile - 1;  // the nearest predecessor not yet counted
	for (;;) {
		uint32_t v[LB_WIN];
#pragma unroll
		for (int w = 0; w < LB_WIN; ++w)
			v[w] = p - w >= 0 ? lb_load(state + (p - w) * stride) : LB_INC;
		int used = 0;
		bool fin = false;
#pragma unroll
		for (int w = 0; w < LB_WIN; ++w) {
			if (fin || used < w)
				continue;  // stopped at an inclusive or at an unpublished word
			const uint32_t f = v[w] & ~LB_VAL;
			if (f == 0)
				continue;
			excl += v[w] & LB_VAL;
			used = w + 1;
			fin = f == LB_INC;
		}
		if (fin)
			return excl;
		p -= used;
		if (used == 0) {
			if (++spins > spin_limit) {
				atomicOr(error, 1u);
				return excl;
			}
			__builtin_amdgcn_s_sleep(1);
		}
	}
}

// Two counters per tile in one 64-bit status word (the select's unique and
// kept counts, 31 bits each, below a 2-bit flag), looked back by a whole
// wave: lane l reads the word of tile p - l, so 64 predecessors per round.
// Every lane of the wave calls it; every lane gets both sums.
constexpr uint64_t LB2_AGG = 1ull << 62;
constexpr uint64_t LB2_INC = 2ull << 62;
constexpr uint64_t LB2_VAL = (1ull << 62) - 1ull;
constexpr int LB2_SHIFT = 31;  // unique count above the kept count

__device__ __forceinline__ void lb_store64(uint64_t* p, uint64_t v) {
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lb_load64(const uint64_t* p) {
	return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lookback_wave(const uint64_t* state, int64_t tile, uint32_t spin_limit,
                                                  uint32_t* error) {
	const int lane = lane_id();
	uint64_t excl = 0;
	uint32_t spins = 0;
	int64_t p = tile - 1;
	for (;;) {
		const int64_t q = p - lane;
		const uint64_t v = q >= 0 ? lb_load64(state + q) : LB2_INC;
		const uint64_t f = v & ~LB2_VAL;
		const uint64_t inc = __ballot(f == LB2_INC);
		const uint64_t stop = inc | __ballot(f == 0);
		const int first = stop ? __ffsll((unsigned long long)stop) - 1 : WAVE;  // nearest inclusive or unpublished
		const bool take = lane < first || (lane == first && f == LB2_INC);
		// the two 31-bit fields add without carrying into each other: every
		// sum is at most the candidate count (< 2^30)
		const uint64_t part = wave_incl_scan(take ? (v & LB2_VAL) : 0ull);
		excl += __shfl(part, WAVE - 1, WAVE);
		if (first < WAVE && ((inc >> first) & 1ull))
			return excl;
		p -= first;
		if (first == 0) {
			if (++spins > spin_limit) {
				if (lane == 0)
					atomicOr(error, 1u);
				return excl;
			}
			__builtin_amdgcn_s_sleep(1);
		}
	}
}

__device__ __forceinline__ int64_t tail_count(const unsigned long long* n_dev, int64_t cap) {
	const int64_t n = (int64_t)*n_dev;
	return n < cap ? n : cap;
}

constexpr int TH_NT = 512;
constexpr int TH_IT = 8;

// Global digit counts of the first min(*n_dev, cap) keys for `passes` 8-bit
// digits (digit p = bits [lbits + 8p, lbits + 8p + 8)), grid-stride over
// tiles; also zeroes
// zero[0, nzero) (the look-back state of the passes and the select).
__global__ __launch_bounds__(TH_NT) void k_tail_hist(const uint64_t* __restrict__ keys,
                                                     const unsigned long long* __restrict__ n_dev,
                                                     int64_t cap, int lbits, int passes, TailCtl* __restrict__ ctl,
                                                     uint32_t* __restrict__ zero, int64_t nzero) {
	__shared__ uint32_t s_h[TAIL_MAX_PASSES][RS_BINS];
	const int t = threadIdx.x;
	const int64_t gsize = (int64_t)gridDim.x * TH_NT;
	for (int64_t i = (int64_t)blockIdx.x * TH_NT + t; i < nzero; i += gsize)
		zero[i] = 0u;
	for (int i = t; i < passes * RS_BINS; i += TH_NT)
		(&s_h[0][0])[i] = 0u;
	__syncthreads();
	const int64_t n = tail_count(n_dev, cap);
	constexpr int TILE = TH_NT * TH_IT;
	for (int64_t base = (int64_t)blockIdx.x * TILE; base < n; base += (int64_t)gridDim.x * TILE) {
		uint64_t k[TH_IT];
#pragma unroll
		for (int r = 0; r < TH_IT; ++r) {
			const int64_t q = base + (int64_t)r * TH_NT + t;
			k[r] = q < n ? keys[q] : 0ull;
		}
#pragma unroll
		for (int r = 0; r < TH_IT; ++r) {
			const int64_t q = base + (int64_t)r * TH_NT + t;
			if (q < n)
				for (int p = 0; p < passes; ++p)
					atomicAdd(&s_h[p][(uint32_t)(k[r] >> (lbits + 8 * p)) & 0xFFu], 1u);
		}
	}
	__syncthreads();
	uint32_t* gh = &ctl->ghist[blockIdx.x % TH_COPIES][0][0];
	for (int i = t; i < passes * RS_BINS; i += TH_NT) {
		const uint32_t c = (&s_h[0][0])[i];
		if (c)
			atomicAdd(&gh[i], c);
	}
}

// One stable LSD pass over the digit at bits lbits + 8 * pass:
// keys_in -> keys_out. state: [tiles][RS_BINS] status words of this pass,
// zeroed. The grid covers the capacity; workgroups past the actual count
// exit at once.
// NT threads of IT keys each: a tile of NT * IT keys (4096 or 8192; fewer,
// larger tiles make the look-back chain shorter)
template <int NT, int IT>
__global__ __launch_bounds__(NT) void k_onesweep(const uint64_t* __restrict__ keys_in,
                                                 uint64_t* __restrict__ keys_out,
                                                 const unsigned long long* __restrict__ n_dev, int64_t cap,
                                                 int lbits, int pass, TailCtl* __restrict__ ctl,
                                                 uint32_t* __restrict__ state, uint32_t spin_limit) {
	constexpr int TILE = NT * IT;
	constexpr int SEG = IT * WAVE;
	__shared__ uint64_t s_keys[TILE];
	__shared__ uint32_t s_cnt[NT / WAVE][RS_BINS];
	__shared__ uint32_t s_dpre[RS_BINS + 1];
	__shared__ uint32_t s_gbase[RS_BINS];
	__shared__ uint32_t s_scan[NT / WAVE + 1];

	const int t = threadIdx.x;
	const int lane = lane_id();
	const int w = t / WAVE;
	const int shift = lbits + 8 * pass;
	const int64_t n = tail_count(n_dev, cap);
	const int64_t tile = blockIdx.x;
	if (tile * TILE >= n)  // uniform
		return;
	const int64_t base = tile * TILE;
	const int tile_n = (int)(n - base < TILE ? n - base : TILE);
	const int64_t seg = base + (int64_t)w * SEG;
	const int wrem = tile_n - w * SEG;
	const int w_valid = wrem <= 0 ? 0 : (wrem < SEG ? wrem : SEG);
	uint64_t k[IT];
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		const bool valid = r * WAVE + lane < w_valid;
		k[r] = valid ? keys_in[seg + (int64_t)r * WAVE + lane] : 0ull;
	}
	uint32_t lp[IT];
	rank_tile_g<NT, IT>(k, w_valid, shift, lp, s_cnt, s_dpre, s_scan);
	if (t == 0)
		s_dpre[RS_BINS] = (uint32_t)tile_n;
	__syncthreads();
	uint32_t cnt = 0;
	if (t < RS_BINS) {
		cnt = s_dpre[t + 1] - s_dpre[t];
		lb_store(state + tile * RS_BINS + t, (tile == 0 ? LB_INC : LB_AGG) | cnt);
	}
	// the digit's first position overall: keys of smaller digits, then the
	// keys of this digit in earlier tiles
	uint32_t g = 0;
	if (t < RS_BINS)
#pragma unroll
		for (int c = 0; c < TH_COPIES; ++c)
			g += ctl->ghist[c][pass][t];
	const uint32_t dstart = block_excl_scan<NT>(g, s_scan, (uint32_t*)nullptr);
	if (t < RS_BINS) {
		const uint32_t excl = tile == 0 ? 0u : lookback(state + t, RS_BINS, tile, spin_limit, &ctl->error);
		if (tile > 0)
			lb_store(state + tile * RS_BINS + t, LB_INC | (excl + cnt));
		s_gbase[t] = dstart + excl;
	}
#pragma unroll
	for (int r = 0; r < IT; ++r)
		if (lp[r] != 0xFFFFFFFFu)
			s_keys[lp[r]] = k[r];
	__syncthreads();
	for (int i = t; i < tile_n; i += NT) {
		const uint64_t key = s_keys[i];
		const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
		const uint32_t pos = s_gbase[d] + ((uint32_t)i - s_dpre[d]);
		if ((int64_t)pos < n)
			keys_out[pos] = key;
	}
}

constexpr int SL_IT = 8;  // sorted candidates per select thread

struct SelectArgs {
	const uint64_t* keys;  // sorted candidates
	const unsigned long long* n_dev;
	int64_t cap;
	int32_t idb;
	int32_t distance;
	int32_t measure;
	uint32_t row_base;
	int32_t low_bits;  // the sort left the bits below this unsorted (k_select_grouped); 0: fully sorted
	const uint8_t* chars;
	const uint32_t* offs;
	const uint32_t* seq_map;  // nullable: id -> row
	const uint32_t* set_of;   // nullable: id -> query set (batched form)
	const uint8_t* chars_b;   // nullable: the second id's strings live here
	const uint32_t* offs_b;   // (the served form: query rows against index rows)
	int2* out_pairs;
	int8_t* out_dist;
	TailCtl* ctl;
	uint64_t* state;  // one status word per tile, zeroed
	uint32_t spin_limit;
};

// Ordered ranks over R * NT positions laid out in rows, position
// q = r * NT + threadIdx.x (row r: NT consecutive positions, one per
// thread, so a wave touches consecutive LDS words). flags bit r: this
// thread's position in row r is marked. ex[r] = marked positions before it,
// in position order; returns the total. A ballot per row and wave gives the
// count below each lane, one wave scans the per-(row, wave) counts in order.
// s_cc: R * NT / WAVE + 1 words. Every thread calls it; it ends in a barrier.
template <int NT, int R>
__device__ __forceinline__ uint32_t striped_rank(uint32_t flags, uint32_t (&ex)[R], uint32_t* s_cc) {
	constexpr int NW = NT / WAVE;
	constexpr int E = R * NW;
	constexpr int C = (E + WAVE - 1) / WAVE;
	const int lane = lane_id();
	const int w = threadIdx.x / WAVE;
	const uint64_t below = (1ull << lane) - 1ull;
	uint32_t within[R];
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const uint64_t b = __ballot((flags >> r) & 1u);
		within[r] = (uint32_t)__popcll(b & below);
		if (lane == 0)
			s_cc[r * NW + w] = (uint32_t)__popcll(b);
	}
	__syncthreads();
	if (w == 0) {
		uint32_t v[C];
		uint32_t sum = 0;
#pragma unroll
		for (int c = 0; c < C; ++c) {
			const int idx = lane * C + c;
			v[c] = idx < E ? s_cc[idx] : 0u;
			sum += v[c];
		}
		const uint32_t inc = wave_incl_scan(sum);
		uint32_t e = inc - sum;
#pragma unroll
		for (int c = 0; c < C; ++c) {
			const int idx = lane * C + c;
			if (idx < E)
				s_cc[idx] = e;
			e += v[c];
		}
		if (lane == WAVE - 1)
			s_cc[E] = inc;
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < R; ++r)
		ex[r] = s_cc[r * NW + w] + within[r];
	const uint32_t total = s_cc[E];
	__syncthreads();
	return total;
}

// the unique pairs s_u[0, nu) of a tile, verified one pair per lane: s_d[q]
// = the distance, or -1 past it (or across query sets)
template <int D, int NT>
__device__ __forceinline__ void select_verify(const SelectArgs& a, const uint64_t* s_u, int8_t* s_d, uint32_t nu) {
	const uint64_t idmask = (1ull << a.idb) - 1ull;
	for (uint32_t q = threadIdx.x; q < nu; q += NT) {
		const uint64_t p = s_u[q];
		const uint32_t ia = (uint32_t)(p >> a.idb), ib = (uint32_t)(p & idmask);
		int8_t dv = -1;
		if (!a.set_of || a.set_of[ia] == a.set_of[ib]) {
			const uint32_t oa = a.seq_map ? a.seq_map[ia] : a.row_base + ia;
			const uint32_t ob = a.seq_map ? a.seq_map[ib] : a.row_base + ib;
			const uint8_t* cb = a.chars_b ? a.chars_b : a.chars;
			const uint32_t* ob_offs = a.offs_b ? a.offs_b : a.offs;
			const uint32_t sa = a.offs[oa], sb = ob_offs[ob];
			const int d = pair_distance_d<D>(a.chars + sa, (int)(a.offs[oa + 1] - sa), cb + sb,
			                                 (int)(ob_offs[ob + 1] - sb), a.distance, a.measure);
			if (d <= a.distance)
				dv = (int8_t)d;
		}
		s_d[q] = dv;
	}
	__syncthreads();
}

// The tile's unique and kept counts published and the tiles before it looked
// back on, then the kept pairs of s_u[0, nu) written in order at the tile's
// offset (rows of NT positions, so consecutive lanes write consecutive
// slots). The last tile writes the totals.
template <int NT, int R>
__device__ __forceinline__ void select_emit(const SelectArgs& a, int64_t tile, int64_t ntiles, const uint64_t* s_u,
                                            const int8_t* s_d, uint32_t nu, uint32_t* s_cc, uint32_t* s_excl) {
	const int t = threadIdx.x;
	uint32_t kflags = 0;
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const uint32_t q = (uint32_t)(r * NT + t);
		if (q < nu && s_d[q] >= 0)
			kflags |= 1u << r;
	}
	uint32_t kpos[R];
	const uint32_t nk = striped_rank<NT, R>(kflags, kpos, s_cc);
	if (t < WAVE) {  // the first wave publishes and looks back
		uint64_t* st = a.state + tile;
		const uint64_t mine = ((uint64_t)nu << LB2_SHIFT) | nk;
		uint64_t excl = 0;
		if (tile == 0) {
			if (t == 0)
				lb_store64(st, LB2_INC | mine);
		} else {
			if (t == 0)
				lb_store64(st, LB2_AGG | mine);
			excl = lookback_wave(a.state, tile, a.spin_limit, &a.ctl->error);
			if (t == 0)
				lb_store64(st, LB2_INC | (excl + mine));
		}
		if (t == 0) {
			*s_excl = (uint32_t)(excl & ((1ull << LB2_SHIFT) - 1ull));
			if (tile == ntiles - 1) {
				const uint64_t tot = excl + mine;
				a.ctl->n_unique = tot >> LB2_SHIFT;
				a.ctl->n_out = tot & ((1ull << LB2_SHIFT) - 1ull);
			}
		}
	}
	__syncthreads();
	const uint64_t idmask = (1ull << a.idb) - 1ull;
	const uint64_t o = *s_excl;
#pragma unroll
	for (int r = 0; r < R; ++r)
		if (kflags & (1u << r)) {
			const uint32_t q = (uint32_t)(r * NT + t);
			const uint64_t p = s_u[q];
			a.out_pairs[o + kpos[r]] = make_int2((int)(p >> a.idb), (int)(p & idmask));
			a.out_dist[o + kpos[r]] = s_d[q];
		}
}

// Workgroup b handles tile b of SL_TILE fully sorted candidates: the first of
// each run of equal pairs (the key before the tile decides for its first key)
// is compacted in LDS in order, verified one pair per lane, and the pairs
// within the distance go out at the tile's offset (look-back over the kept
// counts of the tiles before it) plus their rank in the tile, so the output
// stays sorted. D: the distance's band fixed at compile time
// (pair_distance_d). NT threads, a tile of NT * SL_IT candidates.
template <int D, int NT>
__global__ __launch_bounds__(NT) void k_select(SelectArgs a) {
	constexpr int SL_TILE = NT * SL_IT;
	__shared__ uint64_t s_k[SL_TILE];  // the tile as loaded
	__shared__ uint64_t s_u[SL_TILE];  // its first-of-run pairs, in order
	__shared__ int8_t s_d[SL_TILE];
	__shared__ uint32_t s_cc[SL_IT * NT / WAVE + 1];
	__shared__ uint32_t s_excl;
	const int t = threadIdx.x;
	const int64_t n = tail_count(a.n_dev, a.cap);
	const int64_t ntiles = (n + SL_TILE - 1) / SL_TILE;
	const int64_t tile = blockIdx.x;
	if (tile >= ntiles)  // uniform
		return;
	const int64_t base = tile * SL_TILE;
	uint64_t k[SL_IT];
#pragma unroll
	for (int r = 0; r < SL_IT; ++r) {
		const int64_t q = base + (int64_t)r * NT + t;
		k[r] = q < n ? a.keys[q] : ~0ull;
		s_k[r * NT + t] = k[r];
	}
	const uint64_t before = base > 0 ? a.keys[base - 1] : ~0ull;
	__syncthreads();
	uint32_t uflags = 0;
#pragma unroll
	for (int r = 0; r < SL_IT; ++r) {
		const int q = r * NT + t;
		const uint64_t prev = q == 0 ? before : s_k[q - 1];
		if (base + q < n && k[r] != prev)
			uflags |= 1u << r;
	}
	uint32_t pos[SL_IT];
	const uint32_t nu = striped_rank<NT, SL_IT>(uflags, pos, s_cc);
#pragma unroll
	for (int r = 0; r < SL_IT; ++r)
		if (uflags & (1u << r))
			s_u[pos[r]] = k[r];
	__syncthreads();
	select_verify<D, NT>(a, s_u, s_d, nu);
	select_emit<NT, SL_IT>(a, tile, ntiles, s_u, s_d, nu, s_cc, &s_excl);
}

// Bitonic sort of s[0, p2) in LDS (p2 a power of two), ascending
template <int NT>
__device__ __forceinline__ void lds_bitonic(uint64_t* s, int p2) {
	for (int k = 2; k <= p2; k <<= 1)
		for (int j = k >> 1; j > 0; j >>= 1) {
			for (int i = threadIdx.x; i < (p2 >> 1); i += NT) {
				const int x = ((i & ~(j - 1)) << 1) | (i & (j - 1));
				const int y = x + j;
				const uint64_t u = s[x], v = s[y];
				if ((u > v) == ((x & k) == 0)) {
					s[x] = v;
					s[y] = u;
				}
			}
			__syncthreads();
		}
}

constexpr int SG_SCAN = 96;  // steps a key's look for an earlier equal key may take

// The same over candidates sorted only by their bits above the low a.low_bits
// (a group: the candidates of equal key >> low_bits, in no order inside).
// Tile b owns the groups that start in [b * T, (b + 1) * T) (T = NT * SL_IT)
// and reads on past its end to the next group start, up to a window of 2T;
// a group range that does not fit raises the overflow flag (the host then
// redoes the tail fully sorted) and the tile counts nothing. Positions are
// laid out in rows of NT (position r * NT + thread).
//   1. the groups of the range are numbered by an ordered rank of their starts
//   2. a key is the first of its value when no earlier key of its group
//      equals it; the look goes backwards from the key (duplicates of a pair
//      interleave with the few other pairs of the group, so an equal key is
//      usually a few steps away) and gives up after SG_SCAN steps, counting
//      the key as a first and the tile as hard
//   3. the firsts are compacted in position order, so each group's firsts
//      are contiguous; in an easy tile a first's place is its group's start
//      plus the smaller firsts of the group (at most SG_SCAN of them)
//   4. a hard tile sorts its compacted firsts whole in LDS (bitonic) and
//      keeps the first of each run
// Either way the unique pairs then sit in LDS in ascending order, as after a
// full sort.
template <int D, int NT>
__global__ __launch_bounds__(NT) void k_select_grouped(SelectArgs a) {
	constexpr int T = NT * SL_IT;
	constexpr int WIN = 2 * T;
	constexpr int R = WIN / NT;  // rows of the window
	constexpr uint32_t NONE = 0xFFFFFFFFu;
	__shared__ uint64_t s_k[WIN];       // the window as loaded, then its firsts, then its unique pairs in order
	__shared__ uint16_t s_gs[WIN + 1];  // group g starts at s_gs[g] (s_gs[groups] = end), then its firsts do
	__shared__ int8_t s_d[WIN];
	__shared__ uint32_t s_cc[R * NT / WAVE + 1];
	__shared__ uint32_t s_excl, s_start, s_end, s_hard;
	static_assert(WIN <= 65535, "positions in 16 bits");
	const int t = threadIdx.x;
	const int L = a.low_bits;
	const int64_t n = tail_count(a.n_dev, a.cap);
	const int64_t ntiles = (n + T - 1) / T;
	const int64_t tile = blockIdx.x;
	if (tile >= ntiles)  // uniform
		return;
	const int64_t base = tile * T;
	const int tile_n = (int)(n - base < T ? n - base : T);
	if (t == 0) {
		s_start = NONE;
		s_end = base + T >= n ? (uint32_t)tile_n : NONE;
		s_hard = 0;
	}
	uint64_t k[R];
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const int q = r * NT + t;
		k[r] = r < SL_IT && q < tile_n ? a.keys[base + q] : ~0ull;
		if (r < SL_IT)
			s_k[q] = k[r];
	}
	__syncthreads();
	// the range's end: the first group start at or after (b + 1) T, read on a
	// row at a time
	if (base + T < n) {
#pragma unroll
		for (int r = SL_IT; r < R; ++r) {
			const int e = r * NT + t;
			const int64_t qg = base + e;
			bool stop = qg >= n;
			if (!stop) {
				k[r] = a.keys[qg];
				stop = (k[r] >> L) != (a.keys[qg - 1] >> L);
			}
			s_k[e] = k[r];
			if (stop)
				atomicMin(&s_end, qg < n ? (uint32_t)e : (uint32_t)(n - base));
			if (__syncthreads_or(stop))
				break;
		}
	}
	// the range's start: the first group start in [b T, (b + 1) T)
	const uint64_t before = base > 0 ? a.keys[base - 1] : 0ull;
#pragma unroll
	for (int r = 0; r < SL_IT; ++r) {
		const int q = r * NT + t;
		if (q < tile_n) {
			const bool bnd = q == 0 ? (base == 0 || (k[0] >> L) != (before >> L)) : (k[r] >> L) != (s_k[q - 1] >> L);
			if (bnd) {
				atomicMin(&s_start, (uint32_t)q);
				break;
			}
		}
	}
	__syncthreads();
	int lo = (int)s_start, hi = (int)s_end;
	if (s_start == NONE) {
		lo = hi = 0;  // every key here belongs to a group of an earlier tile
	} else if (s_end == NONE) {
		if (t == 0)
			atomicOr(&a.ctl->overflow, 1u);
		lo = hi = 0;
	}
	// 1. group starts, numbered in order; gid[r] = the group of position r
	uint32_t in = 0, bflags = 0;
#pragma unroll
	for (int r = 0; r < R; ++r) {
		const int q = r * NT + t;
		if (q >= lo && q < hi) {
			in |= 1u << r;
			if (q == lo || (k[r] >> L) != (s_k[q - 1] >> L))
				bflags |= 1u << r;
		}
	}
	uint32_t gid[R];
	const uint32_t ngroups = striped_rank<NT, R>(bflags, gid, s_cc);
#pragma unroll
	for (int r = 0; r < R; ++r) {
		if (bflags & (1u << r))
			s_gs[gid[r]] = (uint16_t)(r * NT + t);
		else
			gid[r] -= 1u;  // a position after its group's start
	}
	if (t == 0)
		s_gs[ngroups] = (uint16_t)hi;
	__syncthreads();
	// 2. firsts of their value
	uint32_t fl = 0;
	bool hard = false;
#pragma unroll
	for (int r = 0; r < R; ++r) {
		if (in & (1u << r)) {
			const int q = r * NT + t;
			const int gs = s_gs[gid[r]];
			const int stop = q - SG_SCAN > gs ? q - SG_SCAN : gs;
			int p = q - 1;
			while (p >= stop && s_k[p] != k[r])
				--p;
			if (p < stop) {  // no equal key seen
				hard = hard || stop > gs;
				fl |= 1u << r;
			}
		}
	}
	if (hard)
		s_hard = 1;
	// 3. the firsts compacted in position order; group g's firsts start at
	// s_gs[g] from here on
	uint32_t c[R];
	const uint32_t nc = striped_rank<NT, R>(fl, c, s_cc);
#pragma unroll
	for (int r = 0; r < R; ++r) {
		if (bflags & (1u << r))
			s_gs[gid[r]] = (uint16_t)c[r];  // firsts before the group's start
		if (fl & (1u << r))
			s_k[c[r]] = k[r];
	}
	if (t == 0)
		s_gs[ngroups] = (uint16_t)nc;
	__syncthreads();
	{
		bool big = false;  // a group of many firsts makes the tile hard as well
#pragma unroll
		for (int r = 0; r < R; ++r)
			if (bflags & (1u << r))
				big = big || (int)s_gs[gid[r] + 1] - (int)s_gs[gid[r]] > SG_SCAN;
		if (big)
			s_hard = 1;
	}
	__syncthreads();
	uint32_t nu = nc;
	if (!s_hard) {  // uniform
#pragma unroll
		for (int r = 0; r < R; ++r) {
			if (fl & (1u << r)) {
				const int cs = s_gs[gid[r]], ce = s_gs[gid[r] + 1];
				uint32_t less = 0;
				for (int p = cs; p < ce; ++p)
					less += s_k[p] < k[r] ? 1u : 0u;
				c[r] = (uint32_t)cs + less;
			}
		}
		__syncthreads();
#pragma unroll
		for (int r = 0; r < R; ++r)
			if (fl & (1u << r))
				s_k[c[r]] = k[r];
		__syncthreads();
	} else {
		// 4. the firsts padded to a power of two with keys above every pair,
		// sorted whole, the first of each run kept
		int p2 = 2;
		while (p2 < (int)nc)
			p2 <<= 1;
		for (int q = (int)nc + t; q < p2; q += NT)
			s_k[q] = ~0ull;
		__syncthreads();
		lds_bitonic<NT>(s_k, p2);
		uint32_t ufl = 0;
#pragma unroll
		for (int r = 0; r < R; ++r) {
			const int q = r * NT + t;
			if (q < (int)nc) {
				k[r] = s_k[q];
				if (q == 0 || k[r] != s_k[q - 1])
					ufl |= 1u << r;
			}
		}
		nu = striped_rank<NT, R>(ufl, c, s_cc);
#pragma unroll
		for (int r = 0; r < R; ++r)
			if (ufl & (1u << r))
				s_k[c[r]] = k[r];
		__syncthreads();
	}
	select_verify<D, NT>(a, s_k, s_d, nu);
	select_emit<NT, R>(a, tile, ntiles, s_k, s_d, nu, s_cc, &s_excl);
}

// the select kernel of a run: the band of its distance (Hamming needs none),
// NT threads per tile, over fully sorted or grouped candidates
typedef void (*SelectFn)(SelectArgs);
template <int NT, bool G>
inline SelectFn select_kernel_nt(int distance, int measure) {
	if (measure == 1)
		return G ? k_select_grouped<1, NT> : k_select<1, NT>;
	switch (distance) {
	case 1: return G ? k_select_grouped<1, NT> : k_select<1, NT>;
	case 2: return G ? k_select_grouped<2, NT> : k_select<2, NT>;
	case 3: return G ? k_select_grouped<3, NT> : k_select<3, NT>;
	case 4: return G ? k_select_grouped<4, NT> : k_select<4, NT>;
	case 5: return G ? k_select_grouped<5, NT> : k_select<5, NT>;
	case 6: return G ? k_select_grouped<6, NT> : k_select<6, NT>;
	case 7: return G ? k_select_grouped<7, NT> : k_select<7, NT>;
	case 8: return G ? k_select_grouped<8, NT> : k_select<8, NT>;
	default: return G ? k_select_grouped<0, NT> : k_select<0, NT>;
	}
}

inline SelectFn select_kernel(int distance, int measure, int nt, bool grouped) {
	if (grouped)
		return nt == 512 ? select_kernel_nt<512, true>(distance, measure) : select_kernel_nt<256, true>(distance, measure);
	return nt == 512 ? select_kernel_nt<512, false>(distance, measure) : select_kernel_nt<256, false>(distance, measure);
}

// ------------------------------------------- packing with its offsets --

// k_pack, the exclusive scan of the variant counts into voff and the
// generator's tile table in one launch: workgroup b packs sequences
// [PK_SEQ b, PK_SEQ b + PK_SEQ), two consecutive ones per thread (so that one
// round of workgroups covers a million sequences), scans their counts, and adds the counts of the
// workgroups before it by a wave-wide decoupled look-back over 64-bit status
// words (flag | 62-bit sum; state[] cleared before the launch). Each sequence
// then writes its offset and records the generator tiles whose first record
// it holds (tile_seq, as k_tile_seq); the last one writes voff[n] = V and
// tile_seq[ceil(V / GEN_TILE)], since it holds record V - 1 (every sequence
// has at least one variant, itself). The waits are on lower-numbered
// workgroups, which each XCD dispatches first; a wait past spin_limit
// raises *error and the host redoes the run.
constexpr int PK_NT = 1024;
constexpr int PK_PER = 2;                // consecutive sequences per thread
constexpr int PK_SEQ = PK_NT * PK_PER;  // sequences per workgroup: a short look-back chain

__global__ __launch_bounds__(PK_NT) void k_pack_scan(const uint8_t* __restrict__ chars,
                                                   const uint32_t* __restrict__ offs,
                                                   const uint32_t* __restrict__ seq_map, uint32_t row_base,
                                                   const uint32_t* __restrict__ set_of, int32_t n, int distance,
                                                   uint64_t* __restrict__ pk, uint64_t* __restrict__ voff,
                                                   uint32_t* __restrict__ tile_seq, uint64_t* __restrict__ state,
                                                   uint32_t spin_limit, uint32_t* __restrict__ error) {
	__shared__ uint32_t s_tmp[PK_NT / WAVE + 1];
	__shared__ uint64_t s_prefix;
	const int t = threadIdx.x;
	const int32_t i0 = blockIdx.x * PK_SEQ + PK_PER * t;
	uint32_t c[PK_PER];
	uint32_t csum = 0;
#pragma unroll
	for (int u = 0; u < PK_PER; ++u) {
		const int32_t i = i0 + u;
		c[u] = i < n ? pack_one(chars, offs, seq_map, row_base, set_of, i, distance, pk) : 0u;
		csum += c[u];
	}
	uint32_t tot;
	const uint32_t ex = block_excl_scan<PK_NT>(csum, s_tmp, &tot);  // <= PK_SEQ * 2^18
	const int64_t b = blockIdx.x;
	if (t < WAVE) {
		if (t == 0)
			lb_store64(state + b, (b == 0 ? LB2_INC : LB2_AGG) | (uint64_t)tot);
		const uint64_t prefix = b == 0 ? 0ull : lookback_wave(state, b, spin_limit, error);
		if (t == 0) {
			if (b > 0)
				lb_store64(state + b, LB2_INC | (prefix + tot));
			s_prefix = prefix;
		}
	}
	__syncthreads();
	uint64_t o = s_prefix + ex;
#pragma unroll
	for (int u = 0; u < PK_PER; ++u) {
		const int32_t i = i0 + u;
		if (i >= n)
			return;
		const uint64_t e = o + c[u];
		voff[i] = o;
		for (uint64_t bt = (o + GEN_TILE - 1) / GEN_TILE; bt * GEN_TILE < e; ++bt)
			tile_seq[bt] = (uint32_t)i;
		if (i == n - 1) {
			voff[n] = e;
			tile_seq[(e + GEN_TILE - 1) / GEN_TILE] = (uint32_t)i;
		}
		o = e;
	}
}

}  // namespace xtn
