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
	uint32_t pad0;
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
// digits (digit p = bits [8p, 8p + 8)), grid-stride over tiles; also zeroes
// zero[0, nzero) (the look-back state of the passes and the select).
__global__ __launch_bounds__(TH_NT) void k_tail_hist(const uint64_t* __restrict__ keys,
                                                     const unsigned long long* __restrict__ n_dev,
                                                     int64_t cap, int passes, TailCtl* __restrict__ ctl,
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
					atomicAdd(&s_h[p][(uint32_t)(k[r] >> (8 * p)) & 0xFFu], 1u);
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

// One stable LSD pass over the digit at `shift` (pass index `pass`):
// keys_in -> keys_out. state: [tiles][RS_BINS] status words of this pass,
// zeroed. The grid covers the capacity; workgroups past the actual count
// exit at once.
// NT threads of IT keys each: a tile of NT * IT keys (4096 or 8192; fewer,
// larger tiles make the look-back chain shorter)
template <int NT, int IT>
__global__ __launch_bounds__(NT) void k_onesweep(const uint64_t* __restrict__ keys_in,
                                                 uint64_t* __restrict__ keys_out,
                                                 const unsigned long long* __restrict__ n_dev, int64_t cap,
                                                 int pass, TailCtl* __restrict__ ctl,
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
	const int shift = 8 * pass;
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

// Workgroup b handles tile b of SL_TILE sorted candidates: the first of each
// run of equal pairs (the key before the tile decides for its first key) is
// compacted in LDS in order, verified one pair per lane, and the pairs within
// the distance go out at the tile's offset (look-back over the kept counts of
// the tiles before it) plus their rank in the tile, so the output stays
// sorted. The last tile writes the unique and kept totals.
// D: the distance's band fixed at compile time (pair_distance_d). NT
// threads, a tile of NT * SL_IT candidates.
template <int D, int NT>
__global__ __launch_bounds__(NT) void k_select(SelectArgs a) {
	constexpr int SL_TILE = NT * SL_IT;
	__shared__ uint64_t s_k[SL_TILE];  // the tile as loaded
	__shared__ uint64_t s_u[SL_TILE];  // its first-of-run pairs, in order
	__shared__ int8_t s_d[SL_TILE];
	__shared__ uint32_t s_tmp[NT / WAVE + 1];
	__shared__ uint32_t s_excl;
	const int t = threadIdx.x;
	const int64_t n = tail_count(a.n_dev, a.cap);
	const int64_t ntiles = (n + SL_TILE - 1) / SL_TILE;
	const uint64_t idmask = (1ull << a.idb) - 1ull;
	const int64_t tile = blockIdx.x;
	if (tile >= ntiles)  // uniform
		return;
	const int64_t base = tile * SL_TILE;
#pragma unroll
	for (int r = 0; r < SL_IT; ++r) {
		const int64_t q = base + (int64_t)r * NT + t;
		s_k[r * NT + t] = q < n ? a.keys[q] : ~0ull;
	}
	const uint64_t before = base > 0 ? a.keys[base - 1] : ~0ull;
	__syncthreads();
	// thread t owns tile positions [t * SL_IT, (t + 1) * SL_IT)
	uint32_t uflags = 0, cu = 0;
#pragma unroll
	for (int i = 0; i < SL_IT; ++i) {
		const int q = t * SL_IT + i;
		const uint64_t prev = q == 0 ? before : s_k[q - 1];
		if (base + q < n && s_k[q] != prev) {
			uflags |= 1u << i;
			++cu;
		}
	}
	uint32_t nu;
	uint32_t pos = block_excl_scan<NT>(cu, s_tmp, &nu);
#pragma unroll
	for (int i = 0; i < SL_IT; ++i)
		if (uflags & (1u << i))
			s_u[pos++] = s_k[t * SL_IT + i];
	__syncthreads();
	for (uint32_t q = t; q < nu; q += NT) {
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
	uint32_t kflags = 0, ck = 0;
#pragma unroll
	for (int i = 0; i < SL_IT; ++i) {
		const uint32_t q = (uint32_t)(t * SL_IT + i);
		if (q < nu && s_d[q] >= 0) {
			kflags |= 1u << i;
			++ck;
		}
	}
	uint32_t nk;
	const uint32_t kpos = block_excl_scan<NT>(ck, s_tmp, &nk);
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
			s_excl = (uint32_t)(excl & ((1ull << LB2_SHIFT) - 1ull));
			if (tile == ntiles - 1) {
				const uint64_t tot = excl + mine;
				a.ctl->n_unique = tot >> LB2_SHIFT;
				a.ctl->n_out = tot & ((1ull << LB2_SHIFT) - 1ull);
			}
		}
	}
	__syncthreads();
	uint64_t o = (uint64_t)s_excl + kpos;
#pragma unroll
	for (int i = 0; i < SL_IT; ++i)
		if (kflags & (1u << i)) {
			const uint64_t p = s_u[t * SL_IT + i];
			a.out_pairs[o] = make_int2((int)(p >> a.idb), (int)(p & idmask));
			a.out_dist[o] = s_d[t * SL_IT + i];
			++o;
		}
}

// the select kernel of a run: the band of its distance (Hamming needs none),
// NT threads per tile
typedef void (*SelectFn)(SelectArgs);
template <int NT>
inline SelectFn select_kernel_nt(int distance, int measure) {
	if (measure == 1)
		return k_select<1, NT>;
	switch (distance) {
	case 1: return k_select<1, NT>;
	case 2: return k_select<2, NT>;
	case 3: return k_select<3, NT>;
	case 4: return k_select<4, NT>;
	case 5: return k_select<5, NT>;
	case 6: return k_select<6, NT>;
	case 7: return k_select<7, NT>;
	case 8: return k_select<8, NT>;
	default: return k_select<0, NT>;
	}
}

inline SelectFn select_kernel(int distance, int measure, int nt) {
	return nt == 512 ? select_kernel_nt<512>(distance, measure) : select_kernel_nt<256>(distance, measure);
}

}  // namespace xtn
