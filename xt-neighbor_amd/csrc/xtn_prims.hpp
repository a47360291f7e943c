// xtn_prims.hpp -- device-wide primitives for gfx950 (wave64): exclusive scan
// and a stable LSD radix sort of 64-bit keys (optionally carrying 32-bit
// values). They replace the CUB facade of the reference (cub.cu:80-285:
// DeviceScan, DeviceMergeSort, DeviceRunLengthEncode, DeviceSelect).
//
// Radix sort: reduce-then-scan per 8-bit digit pass.
//   k_radix_hist    one 4096-key tile per workgroup, LDS digit histogram,
//                   written as the tile's row of 256 counts (tile-major, so
//                   every store is one contiguous 1 KB row).
//   k_dscan_*       turn the rows into output bases in place: the base of
//                   (tile, digit) counts every key of a smaller digit plus
//                   the keys of that digit in earlier tiles. Group sums over
//                   DS_GROUP tiles, one exclusive scan per digit column of
//                   the group sums, then each group rewrites its rows.
//   k_radix_scatter the tile is ranked stably inside each wave with eight
//                   ballots (the lanes holding the same digit), wave-private
//                   LDS counters carry the rank across the wave's 16 rounds,
//                   the tile is reordered by digit in LDS and written out in
//                   digit runs (coalesced).
// No kernel here spins or waits on another workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xtn {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
	const int lane = lane_id();
	return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
	const int lane = lane_id();
#pragma unroll
	for (int o = 1; o < WAVE; o <<= 1) {
		T u = __shfl_up(v, o, WAVE);
		if (lane >= o)
			v += u;
	}
	return v;
}

// Exclusive scan across a workgroup of NT threads. s_tmp needs NT/WAVE+1
// slots. Returns the thread's exclusive prefix; *total gets the sum.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* s_tmp, T* total) {
	constexpr int NW = NT / WAVE;
	const int lane = lane_id();
	const int wid = threadIdx.x / WAVE;
	T inc = wave_incl_scan(v);
	if (lane == WAVE - 1)
		s_tmp[wid] = inc;
	__syncthreads();
	if (wid == 0) {
		T w = lane < NW ? s_tmp[lane] : T(0);
		T wi = wave_incl_scan(w);
		if (lane < NW)
			s_tmp[lane] = wi - w;
		if (lane == NW - 1)
			s_tmp[NW] = wi;
	}
	__syncthreads();
	T r = s_tmp[wid] + inc - v;
	if (total)
		*total = s_tmp[NW];
	__syncthreads();
	return r;
}

// Appends `count` items per lane to a global counter with one atomic per wave.
// Returns this lane's first slot. Every lane of the wave must call it.
__device__ __forceinline__ unsigned long long wave_append(unsigned long long* ctr, uint32_t count) {
	const int lane = lane_id();
	uint32_t inc = wave_incl_scan(count);
	uint32_t total = __shfl(inc, WAVE - 1, WAVE);
	unsigned long long base = 0;
	if (lane == 0 && total > 0)
		base = atomicAdd(ctr, (unsigned long long)total);
	base = __shfl(base, 0, WAVE);
	return base + (inc - count);
}

// Appends `count` items per thread to a global counter with one atomic per
// workgroup of NT threads. Returns the thread's first slot. Every thread of
// the workgroup must call it (it contains barriers).
template <int NT>
__device__ __forceinline__ unsigned long long block_append(unsigned long long* ctr, uint32_t count,
                                                           uint32_t* s_tmp,
                                                           unsigned long long* s_base) {
	uint32_t total;
	const uint32_t ex = block_excl_scan<NT>(count, s_tmp, &total);
	if (threadIdx.x == 0)
		*s_base = total ? atomicAdd(ctr, (unsigned long long)total) : 0ull;
	__syncthreads();
	const unsigned long long b = *s_base;
	__syncthreads();
	return b + ex;
}

// adds v over the workgroup into *ctr with one atomic (all threads call it)
template <int NT>
__device__ __forceinline__ void block_add(unsigned long long* ctr, unsigned long long v,
                                          unsigned long long* s_tmp) {
	unsigned long long total;
	block_excl_scan<NT>(v, s_tmp, &total);
	if (threadIdx.x == 0 && total)
		atomicAdd(ctr, total);
}

// ---------------------------------------------------------------- scan --

constexpr int SC_NT = 256;
constexpr int SC_IT = 8;
constexpr int SC_TILE = SC_NT * SC_IT;

template <typename T>
__global__ __launch_bounds__(SC_NT) void k_scan_sums(const T* __restrict__ in, int64_t n,
                                                     T* __restrict__ sums) {
	__shared__ T s_tmp[SC_NT / WAVE + 1];
	const int64_t base = (int64_t)blockIdx.x * SC_TILE;
	T acc = 0;
#pragma unroll
	for (int i = 0; i < SC_IT; ++i) {
		int64_t idx = base + (int64_t)i * SC_NT + threadIdx.x;
		if (idx < n)
			acc += in[idx];
	}
	T tot;
	block_excl_scan<SC_NT>(acc, s_tmp, &tot);
	if (threadIdx.x == 0)
		sums[blockIdx.x] = tot;
}

// exclusive scan of data[0..n) in place by a single workgroup, the total into
// *total; with n_dev, n shrinks to ceil(*n_dev / per) (entries past it are
// zero and are left as they are)
template <typename T>
__global__ __launch_bounds__(SC_NT) void k_scan_single(T* __restrict__ data, int64_t n,
                                                       T* __restrict__ total,
                                                       const uint32_t* __restrict__ n_dev, int per) {
	__shared__ T s_tmp[SC_NT / WAVE + 1];
	if (n_dev) {  // only the first ceil(*n_dev / per) entries can be nonzero
		const int64_t m = ((int64_t)*n_dev + per - 1) / per;
		n = m < n ? m : n;
	}
	T carry = 0;
	for (int64_t base = 0; base < n; base += SC_TILE) {
		T vals[SC_IT];
		T sum = 0;
#pragma unroll
		for (int i = 0; i < SC_IT; ++i) {
			int64_t idx = base + (int64_t)threadIdx.x * SC_IT + i;
			vals[i] = idx < n ? data[idx] : T(0);
			sum += vals[i];
		}
		T tot;
		T run = carry + block_excl_scan<SC_NT>(sum, s_tmp, &tot);
#pragma unroll
		for (int i = 0; i < SC_IT; ++i) {
			int64_t idx = base + (int64_t)threadIdx.x * SC_IT + i;
			if (idx < n)
				data[idx] = run;
			run += vals[i];
		}
		carry += tot;
	}
	if (threadIdx.x == 0 && total)
		*total = carry;
}

template <typename T>
__global__ __launch_bounds__(SC_NT) void k_scan_apply(const T* in, T* out, int64_t n,
                                                      const T* __restrict__ sums) {
	__shared__ T s_tmp[SC_NT / WAVE + 1];
	const int64_t base = (int64_t)blockIdx.x * SC_TILE + (int64_t)threadIdx.x * SC_IT;
	T vals[SC_IT];
	T sum = 0;
#pragma unroll
	for (int i = 0; i < SC_IT; ++i) {
		int64_t idx = base + i;
		vals[i] = idx < n ? in[idx] : T(0);
		sum += vals[i];
	}
	T run = sums[blockIdx.x] + block_excl_scan<SC_NT>(sum, s_tmp, (T*)nullptr);
#pragma unroll
	for (int i = 0; i < SC_IT; ++i) {
		int64_t idx = base + i;
		if (idx < n)
			out[idx] = run;
		run += vals[i];
	}
}

// ---------------------------------------------------------- radix sort --

constexpr int RS_NT = 256;
constexpr int RS_IT = 16;
constexpr int RS_TILE = RS_NT * RS_IT;  // 4096 keys per workgroup
constexpr int RS_BINS = 256;
constexpr int RS_WAVES = RS_NT / WAVE;
constexpr int RS_SEG = RS_IT * WAVE;  // keys per wave segment

static_assert(RS_NT == RS_BINS, "one thread per digit in the prefix step");

__global__ __launch_bounds__(RS_NT) void k_radix_hist(const uint64_t* __restrict__ keys, int64_t n,
                                                      int shift, uint32_t* __restrict__ hist,
                                                      int64_t ntiles) {
	__shared__ uint32_t s_h[RS_WAVES][RS_BINS];
	for (int i = threadIdx.x; i < RS_WAVES * RS_BINS; i += RS_NT)
		(&s_h[0][0])[i] = 0;
	__syncthreads();
	const int w = threadIdx.x / WAVE;
	const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
	for (int i = 0; i < RS_IT; ++i) {
		int64_t idx = base + (int64_t)i * RS_NT + threadIdx.x;
		if (idx < n) {
			uint32_t d = (uint32_t)(keys[idx] >> shift) & 0xFFu;
			atomicAdd(&s_h[w][d], 1u);
		}
	}
	__syncthreads();
	const int d = threadIdx.x;
	uint32_t c = 0;
#pragma unroll
	for (int ww = 0; ww < RS_WAVES; ++ww)
		c += s_h[ww][d];
	hist[(int64_t)blockIdx.x * RS_BINS + d] = c;
}

// ------------------------------------------------- digit-offset scan --

constexpr int DS_GROUP = 32;  // tiles per group

// gsum[g][d] = sum of hist[t][d] over the tiles t of group g
__global__ __launch_bounds__(RS_BINS) void k_dscan_group(const uint32_t* __restrict__ hist,
                                                         int64_t ntiles, uint32_t* __restrict__ gsum) {
	const int d = threadIdx.x;
	const int64_t t0 = (int64_t)blockIdx.x * DS_GROUP;
	const int nt = ntiles - t0 < DS_GROUP ? (int)(ntiles - t0) : DS_GROUP;
	uint32_t acc = 0;
#pragma unroll 8
	for (int i = 0; i < nt; ++i)
		acc += hist[(t0 + i) * RS_BINS + d];
	gsum[(int64_t)blockIdx.x * RS_BINS + d] = acc;
}

// one workgroup per digit d: exclusive scan of column d of gsum in place,
// the column total into dtot[d]
__global__ __launch_bounds__(SC_NT) void k_dscan_cols(uint32_t* __restrict__ gsum, int64_t ngroups,
                                                      uint32_t* __restrict__ dtot) {
	__shared__ uint32_t s_tmp[SC_NT / WAVE + 1];
	const int d = blockIdx.x;
	uint32_t carry = 0;
	for (int64_t base = 0; base < ngroups; base += SC_TILE) {
		uint32_t vals[SC_IT];
		uint32_t sum = 0;
#pragma unroll
		for (int i = 0; i < SC_IT; ++i) {
			const int64_t g = base + (int64_t)threadIdx.x * SC_IT + i;
			vals[i] = g < ngroups ? gsum[g * RS_BINS + d] : 0u;
			sum += vals[i];
		}
		uint32_t tot;
		uint32_t run = carry + block_excl_scan<SC_NT>(sum, s_tmp, &tot);
#pragma unroll
		for (int i = 0; i < SC_IT; ++i) {
			const int64_t g = base + (int64_t)threadIdx.x * SC_IT + i;
			if (g < ngroups)
				gsum[g * RS_BINS + d] = run;
			run += vals[i];
		}
		carry += tot;
	}
	if (threadIdx.x == 0)
		dtot[d] = carry;
}

// rows of group g rewritten in place as output bases
__global__ __launch_bounds__(RS_BINS) void k_dscan_apply(uint32_t* __restrict__ hist, int64_t ntiles,
                                                         const uint32_t* __restrict__ gsum,
                                                         const uint32_t* __restrict__ dtot) {
	__shared__ uint32_t s_tmp[RS_BINS / WAVE + 1];
	const int d = threadIdx.x;
	uint32_t run = block_excl_scan<RS_BINS>(dtot[d], s_tmp, (uint32_t*)nullptr) +
	               gsum[(int64_t)blockIdx.x * RS_BINS + d];
	const int64_t t0 = (int64_t)blockIdx.x * DS_GROUP;
	const int nt = ntiles - t0 < DS_GROUP ? (int)(ntiles - t0) : DS_GROUP;
	uint32_t v[DS_GROUP];
#pragma unroll
	for (int i = 0; i < DS_GROUP; ++i)
		v[i] = i < nt ? hist[(t0 + i) * RS_BINS + d] : 0u;
#pragma unroll
	for (int i = 0; i < DS_GROUP; ++i)
		if (i < nt) {
			hist[(t0 + i) * RS_BINS + d] = run;
			run += v[i];
		}
}

// lanes of the wave whose digit equals this lane's (among valid lanes)
// Per bit b the ballot m_b of the lanes with that bit set, xored with this
// lane's bit spread over a word (0 or ~0), marks the lanes whose bit b
// differs; or-ing those marks over the 8 bits and complementing once gives
// the lanes of the same digit: five VALU instructions per bit, against nine
// for selecting m_b or ~m_b per lane.
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, bool valid) {
	uint32_t dlo = 0, dhi = 0;  // lanes whose digit differs from this lane's
#pragma unroll
	for (int b = 0; b < 8; ++b) {
		const uint32_t t = (uint32_t)((int32_t)(d << (31 - b)) >> 31);  // bit b spread: 0 or ~0
		const uint64_t m = __ballot(t != 0u);
		dlo |= (uint32_t)m ^ t;
		dhi |= (uint32_t)(m >> 32) ^ t;
	}
	return __ballot(valid) & ~(((uint64_t)dhi << 32) | dlo);
}

// Tiles are handed out so that each of the 8 XCDs (workgroups are dispatched
// to XCDs round-robin) owns one contiguous range of tiles: consecutive tiles
// then write their digit runs into the same XCD's L2, where the partial lines
// of neighbouring tiles merge before write-back.
__device__ __forceinline__ int64_t xcd_tile(int64_t i, int64_t ntiles) {
	constexpr int64_t NX = 8;
	const int64_t xcd = i % NX, local = i / NX;
	const int64_t q = ntiles / NX, r = ntiles % NX;
	return xcd * q + (xcd < r ? xcd : r) + local;
}

// Stable rank of a tile of up to NT * IT keys by the 8-bit digit at `shift`
// (NT threads, NT / 64 waves). Wave w holds the keys of tile positions
// [w*IT*64, (w+1)*IT*64),
// round r lane l holding position w*IT*64 + r*64 + l; only the first w_valid
// positions of a wave are keys. On return lp[r] is the key's position in the
// tile sorted by digit (stable), or 0xFFFFFFFF when invalid, and s_dpre[d]
// the first position of digit d. Every thread must call it.
template <int NT, int IT>
__device__ __forceinline__ void rank_tile_g(const uint64_t (&k)[IT], int w_valid, int shift,
                                            uint32_t (&lp)[IT], uint32_t (*s_cnt)[RS_BINS],
                                            uint32_t* s_dpre, uint32_t* s_scan) {
	static_assert(NT % WAVE == 0 && NT >= RS_BINS, "one thread per digit in the prefix step");
	constexpr int NW = NT / WAVE;
	const int t = threadIdx.x;
	const int lane = lane_id();
	const int w = t / WAVE;
	for (int i = t; i < NW * RS_BINS; i += NT)
		(&s_cnt[0][0])[i] = 0;
	__syncthreads();
	const uint64_t lt = lanemask_lt();
	uint32_t rk[IT];
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		rk[r] = 0xFFFFFFFFu;
		if (r * WAVE >= w_valid)  // uniform over the wave: no key in this round
			continue;
		const bool valid = r * WAVE + lane < w_valid;
		const uint32_t d = (uint32_t)(k[r] >> shift) & 0xFFu;
		const uint64_t peers = match_digit8(d, valid);
		const uint32_t before = s_cnt[w][d];
		const uint32_t rank = before + (uint32_t)__popcll(peers & lt);
		if (valid && (__ffsll((unsigned long long)peers) - 1) == lane)
			s_cnt[w][d] = before + (uint32_t)__popcll(peers);
		rk[r] = valid ? ((rank << 8) | d) : 0xFFFFFFFFu;
	}
	__syncthreads();
	{
		const int d = t;
		uint32_t c[NW];
		uint32_t tot = 0;
#pragma unroll
		for (int ww = 0; ww < NW; ++ww) {
			c[ww] = d < RS_BINS ? s_cnt[ww][d] : 0u;
			tot += c[ww];
		}
		const uint32_t ex = block_excl_scan<NT>(tot, s_scan, (uint32_t*)nullptr);
		if (d < RS_BINS) {
			s_dpre[d] = ex;
			uint32_t run = ex;
#pragma unroll
			for (int ww = 0; ww < NW; ++ww) {
				s_cnt[ww][d] = run;
				run += c[ww];
			}
		}
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < IT; ++r)
		lp[r] = rk[r] == 0xFFFFFFFFu ? 0xFFFFFFFFu : s_cnt[w][rk[r] & 0xFFu] + (rk[r] >> 8);
}

// The same outputs without stability: one LDS atomic per key on a single
// tile histogram (s_cnt[0]) gives its rank inside its digit, in whatever
// order the atomics land. For passes whose output order inside a digit run
// does not matter; about a tenth of rank_tile_g's VALU instructions.
template <int NT, int IT>
__device__ __forceinline__ void rank_tile_u(const uint64_t (&k)[IT], int w_valid, int shift,
                                            uint32_t (&lp)[IT], uint32_t (*s_cnt)[RS_BINS],
                                            uint32_t* s_dpre, uint32_t* s_scan) {
	static_assert(NT % WAVE == 0 && NT >= RS_BINS && NT / WAVE >= 2,
	              "one thread per digit in the prefix step; s_cnt[1][0] is the spare bin");
	const int t = threadIdx.x;
	const int lane = lane_id();
	uint32_t* s_h = s_cnt[0];
	for (int i = t; i <= RS_BINS; i += NT)
		s_h[i] = 0;
	__syncthreads();
	// no branch around the atomics, so they are all in flight together:
	// lanes without a key count into the spare bin s_h[RS_BINS]
	uint32_t rk[IT];
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		const bool valid = r * WAVE + lane < w_valid;
		const uint32_t d = valid ? (uint32_t)(k[r] >> shift) & 0xFFu : (uint32_t)RS_BINS;
		rk[r] = atomicAdd(&s_h[d], 1u);
		rk[r] = valid ? ((rk[r] << 8) | d) : 0xFFFFFFFFu;
	}
	__syncthreads();
	{
		const int d = t;
		const uint32_t c = d < RS_BINS ? s_h[d] : 0u;
		const uint32_t ex = block_excl_scan<NT>(c, s_scan, (uint32_t*)nullptr);
		if (d < RS_BINS)
			s_dpre[d] = ex;
	}
	__syncthreads();
#pragma unroll
	for (int r = 0; r < IT; ++r)
		lp[r] = rk[r] == 0xFFFFFFFFu ? 0xFFFFFFFFu : s_dpre[rk[r] & 0xFFu] + (rk[r] >> 8);
}

template <int IT>
__device__ __forceinline__ void rank_tile_n(const uint64_t (&k)[IT], int w_valid, int shift,
                                            uint32_t (&lp)[IT], uint32_t (*s_cnt)[RS_BINS],
                                            uint32_t* s_dpre, uint32_t* s_scan) {
	rank_tile_g<RS_NT, IT>(k, w_valid, shift, lp, s_cnt, s_dpre, s_scan);
}

__device__ __forceinline__ void rank_tile(const uint64_t (&k)[RS_IT], int w_valid, int shift,
                                          uint32_t (&lp)[RS_IT], uint32_t (*s_cnt)[RS_BINS],
                                          uint32_t* s_dpre, uint32_t* s_scan) {
	rank_tile_n<RS_IT>(k, w_valid, shift, lp, s_cnt, s_dpre, s_scan);
}

// STREAM: the keys are read once and never again from this pass (the
// record path's first pass), so the loads carry the non-temporal hint
template <bool KV, int NT = RS_NT, bool STREAM = false>
__global__ __launch_bounds__(NT) void k_radix_scatter(const uint64_t* __restrict__ keys_in,
                                                         const uint32_t* __restrict__ vals_in,
                                                         uint64_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ vals_out, int64_t n,
                                                         int shift, const uint32_t* __restrict__ offs,
                                                         int64_t ntiles) {
	constexpr int IT = RS_TILE / NT;   // keys per thread
	constexpr int SEG = IT * WAVE;     // keys per wave
	__shared__ uint64_t s_keys[RS_TILE];
	__shared__ uint32_t s_vals[KV ? RS_TILE : 1];
	__shared__ uint32_t s_cnt[NT / WAVE][RS_BINS];
	__shared__ uint32_t s_dpre[RS_BINS];
	__shared__ uint32_t s_gbase[RS_BINS];
	__shared__ uint32_t s_scan[NT / WAVE + 1];

	const int t = threadIdx.x;
	const int lane = lane_id();
	const int w = t / WAVE;
	const int64_t tile = xcd_tile(blockIdx.x, ntiles);
	const int64_t base = tile * RS_TILE;
	const int64_t seg = base + (int64_t)w * SEG;
	const int64_t wrem = n - seg;
	const int w_valid = wrem <= 0 ? 0 : (wrem < SEG ? (int)wrem : SEG);
	uint64_t k[IT];
	uint32_t v[IT];
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		const bool valid = r * WAVE + lane < w_valid;
		const int64_t idx = seg + (int64_t)r * WAVE + lane;
		if constexpr (STREAM)
			k[r] = valid ? __builtin_nontemporal_load(&keys_in[idx]) : 0ull;
		else
			k[r] = valid ? keys_in[idx] : 0ull;
		if (KV)
			v[r] = valid ? vals_in[idx] : 0u;
	}
	if (t < RS_BINS)
		s_gbase[t] = offs[tile * RS_BINS + t];
	uint32_t lp[IT];
	rank_tile_g<NT, IT>(k, w_valid, shift, lp, s_cnt, s_dpre, s_scan);
#pragma unroll
	for (int r = 0; r < IT; ++r) {
		if (lp[r] != 0xFFFFFFFFu) {
			s_keys[lp[r]] = k[r];
			if (KV)
				s_vals[lp[r]] = v[r];
		}
	}
	__syncthreads();

	const int64_t rem = n - base;
	const int tile_n = rem < RS_TILE ? (int)rem : RS_TILE;
	for (int i = t; i < tile_n; i += NT) {
		const uint64_t key = s_keys[i];
		const uint32_t d = (uint32_t)(key >> shift) & 0xFFu;
		const uint32_t g = s_gbase[d] + ((uint32_t)i - s_dpre[d]);
		if ((int64_t)g < n) {
			keys_out[g] = key;
			if (KV)
				vals_out[g] = s_vals[i];
		}
	}
}


}  // namespace xtn
