// xtn_overlap.hpp -- repertoire-overlap kernels (xtn_overlap_inner.cu).
//
// Flow on the device:
//   k_hash_rows        64-bit hash of every full string, row id in the low bits
//   radix sort         equal hashes adjacent, rows ascending inside (stable)
//   k_leaders          inside each run of equal hash bits, every row points at
//                      the first row with an identical string (exact compare)
//   radix sort (k,v)   by leader position: duplicates contiguous, rows
//                      ascending (the stable full-string sort of
//                      xtn_overlap_inner.cu:57-73 gives the same groups)
//   k_group_flags/_starts  group starts, one representative row per group
//   k_exact_term       f_a*f_b into (rep_a, rep_b) for a <= b of a group
//                      (init_overlap_output, kernel.cu:565-584)
//   NN search over the representatives
//   k_neighbor_term    f_a*f_b (x2 inside one repertoire) into (min, max) rep
//                      for every member pair of a verified pair
//                      (pair2rep, kernel.cu:405-435)
//   k_matrix_*         present entries of the R x R u64 matrix in (A, B)
//                      order (sort + sum_by_key of mergeOutput,
//                      xtn_overlap_inner.cu:212-245)
// The matrix sums are mod 2^64 like the reference's size_t arithmetic; an
// entry is reported when any term was added to it, even if the sum wraps to 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xtn_prims.hpp"

namespace xtn {

constexpr int OV_LDS_REPS = 64;  // R <= 64: per-workgroup LDS copy of the matrix

// The caller's rows ({frequency, repertoire, original_index}, 12 bytes each)
// split into a frequency and a repertoire array, with the repertoire count
// (largest index + 1) into ctl[0] and the first row holding a negative
// repertoire index into ctl[1] (ctl zeroed / all ones beforehand); one
// atomic per wave for each.
__global__ __launch_bounds__(256) void k_split_info(const int32_t* __restrict__ rows, int32_t n,
                                                    int32_t* __restrict__ freq, int32_t* __restrict__ rep,
                                                    uint32_t* __restrict__ ctl) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	int32_t r = 0;
	if (i < n) {
		freq[i] = rows[3 * (int64_t)i];
		r = rows[3 * (int64_t)i + 1];
		rep[i] = r;
	}
	uint32_t top = i < n && r >= 0 ? (uint32_t)r + 1u : 0u;
	uint32_t bad = i < n && r < 0 ? (uint32_t)i : 0xFFFFFFFFu;
#pragma unroll
	for (int o = WAVE / 2; o > 0; o >>= 1) {
		const uint32_t t2 = (uint32_t)__shfl_xor((int)top, o);
		const uint32_t b2 = (uint32_t)__shfl_xor((int)bad, o);
		top = t2 > top ? t2 : top;
		bad = b2 < bad ? b2 : bad;
	}
	// atomics on one address serialise: a wave only updates a bound it
	// would move (a stale read costs one atomic more, never a wrong bound)
	if ((threadIdx.x & (WAVE - 1)) == 0) {
		if (top && top > __hip_atomic_load(&ctl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
			atomicMax(&ctl[0], top);
		if (bad != 0xFFFFFFFFu && bad < __hip_atomic_load(&ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
			atomicMin(&ctl[1], bad);
	}
}

__device__ __forceinline__ uint64_t hash_bytes(const uint8_t* __restrict__ s, uint32_t len) {
	uint64_t h = 0xCBF29CE484222325ull ^ ((uint64_t)len * 0x9E3779B97F4A7C15ull);
	for (uint32_t i = 0; i < len; ++i)
		h = (h ^ s[i]) * 0x100000001B3ull;
	h ^= h >> 32;
	h *= 0xD6E8FEB86659FD93ull;
	h ^= h >> 32;
	return h;
}

__global__ __launch_bounds__(256) void k_hash_rows(const uint8_t* __restrict__ chars,
                                                   const uint32_t* __restrict__ offs, int32_t n,
                                                   int idb, uint64_t* __restrict__ recs) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	const uint32_t s = offs[i];
	const uint64_t h = hash_bytes(chars + s, offs[i + 1] - s);
	const uint64_t mask = (1ull << idb) - 1ull;
	recs[i] = (h & ~mask) | (uint64_t)i;
}

__device__ __forceinline__ bool same_string(const uint8_t* __restrict__ chars,
                                            const uint32_t* __restrict__ offs, uint32_t a,
                                            uint32_t b) {
	const uint32_t sa = offs[a], sb = offs[b];
	const uint32_t la = offs[a + 1] - sa;
	if (offs[b + 1] - sb != la)
		return false;
	for (uint32_t k = 0; k < la; ++k)
		if (chars[sa + k] != chars[sb + k])
			return false;
	return true;
}

// One thread per run head of equal (rec >> rs): walks its run in order and
// gives every element the sorted position of the first element holding the
// same string. key_out[e] = leader position, val_out[e] = row.
__global__ __launch_bounds__(256) void k_leaders(const uint64_t* __restrict__ recs, int32_t n, int rs,
                                                 int idb, const uint8_t* __restrict__ chars,
                                                 const uint32_t* __restrict__ offs,
                                                 uint64_t* __restrict__ key_out,
                                                 uint32_t* __restrict__ val_out) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n)
		return;
	const uint64_t cur = recs[i] >> rs;
	if (i > 0 && (recs[i - 1] >> rs) == cur)
		return;  // not a head
	const uint64_t mask = (1ull << idb) - 1ull;
	int32_t end = i + 1;
	while (end < n && (recs[end] >> rs) == cur)
		++end;
	const uint32_t head_row = (uint32_t)(recs[i] & mask);
	for (int32_t e = i; e < end; ++e) {
		const uint32_t row = (uint32_t)(recs[e] & mask);
		int32_t lead = i;
		if (e != i && !same_string(chars, offs, row, head_row)) {
			// a different string with the same hash bits: first earlier
			// element of this run with the same string, else itself
			lead = e;
			for (int32_t p = i + 1; p < e; ++p)
				if (key_out[p] == (uint64_t)p && same_string(chars, offs, row, (uint32_t)(recs[p] & mask))) {
					lead = p;
					break;
				}
		}
		key_out[e] = (uint64_t)lead;
		val_out[e] = row;
	}
}

__global__ __launch_bounds__(256) void k_group_flags(const uint64_t* __restrict__ keys, int32_t n,
                                                     uint32_t* __restrict__ flags) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i < n)
		flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// gstart[g] = first member position of group g, umap[g] = its first row
__global__ __launch_bounds__(256) void k_group_starts(const uint32_t* __restrict__ flags,
                                                      const uint32_t* __restrict__ gid,
                                                      const uint32_t* __restrict__ rows, int32_t n,
                                                      uint32_t* __restrict__ gstart,
                                                      uint32_t* __restrict__ umap) {
	const int32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n || !flags[i])
		return;
	gstart[gid[i]] = (uint32_t)i;
	umap[gid[i]] = rows[i];
}

struct OvMatrix {
	unsigned long long* sum;  // R*R
	uint32_t* present;        // R*R
	int32_t R;
};

// accumulate into the workgroup's LDS copy when R <= OV_LDS_REPS
struct OvAcc {
	unsigned long long* s_sum;
	uint32_t* s_present;
	bool lds;
	OvMatrix m;
	__device__ __forceinline__ void add(int a, int b, unsigned long long v) {
		const int idx = a * m.R + b;
		if (lds) {
			atomicAdd(&s_sum[idx], v);
			s_present[idx] = 1u;
		} else {
			atomicAdd(&m.sum[idx], v);
			m.present[idx] = 1u;
		}
	}
};

__device__ __forceinline__ void ov_lds_init(unsigned long long* s_sum, uint32_t* s_present, int RR) {
	for (int k = threadIdx.x; k < RR; k += blockDim.x) {
		s_sum[k] = 0ull;
		s_present[k] = 0u;
	}
	__syncthreads();
}

__device__ __forceinline__ void ov_lds_flush(const unsigned long long* s_sum,
                                             const uint32_t* s_present, int RR, OvMatrix m) {
	__syncthreads();
	for (int k = threadIdx.x; k < RR; k += blockDim.x)
		if (s_present[k]) {
			atomicAdd(&m.sum[k], s_sum[k]);
			m.present[k] = 1u;
		}
}

__device__ __forceinline__ unsigned long long freq_product(int32_t a, int32_t b) {
	return (unsigned long long)(int64_t)a * (unsigned long long)(int64_t)b;
}

// exact-match term, one thread per group (kernel.cu:565-584), grid-stride
// over a grid of a few workgroups per CU: each workgroup clears and flushes
// its LDS matrix once, so the flush's atomics (one per entry it touched,
// mostly the same diagonal entries in every workgroup) stay few
__global__ __launch_bounds__(256) void k_exact_term(const uint32_t* __restrict__ gstart, int32_t ngroups,
                                                    const uint32_t* __restrict__ rows,
                                                    const int32_t* __restrict__ freq,
                                                    const int32_t* __restrict__ rep, OvMatrix m) {
	__shared__ unsigned long long s_sum[OV_LDS_REPS * OV_LDS_REPS];
	__shared__ uint32_t s_present[OV_LDS_REPS * OV_LDS_REPS];
	const bool lds = m.R <= OV_LDS_REPS;
	if (lds)
		ov_lds_init(s_sum, s_present, m.R * m.R);
	OvAcc acc{s_sum, s_present, lds, m};
	for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < ngroups; g += (int64_t)gridDim.x * 256) {
		const uint32_t b0 = gstart[g], b1 = gstart[g + 1];
		for (uint32_t a = b0; a < b1; ++a) {
			const uint32_t ra = rows[a];
			const int32_t fa = freq[ra], pa = rep[ra];
			for (uint32_t b = a; b < b1; ++b) {
				const uint32_t rb = rows[b];
				acc.add(pa, rep[rb], freq_product(fa, freq[rb]));
			}
		}
	}
	if (lds)
		ov_lds_flush(s_sum, s_present, m.R * m.R, m);
}

// neighbour term, one thread per verified pair of groups (kernel.cu:405-435),
// grid-stride as the exact term
__global__ __launch_bounds__(256) void k_neighbor_term(const int2* __restrict__ pairs, int64_t npairs,
                                                       const uint32_t* __restrict__ gstart,
                                                       const uint32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ freq,
                                                       const int32_t* __restrict__ rep, OvMatrix m) {
	__shared__ unsigned long long s_sum[OV_LDS_REPS * OV_LDS_REPS];
	__shared__ uint32_t s_present[OV_LDS_REPS * OV_LDS_REPS];
	const bool lds = m.R <= OV_LDS_REPS;
	if (lds)
		ov_lds_init(s_sum, s_present, m.R * m.R);
	OvAcc acc{s_sum, s_present, lds, m};
	for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < npairs; q += (int64_t)gridDim.x * 256) {
		const int2 p = pairs[q];
		const uint32_t x0 = gstart[p.x], x1 = gstart[p.x + 1];
		const uint32_t y0 = gstart[p.y], y1 = gstart[p.y + 1];
		for (uint32_t a = x0; a < x1; ++a) {
			const uint32_t ra = rows[a];
			const int32_t fa = freq[ra], pa = rep[ra];
			for (uint32_t b = y0; b < y1; ++b) {
				const uint32_t rb = rows[b];
				const int32_t pb = rep[rb];
				unsigned long long v = freq_product(fa, freq[rb]);
				if (pa == pb)
					v *= 2ull;
				acc.add(pa < pb ? pa : pb, pa < pb ? pb : pa, v);
			}
		}
	}
	if (lds)
		ov_lds_flush(s_sum, s_present, m.R * m.R, m);
}

__global__ __launch_bounds__(256) void k_matrix_emit(const uint32_t* __restrict__ present,
                                                     const uint32_t* __restrict__ pos,
                                                     const unsigned long long* __restrict__ sum,
                                                     int64_t RR, int32_t R, int2* __restrict__ out_pairs,
                                                     unsigned long long* __restrict__ out_sum) {
	const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (k >= RR || !present[k])
		return;
	const uint32_t o = pos[k];
	out_pairs[o] = make_int2((int)(k / R), (int)(k % R));
	out_sum[o] = sum[k];
}

}  // namespace xtn
