// xtn_serve.hpp -- kernels of the served form: queries against a resident,
// fully sorted deletion index (no counterpart in the reference, which runs
// one input per process; the index is the reference's stage-1 output,
// xtn_inner.cu:329-358, kept in HBM).
//
//   index:  rec = hash(variant)[63:idb] | index row, sorted by the hash bits,
//           plus a directory dir[b] = first record whose top DIR_BITS are >= b
//   query:  the query's variant records, each looked up in the index
//           (directory + binary search), one candidate pair per match
//           p = query row << idb | index row, then sort, unique, verification
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xtn_kernels.hpp"

namespace xtn {

constexpr int DIR_BITS = 20;

__global__ __launch_bounds__(256) void k_dir_bounds(const uint64_t* __restrict__ idx, int64_t n,
                                                    uint32_t* __restrict__ dir) {
	const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (b > (1ll << DIR_BITS))
		return;
	int64_t lo = 0, hi = n;
	while (lo < hi) {
		const int64_t mid = (lo + hi) >> 1;
		if ((int64_t)(idx[mid] >> (64 - DIR_BITS)) < b)
			lo = mid + 1;
		else
			hi = mid;
	}
	dir[b] = (uint32_t)lo;
}

// one thread per query record: the index records with the same hash bits
__global__ __launch_bounds__(256) void k_match(const uint64_t* __restrict__ qrec, int64_t nq,
                                               const uint64_t* __restrict__ idx,
                                               const uint32_t* __restrict__ dir, int idb,
                                               uint64_t* __restrict__ pairs, int64_t cap,
                                               unsigned long long* __restrict__ ctr) {
	__shared__ uint32_t s_tmp[256 / WAVE + 1];
	__shared__ unsigned long long s_base;
	const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
	uint32_t lo = 0, count = 0;
	uint64_t key = 0, qid = 0;
	if (r < nq) {
		const uint64_t q = qrec[r];
		key = q >> idb;
		qid = q & ((1ull << idb) - 1ull);
		const uint32_t b = (uint32_t)(q >> (64 - DIR_BITS));
		lo = dir[b];
		uint32_t hi = dir[b + 1];
		while (lo < hi) {
			const uint32_t mid = (lo + hi) >> 1;
			if ((idx[mid] >> idb) < key)
				lo = mid + 1;
			else
				hi = mid;
		}
		const uint32_t end = dir[b + 1];
		for (uint32_t m = lo; m < end && (idx[m] >> idb) == key; ++m)
			++count;
	}
	unsigned long long slot = block_append<256>(ctr, count, s_tmp, &s_base);
	const uint64_t mask = (1ull << idb) - 1ull;
	for (uint32_t m = lo; m < lo + count; ++m, ++slot)
		if ((int64_t)slot < cap)
			pairs[slot] = (qid << idb) | (idx[m] & mask);
}

// one lane per sorted candidate (query row, index row): duplicates skipped,
// the rest verified between the query string and the index string
template <bool WIDE>
__global__ __launch_bounds__(256) void k_verify_cross(const uint64_t* __restrict__ pairs, int64_t n,
                                                      int idb, const uint8_t* __restrict__ qchars,
                                                      const uint32_t* __restrict__ qoffs,
                                                      const uint8_t* __restrict__ ichars,
                                                      const uint32_t* __restrict__ ioffs, int distance,
                                                      int measure, uint32_t* __restrict__ flags,
                                                      int8_t* __restrict__ dist) {
	const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (q >= n)
		return;
	const uint64_t p = pairs[q];
	uint32_t keep = 0;
	int8_t dv = 0;
	if (q == 0 || pairs[q - 1] != p) {
		const uint32_t a = (uint32_t)(p >> idb), b = (uint32_t)(p & ((1ull << idb) - 1ull));
		const uint32_t sa = qoffs[a], sb = ioffs[b];
		const int d = pair_distance<WIDE>(qchars + sa, (int)(qoffs[a + 1] - sa), ichars + sb,
		                            (int)(ioffs[b + 1] - sb), distance, measure);
		keep = d <= distance ? 1u : 0u;
		dv = (int8_t)d;
	}
	flags[q] = keep;
	dist[q] = dv;
}

}  // namespace xtn
