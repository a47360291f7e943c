// xtn_serve.hpp -- kernels of the served form: queries against a resident,
// fully sorted deletion index (no counterpart in the reference, which runs
// one input per process; the index is the reference's stage-1 output,
// xtn_inner.cu:329-358, kept in HBM).
//
//   index:  rec = hash(variant)[63:idb] | index row, sorted by the hash bits,
//           plus a directory dir[b] = first record whose top dir_bits are
//           >= b (about eight records per directory entry)
//   query:  the query's variant records, each looked up in the index
//           (directory, then a short binary search), one candidate pair per
//           match p = query row << idb | index row; the candidates go
//           through the same tail as a search (sort, unique, verification
//           of query string against index string, ordered select)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xtn_kernels.hpp"

namespace xtn {

constexpr int DIR_BITS_MAX = 24;

// Directory of a sorted index of n records over its top dir_bits bits, in
// one pass: the thread of record r writes the entries b in (top(r - 1),
// top(r)], so each of the 2^dir_bits + 1 entries is written exactly once
// (record n stands for "past the end", top = 2^dir_bits).
__global__ __launch_bounds__(256) void k_dir_fill(const uint64_t* __restrict__ idx, int64_t n, int dir_bits,
                                                  uint32_t* __restrict__ dir) {
	const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
	if (r > n)
		return;
	const int sh = 64 - dir_bits;
	const int64_t prev = r == 0 ? -1 : (int64_t)(idx[r - 1] >> sh);
	const int64_t cur = r == n ? ((int64_t)1 << dir_bits) : (int64_t)(idx[r] >> sh);
	for (int64_t b = prev + 1; b <= cur; ++b)
		dir[b] = (uint32_t)r;
}

constexpr int MATCH_NT = 1024;  // one append per workgroup: fewer atomics on the counter

// one thread per query record: the index records with the same hash bits,
// appended as candidate pairs (one atomic per workgroup)
__global__ __launch_bounds__(MATCH_NT) void k_match(const uint64_t* __restrict__ qrec, int64_t nq,
                                                    const uint64_t* __restrict__ idx,
                                                    const uint32_t* __restrict__ dir, int dir_bits, int idb,
                                                    uint64_t* __restrict__ pairs, int64_t cap,
                                                    unsigned long long* __restrict__ ctr) {
	__shared__ uint32_t s_tmp[MATCH_NT / WAVE + 1];
	__shared__ unsigned long long s_base;
	const int64_t r = (int64_t)blockIdx.x * MATCH_NT + threadIdx.x;
	uint32_t lo = 0, count = 0;
	uint64_t qid = 0;
	if (r < nq) {
		const uint64_t q = qrec[r];
		const uint64_t key = q >> idb;
		qid = q & ((1ull << idb) - 1ull);
		const uint32_t b = (uint32_t)(q >> (64 - dir_bits));
		lo = dir[b];
		const uint32_t end = dir[b + 1];
		uint32_t hi = end;
		while (lo < hi) {
			const uint32_t mid = (lo + hi) >> 1;
			if ((idx[mid] >> idb) < key)
				lo = mid + 1;
			else
				hi = mid;
		}
		for (uint32_t m = lo; m < end && (idx[m] >> idb) == key; ++m)
			++count;
	}
	unsigned long long slot = block_append<MATCH_NT>(ctr, count, s_tmp, &s_base);
	const uint64_t mask = (1ull << idb) - 1ull;
	for (uint32_t m = lo; m < lo + count; ++m, ++slot)
		if ((int64_t)slot < cap)
			pairs[slot] = (qid << idb) | (idx[m] & mask);
}

}  // namespace xtn
