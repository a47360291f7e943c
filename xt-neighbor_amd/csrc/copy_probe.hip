// copy_probe.hip -- a plain device-to-device copy kernel of known size.
// Its FETCH_SIZE / WRITE_SIZE counter readings calibrate the per-dispatch
// HBM traffic reported for the engine's kernels (bench.py --probe-copy).
#include <hip/hip_runtime.h>
#include <stdint.h>

// Each workgroup copies one 16 KB tile, four independent 16-byte loads per
// thread in flight before the stores: of the shapes measured on MI355X
// (grid-stride loops of 4096 and 32768 workgroups, tiles of 4, 8 and 16
// loads per thread, hipMemcpy) the fastest, ~5.8 TB/s for 1 GiB.
constexpr int CP_U = 4;

__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    int64_t n) {
	const int64_t base = (int64_t)blockIdx.x * 256 * CP_U + threadIdx.x;
	uint4 v[CP_U];
#pragma unroll
	for (int u = 0; u < CP_U; ++u) {
		const int64_t i = base + (int64_t)u * 256;
		v[u] = i < n ? src[i] : make_uint4(0, 0, 0, 0);
	}
#pragma unroll
	for (int u = 0; u < CP_U; ++u) {
		const int64_t i = base + (int64_t)u * 256;
		if (i < n)
			dst[i] = v[u];
	}
}

// copies `bytes` (a multiple of 16) `iters` times; returns the mean kernel
// time in ms, or a negative value on error
extern "C" double xtn_probe_copy(int64_t bytes, int32_t iters) {
	if (bytes <= 0 || bytes % 16 || iters <= 0)
		return -1.0;
	void *a = nullptr, *b = nullptr;
	if (hipMalloc(&a, (size_t)bytes) != hipSuccess || hipMalloc(&b, (size_t)bytes) != hipSuccess)
		return -2.0;
	(void)hipMemset(a, 1, (size_t)bytes);
	hipEvent_t e0, e1;
	(void)hipEventCreate(&e0);
	(void)hipEventCreate(&e1);
	const int64_t n = bytes / 16;
	const unsigned grid = (unsigned)((n + 256 * CP_U - 1) / (256 * CP_U));
	k_copy_probe<<<grid, 256>>>((const uint4*)a, (uint4*)b, n);
	(void)hipEventRecord(e0, nullptr);
	for (int k = 0; k < iters; ++k)
		k_copy_probe<<<grid, 256>>>((const uint4*)a, (uint4*)b, n);
	(void)hipEventRecord(e1, nullptr);
	(void)hipEventSynchronize(e1);
	float ms = 0.f;
	(void)hipEventElapsedTime(&ms, e0, e1);
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	(void)hipFree(a);
	(void)hipFree(b);
	return hipGetLastError() == hipSuccess ? (double)ms / iters : -3.0;
}
