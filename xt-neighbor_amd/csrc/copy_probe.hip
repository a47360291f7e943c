// copy_probe.hip -- a plain device-to-device copy kernel of known size.
// Its FETCH_SIZE / WRITE_SIZE counter readings calibrate the per-dispatch
// HBM traffic reported for the engine's kernels (bench.py --probe-copy).
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ __launch_bounds__(256) void k_copy_probe(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    int64_t n) {
	const int64_t stride = (int64_t)gridDim.x * 256;
	for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
		dst[i] = src[i];
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
	k_copy_probe<<<4096, 256>>>((const uint4*)a, (uint4*)b, n);
	(void)hipEventRecord(e0, nullptr);
	for (int k = 0; k < iters; ++k)
		k_copy_probe<<<4096, 256>>>((const uint4*)a, (uint4*)b, n);
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
