// xtn_lib.hip -- host side of libxtn.so: handle, chunk scheduler, pipeline
// orchestration and the C ABI declared in include/xtn.h.
//
// Pipeline for one run (NN mode; overlap mode runs it on the unique strings):
//   k_pack + scan          sequences -> packed prefixes, variant offsets (V)
//   per key-space chunk:   k_gen (records, first digit histogram) -> two
//                          radix passes over the bucket bits -> k_bucket
//                          (candidate pairs from LDS bins; buckets too big
//                          for LDS go through k_runs / k_pairs_*)
//   radix sort of the candidate pairs -> unique (k_uniq_*) -> distances of
//   the unique pairs (k_verify_count) -> the pairs within d (k_keep_write)
// The reference runs the same stages as stream 1..4 of xtn_perform
// (xtn.cu:267-509) with every intermediate staged through pinned host memory;
// here everything between input upload and output download stays in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <cstddef>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "../../include/xtn.h"
#include "xtn_kernels.hpp"
#include "xtn_overlap.hpp"
#include "xtn_serve.hpp"
#include "xtn_tail.hpp"

using namespace xtn;

namespace {

struct DBuf {
	void* p = nullptr;
	size_t cap = 0;
	template <typename T>
	T* as() const {
		return static_cast<T*>(p);
	}
};

struct ProfAcc {
	int64_t launches = 0;
	double ms = 0.0;
	double bytes = 0.0;
};

struct PendingEvent {
	std::string name;
	hipEvent_t start, stop;
	double bytes;
};

}  // namespace

struct xtn_handle {
	xtn_config cfg;
	hipStream_t stream = nullptr;
	hipStream_t copy_stream = nullptr;
	std::string err;

	// device-resident input (xtn_set_sequences)
	DBuf chars, offs;
	int32_t n = 0;
	int64_t n_chars = 0;

	// workspaces (grow-only)
	DBuf bcount, ctl;  // region layout: bucket cursors, coarse bucket/tile table
	DBuf recs_all, recs_part;  // chunked runs: every record once, then ordered by part
	DBuf seg_cur, seg_ctl;  // the generator's coarse pass: sub-range cursors, digit tiling
	DBuf pk, voff, recs, recs_alt, hist, scan_tmp, runs_small, runs_big, big_pre, pairs, pairs_alt,
	    flags, pos, dist, out_pairs, out_dist, ctrs, part_hist, scalar;
	// overlap-mode workspaces
	DBuf ov_keys, ov_keys_alt, ov_vals, ov_vals_alt, ov_info, ov_leader, ov_ustart, ov_umap,
	    ov_cnt, ov_matrix, ov_present;

	int64_t out_len = 0;
	xtn_stats stats;

	std::vector<PendingEvent> pending;
	std::vector<hipEvent_t> event_pool;
	std::map<std::string, ProfAcc> prof;
	std::vector<PendingEvent> spans;  // stage timing, always on
	DBuf bstart, ovf, set_of, set_offs, tseq, tail_state, pack_state;
	int64_t last_pairs = 0;
	int64_t last_unique = 0, last_out = 0;  // the previous run's unique candidates and outputs
	                                         // (byte counts of the output select's profile)
	// host copy of the prefix lengths, and per-row prefix sums of the
	// variant counts for one distance: V of any row range without a readback
	int n_cu = 256;
	// the single-chunk run captured as a graph: replayed while a run's
	// signature (sizes, pointers, the allocation generation) matches
	struct GraphSig {
		int64_t nseq, row_base, V, pcap, alloc_gen;
		const void *seq_map, *set_of, *chars, *offs;
		int32_t d, measure, bucket_cap, target, sort_nt, sort_tile, select_nt, low, coarse, pad;
		bool operator==(const GraphSig& o) const { return memcmp(this, &o, sizeof(*this)) == 0; }
	};
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	GraphSig graph_sig{}, graph_want{};
	bool graph_want_set = false;
	int64_t graph_launches = 0;  // kernels the graph holds
	bool graph_coarse = false;   // the graph's generator does the coarse pass
	// engine parameters the autotuner chooses (xtn_set_tuning / xtn_autotune)
	int tune_cap = 3072;      // bucket kernel instantiation: 2048, 3072 or 4096 records
	int tune_target = 2200;   // mean records per bucket
	int tune_sort_nt = 512;   // threads of a pair-sort tile: 256, 512 or 1024
	int tune_sort_tile = 4096;  // keys of a pair-sort tile: 4096 or 8192
	int tune_select_nt = 256;   // threads of a select tile (8 candidates each): 256 or 512
	int tune_low = 0;           // pair bits left unsorted for the select's in-group ranking: 0, 8, 16, 24
	int tune_coarse_sep = 0;    // 1: the records' first ordering pass as a pass of its own (tuned)
	bool tail_full = false;     // a group outgrew the select's window: full sorts until new sequences
	bool coarse_off = false;    // a sub-range of the generator's coarse pass overflowed: the separate
	                            // pass until new sequences
	double tune_ms = 0, tune_ms_default = 0;
	bool capturing = false;
	bool no_alloc = false;  // xtn_nn_noalloc: every buffer is a slice of the static scratch
	int64_t alloc_gen = 0;  // bumped by every device allocation of the handle
	void* pin = nullptr;  // pinned upload staging (xtn_set_sequences, a server worker's queries)
	size_t pin_bytes = 0;
	hipEvent_t pin_ev[2] = {nullptr, nullptr};
	void* out_pin = nullptr;  // pinned output staging of xtn_perform (two chunks of pairs and distances)
	size_t out_pin_bytes = 0;
	double serve_ratio = 0;  // a server worker's candidates per query record so far
	bool tail_safe = false;  // a look-back wait hit its limit once: host-checked tail from then on
	uint32_t spin_limit = LB_SPIN_LIMIT;  // XTN_LOOKBACK_SPIN_LIMIT overrides (tests of the fallback)
	size_t free_cache = 0;   // hipMemGetInfo's free bytes, re-read after an allocation
	bool free_valid = false;
	std::vector<uint8_t> prefix_len;
	std::vector<uint64_t> vprefix;
	int vprefix_d = 0;
};

namespace {

int fail(xtn_handle* h, int code, const char* fmt, ...) {
	char buf[512];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	if (h)
		h->err = buf;
	return code;
}

#define HIP_TRY(h, expr)                                                                      \
	do {                                                                                      \
		hipError_t _e = (expr);                                                               \
		if (_e != hipSuccess)                                                                 \
			return fail((h), XTN_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
			            __FILE__, __LINE__);                                                  \
	} while (0)

#define HIP_LAUNCH_CHECK(h)                                                                   \
	do {                                                                                      \
		++(h)->stats.n_launches;                                                              \
		hipError_t _e = hipGetLastError();                                                    \
		if (_e != hipSuccess)                                                                 \
			return fail((h), XTN_ERR_HIP, "kernel launch failed: %s (%s:%d)",                 \
			            hipGetErrorString(_e), __FILE__, __LINE__);                            \
	} while (0)

// a host wait on a stream, counted in the run's stats
#define HIP_SYNC(h, stream)                            \
	do {                                               \
		++(h)->stats.n_syncs;                          \
		HIP_TRY(h, hipStreamSynchronize(stream));      \
	} while (0)

#define TRY(expr)                 \
	do {                          \
		int _rc = (expr);         \
		if (_rc != XTN_OK)        \
			return _rc;           \
	} while (0)

int ensure(xtn_handle* h, DBuf& b, size_t bytes) {
	if (bytes == 0)
		bytes = 256;
	if (bytes <= b.cap)
		return XTN_OK;
	if (h && h->capturing)  // the caller abandons the capture and issues the run eagerly
		return fail(h, XTN_ERR_STATE, "allocation during graph capture");
	if (h && h->no_alloc)  // buffers are fixed slices of the static scratch
		return fail(h, XTN_ERR_CAPACITY,
		            "the search exceeds the no-allocation scratch (a buffer of %zu bytes is needed where %zu are "
		            "set aside)",
		            bytes, b.cap);
	if (b.p) {
		HIP_SYNC(h, h->stream);
		HIP_TRY(h, hipFree(b.p));
		b.p = nullptr;
		b.cap = 0;
	}
	size_t want = bytes + bytes / 8;
	want = (want + 4095) & ~(size_t)4095;
	hipError_t e = hipMalloc(&b.p, want);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		want = (bytes + 4095) & ~(size_t)4095;
		e = hipMalloc(&b.p, want);
		if (e != hipSuccess) {
			(void)hipGetLastError();
			b.p = nullptr;
			return fail(h, XTN_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
		}
	}
	b.cap = want;
	if (h) {
		h->free_valid = false;
		++h->alloc_gen;
	}
	return XTN_OK;
}

void release(DBuf& b) {
	if (b.p)
		(void)hipFree(b.p);
	b.p = nullptr;
	b.cap = 0;
}

// grow a buffer keeping its first `keep` bytes
int grow_keep(xtn_handle* h, DBuf& b, size_t bytes, size_t keep) {
	if (bytes <= b.cap)
		return XTN_OK;
	DBuf nb;
	TRY(ensure(h, nb, bytes));
	if (keep)
		HIP_TRY(h, hipMemcpyAsync(nb.p, b.p, keep, hipMemcpyDeviceToDevice, h->stream));
	HIP_SYNC(h, h->stream);
	release(b);
	b = nb;
	return XTN_OK;
}

// sum_{k <= min(d, L)} C(L, k) for the 18-residue prefix of a len-residue
// sequence (the count k_pack computes on the device)
uint64_t host_variant_count(uint32_t len, int d) {
	const int L = len < 18 ? (int)len : 18;
	const int m = d < L ? d : L;
	uint64_t c = 1, total = 1;
	for (int j = 1; j <= m; ++j) {
		c = c * (uint64_t)(L - j + 1) / (uint64_t)j;
		total += c;
	}
	return total;
}

inline unsigned grid_for(int64_t n, int per_block) {
	int64_t g = (n + per_block - 1) / per_block;
	return (unsigned)(g > 0 ? g : 1);
}

// the generator's table of each tile's first sequence, over the voff of n
// sequences holding V >= 1 records
static int tile_seq_table(xtn_handle* h, hipStream_t s, int32_t n, uint64_t V, const uint32_t** out) {
	const uint64_t ntiles = (V + GEN_TILE - 1) / GEN_TILE;
	TRY(ensure(h, h->tseq, sizeof(uint32_t) * (size_t)(ntiles + 1)));
	*out = h->tseq.as<uint32_t>();
	k_tile_seq<<<grid_for(n, 256), 256, 0, s>>>(h->voff.as<uint64_t>(), n, V, (uint32_t)ntiles,
	                                            h->tseq.as<uint32_t>());
	HIP_LAUNCH_CHECK(h);
	*out = h->tseq.as<uint32_t>();
	return XTN_OK;
}

// ------------------------------------------------------------ profiling --

hipEvent_t take_event(xtn_handle* h) {
	if (!h->event_pool.empty()) {
		hipEvent_t e = h->event_pool.back();
		h->event_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	if (hipEventCreate(&e) != hipSuccess)
		return nullptr;
	return e;
}

struct ProfScope {
	xtn_handle* h;
	PendingEvent pe;
	bool on;
	ProfScope(xtn_handle* hh, const char* name, double bytes) : h(hh), on(hh->cfg.profile != 0) {
		if (!on)
			return;
		pe.name = name;
		pe.bytes = bytes;
		pe.start = take_event(h);
		pe.stop = take_event(h);
		if (!pe.start || !pe.stop) {
			on = false;
			return;
		}
		(void)hipEventRecord(pe.start, h->stream);
	}
	~ProfScope() {
		if (!on)
			return;
		(void)hipEventRecord(pe.stop, h->stream);
		h->pending.push_back(pe);
	}
};

void collect_profile(xtn_handle* h) {
	for (auto& pe : h->pending) {
		float ms = 0.f;
		if (hipEventSynchronize(pe.stop) == hipSuccess && hipEventElapsedTime(&ms, pe.start, pe.stop) == hipSuccess) {
			ProfAcc& acc = h->prof[pe.name];
			acc.launches += 1;
			acc.ms += ms;
			acc.bytes += pe.bytes;
		}
		h->event_pool.push_back(pe.start);
		h->event_pool.push_back(pe.stop);
	}
	h->pending.clear();
}

// stage spans: a pair of events per stage and chunk, summed after the run;
// recorded only in verbose or profile mode (an event record between two
// kernels leaves the GPU idle for several microseconds)
struct StageScope {
	xtn_handle* h;
	PendingEvent pe;
	bool on = false;
	StageScope(xtn_handle* hh, const char* stage) : h(hh) {
		if (!hh->cfg.verbose && !hh->cfg.profile)
			return;
		pe.name = stage;
		pe.bytes = 0;
		pe.start = take_event(h);
		pe.stop = take_event(h);
		on = pe.start && pe.stop;
		if (on)
			(void)hipEventRecord(pe.start, h->stream);
	}
	~StageScope() {
		if (!on)
			return;
		(void)hipEventRecord(pe.stop, h->stream);
		h->spans.push_back(pe);
	}
};

void collect_stages(xtn_handle* h) {
	for (auto& pe : h->spans) {
		float ms = 0.f;
		if (hipEventSynchronize(pe.stop) == hipSuccess &&
		    hipEventElapsedTime(&ms, pe.start, pe.stop) == hipSuccess) {
			if (pe.name == "generate")
				h->stats.ms_generate += ms;
			else if (pe.name == "sort")
				h->stats.ms_sort += ms;
			else if (pe.name == "pairs")
				h->stats.ms_pairs += ms;
			else if (pe.name == "verify")
				h->stats.ms_verify += ms;
		}
		h->event_pool.push_back(pe.start);
		h->event_pool.push_back(pe.stop);
	}
	h->spans.clear();
}

// ------------------------------------------------------------ primitives --

template <typename T>
int scan_excl(xtn_handle* h, const T* in, T* out, int64_t n, T* d_total) {
	hipStream_t s = h->stream;
	if (n <= 0) {
		if (d_total)
			HIP_TRY(h, hipMemsetAsync(d_total, 0, sizeof(T), s));
		return XTN_OK;
	}
	const int64_t nt = (n + SC_TILE - 1) / SC_TILE;
	TRY(ensure(h, h->scan_tmp, sizeof(T) * (size_t)nt));
	T* tmp = h->scan_tmp.as<T>();
	ProfScope ps(h, "scan", 3.0 * sizeof(T) * (double)n);
	k_scan_sums<T><<<grid_for(nt, 1), SC_NT, 0, s>>>(in, n, tmp);
	HIP_LAUNCH_CHECK(h);
	k_scan_single<T><<<1, SC_NT, 0, s>>>(tmp, nt, d_total, nullptr, 1);
	HIP_LAUNCH_CHECK(h);
	k_scan_apply<T><<<grid_for(nt, 1), SC_NT, 0, s>>>(in, out, n, tmp);
	HIP_LAUNCH_CHECK(h);
	return XTN_OK;
}

// threads of the key-only scatter passes of radix_sort (4096-key tiles, 8
// keys per thread): measured faster on the small pair sorts; the record
// passes (k_radix_scatter / k_region_scatter over ~1e8 records) run 256
// threads, measured faster there
constexpr int SCATTER_NT = 512;

// Per-tile digit counts (rows of RS_BINS) -> output bases, in place.
int digit_scan(xtn_handle* h, uint32_t* hist, int64_t ntiles, uint32_t** dtot_out = nullptr) {
	hipStream_t s = h->stream;
	const int64_t ngroups = (ntiles + DS_GROUP - 1) / DS_GROUP;
	TRY(ensure(h, h->scan_tmp, sizeof(uint32_t) * (size_t)(RS_BINS * (ngroups + 1))));
	uint32_t* gsum = h->scan_tmp.as<uint32_t>();
	uint32_t* dtot = gsum + RS_BINS * ngroups;
	ProfScope ps(h, "scan", 3.0 * 4.0 * RS_BINS * (double)ntiles);
	k_dscan_group<<<(unsigned)ngroups, RS_BINS, 0, s>>>(hist, ntiles, gsum);
	HIP_LAUNCH_CHECK(h);
	k_dscan_cols<<<RS_BINS, SC_NT, 0, s>>>(gsum, ngroups, dtot);
	HIP_LAUNCH_CHECK(h);
	k_dscan_apply<<<(unsigned)ngroups, RS_BINS, 0, s>>>(hist, ntiles, gsum, dtot);
	HIP_LAUNCH_CHECK(h);
	if (dtot_out)  // the digit totals, valid until the next scan
		*dtot_out = dtot;
	return XTN_OK;
}

// Sorts n keys (and values when vals != nullptr) by bits [b0, b1) with one
// stable 8-bit LSD pass per digit, ping-ponging between the pointer pairs; on
// return `keys`/`vals` point at the buffers holding the result.
int radix_sort_ptr(xtn_handle* h, uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals,
                   uint32_t*& vals_alt, int64_t n, int b0, int b1, const char* prof_name,
                   bool first_hist_ready = false) {
	if (n <= 1 || b0 >= b1)
		return XTN_OK;
	if (n > (int64_t)INT32_MAX)
		return fail(h, XTN_ERR_CAPACITY, "radix sort of %lld keys exceeds 2^31", (long long)n);
	hipStream_t s = h->stream;
	const int64_t ntiles = (n + RS_TILE - 1) / RS_TILE;
	TRY(ensure(h, h->hist, sizeof(uint32_t) * (size_t)(RS_BINS * ntiles)));
	const bool kv = vals != nullptr;
	const std::string hist_name = std::string(prof_name) + "_hist";
	const std::string scatter_name = std::string(prof_name) + "_scatter";
	for (int shift = b0; shift < b1; shift += 8) {
		const double key_bytes = (double)n * (8.0 + (kv ? 4.0 : 0.0));
		if (!(first_hist_ready && shift == b0)) {
			ProfScope ps(h, hist_name.c_str(), (double)n * 8.0);
			k_radix_hist<<<(unsigned)ntiles, RS_NT, 0, s>>>(keys, n, shift, h->hist.as<uint32_t>(), ntiles);
			HIP_LAUNCH_CHECK(h);
		}
		TRY(digit_scan(h, h->hist.as<uint32_t>(), ntiles));
		{
			ProfScope ps(h, scatter_name.c_str(), 2.0 * key_bytes);
			if (kv)
				k_radix_scatter<true><<<(unsigned)ntiles, RS_NT, 0, s>>>(
				    keys, vals, keys_alt, vals_alt, n, shift, h->hist.as<uint32_t>(), ntiles);
			else
				k_radix_scatter<false, SCATTER_NT><<<(unsigned)ntiles, SCATTER_NT, 0, s>>>(
				    keys, nullptr, keys_alt, nullptr, n, shift, h->hist.as<uint32_t>(), ntiles);
			HIP_LAUNCH_CHECK(h);
		}
		std::swap(keys, keys_alt);
		if (kv)
			std::swap(vals, vals_alt);
	}
	return XTN_OK;
}

// the same over whole buffers: the DBufs are swapped so that `keys`/`vals`
// name the result
int radix_sort(xtn_handle* h, DBuf& keys, DBuf& keys_alt, DBuf* vals, DBuf* vals_alt, int64_t n,
               int b0, int b1, const char* prof_name, bool first_hist_ready = false) {
	uint64_t* k = keys.as<uint64_t>();
	uint64_t* ka = keys_alt.as<uint64_t>();
	uint32_t* v = vals ? vals->as<uint32_t>() : nullptr;
	uint32_t* va = vals_alt ? vals_alt->as<uint32_t>() : nullptr;
	TRY(radix_sort_ptr(h, k, ka, v, va, n, b0, b1, prof_name, first_hist_ready));
	if (k != keys.as<uint64_t>()) {
		std::swap(keys, keys_alt);
		if (vals)
			std::swap(*vals, *vals_alt);
	}
	return XTN_OK;
}

int read_ctrs(xtn_handle* h, RunCtrs* out) {
	HIP_TRY(h, hipMemcpyAsync(out, h->ctrs.p, sizeof(RunCtrs), hipMemcpyDeviceToHost, h->stream));
	HIP_SYNC(h, h->stream);
	return XTN_OK;
}

int bits_for(int64_t maxval) {
	int b = 1;
	while (b < 62 && (1ll << b) <= maxval)
		++b;
	return b;
}

double now_ms() {
	using namespace std::chrono;
	return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ----------------------------------------------------------- NN pipeline --

constexpr int KEY_BITS = 32;             // hash bits that order records into runs

// device control block: the counters of the candidate search, then the
// tail's block
struct CtlBlock {
	RunCtrs rc;
	TailCtl tc;
};
// what a run reads back: the search counters and the tail's scalars
struct CtlHead {
	RunCtrs rc;
	uint32_t error, overflow;
	unsigned long long n_unique, n_out;
};
static_assert(offsetof(CtlBlock, tc) == sizeof(RunCtrs), "the tail block follows the counters");
static_assert(sizeof(CtlHead) == sizeof(RunCtrs) + offsetof(TailCtl, pad1), "head of the control block");
static_assert(sizeof(CtlBlock) % 8 == 0 && sizeof(RunCtrs) % 8 == 0, "filled in u64 words");

inline CtlBlock* ctl_block(xtn_handle* h) { return h->ctrs.as<CtlBlock>(); }

inline unsigned long long* ctr_field(xtn_handle* h, size_t offset) {
	return reinterpret_cast<unsigned long long*>(static_cast<char*>(h->ctrs.p) + offset);
}

int set_ctr(xtn_handle* h, size_t offset, unsigned long long v) {
	k_fill_u64<<<1, 256, 0, h->stream>>>(ctr_field(h, offset), 1, v);
	HIP_LAUNCH_CHECK(h);
	return XTN_OK;
}

// zeroes the search counters and/or the tail block with one fill launch (a
// kernel, not a memset, so that the sequence can be captured in a graph)
int zero_ctl(xtn_handle* h, bool counters, bool tail) {
	const size_t b0 = counters ? 0 : sizeof(RunCtrs);
	const size_t b1 = tail ? sizeof(CtlBlock) : sizeof(RunCtrs);
	const int64_t words = (int64_t)(b1 - b0) / 8;
	k_fill_u64<<<grid_for(words, 256), 256, 0, h->stream>>>(ctr_field(h, b0), words, 0ull);
	HIP_LAUNCH_CHECK(h);
	return XTN_OK;
}

int read_head(xtn_handle* h, CtlHead* out) {
	HIP_TRY(h, hipMemcpyAsync(out, h->ctrs.p, sizeof(CtlHead), hipMemcpyDeviceToHost, h->stream));
	HIP_SYNC(h, h->stream);
	return XTN_OK;
}

// Global path for one bucket that does not fit the LDS tile: sort the
// segment by the run bits, list its runs, emit the pairs of every run whose
// smaller id lies in [lo, hi). *split is set (and nothing emitted) when the
// segment's pairs could exceed pmax candidates in all.
int bucket_fallback(xtn_handle* h, int64_t seg_start, int64_t seg_len, int rs, int sort_hi, int idb,
                    uint32_t lo, uint32_t hi, int64_t pmax, int64_t* n_pairs, bool* split) {
	hipStream_t s = h->stream;
	uint64_t* seg = h->recs.as<uint64_t>() + seg_start;
	uint64_t* k = seg;
	uint64_t* ka = h->recs_alt.as<uint64_t>() + seg_start;
	uint32_t* nv = nullptr;
	uint32_t* nva = nullptr;
	TRY(radix_sort_ptr(h, k, ka, nv, nva, seg_len, rs, sort_hi, "sort_seg"));
	if (k != seg)
		HIP_TRY(h, hipMemcpyAsync(seg, k, sizeof(uint64_t) * (size_t)seg_len, hipMemcpyDeviceToDevice, s));
	k_fill_u64<<<1, 256, 0, s>>>(reinterpret_cast<unsigned long long*>(h->ctrs.p), 4, 0ull);
	TRY(set_ctr(h, offsetof(RunCtrs, n_cand), (unsigned long long)*n_pairs));
	const int64_t small_cap = seg_len / 2 + 1, big_cap = seg_len / (SMALL_RUN_MAX + 1) + 1;
	{
		ProfScope ps(h, "runs", (double)seg_len * 8.0);
		k_runs<<<grid_for(seg_len, RN_TILE), RN_NT, 0, s>>>(seg, seg_len, rs, h->runs_small.as<Run>(), small_cap,
		                                                   h->runs_big.as<Run>(), big_cap, h->ctrs.as<RunCtrs>());
		HIP_LAUNCH_CHECK(h);
	}
	RunCtrs rc{};
	TRY(read_ctrs(h, &rc));
	if ((int64_t)rc.n_small > small_cap || (int64_t)rc.n_big > big_cap)
		return fail(h, XTN_ERR_HIP, "run list overflow (%llu small, %llu big)",
		            (unsigned long long)rc.n_small, (unsigned long long)rc.n_big);
	const uint64_t bound = rc.pairs_small + rc.pairs_big;
	if ((uint64_t)*n_pairs + bound > (uint64_t)pmax) {
		*split = true;
		return XTN_OK;
	}
	const int64_t need = *n_pairs + (int64_t)bound;
	if ((size_t)need * sizeof(uint64_t) > h->pairs.cap)
		TRY(grow_keep(h, h->pairs, sizeof(uint64_t) * (size_t)need, sizeof(uint64_t) * (size_t)*n_pairs));
	const int64_t pcap = (int64_t)(h->pairs.cap / sizeof(uint64_t));
	if (rc.n_small) {
		ProfScope ps(h, "pairs_small", (double)rc.n_small * 8.0);
		k_pairs_small<<<grid_for((int64_t)rc.n_small, 256), 256, 0, s>>>(
		    seg, h->runs_small.as<Run>(), (int64_t)rc.n_small, idb, lo, hi, h->pairs.as<uint64_t>(), pcap,
		    h->ctrs.as<RunCtrs>());
		HIP_LAUNCH_CHECK(h);
	}
	if (rc.n_big) {
		TRY(ensure(h, h->big_pre, sizeof(uint64_t) * ((size_t)rc.n_big + 1)));
		k_big_counts<<<grid_for((int64_t)rc.n_big, 256), 256, 0, s>>>(h->runs_big.as<Run>(), (int64_t)rc.n_big,
		                                                              h->big_pre.as<uint64_t>());
		HIP_LAUNCH_CHECK(h);
		TRY(scan_excl<uint64_t>(h, h->big_pre.as<uint64_t>(), h->big_pre.as<uint64_t>(), (int64_t)rc.n_big,
		                        h->big_pre.as<uint64_t>() + rc.n_big));
		const unsigned g = (unsigned)std::min<uint64_t>((rc.pairs_big + 255) / 256, 16384);
		ProfScope ps(h, "pairs_big", (double)rc.pairs_big * 16.0);
		k_pairs_big<<<g > 0 ? g : 1, 256, 0, s>>>(seg, h->runs_big.as<Run>(), h->big_pre.as<uint64_t>(),
		                                          (int64_t)rc.n_big, idb, lo, hi, h->pairs.as<uint64_t>(), pcap,
		                                          h->ctrs.as<RunCtrs>());
		HIP_LAUNCH_CHECK(h);
	}
	TRY(read_ctrs(h, &rc));
	if ((int64_t)rc.n_cand > pcap)
		return fail(h, XTN_ERR_HIP, "pair buffer overflow (%llu > %lld)", (unsigned long long)rc.n_cand,
		            (long long)pcap);
	h->stats.n_runs += (int64_t)(rc.n_small + rc.n_big);
	*n_pairs = (int64_t)rc.n_cand;
	return XTN_OK;
}

// ---- the candidate tail (xtn_tail.hpp): sort, unique, verify, select --

struct TailPlan {
	int64_t pcap = 0;  // candidates the launches are sized for
	int passes = 0;    // 8-bit LSD digits over the pair bits from low_bits up to 2 * idb
	int low_bits = 0;  // > 0: the select ranks the bits below in LDS (k_select_grouped)
	int64_t sort_tiles = 0, sel_tiles = 0;
};

// LSD passes of the next candidate sort over 2 * idb pair bits: the sort
// stops short of the low bits when asked (a key's group is its bits above
// them); the bits left over go to the last digit's side
int tail_passes(const xtn_handle* h, int idb) {
	const int want = h->tail_full ? 0 : h->tune_low;
	return std::max(1, (2 * idb - want + 7) / 8);
}

// buffers of a tail over up to pcap candidates whose output goes after the
// out_keep pairs already held: the sort's second buffer, the outputs and the
// look-back state
int tail_prepare(xtn_handle* h, int64_t pcap, int idb, int64_t out_keep, TailPlan* tp) {
	tp->pcap = pcap;
	tp->passes = tail_passes(h, idb);
	tp->low_bits = std::max(0, 2 * idb - 8 * tp->passes);
	const int64_t sort_tile = h->tune_sort_tile, sel_tile = (int64_t)h->tune_select_nt * SL_IT;
	tp->sort_tiles = (pcap + sort_tile - 1) / sort_tile;
	tp->sel_tiles = (pcap + sel_tile - 1) / sel_tile;
	TRY(ensure(h, h->pairs_alt, sizeof(uint64_t) * (size_t)pcap));
	TRY(grow_keep(h, h->out_pairs, sizeof(int2) * (size_t)(out_keep + pcap), sizeof(int2) * (size_t)out_keep));
	TRY(grow_keep(h, h->out_dist, (size_t)(out_keep + pcap), (size_t)out_keep));
	// the sort's 32-bit words, then the select's 64-bit words (the sort part
	// is a multiple of RS_BINS words, so the select's start is 8-aligned)
	TRY(ensure(h, h->tail_state,
	           sizeof(uint32_t) * (size_t)((int64_t)tp->passes * tp->sort_tiles * RS_BINS + 2 * tp->sel_tiles)));
	return XTN_OK;
}

// the tail's launches over the candidates in `pairs` (count: the search
// counter), outputs written from out_pairs[out_off]; the tail block must be
// zero (zero_ctl)
int issue_tail(xtn_handle* h, const TailPlan& tp, int idb, const uint32_t* seq_map, uint32_t row_base,
               const uint32_t* set_of, int64_t out_off, const uint8_t* chars_b = nullptr,
               const uint32_t* offs_b = nullptr) {
	hipStream_t s = h->stream;
	CtlBlock* cb = ctl_block(h);
	const unsigned long long* n_dev = &cb->rc.n_cand;
	uint32_t* state = h->tail_state.as<uint32_t>();
	const int64_t nsort = (int64_t)tp.passes * tp.sort_tiles * RS_BINS;
	// profile byte counts: the candidates of the previous run (the count of
	// this one is only on the device)
	const double cand = (double)h->last_pairs;
	{
		ProfScope ps(h, "tail_hist", 8.0 * cand);
		const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(tp.sort_tiles, 256));
		k_tail_hist<<<g, TH_NT, 0, s>>>(h->pairs.as<uint64_t>(), n_dev, tp.pcap, tp.low_bits, tp.passes, &cb->tc,
		                                state, nsort + 2 * tp.sel_tiles);
		HIP_LAUNCH_CHECK(h);
	}
	uint64_t* a = h->pairs.as<uint64_t>();
	uint64_t* b = h->pairs_alt.as<uint64_t>();
	const unsigned sort_grid = (unsigned)std::max<int64_t>(tp.sort_tiles, 1);
	for (int p = 0; p < tp.passes; ++p) {
		ProfScope ps(h, "sort_pairs", 16.0 * cand);
		uint32_t* st = state + (int64_t)p * tp.sort_tiles * RS_BINS;
		const int nt = h->tune_sort_nt, tile = h->tune_sort_tile, lb = tp.low_bits;
		if (nt == 256)
			k_onesweep<256, 16><<<sort_grid, 256, 0, s>>>(a, b, n_dev, tp.pcap, lb, p, &cb->tc, st, h->spin_limit);
		else if (nt == 512 && tile == 4096)
			k_onesweep<512, 8><<<sort_grid, 512, 0, s>>>(a, b, n_dev, tp.pcap, lb, p, &cb->tc, st, h->spin_limit);
		else if (nt == 512)
			k_onesweep<512, 16><<<sort_grid, 512, 0, s>>>(a, b, n_dev, tp.pcap, lb, p, &cb->tc, st, h->spin_limit);
		else
			k_onesweep<1024, 8><<<sort_grid, 1024, 0, s>>>(a, b, n_dev, tp.pcap, lb, p, &cb->tc, st, h->spin_limit);
		HIP_LAUNCH_CHECK(h);
		std::swap(a, b);
	}
	SelectArgs sa{};
	sa.keys = a;
	sa.n_dev = n_dev;
	sa.cap = tp.pcap;
	sa.idb = idb;
	sa.distance = h->cfg.distance;
	sa.measure = h->cfg.measure;
	sa.row_base = row_base;
	sa.low_bits = tp.low_bits;
	sa.chars = h->chars.as<uint8_t>();
	sa.offs = h->offs.as<uint32_t>();
	sa.seq_map = seq_map;
	sa.set_of = set_of;
	sa.chars_b = chars_b;
	sa.offs_b = offs_b;
	sa.out_pairs = h->out_pairs.as<int2>() + out_off;
	sa.out_dist = h->out_dist.as<int8_t>() + out_off;
	sa.ctl = &cb->tc;
	sa.state = reinterpret_cast<uint64_t*>(state + nsort);
	sa.spin_limit = h->spin_limit;
	{
		// candidates read once, the unique ones' two strings and offsets, the
		// kept triples written
		ProfScope ps(h, "select", 8.0 * cand + 40.0 * (double)h->last_unique + 9.0 * (double)h->last_out);
		const unsigned g = (unsigned)std::max<int64_t>(1, tp.sel_tiles);
		select_kernel(h->cfg.distance, h->cfg.measure, h->tune_select_nt, tp.low_bits > 0)<<<g, h->tune_select_nt, 0,
		                                                                                   s>>>(sa);
		HIP_LAUNCH_CHECK(h);
	}
	return XTN_OK;
}

// The same tail with the host in the loop, for a run whose look-back waits
// hit their limit: the candidate count read back, an LSD sort of 8-bit
// reduce-then-scan passes, then unique, verification and the order-preserving
// select as separate launches. Fills head->n_unique and head->n_out.
int tail_rts(xtn_handle* h, int64_t n_pairs, int idb, const uint32_t* seq_map, uint32_t row_base,
             const uint32_t* set_of, int64_t out_off, CtlHead* head) {
	hipStream_t s = h->stream;
	head->n_unique = head->n_out = 0;
	if (n_pairs <= 0)
		return XTN_OK;
	TRY(ensure(h, h->pairs_alt, sizeof(uint64_t) * (size_t)n_pairs));
	TRY(radix_sort(h, h->pairs, h->pairs_alt, nullptr, nullptr, n_pairs, 0, 2 * idb, "sort_pairs_rts"));
	const int64_t nt = (n_pairs + SEL_TILE - 1) / SEL_TILE;  // unique step tiles
	const int64_t nv = (n_pairs + SEL_NT - 1) / SEL_NT;      // verify step tiles
	TRY(ensure(h, h->pos, sizeof(uint32_t) * (size_t)(nt + 1 + nv + 1)));
	TRY(ensure(h, h->dist, (size_t)n_pairs));
	TRY(grow_keep(h, h->out_pairs, sizeof(int2) * (size_t)(out_off + n_pairs), sizeof(int2) * (size_t)out_off));
	TRY(grow_keep(h, h->out_dist, (size_t)(out_off + n_pairs), (size_t)out_off));
	uint32_t* ucnt = h->pos.as<uint32_t>();  // nt counts, total at [nt]
	uint32_t* kcnt = ucnt + nt + 1;          // nv counts of kept pairs, total at [nv]
	uint64_t* upairs = h->pairs_alt.as<uint64_t>();
	k_uniq_count<<<(unsigned)nt, SEL_NT, 0, s>>>(h->pairs.as<uint64_t>(), n_pairs, ucnt);
	HIP_LAUNCH_CHECK(h);
	k_scan_single<uint32_t><<<1, SC_NT, 0, s>>>(ucnt, nt, ucnt + nt, nullptr, 1);
	HIP_LAUNCH_CHECK(h);
	k_uniq_write<<<(unsigned)nt, SEL_NT, 0, s>>>(h->pairs.as<uint64_t>(), n_pairs, ucnt, upairs);
	HIP_LAUNCH_CHECK(h);
	auto verify = h->cfg.measure == 0 && h->cfg.distance > 8 ? k_verify_count<true> : k_verify_count<false>;
	verify<<<(unsigned)nv, SEL_NT, 0, s>>>(upairs, ucnt + nt, idb, h->chars.as<uint8_t>(), h->offs.as<uint32_t>(),
	                                       seq_map, row_base, set_of, h->cfg.distance, h->cfg.measure,
	                                       h->dist.as<int8_t>(), kcnt);
	HIP_LAUNCH_CHECK(h);
	k_scan_single<uint32_t><<<1, SC_NT, 0, s>>>(kcnt, nv, kcnt + nv, ucnt + nt, SEL_NT);
	HIP_LAUNCH_CHECK(h);
	k_keep_write<<<(unsigned)nv, SEL_NT, 0, s>>>(upairs, ucnt + nt, idb, h->dist.as<int8_t>(), kcnt,
	                                             h->out_pairs.as<int2>() + out_off, h->out_dist.as<int8_t>() + out_off);
	HIP_LAUNCH_CHECK(h);
	uint32_t totals[2] = {0, 0};
	HIP_TRY(h, hipMemcpyAsync(&totals[0], ucnt + nt, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
	HIP_TRY(h, hipMemcpyAsync(&totals[1], kcnt + nv, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
	HIP_SYNC(h, s);
	head->n_unique = totals[0];
	head->n_out = totals[1];
	return XTN_OK;
}

// ---- grouping of one key-space chunk's records into buckets --

struct Grouping {
	int64_t nrec = 0;
	int64_t nseq = 0;  // ids of the run: a round over [0, nseq) is not ranged
	int kernel_cap = BK_CAP;  // the bucket kernel instantiation (2048, 3072, 4096)
	int B = 0, SB = 0, nb = 1;
	int hi = 64, rs = 0;  // the chunk's hash bits end at hi; runs are equal rec >> rs
	bool region = false;  // fixed bucket regions (more than 8 bucket bits)
	bool fused_hist = false;
	bool coarse = false;  // region layout whose coarse pass the generator does (segment layout in recs_alt)
	const uint64_t* src = nullptr;  // region layout: the chunk's records read from here, not from recs
	                                // (its slice of the part-ordered records, left intact for later rounds)
	int64_t seg_t0 = -1, seg_nt = 0;  // region layout: the chunk's records are coarse tiles
	                                  // [seg_t0, seg_t0 + seg_nt) of the generator's sub-ranges (recs_all)
	int nseg = 0;
	int64_t sub_cap = 0;
	int64_t rtiles = 0;
	BucketArgs ba{};
};

// sub-ranges per coarse digit in the generator's coarse pass: a generator
// tile takes the one of its index mod 64 (tiles of one sub-range run on one
// XCD), so each slot cursor sees 1/16384 of the tiles' claims (same-address
// atomics serialise)
constexpr int COARSE_SUB_BITS = 6;

// records in recs (written by the generator) -> the bucket layout ba names:
// region layout (exact pass over the top 8 bucket bits into recs_alt, then
// the region scatter back into recs) or, with at most 8 bucket bits, the
// sorted layout with bucket bounds
int issue_grouping(xtn_handle* h, Grouping& g) {
	hipStream_t s = h->stream;
	BucketArgs& ba = g.ba;
	if (g.region && g.coarse) {
		StageScope sg(h, "sort");
		k_seg_tiles<1 << COARSE_SUB_BITS><<<1, SEG_NT, 0, s>>>(h->seg_cur.as<uint32_t>(), g.sub_cap,
		                                                       h->seg_ctl.as<uint32_t>(), h->bcount.as<uint32_t>(), g.nb,
		                                                       ctr_field(h, offsetof(RunCtrs, n_records)));
		HIP_LAUNCH_CHECK(h);
		RegionArgs ra{};
		ra.in = h->recs_alt.as<uint64_t>();
		ra.out = h->recs.as<uint64_t>();
		ra.ctl = h->seg_ctl.as<uint32_t>();
		ra.count = h->bcount.as<uint32_t>();
		ra.stride = ba.cap;
		ra.shift = g.hi - g.B;
		ra.fine_bits = g.B - 8;
		ra.sub_cap = g.sub_cap;
		ra.sub_bits = COARSE_SUB_BITS;
		{
			ProfScope ps(h, "region", 16.0 * (double)g.nrec);
			k_region_scatter<RS_NT, true><<<(unsigned)(g.rtiles + RS_BINS), RS_NT, 0, s>>>(ra);
			HIP_LAUNCH_CHECK(h);
		}
		ba.rec = h->recs.as<uint64_t>();
		ba.bcount = h->bcount.as<uint32_t>();
		ba.bstride = ba.cap;
		ba.bstart = nullptr;
		return XTN_OK;
	}
	if (g.region) {
		StageScope sg(h, "sort");
		uint32_t* hist = h->hist.as<uint32_t>();
		const uint64_t* in = g.src ? g.src : h->recs.as<uint64_t>();
		if (g.seg_t0 >= 0) {
			// the part's digits read from the generator's sub-ranges
			TRY(ensure(h, h->hist, sizeof(uint32_t) * (size_t)(RS_BINS * g.seg_nt)));
			hist = h->hist.as<uint32_t>();
			{
				ProfScope ps(h, "sort_rec_hist", (double)g.nrec * 8.0);
				k_seg_hist<<<(unsigned)g.seg_nt, RS_NT, 0, s>>>(h->recs_all.as<uint64_t>(), h->seg_ctl.as<uint32_t>(),
				                                                g.seg_t0, COARSE_SUB_BITS, g.sub_cap, g.hi - 8, hist);
				HIP_LAUNCH_CHECK(h);
			}
			uint32_t* dtot = nullptr;
			TRY(digit_scan(h, hist, g.seg_nt, &dtot));
			k_coarse_tiles<<<1, RS_BINS, 0, s>>>(dtot, h->ctl.as<uint32_t>());
			HIP_LAUNCH_CHECK(h);
			ProfScope ps(h, "sort_rec_scatter", 16.0 * (double)g.nrec);
			k_seg_scatter<<<(unsigned)g.seg_nt, RS_NT, 0, s>>>(h->recs_all.as<uint64_t>(), h->seg_ctl.as<uint32_t>(),
			                                                   g.seg_t0, g.seg_nt, COARSE_SUB_BITS, g.sub_cap,
			                                                   g.hi - 8, hist, h->recs_alt.as<uint64_t>(), g.nrec);
			HIP_LAUNCH_CHECK(h);
		} else {
			if (!g.fused_hist) {
				ProfScope ps(h, "sort_rec_hist", (double)g.nrec * 8.0);
				k_radix_hist<<<(unsigned)g.rtiles, RS_NT, 0, s>>>(in, g.nrec, g.hi - 8, hist, g.rtiles);
				HIP_LAUNCH_CHECK(h);
			}
			uint32_t* dtot = nullptr;
			TRY(digit_scan(h, hist, g.rtiles, &dtot));
			k_coarse_tiles<<<1, RS_BINS, 0, s>>>(dtot, h->ctl.as<uint32_t>());
			HIP_LAUNCH_CHECK(h);
			ProfScope ps(h, "sort_rec_scatter", 16.0 * (double)g.nrec);
			k_radix_scatter<false, RS_NT, true><<<(unsigned)g.rtiles, RS_NT, 0, s>>>(
			    in, nullptr, h->recs_alt.as<uint64_t>(), nullptr, g.nrec, g.hi - 8, hist, g.rtiles);
			HIP_LAUNCH_CHECK(h);
		}
		k_fill_u64<<<grid_for(g.nb / 2, 256), 256, 0, s>>>(h->bcount.as<unsigned long long>(), g.nb / 2, 0ull);
		HIP_LAUNCH_CHECK(h);
		RegionArgs ra{};
		ra.in = h->recs_alt.as<uint64_t>();
		ra.out = h->recs.as<uint64_t>();
		ra.ctl = h->ctl.as<uint32_t>();
		ra.count = h->bcount.as<uint32_t>();
		ra.stride = ba.cap;
		ra.shift = g.hi - g.B;
		ra.fine_bits = g.B - 8;
		{
			ProfScope ps(h, "region", 16.0 * (double)g.nrec);
			k_region_scatter<<<(unsigned)(g.rtiles + RS_BINS), RS_NT, 0, s>>>(ra);
			HIP_LAUNCH_CHECK(h);
		}
		ba.rec = h->recs.as<uint64_t>();
		ba.bcount = h->bcount.as<uint32_t>();
		ba.bstride = ba.cap;
		ba.bstart = nullptr;
		return XTN_OK;
	}
	if (g.B > 0) {
		StageScope sg(h, "sort");
		TRY(radix_sort(h, h->recs, h->recs_alt, nullptr, nullptr, g.nrec, g.hi - g.SB, g.hi, "sort_rec",
		               g.fused_hist));
	}
	ba.rec = h->recs.as<uint64_t>();
	ba.bcount = nullptr;
	ba.bstart = h->bstart.as<uint32_t>();
	k_bucket_bounds<<<grid_for(g.nb + 1, 256), 256, 0, s>>>(ba.rec, g.nrec, g.B > 0 ? g.hi - g.B : 0,
	                                                        (uint32_t)(g.nb - 1), g.nb, h->bstart.as<uint32_t>());
	HIP_LAUNCH_CHECK(h);
	return XTN_OK;
}

// a bucket outgrew its region: the chunk's records, ordered by the top 8
// bucket bits in recs_alt by the first pass, are sorted exactly by every
// bucket bit and bucketed by bounds
int regroup_exact(xtn_handle* h, Grouping& g) {
	if (g.coarse) {
		// the generator's segment layout, compacted into recs; it holds the
		// digits' totals rather than nrec records (the generator writes no
		// record for a repeated variant of a sequence). The caller has just
		// synchronised: the totals are read back here
		const int nsub = 1 << COARSE_SUB_BITS;
		std::vector<uint32_t> ctl((size_t)seg_words(nsub));
		HIP_TRY(h, hipMemcpyAsync(ctl.data(), h->seg_ctl.p, sizeof(uint32_t) * ctl.size(), hipMemcpyDeviceToHost,
		                          h->stream));
		HIP_SYNC(h, h->stream);
		int64_t total = 0;
		for (int c = 0; c < RS_BINS; ++c)
			total += ctl[(size_t)SEG_PRE + (size_t)c * seg_stride(nsub) + (size_t)nsub];
		if (total > g.nrec)
			return fail(h, XTN_ERROR, "sub-range totals %lld exceed the %lld records generated", (long long)total,
			            (long long)g.nrec);
		k_seg_compact<<<(unsigned)(g.rtiles + RS_BINS), 256, 0, h->stream>>>(
		    h->recs_alt.as<uint64_t>(), h->recs.as<uint64_t>(), h->seg_ctl.as<uint32_t>(), COARSE_SUB_BITS, g.sub_cap,
		    0, -1);
		HIP_LAUNCH_CHECK(h);
		g.nrec = total;
		g.coarse = false;
	} else {
		std::swap(h->recs, h->recs_alt);
	}
	g.src = nullptr;
	g.seg_t0 = -1;
	g.region = false;
	g.fused_hist = false;
	return issue_grouping(h, g);
}

int issue_buckets(xtn_handle* h, Grouping& g, uint32_t lo, uint32_t hi, int64_t pcap) {
	BucketArgs& ba = g.ba;
	ba.pairs = h->pairs.as<uint64_t>();
	ba.pcap = pcap;
	ba.id_lo = lo;
	ba.id_hi = hi;
	const bool ranged = lo > 0 || (int64_t)hi < g.nseq;
	StageScope sg(h, "pairs");
	ProfScope ps(h, "bucket", (double)g.nrec * 8.0);
	// persistent workgroups, as many per CU as the LDS of the instantiation
	// leaves room for (160 KB per CU)
	const int kcap = g.kernel_cap;
	const int per_cu = kcap <= 3072 ? 3 : 2;
	const unsigned grid = (unsigned)std::min<int64_t>(g.nb, (int64_t)h->n_cu * per_cu);
	hipStream_t s = h->stream;
	if (kcap <= 2048)
		ranged ? k_bucket<true, 2048><<<grid, BK_NT, 0, s>>>(ba) : k_bucket<false, 2048><<<grid, BK_NT, 0, s>>>(ba);
	else if (kcap <= 3072)
		ranged ? k_bucket<true, 3072><<<grid, BK_NT, 0, s>>>(ba) : k_bucket<false, 3072><<<grid, BK_NT, 0, s>>>(ba);
	else
		ranged ? k_bucket<true, 4096><<<grid, BK_NT, 0, s>>>(ba) : k_bucket<false, 4096><<<grid, BK_NT, 0, s>>>(ba);
	HIP_LAUNCH_CHECK(h);
	return XTN_OK;
}

// The candidates of one chunk with the host checking each step: the bucket
// kernel (candidate counter starting at *n_pairs), the pair buffer regrown
// when it overflowed (keeping the *n_pairs already held), a region layout
// whose bucket overflowed regrouped exactly, buckets too large for LDS
// through the global path. *split is set when the round's candidates would
// exceed pmax; *pcap follows the buffer.
int checked_candidates(xtn_handle* h, Grouping& g, uint32_t lo, uint32_t hi, int64_t pmax, int64_t* pcap,
                       int64_t* n_pairs, bool* split) {
	hipStream_t s = h->stream;
	const int64_t n0 = *n_pairs;
	for (int attempt = 0; attempt < 6; ++attempt) {
		TRY(set_ctr(h, offsetof(RunCtrs, append), 0ull));
		TRY(set_ctr(h, offsetof(RunCtrs, n_cand), (unsigned long long)n0));
		TRY(issue_buckets(h, g, lo, hi, *pcap));
		RunCtrs rc{};
		TRY(read_ctrs(h, &rc));
		if ((int64_t)rc.n_cand > *pcap) {
			const int64_t need = (int64_t)rc.n_cand + (int64_t)rc.n_cand / 4;
			if ((int64_t)rc.n_cand > pmax) {
				*split = true;
				return XTN_OK;
			}
			const int64_t grow = std::min(need, pmax);
			TRY(grow_keep(h, h->pairs, sizeof(uint64_t) * (size_t)grow, sizeof(uint64_t) * (size_t)n0));
			*pcap = grow;
			continue;
		}
		if (rc.append > 0 && g.region) {
			TRY(regroup_exact(h, g));
			continue;
		}
		int64_t np = (int64_t)rc.n_cand;
		if (rc.append > 0) {
			StageScope sg(h, "pairs");
			const int64_t n_ovf = (int64_t)rc.append;
			std::vector<uint32_t> ids((size_t)n_ovf), bs((size_t)g.nb + 1);
			HIP_TRY(h, hipMemcpyAsync(ids.data(), h->ovf.p, sizeof(uint32_t) * (size_t)n_ovf, hipMemcpyDeviceToHost, s));
			HIP_TRY(h, hipMemcpyAsync(bs.data(), h->bstart.p, sizeof(uint32_t) * ((size_t)g.nb + 1),
			                          hipMemcpyDeviceToHost, s));
			HIP_SYNC(h, s);
			for (uint32_t b : ids) {
				TRY(bucket_fallback(h, bs[b], (int64_t)bs[b + 1] - bs[b], g.rs, g.hi - g.B, g.ba.idb, lo, hi, pmax,
				                    &np, split));
				if (*split)
					return XTN_OK;
			}
			*pcap = std::max<int64_t>(*pcap, np);
		}
		*n_pairs = np;
		return XTN_OK;
	}
	return fail(h, XTN_ERR_HIP, "candidate search did not settle");
}

// Runs the candidate search and verification over `nseq` ids. seq_map (device,
// nullable) maps an id to its input row. Leaves sorted (x, y, distance)
// triples in out_pairs/out_dist and their count in h->out_len.
//
// One key-space chunk (the usual case on a 288 GB device): generation,
// grouping, the bucket kernel and the whole tail are issued back to back and
// the host reads the counters once at the end. Only a pair buffer that was
// too small or a bucket too large for LDS (both seen in those counters) send
// the round through checked_candidates. Candidates beyond the round budget
// (pmax) split the smaller-id range into rounds whose outputs follow each
// other in order (the reference's lower-bound rounds, xtn.cu:385-470).
// the generator with the records' first ordering pass; distances up to 3 take
// the instantiation without the unranking of larger deletion sets
void launch_gen_coarse(const GenArgs& ga, unsigned grid, hipStream_t s) {
	if (ga.distance <= 3)
		k_gen<GEN_COARSE, GEN_COARSE_NT, true><<<grid, GEN_COARSE_NT, 0, s>>>(ga);
	else
		k_gen<GEN_COARSE, GEN_COARSE_NT><<<grid, GEN_COARSE_NT, 0, s>>>(ga);
}

int nn_pipeline(xtn_handle* h, int32_t nseq, const uint32_t* seq_map, uint32_t row_base = 0,
                const uint32_t* set_of = nullptr) {
	hipStream_t s = h->stream;
	const int d = h->cfg.distance;
	xtn_stats& st = h->stats;
	const double t0 = now_ms();
	h->out_len = 0;
	st.n_sequences = nseq;
	if (nseq < 2) {
		HIP_SYNC(h, s);
		st.ms_total = now_ms() - t0;
		return XTN_OK;
	}
	const int idb = bits_for(nseq - 1);
	if (2 * idb > 62)
		return fail(h, XTN_ERR_CAPACITY, "too many sequences (%d)", nseq);

	// pack + variant offsets
	TRY(ensure(h, h->pk, sizeof(uint64_t) * 2 * (size_t)nseq));
	TRY(ensure(h, h->voff, sizeof(uint64_t) * ((size_t)nseq + 1)));
	TRY(ensure(h, h->ctrs, sizeof(CtlBlock)));
	uint64_t V = 0;
	if (!seq_map && (int64_t)row_base + nseq <= (int64_t)h->prefix_len.size()) {
		// the closed form the device scan sums
		if (h->vprefix_d != d || h->vprefix.size() != h->prefix_len.size() + 1) {
			uint64_t per[19];
			for (int L = 0; L <= 18; ++L)
				per[L] = host_variant_count((uint32_t)L, d);
			h->vprefix.assign(h->prefix_len.size() + 1, 0);
			for (size_t i = 0; i < h->prefix_len.size(); ++i)
				h->vprefix[i + 1] = h->vprefix[i] + per[h->prefix_len[i]];
			h->vprefix_d = d;
		}
		V = h->vprefix[row_base + nseq] - h->vprefix[row_base];
	}
	auto issue_pack = [&]() -> int {
		StageScope sg(h, "generate");
		{
			ProfScope ps(h, "pack", (double)nseq * (16.0 + 16.0 + 8.0));
			k_pack<<<grid_for(nseq, 256), 256, 0, s>>>(h->chars.as<uint8_t>(), h->offs.as<uint32_t>(), seq_map,
			                                            row_base, set_of, nseq, d, h->pk.as<uint64_t>(),
			                                            h->voff.as<uint64_t>());
			HIP_LAUNCH_CHECK(h);
		}
		TRY(scan_excl<uint64_t>(h, h->voff.as<uint64_t>(), h->voff.as<uint64_t>(), nseq,
		                        h->voff.as<uint64_t>() + nseq));
		return XTN_OK;
	};
	const bool v_known = V > 0;
	if (!v_known) {
		TRY(issue_pack());
		HIP_TRY(h, hipMemcpyAsync(&V, h->voff.as<uint64_t>() + nseq, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
		HIP_SYNC(h, s);
	}
	st.n_variants = (int64_t)V;

	// chunk plan: the records of one chunk (plus the sort buffer and the
	// bucket regions, ~40 B per record in all) must fit the HBM budget, and
	// its buckets must stay near the target record count
	if (!h->free_valid) {
		size_t total_b = 0;
		HIP_TRY(h, hipMemGetInfo(&h->free_cache, &total_b));
		h->free_valid = true;
	}
	const size_t free_b = h->free_cache;
	int64_t budget = h->cfg.max_records_per_chunk;
	if (budget <= 0) {
		budget = (int64_t)((double)free_b * 0.6 / 40.0);
		budget = std::min<int64_t>(budget, (int64_t)65536 * h->tune_target);
	}
	budget = std::max<int64_t>(budget, 1 << 16);
	int pbits = 0;
	while (((int64_t)V >> pbits) > budget && pbits < 12)
		++pbits;
	// partitions are hash buckets: give them 25% headroom over the mean
	while (pbits > 0 && pbits < 12 && ((int64_t)V >> pbits) + ((int64_t)V >> (pbits + 2)) > budget)
		++pbits;
	const int nparts = 1 << pbits;
	const int hash_bits = 64 - idb;
	const int K = std::min(KEY_BITS, hash_bits - pbits);
	if (K < 16)
		return fail(h, XTN_ERR_CAPACITY, "too few key bits (%d sequences, %d chunks)", nseq, nparts);
	const int hi = 64 - pbits;
	const int rs = hi - K;
	st.n_chunks = nparts;

	// candidates per round: the look-back counts (30 bits) and ~30 bytes of
	// HBM per candidate (two sort buffers, the outputs, the state)
	int64_t pmax = h->cfg.max_pairs_per_round > 0 ? h->cfg.max_pairs_per_round
	                                               : (int64_t)((double)free_b * 0.3 / 30.0);
	pmax = std::max<int64_t>(std::min<int64_t>(pmax, (int64_t)LB_VAL), 1024);

	GenArgs ga{};
	ga.pk = h->pk.as<uint64_t>();
	ga.voff = h->voff.as<uint64_t>();
	ga.n = nseq;
	ga.distance = d;
	ga.idb = idb;
	ga.part_bits = pbits;
	ga.V = (int64_t)V;
	ga.set_of = set_of;
	// packing (unless done above to learn V) and the generator's tile table
	bool prologue_done = false;
	auto prologue = [&]() -> int {
		if (prologue_done)
			return XTN_OK;
		if (v_known && nparts == 1 && !h->tail_safe) {
			// packing, offsets and tile table in one launch (its look-back
			// shares the error flag the single-chunk run reads at its end)
			const int64_t nb = (nseq + PK_SEQ - 1) / PK_SEQ;
			const uint64_t ntiles = (V + GEN_TILE - 1) / GEN_TILE;
			TRY(ensure(h, h->pack_state, sizeof(uint64_t) * (size_t)nb));
			TRY(ensure(h, h->tseq, sizeof(uint32_t) * (size_t)(ntiles + 1)));
			StageScope sg(h, "generate");
			k_fill_u64<<<grid_for(nb, 256), 256, 0, s>>>(h->pack_state.as<unsigned long long>(), nb, 0ull);
			HIP_LAUNCH_CHECK(h);
			ProfScope ps(h, "pack", (double)nseq * (16.0 + 16.0 + 8.0));
			k_pack_scan<<<(unsigned)nb, PK_NT, 0, s>>>(h->chars.as<uint8_t>(), h->offs.as<uint32_t>(), seq_map, row_base,
			                                         set_of, nseq, d, h->pk.as<uint64_t>(), h->voff.as<uint64_t>(),
			                                         h->tseq.as<uint32_t>(), h->pack_state.as<uint64_t>(),
			                                         LB_SPIN_LIMIT, &ctl_block(h)->tc.error);
			HIP_LAUNCH_CHECK(h);
			ga.tile_seq = h->tseq.as<uint32_t>();
		} else {
			if (v_known)
				TRY(issue_pack());
			TRY(tile_seq_table(h, s, nseq, V, &ga.tile_seq));
		}
		prologue_done = true;
		return XTN_OK;
	};
	const unsigned gen_grid = grid_for((int64_t)V, GEN_TILE);

	std::vector<unsigned long long> part_count(nparts, 0);
	std::vector<int64_t> part_start(nparts + 1, 0);
	int64_t cap_recs = (int64_t)V;
	// Chunks cut by bucket size rather than by memory (HBM holds two full
	// record buffers besides a chunk's working set): every record is
	// generated once, ordered by its top 8 hash bits (the part bits lead
	// them), and each chunk copies its part's slice instead of generating
	// all V records again to keep its own
	const bool gen_once = nparts > 1 && pbits <= 8 && h->cfg.max_records_per_chunk <= 0 && !h->no_alloc &&
	                      (double)V * 16.0 + (double)(V >> pbits) * 1.25 * 40.0 <= (double)free_b * 0.6;
	// generated once, through the generator's coarse pass when it applies:
	// the records land in sub-ranges of their top 8 hash bits, and each part
	// (a range of those digits) is read from there by its first pass
	bool gen_seg = false;
	std::vector<int64_t> part_t0(nparts + 1, 0);
	int64_t seg_sub_cap = 0;
	if (gen_once && h->cfg.coarse_separate == 0 && h->tune_coarse_sep == 0 && !h->coarse_off) {
		TRY(prologue());
		const int nseg = RS_BINS << COARSE_SUB_BITS;
		const double mean = (double)V / nseg;
		seg_sub_cap = (int64_t)(mean * 1.1 + 8.0 * std::sqrt(mean) + 64.0) | 1;
		TRY(ensure(h, h->recs_all, sizeof(uint64_t) * (size_t)(seg_sub_cap * nseg)));
		TRY(ensure(h, h->seg_cur, sizeof(uint32_t) * (size_t)nseg));
		TRY(ensure(h, h->seg_ctl, sizeof(uint32_t) * (size_t)seg_words(1 << COARSE_SUB_BITS)));
		TRY(set_ctr(h, offsetof(RunCtrs, coarse_ovf), 0ull));
		k_fill_u64<<<grid_for(nseg / 2, 256), 256, 0, s>>>(h->seg_cur.as<unsigned long long>(), nseg / 2, 0ull);
		HIP_LAUNCH_CHECK(h);
		ga.out = h->recs_all.as<uint64_t>();
		ga.cap = seg_sub_cap * nseg;
		ga.digit_hist = nullptr;
		ga.digit_shift = 56;
		ga.sub_bits = COARSE_SUB_BITS;
		ga.cursor = h->seg_cur.as<uint32_t>();
		ga.coarse_ovf = ctr_field(h, offsetof(RunCtrs, coarse_ovf));
		ga.sub_cap = seg_sub_cap;
		{
			StageScope sg(h, "generate");
			ProfScope ps(h, "gen_coarse", (double)V * 8.0 + (double)nseq * 24.0);
			launch_gen_coarse(ga, gen_grid, s);
			HIP_LAUNCH_CHECK(h);
		}
		k_seg_tiles<1 << COARSE_SUB_BITS><<<1, SEG_NT, 0, s>>>(h->seg_cur.as<uint32_t>(), seg_sub_cap,
		                                                       h->seg_ctl.as<uint32_t>(), nullptr, 0, nullptr);
		HIP_LAUNCH_CHECK(h);
		std::vector<uint32_t> segh((size_t)seg_words(1 << COARSE_SUB_BITS));
		HIP_TRY(h, hipMemcpyAsync(segh.data(), h->seg_ctl.p, sizeof(uint32_t) * segh.size(), hipMemcpyDeviceToHost, s));
		RunCtrs rc{};
		TRY(read_ctrs(h, &rc));
		if (rc.coarse_ovf) {
			h->coarse_off = true;  // a sub-range overflowed: the separate pass below
		} else {
			gen_seg = true;
			const int per = RS_BINS >> pbits;  // top-8-bit digits per part
			const int nsub = 1 << COARSE_SUB_BITS;
			cap_recs = 0;
			for (int p = 0; p < nparts; ++p) {
				unsigned long long c = 0;
				for (int k = 0; k < per; ++k)
					c += segh[SEG_PRE + (size_t)(p * per + k) * seg_stride(nsub) + nsub];
				part_count[p] = c;
				part_t0[p] = segh[SEG_TSTART + p * per];
				cap_recs = std::max<int64_t>(cap_recs, (int64_t)c);
			}
			part_t0[nparts] = segh[SEG_TSTART + RS_BINS];
		}
	}
	if (gen_once && !gen_seg) {
		TRY(prologue());
		const int64_t vt = ((int64_t)V + RS_TILE - 1) / RS_TILE;
		TRY(ensure(h, h->recs_all, sizeof(uint64_t) * (size_t)V));
		TRY(ensure(h, h->recs_part, sizeof(uint64_t) * (size_t)V));
		TRY(ensure(h, h->hist, sizeof(uint32_t) * (size_t)(RS_BINS * vt)));
		uint32_t* hist = h->hist.as<uint32_t>();
		ga.out = h->recs_all.as<uint64_t>();
		ga.cap = (int64_t)V;
		ga.digit_hist = hist;
		ga.ntiles = vt;
		ga.digit_shift = 56;
		{
			StageScope sg(h, "generate");
			ProfScope ps(h, "gen", (double)V * 8.0 + (double)nseq * 24.0 + 1024.0 * (double)vt);
			k_gen<GEN_DIRECT><<<gen_grid, GEN_NT, 0, s>>>(ga);
			HIP_LAUNCH_CHECK(h);
		}
		ga.digit_hist = nullptr;
		uint32_t* dtot = nullptr;
		TRY(digit_scan(h, hist, vt, &dtot));
		{
			StageScope sg(h, "sort");
			ProfScope ps(h, "part_scatter", 16.0 * (double)V);
			k_radix_scatter<false, RS_NT, true><<<(unsigned)vt, RS_NT, 0, s>>>(
			    h->recs_all.as<uint64_t>(), nullptr, h->recs_part.as<uint64_t>(), nullptr, (int64_t)V, 56, hist, vt);
			HIP_LAUNCH_CHECK(h);
		}
		uint32_t tot[RS_BINS];
		HIP_TRY(h, hipMemcpyAsync(tot, dtot, sizeof(tot), hipMemcpyDeviceToHost, s));
		HIP_SYNC(h, s);
		const int per = RS_BINS >> pbits;  // top-8-bit digits per part
		for (int p = 0; p < nparts; ++p) {
			unsigned long long c = 0;
			for (int k = 0; k < per; ++k)
				c += tot[p * per + k];
			part_count[p] = c;
			part_start[p + 1] = part_start[p] + (int64_t)c;
		}
		cap_recs = 0;
		for (auto c : part_count)
			cap_recs = std::max<int64_t>(cap_recs, (int64_t)c);
	} else if (gen_seg) {
		// part counts and tile ranges read back above
	} else if (nparts > 1) {
		TRY(prologue());
		TRY(ensure(h, h->part_hist, sizeof(unsigned long long) * nparts));
		k_fill_u64<<<grid_for(nparts, 256), 256, 0, s>>>(h->part_hist.as<unsigned long long>(), nparts, 0ull);
		ga.part_hist = h->part_hist.as<unsigned long long>();
		{
			StageScope sg(h, "generate");
			ProfScope ps(h, "gen_count", (double)nseq * 16.0);
			k_gen<GEN_HIST><<<gen_grid, GEN_NT, 0, s>>>(ga);
			HIP_LAUNCH_CHECK(h);
		}
		HIP_TRY(h, hipMemcpyAsync(part_count.data(), h->part_hist.p, sizeof(unsigned long long) * nparts,
		                          hipMemcpyDeviceToHost, s));
		HIP_SYNC(h, s);
		cap_recs = 0;
		for (auto c : part_count)
			cap_recs = std::max<int64_t>(cap_recs, (int64_t)c);
	} else {
		part_count[0] = V;
	}
	if (cap_recs > (int64_t)INT32_MAX)
		return fail(h, XTN_ERR_CAPACITY, "chunk of %lld records exceeds 2^31", (long long)cap_recs);
	TRY(ensure(h, h->recs, sizeof(uint64_t) * (size_t)cap_recs));
	TRY(ensure(h, h->recs_alt, sizeof(uint64_t) * (size_t)cap_recs));
	TRY(ensure(h, h->runs_small, sizeof(Run) * (size_t)(cap_recs / 2 + 1)));
	TRY(ensure(h, h->runs_big, sizeof(Run) * (size_t)(cap_recs / (SMALL_RUN_MAX + 1) + 1)));
	TRY(ensure(h, h->bstart, sizeof(uint32_t) * (65536 + 1)));
	TRY(ensure(h, h->bcount, sizeof(uint32_t) * 65536));
	TRY(ensure(h, h->ctl, sizeof(uint32_t) * CTL_WORDS));
	TRY(ensure(h, h->ovf, sizeof(uint32_t) * 65536));
	const int bucket_cap = h->cfg.bucket_capacity > 0 ? std::min(h->cfg.bucket_capacity, h->tune_cap) : h->tune_cap;

	// the grouping plan of a chunk of nrec records
	auto plan_chunk = [&](int64_t nrec, Grouping& g) -> int {
		g = Grouping{};
		g.nrec = nrec;
		g.nseq = nseq;
		g.kernel_cap = h->tune_cap;
		// bucket bits: the B hash bits right below the chunk bits, B such that
		// buckets hold about tune_target records; the sort orders the
		// records by the next multiple of 8 bits
		int B = 0;
		while (B < 16 && (nrec >> B) > h->tune_target)
			++B;
		g.B = std::min(B, K);
		g.SB = (g.B + 7) / 8 * 8;
		g.nb = 1 << g.B;
		g.hi = hi;
		g.rs = rs;
		// more than 8 bucket bits: an exact pass over the top 8, then the
		// region scatter (fixed regions of bucket_cap slots in recs)
		// (the no-allocation form keeps to the sorted layout: its scratch holds
		// two record buffers, not the regions)
		g.region = g.B > 8 && !h->no_alloc;
		TRY(ensure(h, h->recs, sizeof(uint64_t) * (size_t)cap_recs));
		TRY(ensure(h, h->recs_alt, sizeof(uint64_t) * (size_t)cap_recs));
		if (g.region)
			TRY(ensure(h, h->recs, sizeof(uint64_t) * (size_t)((int64_t)bucket_cap << g.B)));
		// a single chunk on the region layout: the generator does the coarse
		// pass itself, into 2^COARSE_SUB_BITS sub-ranges per coarse digit of
		// sub_cap slots (the mean with 10 % and eight standard deviations of
		// headroom; a sub-range that still overflows sends the run back to the
		// separate pass)
		g.coarse = g.region && nparts == 1 && !h->coarse_off && !h->tail_safe && h->cfg.coarse_separate == 0 &&
		           h->tune_coarse_sep == 0;
		if (g.coarse) {
			g.nseg = RS_BINS << COARSE_SUB_BITS;
			const double mean = (double)nrec / g.nseg;
			g.sub_cap = (int64_t)(mean * 1.1 + 8.0 * std::sqrt(mean) + 64.0) | 1;
			TRY(ensure(h, h->recs_alt, sizeof(uint64_t) * (size_t)(g.sub_cap * g.nseg)));
			TRY(ensure(h, h->seg_cur, sizeof(uint32_t) * (size_t)g.nseg));
			TRY(ensure(h, h->seg_ctl, sizeof(uint32_t) * (size_t)seg_words(1 << COARSE_SUB_BITS)));
		}
		// the single-chunk generator also counts the first sort digit per tile
		g.fused_hist = nparts == 1 && g.B > 0 && !g.coarse;
		g.rtiles = (nrec + RS_TILE - 1) / RS_TILE;
		TRY(ensure(h, h->hist, sizeof(uint32_t) * (size_t)(RS_BINS * g.rtiles)));
		BucketArgs& ba = g.ba;
		ba.nb = g.nb;
		// bins: the (up to) 12 stored hash bits right below the bucket bits
		ba.bin_bits = std::min(BK_BIN_BITS, hi - g.B - idb);
		ba.bin_shift = hi - g.B - ba.bin_bits;
		ba.idb = idb;
		ba.cap = bucket_cap;
		ba.ovf_cap = 65536;
		ba.ctr = h->ctrs.as<RunCtrs>();
		ba.ovf = h->ovf.as<uint32_t>();
		ba.n_ovf = ctr_field(h, offsetof(RunCtrs, append));
		return XTN_OK;
	};
	// the generator launch of one chunk (records into recs, or the segment
	// layout in recs_alt when it does the coarse pass)
	bool used_coarse = false;
	auto issue_gen = [&](int part, Grouping& g) -> int {
		used_coarse = false;
		ga.out = h->recs.as<uint64_t>();
		ga.cap = cap_recs;
		if (g.fused_hist) {
			ga.digit_hist = h->hist.as<uint32_t>();
			ga.ntiles = g.rtiles;
			ga.digit_shift = g.region ? hi - 8 : hi - g.SB;
		} else {
			ga.digit_hist = nullptr;
		}
		StageScope sg(h, "generate");
		if (nparts == 1 && g.coarse) {
			static_assert((RS_BINS << COARSE_SUB_BITS) % 2 == 0, "cursors cleared in u64 words");
			k_fill_u64<<<grid_for(g.nseg / 2, 256), 256, 0, s>>>(h->seg_cur.as<unsigned long long>(), g.nseg / 2, 0ull);
			HIP_LAUNCH_CHECK(h);
			ga.out = h->recs_alt.as<uint64_t>();
			ga.cap = g.sub_cap * g.nseg;
			ga.digit_hist = nullptr;
			ga.digit_shift = hi - 8;
			ga.sub_bits = COARSE_SUB_BITS;
			ga.cursor = h->seg_cur.as<uint32_t>();
			ga.coarse_ovf = ctr_field(h, offsetof(RunCtrs, coarse_ovf));
			ga.sub_cap = g.sub_cap;
			ProfScope ps(h, "gen_coarse", (double)V * 8.0 + (double)nseq * 24.0);
			launch_gen_coarse(ga, gen_grid, s);
			HIP_LAUNCH_CHECK(h);
			used_coarse = true;
		} else if (nparts == 1) {
			ProfScope ps(h, "gen", (double)V * 8.0 + (double)nseq * 24.0 + (g.fused_hist ? 1024.0 * g.rtiles : 0.0));
			k_gen<GEN_DIRECT><<<gen_grid, GEN_NT, 0, s>>>(ga);
			HIP_LAUNCH_CHECK(h);
		} else if (gen_seg && g.region) {
			// the region layout's first pass reads the part's digits from the sub-ranges
			g.seg_t0 = part_t0[part];
			g.seg_nt = part_t0[part + 1] - part_t0[part];
			g.sub_cap = seg_sub_cap;
		} else if (gen_seg) {
			ProfScope ps(h, "part_copy", 16.0 * (double)part_count[part]);
			const int64_t nt = part_t0[part + 1] - part_t0[part];
			if (nt > 0) {
				k_seg_compact<<<(unsigned)nt, 256, 0, s>>>(h->recs_all.as<uint64_t>(), h->recs.as<uint64_t>(),
				                                           h->seg_ctl.as<uint32_t>(), COARSE_SUB_BITS, seg_sub_cap,
				                                           part_t0[part], nt);
				HIP_LAUNCH_CHECK(h);
			}
		} else if (gen_once && g.region) {
			// the region layout's first pass reads the part's slice in place
			g.src = h->recs_part.as<uint64_t>() + part_start[part];
		} else if (gen_once) {
			ProfScope ps(h, "part_copy", 16.0 * (double)part_count[part]);
			if (part_count[part])
				HIP_TRY(h, hipMemcpyAsync(h->recs.p, h->recs_part.as<uint64_t>() + part_start[part],
				                          sizeof(uint64_t) * (size_t)part_count[part], hipMemcpyDeviceToDevice, s));
		} else {
			ga.part_id = (uint32_t)part;
			ga.append = ctr_field(h, offsetof(RunCtrs, append));
			TRY(set_ctr(h, offsetof(RunCtrs, append), 0ull));
			ProfScope ps(h, "gen", (double)part_count[part] * 8.0 + (double)nseq * 24.0);
			k_gen<GEN_APPEND><<<gen_grid, GEN_NT, 0, s>>>(ga);
			HIP_LAUNCH_CHECK(h);
		}
		return XTN_OK;
	};

	// candidate pair capacity: the last run's count with a quarter of
	// headroom (V / 16 before any run); regrown on overflow. The tail's
	// launches are sized by it.
	int64_t pcap = std::min<int64_t>(
	    pmax, std::max<int64_t>(h->last_pairs > 0 ? h->last_pairs + h->last_pairs / 4 : (int64_t)V / 16, 1 << 16));
	if ((size_t)pcap * sizeof(uint64_t) > h->pairs.cap)
		TRY(ensure(h, h->pairs, sizeof(uint64_t) * (size_t)pcap));

	// rounds over the smaller id of a pair, processed in ascending order
	std::deque<std::pair<uint32_t, uint32_t>> rounds;
	rounds.emplace_back(0u, (uint32_t)nseq);
	int64_t out_total = 0, max_round = 0;
	Grouping g1;
	bool grouped = false;  // single chunk: the grouping is done once for every round
	bool no_graph = false;  // a replay needed the host: the rest of this run is eager
	TailPlan tp;
	CtlHead head{};
	while (!rounds.empty()) {
		const uint32_t lo = rounds.front().first, hi_id = rounds.front().second;
		bool split = false;
		int64_t n_pairs = 0;
		bool done = false;
		if (nparts == 1 && h->tail_safe) {
			if (!grouped) {
				TRY(prologue());
				TRY(plan_chunk((int64_t)V, g1));
				TRY(issue_gen(0, g1));
				TRY(issue_grouping(h, g1));
				grouped = true;
			}
			TRY(checked_candidates(h, g1, lo, hi_id, pmax, &pcap, &n_pairs, &split));
		} else if (nparts == 1) {
			// everything issued, one readback: the whole run of a single
			// round is a graph once it has run the same way before
			auto issue_all = [&]() -> int {
				// counters first: the generator's coarse pass may flag an overflow
				TRY(zero_ctl(h, true, true));
				if (!grouped) {
					TRY(prologue());
					TRY(plan_chunk((int64_t)V, g1));
					TRY(issue_gen(0, g1));
					TRY(issue_grouping(h, g1));
					grouped = true;
				}
				TRY(tail_prepare(h, pcap, idb, out_total, &tp));
				TRY(issue_buckets(h, g1, lo, hi_id, pcap));
				StageScope sg(h, "verify");
				TRY(issue_tail(h, tp, idb, seq_map, row_base, set_of, out_total));
				return XTN_OK;
			};
			const bool graph_ok = !grouped && !no_graph && h->cfg.no_graph == 0 && !h->cfg.profile &&
			                      !h->cfg.verbose && v_known && lo == 0 && hi_id == (uint32_t)nseq && out_total == 0;
			xtn_handle::GraphSig sig{};
			if (graph_ok) {
				sig.nseq = nseq;
				sig.row_base = row_base;
				sig.V = (int64_t)V;
				sig.pcap = pcap;
				sig.alloc_gen = h->alloc_gen;
				sig.seq_map = seq_map;
				sig.set_of = set_of;
				sig.chars = h->chars.p;
				sig.offs = h->offs.p;
				sig.d = d;
				sig.measure = h->cfg.measure;
				sig.bucket_cap = bucket_cap;
				sig.target = h->tune_target;
				sig.sort_nt = h->tune_sort_nt;
				sig.sort_tile = h->tune_sort_tile;
				sig.select_nt = h->tune_select_nt;
				sig.low = h->tail_full ? 0 : h->tune_low;
				sig.coarse = h->coarse_off || h->cfg.coarse_separate || h->tune_coarse_sep ? 0 : 1;
				sig.pad = 0;
			}
			bool via_graph = false;
			if (graph_ok && h->graph_exec && sig == h->graph_sig) {
				HIP_TRY(h, hipGraphLaunch(h->graph_exec, s));
				st.n_launches += h->graph_launches;
				used_coarse = h->graph_coarse;
				via_graph = true;
			} else if (graph_ok && h->graph_want_set && sig == h->graph_want) {
				// the previous run had this signature: capture this one (every
				// buffer is already allocated; an allocation attempt during the
				// capture abandons it)
				const int64_t l0 = st.n_launches;
				hipGraph_t graph = nullptr;
				hipGraphExec_t exec = nullptr;
				HIP_TRY(h, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
				h->capturing = true;
				int rc = issue_all();
				h->capturing = false;
				hipError_t e = hipStreamEndCapture(s, &graph);
				if (rc == XTN_OK && e == hipSuccess)
					e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
				if (rc == XTN_OK && e == hipSuccess) {
					if (h->graph_exec)
						(void)hipGraphExecDestroy(h->graph_exec);
					if (h->graph)
						(void)hipGraphDestroy(h->graph);
					h->graph = graph;
					h->graph_exec = exec;
					h->graph_sig = sig;
					h->graph_launches = st.n_launches - l0;
					h->graph_coarse = used_coarse;
					HIP_TRY(h, hipGraphLaunch(exec, s));
					via_graph = true;
				} else {
					// nothing of the capture ran: issue the run eagerly
					(void)hipGetLastError();
					if (graph)
						(void)hipGraphDestroy(graph);
					h->graph_want_set = false;
					grouped = false;
					prologue_done = false;
					st.n_launches = l0;
					TRY(issue_all());
				}
			} else {
				TRY(issue_all());
			}
			TRY(read_head(h, &head));
			if (head.rc.coarse_ovf) {
				// a sub-range of the generator's coarse pass overflowed and lost
				// records: the run again with the separate pass, which these
				// sequences keep from now on
				h->coarse_off = true;
				grouped = false;
				prologue_done = false;
				continue;
			}
			const bool clean =
			    !head.error && !head.overflow && (int64_t)head.rc.n_cand <= pcap && head.rc.append == 0;
			if (head.overflow)
				h->tail_full = true;  // a group outgrew the select's window: full sorts from now on
			if (via_graph) {
				st.graph_replays += 1;
				if (!clean) {
					// redo the run eagerly (the host's plan of the layout is
					// rebuilt with it)
					no_graph = true;
					grouped = false;
					prologue_done = false;
					continue;
				}
			}
			if (head.error) {
				// a look-back wait hit its limit: this handle keeps to the
				// host-checked tail from now on, and the round is redone
				// (packing included, whose offsets share the flag)
				h->tail_safe = true;
				grouped = false;
				prologue_done = false;
				continue;
			}
			if (head.overflow)
				continue;  // the round is redone with the full sort
			if (clean) {
				n_pairs = (int64_t)head.rc.n_cand;
				done = true;
				if (graph_ok && !via_graph) {
					h->graph_want = sig;
					h->graph_want_set = true;
				}
			} else {
				TRY(checked_candidates(h, g1, lo, hi_id, pmax, &pcap, &n_pairs, &split));
			}
		} else {
			TRY(set_ctr(h, offsetof(RunCtrs, n_cand), 0ull));
			for (int part = 0; part < nparts && !split; ++part) {
				const int64_t nrec = (int64_t)part_count[part];
				if (nrec < 2)
					continue;
				Grouping g;
				TRY(plan_chunk(nrec, g));
				TRY(issue_gen(part, g));
				TRY(issue_grouping(h, g));
				TRY(checked_candidates(h, g, lo, hi_id, pmax, &pcap, &n_pairs, &split));
			}
		}
		if (split) {
			// more candidates than a round holds: this id range becomes
			// several, each a round of its own
			const uint32_t span = hi_id - lo;
			if (span <= 1)
				return fail(h, XTN_ERR_CAPACITY, "the candidates of one sequence exceed the round budget (%lld)",
				            (long long)pmax);
			rounds.pop_front();
			const uint32_t mid = lo + span / 2;
			rounds.emplace_front(mid, hi_id);
			rounds.emplace_front(lo, mid);
			continue;
		}
		if (!done && h->tail_safe) {
			StageScope sg(h, "verify");
			TRY(tail_rts(h, n_pairs, idb, seq_map, row_base, set_of, out_total, &head));
		} else if (!done) {
			TRY(tail_prepare(h, pcap, idb, out_total, &tp));
			TRY(set_ctr(h, offsetof(RunCtrs, n_cand), (unsigned long long)n_pairs));
			TRY(zero_ctl(h, false, true));
			{
				StageScope sg(h, "verify");
				TRY(issue_tail(h, tp, idb, seq_map, row_base, set_of, out_total));
			}
			TRY(read_head(h, &head));
			if (head.error) {
				h->tail_safe = true;  // see above; the round is redone
				continue;
			}
			if (head.overflow) {
				h->tail_full = true;  // see above; the round is redone
				continue;
			}
		}
		st.n_candidates += n_pairs;
		st.n_unique += (int64_t)head.n_unique;
		out_total += (int64_t)head.n_out;
		max_round = std::max(max_round, n_pairs);
		++st.n_rounds;
		st.tail_checked = h->tail_safe ? 1 : 0;
		st.tail_passes = h->tail_safe ? (2 * idb + 7) / 8 : tail_passes(h, idb);
		st.coarse_in_generator = used_coarse || gen_seg ? 1 : 0;
		rounds.pop_front();
	}
	// records written: fewer than V when the generator's coarse pass left
	// out repeated variants (its total is in the counters read at the end)
	st.n_records = (int64_t)V;
	if (gen_seg) {
		st.n_records = 0;
		for (int p = 0; p < nparts; ++p)
			st.n_records += (int64_t)part_count[p];
	} else if (used_coarse) {
		if (head.rc.n_records == 0) {  // no tail ran: the counters were not read back
			RunCtrs rc{};
			TRY(read_ctrs(h, &rc));
			head.rc.n_records = rc.n_records;
		}
		st.n_records = (int64_t)head.rc.n_records;
	}
	h->last_pairs = max_round;
	h->last_unique = st.n_unique;
	h->last_out = out_total;
	h->out_len = out_total;
	st.n_output = out_total;
	st.ms_total = now_ms() - t0;
	collect_stages(h);
	collect_profile(h);
	return XTN_OK;
}

// every byte one of the 20 residues ACDEFGHIKLMNPQRSTVWY (codec.cu:25-38):
// branch-free byte compares OR-ed over 64 KiB blocks, which the compiler
// vectorises, with one test per block
bool all_residues(const char* chars, size_t total) {
	const unsigned char* c = (const unsigned char*)chars;
	for (size_t b = 0; b < total; b += 65536) {
		const size_t e = std::min(total, b + 65536);
		uint8_t bad = 0;
		for (size_t k = b; k < e; ++k) {
			const uint8_t x = c[k];
			bad |= (uint8_t)((uint8_t)(x - 'A') > (uint8_t)('Y' - 'A')) | (uint8_t)(x == 'B') |
			       (uint8_t)(x == 'J') | (uint8_t)(x == 'O') | (uint8_t)(x == 'U') | (uint8_t)(x == 'X');
		}
		if (bad)
			return false;
	}
	return true;
}

void print_verbose(xtn_handle* h, const char* mode) {
	if (!h->cfg.verbose)
		return;
	const xtn_stats& st = h->stats;
	printf("mode %s: sequences %lld, variants %lld, chunks %lld, runs %lld, candidates %lld, "
	       "unique %lld, output %lld\n",
	       mode, (long long)st.n_sequences, (long long)st.n_variants, (long long)st.n_chunks,
	       (long long)st.n_runs, (long long)st.n_candidates, (long long)st.n_unique,
	       (long long)st.n_output);
	printf("device ms: total %.3f generate %.3f sort %.3f pairs %.3f verify %.3f output %.3f\n",
	       st.ms_total, st.ms_generate, st.ms_sort, st.ms_pairs, st.ms_verify, st.ms_output);
}

// Repertoire overlap (xtn.cu:281-290 overlap_mode_init, stream 4 overlap and
// mergeOutput): exact-duplicate groups, NN search over one representative
// per group, both overlap terms into a dense R x R matrix, one callback.
int overlap_perform(xtn_handle* h, const xtn_seq_info* info, xtn_callback callback, void* user) {
	hipStream_t s = h->stream;
	const int32_t n = h->n;
	xtn_stats& st = h->stats;
	const double t0 = now_ms();
	if (n == 0) {
		xtn_output out{nullptr, nullptr, nullptr, 0};
		if (callback)
			callback(user, &out);
		return XTN_OK;
	}
	// the rows go up as the caller holds them; the device splits them and
	// finds the repertoire count and any negative repertoire index, read
	// back with the unique count below
	static_assert(sizeof(xtn_seq_info) == 3 * sizeof(int32_t), "packed SeqInfo rows");
	TRY(ensure(h, h->ov_info, sizeof(int32_t) * 2 * (size_t)n));
	TRY(ensure(h, h->ov_keys_alt, sizeof(xtn_seq_info) * (size_t)n));  // staging, reused below
	TRY(ensure(h, h->scalar, sizeof(uint64_t)));
	HIP_TRY(h, hipMemcpyAsync(h->ov_keys_alt.p, info, sizeof(xtn_seq_info) * (size_t)n, hipMemcpyHostToDevice, s));
	k_fill_u64<<<1, 256, 0, s>>>(h->scalar.as<unsigned long long>(), 1, 0xFFFFFFFF00000000ull);
	k_split_info<<<grid_for(n, 256), 256, 0, s>>>(h->ov_keys_alt.as<int32_t>(), n, h->ov_info.as<int32_t>(),
	                                               h->ov_info.as<int32_t>() + n, h->scalar.as<uint32_t>());
	HIP_LAUNCH_CHECK(h);
	const int32_t* freq_d = h->ov_info.as<int32_t>();
	const int32_t* rep_d = freq_d + n;

	// exact duplicates: hash sort, leaders, stable sort by leader position
	const int idb = bits_for(std::max(n - 1, 1));
	const int K = std::min(32, 64 - idb);
	const int rs = 64 - K;
	TRY(ensure(h, h->ov_keys, sizeof(uint64_t) * (size_t)n));
	TRY(ensure(h, h->ov_keys_alt, sizeof(uint64_t) * (size_t)n));
	TRY(ensure(h, h->ov_vals, sizeof(uint32_t) * (size_t)n));
	TRY(ensure(h, h->ov_vals_alt, sizeof(uint32_t) * (size_t)n));
	{
		ProfScope ps(h, "ov_hash", (double)n * 28.0);
		k_hash_rows<<<grid_for(n, 256), 256, 0, s>>>(h->chars.as<uint8_t>(), h->offs.as<uint32_t>(), n, idb,
		                                              h->ov_keys.as<uint64_t>());
		HIP_LAUNCH_CHECK(h);
	}
	TRY(radix_sort(h, h->ov_keys, h->ov_keys_alt, nullptr, nullptr, n, rs, 64, "ov_sort_hash"));
	{
		ProfScope ps(h, "ov_leaders", (double)n * 40.0);
		k_leaders<<<grid_for(n, 256), 256, 0, s>>>(h->ov_keys.as<uint64_t>(), n, rs, idb,
		                                            h->chars.as<uint8_t>(), h->offs.as<uint32_t>(),
		                                            h->ov_keys_alt.as<uint64_t>(), h->ov_vals.as<uint32_t>());
		HIP_LAUNCH_CHECK(h);
	}
	TRY(radix_sort(h, h->ov_keys_alt, h->ov_keys, &h->ov_vals, &h->ov_vals_alt, n, 0, bits_for(n),
	               "ov_sort_leader"));
	const uint64_t* lead = h->ov_keys_alt.as<uint64_t>();
	const uint32_t* rows = h->ov_vals.as<uint32_t>();

	TRY(ensure(h, h->flags, sizeof(uint32_t) * (size_t)n));
	TRY(ensure(h, h->pos, sizeof(uint32_t) * ((size_t)n + 1)));
	k_group_flags<<<grid_for(n, 256), 256, 0, s>>>(lead, n, h->flags.as<uint32_t>());
	HIP_LAUNCH_CHECK(h);
	TRY(scan_excl<uint32_t>(h, h->flags.as<uint32_t>(), h->pos.as<uint32_t>(), n, h->pos.as<uint32_t>() + n));
	uint32_t U = 0, ctl[2] = {0, 0};
	HIP_TRY(h, hipMemcpyAsync(&U, h->pos.as<uint32_t>() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
	HIP_TRY(h, hipMemcpyAsync(ctl, h->scalar.p, sizeof(ctl), hipMemcpyDeviceToHost, s));
	HIP_SYNC(h, s);
	if (ctl[1] != 0xFFFFFFFFu)
		return fail(h, XTN_ERR_ARG, "negative repertoire index at row %u", ctl[1]);
	const int32_t R = (int32_t)ctl[0];
	const int64_t RR = (int64_t)R * R;
	if (RR > ((int64_t)1 << 31))
		return fail(h, XTN_ERR_CAPACITY, "%d repertoires exceed the dense overlap matrix limit (46340)", R);
	TRY(ensure(h, h->ov_ustart, sizeof(uint32_t) * ((size_t)U + 1)));
	TRY(ensure(h, h->ov_umap, sizeof(uint32_t) * (size_t)U));
	k_group_starts<<<grid_for(n, 256), 256, 0, s>>>(h->flags.as<uint32_t>(), h->pos.as<uint32_t>(), rows, n,
	                                                 h->ov_ustart.as<uint32_t>(), h->ov_umap.as<uint32_t>());
	HIP_LAUNCH_CHECK(h);
	const uint32_t nn32 = (uint32_t)n;
	HIP_TRY(h, hipMemcpyAsync(h->ov_ustart.as<uint32_t>() + U, &nn32, sizeof(uint32_t),
	                          hipMemcpyHostToDevice, s));

	// matrix
	TRY(ensure(h, h->ov_matrix, sizeof(unsigned long long) * (size_t)RR));
	TRY(ensure(h, h->ov_present, sizeof(uint32_t) * (size_t)(RR + 2)));
	k_fill_u64<<<grid_for(RR, 256), 256, 0, s>>>(h->ov_matrix.as<unsigned long long>(), RR, 0ull);
	k_fill_u64<<<grid_for((RR + 1) / 2, 256), 256, 0, s>>>(h->ov_present.as<unsigned long long>(), (RR + 1) / 2, 0ull);
	HIP_LAUNCH_CHECK(h);
	OvMatrix m{h->ov_matrix.as<unsigned long long>(), h->ov_present.as<uint32_t>(), R};
	{
		ProfScope ps(h, "ov_exact", (double)n * 16.0);
		const unsigned g = (unsigned)std::min<int64_t>(grid_for(U, 256), (int64_t)h->n_cu * 3);
		k_exact_term<<<g, 256, 0, s>>>(h->ov_ustart.as<uint32_t>(), (int32_t)U, rows, freq_d, rep_d, m);
		HIP_LAUNCH_CHECK(h);
	}
	HIP_SYNC(h, s);
	const double t_dedup = now_ms() - t0;

	// neighbours among the unique strings
	TRY(nn_pipeline(h, (int32_t)U, h->ov_umap.as<uint32_t>()));
	const int64_t npairs = h->out_len;
	if (npairs > 0) {
		ProfScope ps(h, "ov_neighbor", (double)npairs * 40.0);
		const unsigned g = (unsigned)std::min<int64_t>(grid_for(npairs, 256), (int64_t)h->n_cu * 3);
		k_neighbor_term<<<g, 256, 0, s>>>(h->out_pairs.as<int2>(), npairs, h->ov_ustart.as<uint32_t>(), rows,
		                                  freq_d, rep_d, m);
		HIP_LAUNCH_CHECK(h);
	}

	// present entries in (A, B) order
	const double to = now_ms();
	TRY(ensure(h, h->pos, sizeof(uint32_t) * ((size_t)RR + 1)));
	TRY(scan_excl<uint32_t>(h, h->ov_present.as<uint32_t>(), h->pos.as<uint32_t>(), RR,
	                        h->pos.as<uint32_t>() + RR));
	uint32_t nout = 0;
	HIP_TRY(h, hipMemcpyAsync(&nout, h->pos.as<uint32_t>() + RR, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
	HIP_SYNC(h, s);
	TRY(ensure(h, h->out_pairs, sizeof(int2) * (size_t)std::max<uint32_t>(nout, 1)));
	TRY(ensure(h, h->ov_cnt, sizeof(unsigned long long) * (size_t)std::max<uint32_t>(nout, 1)));
	k_matrix_emit<<<grid_for(RR, 256), 256, 0, s>>>(h->ov_present.as<uint32_t>(), h->pos.as<uint32_t>(),
	                                                 h->ov_matrix.as<unsigned long long>(), RR, R,
	                                                 h->out_pairs.as<int2>(), h->ov_cnt.as<unsigned long long>());
	HIP_LAUNCH_CHECK(h);
	std::vector<xtn_int2> hp(std::max<uint32_t>(nout, 1));
	std::vector<uint64_t> hv(std::max<uint32_t>(nout, 1));
	if (nout) {
		HIP_TRY(h, hipMemcpyAsync(hp.data(), h->out_pairs.p, sizeof(xtn_int2) * nout, hipMemcpyDeviceToHost, s));
		HIP_TRY(h, hipMemcpyAsync(hv.data(), h->ov_cnt.p, sizeof(uint64_t) * nout, hipMemcpyDeviceToHost, s));
	}
	HIP_SYNC(h, s);
	collect_profile(h);
	st.n_sequences = U;
	st.n_output = nout;
	st.ms_output = now_ms() - to;
	st.ms_total = now_ms() - t0;
	if (h->cfg.verbose)
		printf("overlap: %d rows, %u unique, %d repertoires, dedup %.3f ms\n", n, U, R, t_dedup);
	print_verbose(h, "overlap");
	xtn_output out;
	out.index_pairs = hp.data();
	out.pairwise_distances = nullptr;
	out.pairwise_frequencies = hv.data();
	out.len = nout;
	if (callback)
		callback(user, &out);
	return XTN_OK;
}

}  // namespace

// =============================================================== C ABI ==

extern "C" {

void xtn_config_init(xtn_config* cfg) {
	if (!cfg)
		return;
	memset(cfg, 0, sizeof(*cfg));
	cfg->distance = 1;
	cfg->measure = XTN_LEVENSHTEIN;
}

static int check_config(xtn_handle* h, const xtn_config* cfg) {
	if (cfg->distance < 1)
		return fail(h, XTN_ERR_ARG, "distance must be a valid number larger than 0");
	if (cfg->distance > 127)
		return fail(h, XTN_ERR_ARG, "distance above 127 is not supported");
	if (cfg->measure != XTN_LEVENSHTEIN && cfg->measure != XTN_HAMMING)
		return fail(h, XTN_ERR_ARG, "invalid measure option");
	return XTN_OK;
}

int xtn_open(xtn_handle** out, const xtn_config* cfg) {
	if (!out)
		return XTN_ERR_ARG;
	*out = nullptr;
	xtn_config c;
	xtn_config_init(&c);
	if (cfg)
		c = *cfg;
	if (check_config(nullptr, &c) != XTN_OK)
		return XTN_ERR_ARG;
	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
		(void)hipGetLastError();
		return XTN_ERR_HIP;
	}
	if (c.device < 0 || c.device >= ndev)
		return XTN_ERR_ARG;
	if (hipSetDevice(c.device) != hipSuccess)
		return XTN_ERR_HIP;
	xtn_handle* h = new xtn_handle();
	h->cfg = c;
	if (const char* e = getenv("XTN_LOOKBACK_SPIN_LIMIT"))
		h->spin_limit = (uint32_t)strtoul(e, nullptr, 10);
	int ncu = 0;
	if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c.device) == hipSuccess && ncu > 0)
		h->n_cu = ncu;
	memset(&h->stats, 0, sizeof(h->stats));
	if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking) != hipSuccess) {
		delete h;
		return XTN_ERR_HIP;
	}
	*out = h;
	return XTN_OK;
}

void xtn_close(xtn_handle* h) {
	if (!h)
		return;
	(void)hipSetDevice(h->cfg.device);
	(void)hipStreamSynchronize(h->stream);
	DBuf* bufs[] = {&h->recs_all, &h->recs_part, &h->seg_cur, &h->seg_ctl, &h->bcount, &h->ctl, &h->chars, &h->offs, &h->pk, &h->voff,
	                &h->recs, &h->recs_alt, &h->hist,
	                &h->scan_tmp, &h->runs_small, &h->runs_big, &h->big_pre, &h->pairs,
	                &h->pairs_alt, &h->flags, &h->pos, &h->dist, &h->out_pairs, &h->out_dist,
	                &h->ctrs, &h->part_hist, &h->scalar, &h->ov_keys, &h->ov_keys_alt, &h->ov_vals,
	                &h->ov_vals_alt, &h->ov_info, &h->ov_leader, &h->ov_ustart, &h->ov_umap,
	                &h->ov_cnt, &h->ov_matrix, &h->ov_present, &h->bstart, &h->ovf, &h->set_of,
	                &h->set_offs, &h->tseq, &h->tail_state, &h->pack_state};
	for (DBuf* b : bufs)
		release(*b);
	for (auto& pe : h->pending) {
		(void)hipEventDestroy(pe.start);
		(void)hipEventDestroy(pe.stop);
	}
	for (hipEvent_t e : h->event_pool)
		(void)hipEventDestroy(e);
	if (h->pin)
		(void)hipHostFree(h->pin);
	if (h->out_pin)
		(void)hipHostFree(h->out_pin);
	for (hipEvent_t e : h->pin_ev)
		if (e)
			(void)hipEventDestroy(e);
	if (h->graph_exec)
		(void)hipGraphExecDestroy(h->graph_exec);
	if (h->graph)
		(void)hipGraphDestroy(h->graph);
	(void)hipStreamDestroy(h->stream);
	(void)hipStreamDestroy(h->copy_stream);
	delete h;
}

int xtn_set_config(xtn_handle* h, const xtn_config* cfg) {
	if (!h || !cfg)
		return XTN_ERR_ARG;
	TRY(check_config(h, cfg));
	if (cfg->device != h->cfg.device)
		return fail(h, XTN_ERR_ARG, "the device of an open handle cannot change");
	h->cfg = *cfg;
	return XTN_OK;
}

}  // extern "C"

namespace {

// runs fn(t, lo, hi) over [0, n) split into up to 16 ranges on host threads
// (one range below `grain` items per thread)
template <typename Fn>
void host_parallel(int64_t n, int64_t grain, Fn fn) {
	int T = (int)std::thread::hardware_concurrency();
	T = T < 1 ? 1 : (T > 16 ? 16 : T);
	while (T > 1 && n / T < grain)
		--T;
	if (T == 1) {
		fn(0, (int64_t)0, n);
		return;
	}
	std::vector<std::thread> th;
	for (int t = 0; t < T; ++t)
		th.emplace_back(fn, t, n * t / T, n * (t + 1) / T);
	for (auto& x : th)
		x.join();
}

// host bytes through the handle's pinned staging (two halves in turn, so
// the copy into one half overlaps the transfer of the other) to dst on the
// handle's stream; waits for the last transfer
int upload_staged(xtn_handle* h, void* dst, const void* src, size_t bytes) {
	constexpr size_t HALF = (size_t)8 << 20;
	if (bytes == 0)
		return XTN_OK;
	if (h->pin_bytes < 2 * HALF) {
		if (h->pin)
			(void)hipHostFree(h->pin);
		h->pin = nullptr;
		h->pin_bytes = 0;
		if (hipHostMalloc(&h->pin, 2 * HALF, hipHostMallocDefault) != hipSuccess) {
			(void)hipGetLastError();
			h->pin = nullptr;
			// no pinned memory: the runtime's own staging
			HIP_TRY(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
			HIP_SYNC(h, h->stream);
			return XTN_OK;
		}
		h->pin_bytes = 2 * HALF;
	}
	if (!h->pin_ev[0]) {
		for (int k = 0; k < 2; ++k)
			if (hipEventCreateWithFlags(&h->pin_ev[k], hipEventDisableTiming) != hipSuccess)
				return fail(h, XTN_ERR_HIP, "staging event creation failed");
	}
	char* pin = static_cast<char*>(h->pin);
	bool used[2] = {false, false};
	for (size_t off = 0, k = 0; off < bytes; off += HALF, ++k) {
		const int half = (int)(k & 1);
		const size_t len = std::min(HALF, bytes - off);
		if (used[half])
			HIP_TRY(h, hipEventSynchronize(h->pin_ev[half]));  // this half's previous transfer is done
		memcpy(pin + half * HALF, static_cast<const char*>(src) + off, len);
		HIP_TRY(h, hipMemcpyAsync(static_cast<char*>(dst) + off, pin + half * HALF, len, hipMemcpyHostToDevice,
		                          h->stream));
		HIP_TRY(h, hipEventRecord(h->pin_ev[half], h->stream));
		used[half] = true;
	}
	HIP_SYNC(h, h->stream);
	return XTN_OK;
}

}  // namespace

extern "C" {

int xtn_set_sequences(xtn_handle* h, const char* chars, const uint32_t* offsets, int32_t n) {
	if (!h || n < 0 || !offsets || (n > 0 && !chars && offsets[n] > 0))
		return h ? fail(h, XTN_ERR_ARG, "invalid sequence arguments") : XTN_ERR_ARG;
	if (offsets[0] != 0)
		return fail(h, XTN_ERR_ARG, "offsets[0] must be 0");
	HIP_TRY(h, hipSetDevice(h->cfg.device));
	// Inputs up to SET_PIN_MAX go up from one pinned copy made by the same
	// host threads that check them: offsets non-decreasing and the prefix
	// lengths, then every residue one of the 20 (codec.cu:25-38); the first
	// bad row is the one reported. Nothing reaches the device unless every
	// check passes, so a rejected input leaves the handle's sequences as they
	// were. Larger inputs are checked first and then staged in halves.
	constexpr size_t SET_PIN_MAX = (size_t)64 << 20;
	const uint32_t total = offsets[n];
	const size_t cbytes = ((size_t)total + 255) & ~(size_t)255, obytes = sizeof(uint32_t) * ((size_t)n + 1);
	char* pin = nullptr;
	if (cbytes + obytes <= SET_PIN_MAX) {
		if (h->pin_bytes < cbytes + obytes) {
			if (h->pin)
				(void)hipHostFree(h->pin);
			h->pin = nullptr;
			h->pin_bytes = 0;
			const size_t want = std::max(cbytes + obytes, (size_t)16 << 20);
			if (hipHostMalloc(&h->pin, want, hipHostMallocDefault) == hipSuccess)
				h->pin_bytes = want;
			else
				(void)hipGetLastError();
		}
		pin = h->pin_bytes >= cbytes + obytes ? static_cast<char*>(h->pin) : nullptr;
	}
	uint32_t* pin_offs = pin ? reinterpret_cast<uint32_t*>(pin + cbytes) : nullptr;
	std::vector<uint8_t> plen((size_t)n);
	std::vector<int64_t> first_bad(16, INT64_MAX);
	host_parallel(n, 1 << 16, [&](int t, int64_t lo, int64_t hi) {
		for (int64_t i = lo; i < hi; ++i) {
			const uint32_t a = offsets[i], b = offsets[i + 1];
			if (b < a) {
				first_bad[t] = i;
				return;
			}
			const uint32_t len = b - a;
			plen[(size_t)i] = (uint8_t)(len < 18 ? len : 18);
		}
		if (pin_offs)
			memcpy(pin_offs + lo, offsets + lo, sizeof(uint32_t) * (size_t)(hi - lo + (hi == n ? 1 : 0)));
	});
	if (n == 0 && pin_offs)
		pin_offs[0] = 0;
	for (int64_t fb : first_bad)
		if (fb != INT64_MAX) {
			int64_t row = INT64_MAX;
			for (int64_t x : first_bad)
				row = std::min(row, x);
			return fail(h, XTN_ERR_ARG, "offsets must be non-decreasing (row %lld)", (long long)row);
		}
	std::vector<char> bad(16, 0);
	host_parallel((int64_t)total, (int64_t)1 << 20, [&](int t, int64_t lo, int64_t hi) {
		if (pin) {
			memcpy(pin + lo, chars + lo, (size_t)(hi - lo));
			bad[t] = all_residues(pin + lo, (size_t)(hi - lo)) ? 0 : 1;
		} else {
			bad[t] = all_residues(chars + lo, (size_t)(hi - lo)) ? 0 : 1;
		}
	});
	for (char x : bad)
		if (x)
			return fail(h, XTN_ERROR, "invalid char");
	TRY(ensure(h, h->chars, (size_t)total + 16));
	TRY(ensure(h, h->offs, obytes));
	if (pin) {
		if (total)
			HIP_TRY(h, hipMemcpyAsync(h->chars.p, pin, total, hipMemcpyHostToDevice, h->stream));
		HIP_TRY(h, hipMemcpyAsync(h->offs.p, pin_offs, obytes, hipMemcpyHostToDevice, h->stream));
		HIP_SYNC(h, h->stream);
	} else {
		TRY(upload_staged(h, h->chars.p, chars, total));
		TRY(upload_staged(h, h->offs.p, offsets, obytes));
	}
	h->n = n;
	h->n_chars = total;
	h->out_len = 0;
	h->prefix_len.swap(plen);
	h->vprefix.clear();
	h->vprefix_d = 0;
	h->tail_full = false;
	h->coarse_off = false;
	return XTN_OK;
}

int xtn_run_nn_resident(xtn_handle* h, int64_t* out_len) {
	if (!h)
		return XTN_ERR_ARG;
	if (!h->offs.p)
		return fail(h, XTN_ERR_STATE, "no sequences set");
	HIP_TRY(h, hipSetDevice(h->cfg.device));
	memset(&h->stats, 0, sizeof(h->stats));
	TRY(nn_pipeline(h, h->n, nullptr));
	if (out_len)
		*out_len = h->out_len;
	print_verbose(h, "nn");
	return XTN_OK;
}

int xtn_copy_nn_output(xtn_handle* h, xtn_int2* pairs, int8_t* distances, int64_t cap) {
	if (!h || (cap > 0 && (!pairs || !distances)))
		return XTN_ERR_ARG;
	const int64_t n = std::min<int64_t>(cap, h->out_len);
	if (n <= 0)
		return XTN_OK;
	HIP_TRY(h, hipMemcpyAsync(pairs, h->out_pairs.p, sizeof(xtn_int2) * (size_t)n,
	                          hipMemcpyDeviceToHost, h->stream));
	HIP_TRY(h, hipMemcpyAsync(distances, h->out_dist.p, (size_t)n, hipMemcpyDeviceToHost, h->stream));
	HIP_SYNC(h, h->stream);
	return XTN_OK;
}

int xtn_perform(xtn_handle* h, const xtn_seq_info* info, xtn_callback callback, void* user) {
	if (!h)
		return XTN_ERR_ARG;
	if (!h->offs.p)
		return fail(h, XTN_ERR_STATE, "no sequences set");
	HIP_TRY(h, hipSetDevice(h->cfg.device));
	memset(&h->stats, 0, sizeof(h->stats));
	if (info)
		return overlap_perform(h, info, callback, user);

	TRY(nn_pipeline(h, h->n, nullptr));
	print_verbose(h, "nn");
	const double t0 = now_ms();
	const int64_t n = h->out_len;
	int64_t chunk = h->cfg.output_chunk_pairs > 0 ? h->cfg.output_chunk_pairs : ((int64_t)1 << 22);
	if (n == 0) {
		h->stats.ms_output = now_ms() - t0;
		return XTN_OK;
	}
	chunk = std::min(chunk, n);
	// double-buffered pinned staging (kept by the handle from call to call):
	// chunk k+1 is copied on the side stream while the callback consumes
	// chunk k
	xtn_int2* hp[2] = {nullptr, nullptr};
	int8_t* hd[2] = {nullptr, nullptr};
	hipEvent_t ready[2] = {nullptr, nullptr};
	int rc = XTN_OK;
	const size_t pbytes = (sizeof(xtn_int2) * (size_t)chunk + 255) & ~(size_t)255;
	const size_t dbytes = ((size_t)chunk + 255) & ~(size_t)255;
	if (h->out_pin_bytes < 2 * (pbytes + dbytes)) {
		if (h->out_pin)
			(void)hipHostFree(h->out_pin);
		h->out_pin = nullptr;
		h->out_pin_bytes = 0;
		if (hipHostMalloc(&h->out_pin, 2 * (pbytes + dbytes), hipHostMallocDefault) == hipSuccess)
			h->out_pin_bytes = 2 * (pbytes + dbytes);
		else
			rc = fail(h, XTN_ERR_NOMEM, "pinned staging allocation failed");
	}
	for (int b = 0; b < 2 && rc == XTN_OK; ++b) {
		char* base = static_cast<char*>(h->out_pin) + (size_t)b * (pbytes + dbytes);
		hp[b] = reinterpret_cast<xtn_int2*>(base);
		hd[b] = reinterpret_cast<int8_t*>(base + pbytes);
		if (hipEventCreateWithFlags(&ready[b], hipEventDisableTiming) != hipSuccess)
			rc = fail(h, XTN_ERR_HIP, "event creation failed");
	}
	auto issue = [&](int64_t start, int b) -> int {
		const int64_t len = std::min(chunk, n - start);
		HIP_TRY(h, hipMemcpyAsync(hp[b], h->out_pairs.as<int2>() + start, sizeof(xtn_int2) * (size_t)len,
		                          hipMemcpyDeviceToHost, h->copy_stream));
		HIP_TRY(h, hipMemcpyAsync(hd[b], h->out_dist.as<int8_t>() + start, (size_t)len,
		                          hipMemcpyDeviceToHost, h->copy_stream));
		HIP_TRY(h, hipEventRecord(ready[b], h->copy_stream));
		return XTN_OK;
	};
	if (rc == XTN_OK)
		rc = issue(0, 0);
	for (int64_t start = 0, k = 0; rc == XTN_OK && start < n; start += chunk, ++k) {
		const int b = (int)(k & 1);
		if (hipEventSynchronize(ready[b]) != hipSuccess) {
			rc = fail(h, XTN_ERR_HIP, "output copy failed");
			break;
		}
		if (start + chunk < n) {
			rc = issue(start + chunk, b ^ 1);
			if (rc != XTN_OK)
				break;
		}
		xtn_output out;
		out.index_pairs = hp[b];
		out.pairwise_distances = hd[b];
		out.pairwise_frequencies = nullptr;
		out.len = std::min(chunk, n - start);
		if (callback)
			callback(user, &out);
	}
	(void)hipStreamSynchronize(h->copy_stream);
	for (int b = 0; b < 2; ++b)
		if (ready[b])
			(void)hipEventDestroy(ready[b]);
	h->stats.ms_output = now_ms() - t0;
	return rc;
}

static int copy_output(xtn_handle* h, std::vector<xtn_int2>& pairs, std::vector<int8_t>& dist) {
	const int64_t n = h->out_len;
	pairs.resize((size_t)std::max<int64_t>(n, 1));
	dist.resize((size_t)std::max<int64_t>(n, 1));
	if (n > 0) {
		HIP_TRY(h, hipMemcpyAsync(pairs.data(), h->out_pairs.p, sizeof(xtn_int2) * (size_t)n,
		                          hipMemcpyDeviceToHost, h->stream));
		HIP_TRY(h, hipMemcpyAsync(dist.data(), h->out_dist.p, (size_t)n, hipMemcpyDeviceToHost, h->stream));
		HIP_SYNC(h, h->stream);
	}
	return XTN_OK;
}

int xtn_perform_sets(xtn_handle* h, const uint32_t* set_offsets, int32_t n_sets, int32_t batched,
                     xtn_set_callback callback, void* user) {
	if (!h || n_sets < 0 || (n_sets > 0 && !set_offsets))
		return h ? fail(h, XTN_ERR_ARG, "invalid query sets") : XTN_ERR_ARG;
	if (!h->offs.p)
		return fail(h, XTN_ERR_STATE, "no sequences set");
	for (int32_t k = 0; k < n_sets; ++k)
		if (set_offsets[k + 1] < set_offsets[k])
			return fail(h, XTN_ERR_ARG, "set offsets must be non-decreasing");
	if (n_sets > (1 << 30))  // set ids are packed into 30 bits
		return fail(h, XTN_ERR_ARG, "too many query sets (%d)", n_sets);
	if (n_sets > 0 && set_offsets[n_sets] > (uint32_t)h->n)
		return fail(h, XTN_ERR_ARG, "query sets exceed the %d rows", h->n);
	HIP_TRY(h, hipSetDevice(h->cfg.device));
	memset(&h->stats, 0, sizeof(h->stats));
	if (n_sets == 0)
		return XTN_OK;
	const double t0 = now_ms();
	xtn_stats tot;
	memset(&tot, 0, sizeof(tot));
	auto accumulate = [&]() {
		const xtn_stats& st = h->stats;
		tot.n_sequences += st.n_sequences;
		tot.n_variants += st.n_variants;
		tot.n_chunks += st.n_chunks;
		tot.n_runs += st.n_runs;
		tot.n_candidates += st.n_candidates;
		tot.n_unique += st.n_unique;
		tot.n_output += st.n_output;
		tot.ms_generate += st.ms_generate;
		tot.ms_sort += st.ms_sort;
		tot.ms_pairs += st.ms_pairs;
		tot.ms_verify += st.ms_verify;
		tot.n_rounds += st.n_rounds;
		tot.n_launches += st.n_launches;
		tot.n_syncs += st.n_syncs;
		tot.graph_replays += st.graph_replays;
		tot.tail_checked |= st.tail_checked;
		tot.tail_passes = std::max(tot.tail_passes, st.tail_passes);
	};
	std::vector<xtn_int2> hp;
	std::vector<int8_t> hd;
	if (batched) {
		const uint32_t r0 = set_offsets[0], r1 = set_offsets[n_sets];
		// set of every row, on the device from the n_sets + 1 offsets
		TRY(ensure(h, h->set_offs, sizeof(uint32_t) * ((size_t)n_sets + 1)));
		TRY(ensure(h, h->set_of, sizeof(uint32_t) * (size_t)std::max<uint32_t>(r1 - r0, 1)));
		HIP_TRY(h, hipMemcpyAsync(h->set_offs.p, set_offsets, sizeof(uint32_t) * ((size_t)n_sets + 1),
		                          hipMemcpyHostToDevice, h->stream));
		if (r1 > r0) {
			k_set_of<<<grid_for(r1 - r0, 256), 256, 0, h->stream>>>(h->set_offs.as<uint32_t>(), n_sets, r0,
			                                                       r1 - r0, h->set_of.as<uint32_t>());
			HIP_LAUNCH_CHECK(h);
		}
		TRY(nn_pipeline(h, (int32_t)(r1 - r0), nullptr, r0, h->set_of.as<uint32_t>()));
		// the query set holding a row, as k_set_of finds it
		auto set_of = [&](uint32_t row) -> uint32_t {
			return (uint32_t)(std::upper_bound(set_offsets, set_offsets + n_sets, row) - set_offsets - 1);
		};
		accumulate();
		if (callback && h->out_len > 0) {
			TRY(copy_output(h, hp, hd));
			// the sorted output is grouped by set, sets in order
			int64_t i = 0;
			const int64_t n = h->out_len;
			while (i < n) {
				const uint32_t k = set_of(r0 + (uint32_t)hp[i].x);
				const int32_t base = (int32_t)(set_offsets[k] - r0);
				const int32_t end = (int32_t)(set_offsets[k + 1] - r0);
				int64_t j = i;
				while (j < n && hp[j].x < end) {
					hp[j].x -= base;
					hp[j].y -= base;
					++j;
				}
				xtn_output out{hp.data() + i, hd.data() + i, nullptr, j - i};
				callback(user, (int32_t)k, &out);
				i = j;
			}
		}
	} else {
		for (int32_t k = 0; k < n_sets; ++k) {
			const uint32_t a = set_offsets[k], b = set_offsets[k + 1];
			TRY(nn_pipeline(h, (int32_t)(b - a), nullptr, a, nullptr));
			accumulate();
			if (callback && h->out_len > 0) {
				TRY(copy_output(h, hp, hd));
				xtn_output out{hp.data(), hd.data(), nullptr, h->out_len};
				callback(user, k, &out);
			}
		}
	}
	tot.ms_total = now_ms() - t0;
	h->stats = tot;
	return XTN_OK;
}

int xtn_get_tuning(xtn_handle* h, xtn_tuning* out) {
	if (!h || !out)
		return XTN_ERR_ARG;
	memset(out, 0, sizeof(*out));
	out->bucket_capacity = h->tune_cap;
	out->bucket_target = h->tune_target;
	out->sort_threads = h->tune_sort_nt;
	out->sort_tile = h->tune_sort_tile;
	out->select_threads = h->tune_select_nt;
	out->sort_low_bits = h->tune_low;
	out->coarse_separate = h->tune_coarse_sep;
	out->ms = h->tune_ms;
	out->ms_default = h->tune_ms_default;
	return XTN_OK;
}

int xtn_set_tuning(xtn_handle* h, const xtn_tuning* t) {
	if (!h || !t)
		return XTN_ERR_ARG;
	const int c = t->bucket_capacity, nt = t->sort_threads, tile = t->sort_tile;
	const bool sort_ok = (nt == 256 && tile == 4096) || (nt == 512 && (tile == 4096 || tile == 8192)) ||
	                     (nt == 1024 && tile == 8192);
	const int lb = t->sort_low_bits;
	if ((c != 2048 && c != 3072 && c != 4096) || !sort_ok ||
	    (t->select_threads != 256 && t->select_threads != 512) || t->bucket_target < c / 4 ||
	    t->bucket_target > c * 9 / 10 || (lb != 0 && lb != 8 && lb != 16 && lb != 24) ||
	    (t->coarse_separate != 0 && t->coarse_separate != 1))
		return fail(h, XTN_ERR_ARG,
		            "invalid tuning (capacity %d, target %d, sort threads %d, sort tile %d, select threads %d, "
		            "low bits %d, separate coarse pass %d)",
		            c, t->bucket_target, nt, tile, t->select_threads, lb, t->coarse_separate);
	h->tune_cap = c;
	h->tune_target = t->bucket_target;
	h->tune_sort_nt = nt;
	h->tune_sort_tile = tile;
	h->tune_select_nt = t->select_threads;
	h->tune_low = lb;
	h->tune_coarse_sep = t->coarse_separate;
	return XTN_OK;
}

int xtn_get_stats(xtn_handle* h, xtn_stats* out) {
	if (!h || !out)
		return XTN_ERR_ARG;
	*out = h->stats;
	return XTN_OK;
}

int xtn_kernel_profile(xtn_handle* h, const char* kernel, int64_t* launches, double* ms, double* bytes) {
	if (!h || !kernel)
		return XTN_ERR_ARG;
	auto it = h->prof.find(kernel);
	ProfAcc acc;
	if (it != h->prof.end())
		acc = it->second;
	if (launches)
		*launches = acc.launches;
	if (ms)
		*ms = acc.ms;
	if (bytes)
		*bytes = acc.bytes;
	return XTN_OK;
}

int xtn_kernel_profile_reset(xtn_handle* h) {
	if (!h)
		return XTN_ERR_ARG;
	h->prof.clear();
	return XTN_OK;
}

const char* xtn_last_error(xtn_handle* h) { return h ? h->err.c_str() : "null handle"; }

const char* xtn_status_string(int status) {
	switch (status) {
	case XTN_OK: return "ok";
	case XTN_ERROR: return "invalid input";
	case XTN_EXIT: return "exit";
	case XTN_ERR_ARG: return "invalid argument";
	case XTN_ERR_HIP: return "HIP runtime error";
	case XTN_ERR_NOMEM: return "out of memory";
	case XTN_ERR_CAPACITY: return "capacity exceeded";
	case XTN_ERR_STATE: return "invalid call order";
	default: return "unknown status";
	}
}

// ---- autotuner: the workload's signature, the cache file, the search --

static std::string tune_signature(xtn_handle* h) {
	uint64_t per[19];
	for (int L = 0; L <= 18; ++L)
		per[L] = host_variant_count((uint32_t)L, h->cfg.distance);
	uint64_t V = 0;
	for (uint8_t L : h->prefix_len)
		V += per[L];
	const double n = (double)std::max<int32_t>(h->n, 1);
	char buf[128];
	// sequence count in quarter octaves, variants per sequence to two decimals
	snprintf(buf, sizeof(buf), "n%.2f_d%d_m%d_v%.2f", std::round(4.0 * std::log2(n)) / 4.0, h->cfg.distance,
	         h->cfg.measure, (double)V / n);
	return buf;
}

// the cache holds one JSON object per line inside {"xtn_tuning": 1, "entries": [ ... ]}
static bool tune_cache_read(const char* path, const std::string& sig, xtn_tuning* t) {
	FILE* f = fopen(path, "r");
	if (!f)
		return false;
	char line[1024];
	bool found = false;
	const std::string key = "\"workload\": \"" + sig + "\"";
	while (!found && fgets(line, sizeof(line), f)) {
		if (!strstr(line, key.c_str()))
			continue;
		auto num = [&](const char* name, double* v) {
			const char* p = strstr(line, name);
			return p && sscanf(p + strlen(name), " : %lf", v) == 1;
		};
		double c, tg, nt, tl, sn, lb, ms, msd, cs = 0;
		if (num("\"bucket_capacity\"", &c) && num("\"bucket_target\"", &tg) && num("\"sort_threads\"", &nt) &&
		    num("\"sort_tile\"", &tl) && num("\"select_threads\"", &sn) && num("\"sort_low_bits\"", &lb) &&
		    num("\"ms\"", &ms) && num("\"ms_default\"", &msd)) {
			t->bucket_capacity = (int32_t)c;
			t->bucket_target = (int32_t)tg;
			t->sort_threads = (int32_t)nt;
			t->sort_tile = (int32_t)tl;
			t->select_threads = (int32_t)sn;
			t->sort_low_bits = (int32_t)lb;
			t->coarse_separate = num("\"coarse_separate\"", &cs) ? (int32_t)cs : 0;  // absent: older entries
			t->ms = ms;
			t->ms_default = msd;
			found = true;
		}
	}
	fclose(f);
	return found;
}

static bool tune_cache_write(const char* path, const std::string& sig, const xtn_tuning& t) {
	std::vector<std::string> keep;
	if (FILE* f = fopen(path, "r")) {
		char line[1024];
		const std::string key = "\"workload\": \"" + sig + "\"";
		while (fgets(line, sizeof(line), f))
			if (strstr(line, "\"workload\"") && !strstr(line, key.c_str())) {
				std::string l(line);
				while (!l.empty() && (l.back() == '\n' || l.back() == ',' || l.back() == ' '))
					l.pop_back();
				keep.push_back(l);
			}
		fclose(f);
	}
	char buf[512];
	snprintf(buf, sizeof(buf),
	         "  {\"workload\": \"%s\", \"bucket_capacity\": %d, \"bucket_target\": %d, \"sort_threads\": %d, "
	         "\"sort_tile\": %d, \"select_threads\": %d, \"sort_low_bits\": %d, \"coarse_separate\": %d, "
	         "\"ms\": %.4f, \"ms_default\": %.4f}",
	         sig.c_str(), t.bucket_capacity, t.bucket_target, t.sort_threads, t.sort_tile, t.select_threads,
	         t.sort_low_bits, t.coarse_separate, t.ms, t.ms_default);
	keep.push_back(buf);
	const std::string tmp = std::string(path) + ".tmp";
	FILE* f = fopen(tmp.c_str(), "w");
	if (!f)
		return false;
	fprintf(f, "{\"xtn_tuning\": 1, \"entries\": [\n");
	for (size_t i = 0; i < keep.size(); ++i)
		fprintf(f, "%s%s\n", keep[i].c_str(), i + 1 < keep.size() ? "," : "");
	fprintf(f, "]}\n");
	const bool ok = fclose(f) == 0;
	return ok && rename(tmp.c_str(), path) == 0;
}

// FNV-1a over the output triples of the last run, copied to the host
static int output_digest(xtn_handle* h, uint64_t* out) {
	std::vector<xtn_int2> p;
	std::vector<int8_t> d;
	const int64_t n = h->out_len;
	p.resize((size_t)std::max<int64_t>(n, 1));
	d.resize((size_t)std::max<int64_t>(n, 1));
	if (n > 0) {
		HIP_TRY(h, hipMemcpyAsync(p.data(), h->out_pairs.p, sizeof(xtn_int2) * (size_t)n, hipMemcpyDeviceToHost,
		                          h->stream));
		HIP_TRY(h, hipMemcpyAsync(d.data(), h->out_dist.p, (size_t)n, hipMemcpyDeviceToHost, h->stream));
		HIP_SYNC(h, h->stream);
	}
	uint64_t x = 1469598103934665603ull;
	auto mix = [&](const void* b, size_t len) {
		const unsigned char* c = (const unsigned char*)b;
		for (size_t i = 0; i < len; ++i)
			x = (x ^ c[i]) * 1099511628211ull;
	};
	if (n > 0) {
		mix(p.data(), sizeof(xtn_int2) * (size_t)n);
		mix(d.data(), (size_t)n);
	}
	mix(&n, sizeof(n));
	*out = x;
	return XTN_OK;
}

int xtn_autotune(xtn_handle* h, const char* cache_path, int32_t runs, xtn_tuning* out) {
	if (!h || runs < 1)
		return XTN_ERR_ARG;
	if (!h->offs.p)
		return fail(h, XTN_ERR_STATE, "no sequences set");
	HIP_TRY(h, hipSetDevice(h->cfg.device));
	const std::string sig = tune_signature(h);
	xtn_tuning best{};
	if (cache_path && tune_cache_read(cache_path, sig, &best) && xtn_set_tuning(h, &best) == XTN_OK) {
		best.from_cache = 1;
		h->tune_ms = best.ms;
		h->tune_ms_default = best.ms_default;
		if (out)
			*out = best;
		return XTN_OK;
	}
	// one setting: two warm-up runs (an eager run and its graph capture),
	// then the median host time of `runs` runs, each ending in the readback;
	// its output must be the default setting's
	int32_t runs_now = runs;
	struct Setting {
		int cap, target, sort_nt, sort_tile, select_nt, low, sep;
		bool operator==(const Setting& o) const {
			return cap == o.cap && target == o.target && sort_nt == o.sort_nt && sort_tile == o.sort_tile &&
			       select_nt == o.select_nt && low == o.low && sep == o.sep;
		}
	};
	uint64_t dig_def = 0;
	bool have_def = false;
	int tried = 0;
	auto time_setting = [&](const Setting& st, double* ms) -> int {
		h->tune_cap = st.cap;
		h->tune_target = st.target;
		h->tune_sort_nt = st.sort_nt;
		h->tune_sort_tile = st.sort_tile;
		h->tune_select_nt = st.select_nt;
		h->tune_low = st.low;
		h->tune_coarse_sep = st.sep;
		std::vector<double> t;
		for (int r = 0; r < runs_now + 2; ++r) {
			memset(&h->stats, 0, sizeof(h->stats));
			TRY(nn_pipeline(h, h->n, nullptr));
			if (r >= 2)
				t.push_back(h->stats.ms_total);
		}
		std::sort(t.begin(), t.end());
		*ms = t[t.size() / 2];
		uint64_t dig = 0;
		TRY(output_digest(h, &dig));
		if (!have_def) {
			dig_def = dig;
			have_def = true;
		} else if (dig != dig_def) {
			return fail(h, XTN_ERR_HIP,
			            "autotune: capacity %d, target %d, sort %dx%d, select %d, low bits %d, separate coarse pass %d "
			            "changed the output",
			            st.cap, st.target, st.sort_nt, st.sort_tile, st.select_nt, st.low, st.sep);
		}
		++tried;
		return XTN_OK;
	};
	const Setting def{3072, 2200, 512, 4096, 256, 0, 0};
	double ms_def = 0;
	TRY(time_setting(def, &ms_def));
	Setting cur = def;
	double cur_ms = ms_def;
	// coordinate descent: bucket capacity and target, then the sort tile,
	// then the select tile, then the sort's low bits, then the generator's
	// coarse pass on or off, each step keeping the best so far
	const int caps[3] = {3072, 4096, 2048};
	const int targets[3][3] = {{1840, 2200, 2600}, {2450, 2950, 3480}, {1230, 1470, 1740}};
	std::vector<Setting> step1;
	for (int ci = 0; ci < 3; ++ci)
		for (int tg : targets[ci])
			step1.push_back({caps[ci], tg, def.sort_nt, def.sort_tile, def.select_nt, def.low, def.sep});
	for (int step = 1; step <= 5; ++step) {
		std::vector<Setting> cands = step1;
		const Setting c = cur;
		if (step == 2)
			cands = {{c.cap, c.target, 256, 4096, c.select_nt, c.low, c.sep},
			         {c.cap, c.target, 512, 4096, c.select_nt, c.low, c.sep},
			         {c.cap, c.target, 512, 8192, c.select_nt, c.low, c.sep},
			         {c.cap, c.target, 1024, 8192, c.select_nt, c.low, c.sep}};
		if (step == 3)
			cands = {{c.cap, c.target, c.sort_nt, c.sort_tile, 256, c.low, c.sep},
			         {c.cap, c.target, c.sort_nt, c.sort_tile, 512, c.low, c.sep}};
		if (step == 4)
			cands = {{c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, 0, c.sep},
			         {c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, 8, c.sep},
			         {c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, 16, c.sep},
			         {c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, 24, c.sep}};
		if (step == 5)
			cands = {{c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, c.low, 0},
			         {c.cap, c.target, c.sort_nt, c.sort_tile, c.select_nt, c.low, 1}};
		for (const Setting& st : cands) {
			if (st == cur)
				continue;
			double ms = 0;
			TRY(time_setting(st, &ms));
			if (ms < cur_ms) {
				cur = st;
				cur_ms = ms;
			}
		}
	}
	// the best against the defaults once more, interleaved, with twice the
	// runs: a setting is kept only when it still wins by more than 1% (the
	// first timings of a search carry the order they were taken in)
	if (!(cur == def)) {
		double a1 = 0, b1 = 0, a2 = 0, b2 = 0;
		runs_now = 2 * runs;
		TRY(time_setting(def, &a1));
		TRY(time_setting(cur, &b1));
		TRY(time_setting(def, &a2));
		TRY(time_setting(cur, &b2));
		runs_now = runs;
		ms_def = std::min(a1, a2);
		cur_ms = std::min(b1, b2);
		if (cur_ms > 0.99 * ms_def) {
			cur = def;
			cur_ms = ms_def;
		}
	}
	best.bucket_capacity = cur.cap;
	best.bucket_target = cur.target;
	best.sort_threads = cur.sort_nt;
	best.sort_tile = cur.sort_tile;
	best.select_threads = cur.select_nt;
	best.sort_low_bits = cur.low;
	best.coarse_separate = cur.sep;
	best.ms = cur_ms;
	best.ms_default = ms_def;
	best.candidates = tried;
	TRY(xtn_set_tuning(h, &best));
	h->tune_ms = best.ms;
	h->tune_ms_default = ms_def;
	if (cache_path && !tune_cache_write(cache_path, sig, best))
		return fail(h, XTN_ERR_ARG, "autotune: cannot write the cache file %s", cache_path);
	if (out)
		*out = best;
	return XTN_OK;
}

const char* xtn_version(void) { return "2.0.0-mi355x"; }

}  // extern "C"

// ============================================== no-allocation search ==
//
// xtn_nn_noalloc runs the search of xtn_run_nn_resident on the caller's
// device buffers, with every workspace a fixed slice of one file-scope
// device array (`scratch`) instead of handle allocations: no hipMalloc,
// hipFree, stream or event creation. Sizes beyond the slices are rejected.
// The array of a device is shared by every call on it: calls are serialised
// per device by a host mutex, and every kernel clears what it needs cleared.

namespace {

constexpr int64_t NA_MAX_SEQS = 1 << 16;     // sequences
constexpr int64_t NA_MAX_RECORDS = 1 << 20;  // deletion records (sum of the variant counts)
constexpr int64_t NA_MAX_PAIRS = 1 << 20;    // candidate pairs of the search, and output pairs
constexpr size_t NA_SCRATCH_BYTES = (size_t)52 << 20;

__device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

struct NoAllocDevice {
	std::mutex mu;
	uint8_t* ptr = nullptr;  // this device's copy of `scratch` (hipGetSymbolAddress)
	bool ready = false;
};

NoAllocDevice g_noalloc[64];

// carves the handle's buffers out of the scratch; false when the layout does
// not fit (a build-time inconsistency)
bool carve_scratch(xtn_handle* h, uint8_t* base) {
	size_t off = 0;
	auto take = [&](DBuf& b, size_t bytes) {
		bytes = (bytes + 255) & ~(size_t)255;
		b.p = base + off;
		b.cap = bytes;
		off += bytes;
	};
	const int64_t rtiles = NA_MAX_RECORDS / RS_TILE + 1;
	// tiles of at least 4096 sort keys and 2048 select candidates
	const int64_t sort_tiles = NA_MAX_PAIRS / RS_TILE + 1, sel_tiles = NA_MAX_PAIRS / 2048 + 1;
	take(h->pk, 16 * NA_MAX_SEQS);
	take(h->voff, 8 * (NA_MAX_SEQS + 1));
	take(h->pack_state, 8 * (NA_MAX_SEQS / 256 + 1));
	take(h->tseq, 4 * (NA_MAX_RECORDS / GEN_TILE + 2));
	take(h->recs, 8 * NA_MAX_RECORDS);
	take(h->recs_alt, 8 * NA_MAX_RECORDS);
	take(h->hist, 4 * RS_BINS * rtiles);
	take(h->scan_tmp, 64 << 10);
	take(h->runs_small, sizeof(Run) * (NA_MAX_RECORDS / 2 + 1));
	take(h->runs_big, sizeof(Run) * (NA_MAX_RECORDS / (SMALL_RUN_MAX + 1) + 1));
	take(h->big_pre, 8 * (NA_MAX_RECORDS / (SMALL_RUN_MAX + 1) + 2));
	take(h->bstart, 4 * (65536 + 1));
	take(h->bcount, 4 * 65536);
	take(h->ovf, 4 * 65536);
	take(h->ctl, 4 * CTL_WORDS);
	take(h->ctrs, sizeof(CtlBlock));
	take(h->part_hist, 8 * 4096);
	take(h->pairs, 8 * NA_MAX_PAIRS);
	take(h->pairs_alt, 8 * NA_MAX_PAIRS);
	take(h->out_pairs, 8 * NA_MAX_PAIRS);
	take(h->out_dist, NA_MAX_PAIRS);
	take(h->tail_state, 4 * (4 * sort_tiles * RS_BINS + 2 * sel_tiles));  // 2 * idb <= 32 bits: 4 passes
	take(h->pos, 4 * (NA_MAX_PAIRS / SEL_TILE + NA_MAX_PAIRS / SEL_NT + 4));
	take(h->dist, NA_MAX_PAIRS);
	return off <= NA_SCRATCH_BYTES;
}

}  // namespace

extern "C" {

int xtn_nn_noalloc(const char* d_chars, const uint32_t* d_offsets, int32_t n, int32_t distance, int32_t measure,
                   xtn_int2* d_pairs, int8_t* d_dist, int64_t out_cap, int64_t* out_len, void* stream) {
	if (!out_len || n < 0 || (n > 0 && (!d_chars || !d_offsets)) || out_cap < 0 ||
	    (out_cap > 0 && (!d_pairs || !d_dist)))
		return XTN_ERR_ARG;
	*out_len = 0;
	if (distance < 1 || distance > 127 || (measure != XTN_LEVENSHTEIN && measure != XTN_HAMMING))
		return XTN_ERR_ARG;
	if (n > NA_MAX_SEQS)
		return XTN_ERR_CAPACITY;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
		(void)hipGetLastError();
		return XTN_ERR_HIP;
	}
	NoAllocDevice& nd = g_noalloc[dev];
	std::lock_guard<std::mutex> lk(nd.mu);
	if (!nd.ready) {
		// once per device, with the device current: this device's address of
		// the array (the host-side name is no device address)
		void* p = nullptr;
		if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
			(void)hipGetLastError();
			return XTN_ERR_HIP;
		}
		nd.ptr = static_cast<uint8_t*>(p);
		nd.ready = true;
	}
	if (n < 2)
		return XTN_OK;
	xtn_handle h;
	xtn_config_init(&h.cfg);
	h.cfg.distance = distance;
	h.cfg.measure = measure;
	h.cfg.device = dev;
	h.cfg.no_graph = 1;
	h.cfg.max_records_per_chunk = NA_MAX_RECORDS;
	h.cfg.max_pairs_per_round = NA_MAX_PAIRS;
	h.stream = static_cast<hipStream_t>(stream);
	h.no_alloc = true;
	h.free_valid = true;  // no free-memory query: the slices are the budget
	h.free_cache = NA_SCRATCH_BYTES;
	h.last_pairs = NA_MAX_PAIRS;  // the candidate capacity is the whole slice
	int ncu = 0;
	if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
		h.n_cu = ncu;
	if (!carve_scratch(&h, nd.ptr))
		return XTN_ERR_CAPACITY;
	h.chars.p = const_cast<char*>(d_chars);
	h.chars.cap = ~(size_t)0 >> 1;
	h.offs.p = const_cast<uint32_t*>(d_offsets);
	h.offs.cap = ~(size_t)0 >> 1;
	h.n = n;
	int rc = nn_pipeline(&h, n, nullptr);
	if (rc != XTN_OK)
		return rc;
	if (h.out_len > out_cap)
		return XTN_ERR_CAPACITY;
	if (h.out_len > 0) {
		if (hipMemcpyAsync(d_pairs, h.out_pairs.p, sizeof(xtn_int2) * (size_t)h.out_len, hipMemcpyDeviceToDevice,
		                   h.stream) != hipSuccess ||
		    hipMemcpyAsync(d_dist, h.out_dist.p, (size_t)h.out_len, hipMemcpyDeviceToDevice, h.stream) != hipSuccess ||
		    hipStreamSynchronize(h.stream) != hipSuccess) {
			(void)hipGetLastError();
			return XTN_ERR_HIP;
		}
	}
	*out_len = h.out_len;
	return XTN_OK;
}

}  // extern "C"

// ============================================================ server ==

namespace {

struct ServeRequest {
	uint64_t ticket = 0;
	std::vector<char> chars;
	std::vector<uint32_t> offs;
	int32_t n = 0;
	int64_t variants = 0;
	int64_t est_bytes = 0;
	double t_submit = 0;
};

struct ServeResult {
	std::vector<xtn_int2> pairs;
	std::vector<int8_t> dist;
	double latency_ms = 0;
	int status = XTN_OK;
	std::string err;
};

}  // namespace

struct xtn_server {
	xtn_server_config cfg;
	xtn_handle* base = nullptr;  // owns the index: chars/offs, recs (sorted), bstart (directory)
	int32_t n_index = 0;
	int idb = 0;
	int64_t v_index = 0;
	const uint64_t* index = nullptr;
	const uint32_t* dir = nullptr;
	int dir_bits = 8;
	std::vector<xtn_handle*> workers;
	std::vector<std::thread> threads;
	std::mutex mu;
	std::condition_variable cv_work, cv_done;
	std::deque<ServeRequest> queue;
	std::unordered_map<uint64_t, ServeResult> done;
	std::unordered_set<uint64_t> open_tickets;  // issued and not yet claimed by a wait
	uint64_t next_ticket = 1;
	int64_t inflight = 0;
	int64_t budget = 0;
	bool stop = false;
	std::string err;
};

namespace {

// the server's last error, written under its mutex (submit, wait and the
// workers run on different threads)
int sv_fail(xtn_server* sv, int code, const std::string& msg) {
	std::lock_guard<std::mutex> lk(sv->mu);
	sv->err = msg;
	return code;
}

// one query on a worker handle: upload from the worker's pinned staging,
// variants, index lookups appending candidate pairs, then the search's tail
// (sort, unique, verification against the index strings, ordered select)
// issued back to back; the host reads the counters once and copies the
// output. A candidate buffer that was too small is regrown and the lookups
// redone (the records are still in place).
int serve_query(xtn_server* sv, xtn_handle* w, const ServeRequest& rq, ServeResult* res) {
	hipStream_t s = w->stream;
	const int32_t n = rq.n;
	if (n == 0)
		return XTN_OK;
	const int idb = sv->idb;
	const int d = sv->cfg.distance;
	const uint32_t total = rq.offs[n];
	const size_t up = (size_t)total + sizeof(uint32_t) * ((size_t)n + 1);
	if (up > w->pin_bytes) {
		if (w->pin)
			(void)hipHostFree(w->pin);
		w->pin = nullptr;
		w->pin_bytes = 0;
		if (hipHostMalloc(&w->pin, up + up / 4, hipHostMallocDefault) != hipSuccess) {
			(void)hipGetLastError();
			w->pin = nullptr;
			return fail(w, XTN_ERR_NOMEM, "pinned staging allocation failed");
		}
		w->pin_bytes = up + up / 4;
	}
	// the previous query's upload has completed: its output copy, after it
	// on the same stream, was waited on
	char* pin = static_cast<char*>(w->pin);
	uint32_t* pin_offs = reinterpret_cast<uint32_t*>(pin + (((size_t)total + 3) & ~(size_t)3));
	memcpy(pin, rq.chars.data(), total);
	memcpy(pin_offs, rq.offs.data(), sizeof(uint32_t) * ((size_t)n + 1));
	TRY(ensure(w, w->chars, (size_t)total + 16));
	TRY(ensure(w, w->offs, sizeof(uint32_t) * ((size_t)n + 1)));
	if (total)
		HIP_TRY(w, hipMemcpyAsync(w->chars.p, pin, total, hipMemcpyHostToDevice, s));
	HIP_TRY(w, hipMemcpyAsync(w->offs.p, pin_offs, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyHostToDevice, s));
	TRY(ensure(w, w->pk, sizeof(uint64_t) * 2 * (size_t)n));
	TRY(ensure(w, w->voff, sizeof(uint64_t) * ((size_t)n + 1)));
	TRY(ensure(w, w->ctrs, sizeof(CtlBlock)));
	k_pack<<<grid_for(n, 256), 256, 0, s>>>(w->chars.as<uint8_t>(), w->offs.as<uint32_t>(), nullptr, 0, nullptr,
	                                         n, d, w->pk.as<uint64_t>(), w->voff.as<uint64_t>());
	HIP_LAUNCH_CHECK(w);
	TRY(scan_excl<uint64_t>(w, w->voff.as<uint64_t>(), w->voff.as<uint64_t>(), n, w->voff.as<uint64_t>() + n));
	const int64_t V = rq.variants;  // the same closed form the device scan computes
	TRY(ensure(w, w->recs, sizeof(uint64_t) * (size_t)V));
	GenArgs ga{};
	ga.pk = w->pk.as<uint64_t>();
	ga.voff = w->voff.as<uint64_t>();
	ga.n = n;
	ga.distance = d;
	ga.idb = idb;
	ga.V = V;
	ga.out = w->recs.as<uint64_t>();
	ga.cap = V;
	TRY(tile_seq_table(w, s, n, (uint64_t)V, &ga.tile_seq));
	k_gen<GEN_DIRECT><<<grid_for(V, GEN_TILE), GEN_NT, 0, s>>>(ga);
	HIP_LAUNCH_CHECK(w);
	// candidates per record seen so far on this worker, with headroom
	const double ratio = w->serve_ratio > 0 ? w->serve_ratio : 1.0;
	int64_t pcap = std::max<int64_t>((int64_t)((double)V * ratio * 1.5) + 1024, 1 << 16);
	pcap = std::min<int64_t>(pcap, (int64_t)LB_VAL);
	CtlHead head{};
	TailPlan tp;
	bool grown = false;
	w->tail_full = false;
	for (;;) {
		TRY(ensure(w, w->pairs, sizeof(uint64_t) * (size_t)pcap));
		TRY(tail_prepare(w, pcap, idb, 0, &tp));
		TRY(zero_ctl(w, true, true));
		k_match<<<grid_for(V, MATCH_NT), MATCH_NT, 0, s>>>(w->recs.as<uint64_t>(), V, sv->index, sv->dir,
		                                                   sv->dir_bits, idb, w->pairs.as<uint64_t>(), pcap,
		                                                   &ctl_block(w)->rc.n_cand);
		HIP_LAUNCH_CHECK(w);
		TRY(issue_tail(w, tp, idb, nullptr, 0, nullptr, 0, sv->base->chars.as<uint8_t>(),
		               sv->base->offs.as<uint32_t>()));
		TRY(read_head(w, &head));
		if (head.error)
			return fail(w, XTN_ERR_HIP, "a look-back wait of the query's candidate tail exceeded its limit");
		if ((int64_t)head.rc.n_cand > pcap) {
			if (grown || head.rc.n_cand > (unsigned long long)LB_VAL)
				return fail(w, XTN_ERR_CAPACITY, "query candidates exceed the per-query limit (%llu)",
				            (unsigned long long)head.rc.n_cand);
			pcap = (int64_t)head.rc.n_cand + (int64_t)head.rc.n_cand / 4;
			pcap = std::min<int64_t>(pcap, (int64_t)LB_VAL);
			grown = true;
			continue;
		}
		if (!head.overflow)
			break;
		// a group outgrew the select's window: this query's tail again, fully
		// sorted (the full sort never sets the flag)
		w->tail_full = true;
	}
	w->tail_full = false;
	const double r = (double)head.rc.n_cand / (double)std::max<int64_t>(V, 1);
	w->serve_ratio = w->serve_ratio > 0 ? std::max(r, 0.75 * w->serve_ratio + 0.25 * r) : std::max(r, 0.01);
	const int64_t nout = (int64_t)head.n_out;
	res->pairs.resize((size_t)nout);
	res->dist.resize((size_t)nout);
	if (nout) {
		HIP_TRY(w, hipMemcpyAsync(res->pairs.data(), w->out_pairs.p, sizeof(xtn_int2) * (size_t)nout,
		                          hipMemcpyDeviceToHost, s));
		HIP_TRY(w, hipMemcpyAsync(res->dist.data(), w->out_dist.p, (size_t)nout, hipMemcpyDeviceToHost, s));
		HIP_SYNC(w, s);
	}
	return XTN_OK;
}

void serve_worker(xtn_server* sv, int slot) {
	xtn_handle* w = sv->workers[slot];
	(void)hipSetDevice(sv->cfg.device);
	for (;;) {
		ServeRequest rq;
		{
			std::unique_lock<std::mutex> lk(sv->mu);
			// FIFO admission: the head of the queue starts once its estimate
			// fits the budget, or when nothing else is in flight
			sv->cv_work.wait(lk, [&] {
				return sv->stop || (!sv->queue.empty() &&
				                    (sv->inflight == 0 || sv->inflight + sv->queue.front().est_bytes <= sv->budget));
			});
			if (sv->queue.empty())
				return;  // stop requested and nothing left
			rq = std::move(sv->queue.front());
			sv->queue.pop_front();
			sv->inflight += rq.est_bytes;
		}
		ServeResult res;
		res.status = serve_query(sv, w, rq, &res);
		if (res.status != XTN_OK)
			res.err = w->err;
		res.latency_ms = now_ms() - rq.t_submit;
		{
			std::lock_guard<std::mutex> lk(sv->mu);
			sv->inflight -= rq.est_bytes;
			sv->done[rq.ticket] = std::move(res);
		}
		sv->cv_done.notify_all();
		sv->cv_work.notify_all();
	}
}

}  // namespace

extern "C" {

void xtn_server_config_init(xtn_server_config* cfg) {
	if (!cfg)
		return;
	memset(cfg, 0, sizeof(*cfg));
	cfg->distance = 1;
	cfg->measure = XTN_LEVENSHTEIN;
	cfg->slots = 4;
}

int xtn_server_open(xtn_server** out, const xtn_server_config* cfg_in, const char* chars,
                    const uint32_t* offsets, int32_t n) {
	if (!out || !offsets || n < 1)
		return XTN_ERR_ARG;
	*out = nullptr;
	xtn_server_config cfg;
	xtn_server_config_init(&cfg);
	if (cfg_in)
		cfg = *cfg_in;
	if (cfg.slots <= 0)
		cfg.slots = 4;
	if (cfg.slots > 16)
		cfg.slots = 16;
	if (cfg.max_query_rows <= 0)
		cfg.max_query_rows = 1 << 20;
	xtn_config hc;
	xtn_config_init(&hc);
	hc.distance = cfg.distance;
	hc.measure = cfg.measure;
	hc.device = cfg.device;
	xtn_server* sv = new xtn_server();
	sv->cfg = cfg;
	int rc = xtn_open(&sv->base, &hc);
	if (rc != XTN_OK) {
		delete sv;
		return rc;
	}
	xtn_handle* h = sv->base;
	rc = xtn_set_sequences(h, chars, offsets, n);
	if (rc != XTN_OK) {
		sv->err = h->err;
		xtn_server_close(sv);
		return rc;
	}
	sv->n_index = n;
	sv->idb = std::max(bits_for(std::max(n - 1, 1)), bits_for(std::max(cfg.max_query_rows - 1, 1)));
	if (2 * sv->idb > 62 || 64 - sv->idb < DIR_BITS_MAX + 8) {
		xtn_server_close(sv);
		return XTN_ERR_CAPACITY;
	}
	auto build = [&]() -> int {
		hipStream_t s = h->stream;
		TRY(ensure(h, h->pk, sizeof(uint64_t) * 2 * (size_t)n));
		TRY(ensure(h, h->voff, sizeof(uint64_t) * ((size_t)n + 1)));
		k_pack<<<grid_for(n, 256), 256, 0, s>>>(h->chars.as<uint8_t>(), h->offs.as<uint32_t>(), nullptr, 0,
		                                         nullptr, n, cfg.distance, h->pk.as<uint64_t>(),
		                                         h->voff.as<uint64_t>());
		HIP_LAUNCH_CHECK(h);
		TRY(scan_excl<uint64_t>(h, h->voff.as<uint64_t>(), h->voff.as<uint64_t>(), n, h->voff.as<uint64_t>() + n));
		uint64_t V = 0;
		HIP_TRY(h, hipMemcpyAsync(&V, h->voff.as<uint64_t>() + n, sizeof(V), hipMemcpyDeviceToHost, s));
		HIP_SYNC(h, s);
		if (V > (uint64_t)INT32_MAX)
			return fail(h, XTN_ERR_CAPACITY, "index of %llu records exceeds 2^31", (unsigned long long)V);
		TRY(ensure(h, h->recs, sizeof(uint64_t) * V));
		TRY(ensure(h, h->recs_alt, sizeof(uint64_t) * V));
		GenArgs ga{};
		ga.pk = h->pk.as<uint64_t>();
		ga.voff = h->voff.as<uint64_t>();
		ga.n = n;
		ga.distance = cfg.distance;
		ga.idb = sv->idb;
		ga.V = (int64_t)V;
		ga.out = h->recs.as<uint64_t>();
		ga.cap = (int64_t)V;
		if (V > 0)
			TRY(tile_seq_table(h, s, n, V, &ga.tile_seq));
		k_gen<GEN_DIRECT><<<grid_for((int64_t)V, GEN_TILE), GEN_NT, 0, s>>>(ga);
		HIP_LAUNCH_CHECK(h);
		TRY(radix_sort(h, h->recs, h->recs_alt, nullptr, nullptr, (int64_t)V, sv->idb, 64, "index_sort"));
		release(h->recs_alt);
		// directory bits: about eight index records per entry
		int db = 8;
		while (db < DIR_BITS_MAX && ((int64_t)V >> (db + 3)) > 0)
			++db;
		sv->dir_bits = db;
		TRY(ensure(h, h->bstart, sizeof(uint32_t) * (((size_t)1 << db) + 1)));
		k_dir_fill<<<grid_for((int64_t)V + 1, 256), 256, 0, s>>>(h->recs.as<uint64_t>(), (int64_t)V, db,
		                                                          h->bstart.as<uint32_t>());
		HIP_LAUNCH_CHECK(h);
		HIP_SYNC(h, s);
		sv->v_index = (int64_t)V;
		sv->index = h->recs.as<uint64_t>();
		sv->dir = h->bstart.as<uint32_t>();
		size_t fr = 0, tot = 0;
		HIP_TRY(h, hipMemGetInfo(&fr, &tot));
		sv->budget = cfg.hbm_budget_bytes > 0 ? cfg.hbm_budget_bytes : (int64_t)(fr / 4);
		return XTN_OK;
	};
	rc = build();
	if (rc != XTN_OK) {
		sv->err = h->err;
		xtn_server_close(sv);
		return rc;
	}
	for (int k = 0; k < cfg.slots; ++k) {
		xtn_handle* w = nullptr;
		rc = xtn_open(&w, &hc);
		if (rc != XTN_OK) {
			xtn_server_close(sv);
			return rc;
		}
		sv->workers.push_back(w);
	}
	for (int k = 0; k < cfg.slots; ++k)
		sv->threads.emplace_back(serve_worker, sv, k);
	*out = sv;
	return XTN_OK;
}

int xtn_server_submit(xtn_server* sv, const char* chars, const uint32_t* offsets, int32_t n, uint64_t* ticket) {
	if (!sv || !offsets || n < 0 || !ticket)
		return sv ? sv_fail(sv, XTN_ERR_ARG, "invalid query arguments") : XTN_ERR_ARG;
	if (n > sv->cfg.max_query_rows)
		return sv_fail(sv, XTN_ERR_CAPACITY, "query set larger than max_query_rows");
	if (offsets[0] != 0)
		return sv_fail(sv, XTN_ERR_ARG, "offsets[0] must be 0");
	ServeRequest rq;
	rq.n = n;
	for (int32_t i = 0; i < n; ++i)
		if (offsets[i + 1] < offsets[i])
			return sv_fail(sv, XTN_ERR_ARG, "offsets must be non-decreasing");
	const uint32_t total = offsets[n];
	if (total > 0 && !chars)
		return sv_fail(sv, XTN_ERR_ARG, "invalid query arguments");
	if (!all_residues(chars, total))
		return sv_fail(sv, XTN_ERROR, "invalid char");
	rq.chars.assign(chars, chars + total);
	rq.offs.assign(offsets, offsets + n + 1);
	for (int32_t i = 0; i < n; ++i)
		rq.variants += (int64_t)host_variant_count(offsets[i + 1] - offsets[i], sv->cfg.distance);
	if (rq.variants > (int64_t)INT32_MAX)
		return sv_fail(sv, XTN_ERR_CAPACITY, "query variants exceed 2^31");
	// records, candidates (about V/2, sort buffer), flags/positions
	rq.est_bytes = rq.variants * 8 + rq.variants * 12 + (int64_t)n * 64 + (int64_t)total;
	rq.t_submit = now_ms();
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		rq.ticket = sv->next_ticket++;
		*ticket = rq.ticket;
		sv->open_tickets.insert(rq.ticket);
		sv->queue.push_back(std::move(rq));
	}
	sv->cv_work.notify_all();
	return XTN_OK;
}

int xtn_server_wait(xtn_server* sv, uint64_t ticket, xtn_callback callback, void* user, double* latency_ms) {
	if (!sv)
		return XTN_ERR_ARG;
	ServeResult res;
	{
		std::unique_lock<std::mutex> lk(sv->mu);
		// a ticket is collected by exactly one wait: a second wait on it (or
		// one on a ticket never issued) is an error, not a wait forever
		if (sv->open_tickets.erase(ticket) == 0) {
			sv->err = "unknown or already collected ticket";
			return XTN_ERR_ARG;
		}
		sv->cv_done.wait(lk, [&] { return sv->done.count(ticket) > 0; });
		res = std::move(sv->done[ticket]);
		sv->done.erase(ticket);
		if (res.status != XTN_OK)
			sv->err = res.err;
	}
	if (latency_ms)
		*latency_ms = res.latency_ms;
	if (res.status != XTN_OK)
		return res.status;
	xtn_output out{res.pairs.data(), res.dist.data(), nullptr, (int64_t)res.pairs.size()};
	if (callback)
		callback(user, &out);
	return XTN_OK;
}

void xtn_server_close(xtn_server* sv) {
	if (!sv)
		return;
	{
		std::lock_guard<std::mutex> lk(sv->mu);
		sv->stop = true;
	}
	sv->cv_work.notify_all();
	for (auto& t : sv->threads)
		t.join();
	for (xtn_handle* w : sv->workers)
		xtn_close(w);
	xtn_close(sv->base);
	delete sv;
}

// a copy per calling thread, taken under the mutex: the text stays valid
// until this thread's next call
const char* xtn_server_last_error(xtn_server* sv) {
	if (!sv)
		return "null server";
	static thread_local std::string copy;
	std::lock_guard<std::mutex> lk(sv->mu);
	copy = sv->err;
	return copy.c_str();
}

}  // extern "C"
