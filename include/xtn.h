/*
 * xtn.h -- C ABI of the MI355X XTNeighbor-streaming engine (libxtn.so).
 *
 * The reference has no exported library: its algorithm entry is
 *     void xtn_perform(XTNArgs args, SeqArray* seqArr, SeqInfo* seqInfo,
 *                      void callback(XTNOutput))          (xtn.cu:267, xtn.h:124)
 * compiled into the CLI's translation unit. This header is the boundary a
 * caller (the xt_neighbor CLI, a cgo/ctypes binding) binds instead. Each
 * declaration names the reference interface it replaces. All types are plain
 * C; no HIP or torch types cross the boundary. Buffers passed in are host
 * memory owned by the caller; buffers handed to a callback are owned by the
 * library and valid only during that call (xtn.cu:458-459, 478-479).
 *
 * Every function returns an int status (XTN_OK on success). The library never
 * calls exit(); xtn_last_error() describes the most recent failure of a handle.
 * A handle is not safe for concurrent use; distinct handles are independent.
 */
#ifndef XTN_H
#define XTN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* distance measures: LEVENSHTEIN / HAMMING (xtn.h:12-14) */
#define XTN_LEVENSHTEIN 0
#define XTN_HAMMING 1

/* status codes. XTN_OK, XTN_ERROR and XTN_EXIT keep the values of the
 * reference's ReturnCode {SUCCESS, ERROR, EXIT} (xtn.h:93). */
enum {
	XTN_OK = 0,
	XTN_ERROR = 1,         /* invalid input data (bad residue, length mismatch) */
	XTN_EXIT = 2,          /* reserved: the CLI's help/version exit */
	XTN_ERR_ARG = 3,       /* invalid argument (NULL pointer, distance < 1, ...) */
	XTN_ERR_HIP = 4,       /* HIP runtime failure */
	XTN_ERR_NOMEM = 5,     /* device or pinned host allocation failed */
	XTN_ERR_CAPACITY = 6,  /* workload exceeds a documented limit */
	XTN_ERR_STATE = 7      /* call order violated (e.g. run before set_sequences) */
};

/* Int2 (xtn.h:30-36): a pair of 0-based input row indexes, x < y. */
typedef struct {
	int32_t x, y;
} xtn_int2;

/* SeqInfo (xtn.h:41-45): per input row, its frequency (duplicate_count),
 * repertoire index and original row index. */
typedef struct {
	int32_t frequency, repertoire, original_index;
} xtn_seq_info;

/* XTNOutput (xtn.h:83-88). NN mode fills index_pairs + pairwise_distances
 * (ascending (x, y), possibly over several callbacks). Overlap mode fills
 * index_pairs (repertoire pairs, x <= y) + pairwise_frequencies, in exactly
 * one callback. */
typedef struct {
	const xtn_int2* index_pairs;
	const int8_t* pairwise_distances;
	const uint64_t* pairwise_frequencies;
	int64_t len;
} xtn_output;

/* the reference's `void callback(XTNOutput)` (xtn.cu:267) plus a user pointer */
typedef void (*xtn_callback)(void* user, const xtn_output* out);

/* XTNArgs (xtn.h:50-60): distance, measure, verbose. The engine knobs below
 * replace the reference's MemoryContext budgets (xtn.h:65-78), which it sized
 * from cudaMemGetInfo and sysinfo (xtn.cu:35-184). */
typedef struct {
	int32_t distance;              /* -d, >= 1 */
	int32_t measure;               /* -m: XTN_LEVENSHTEIN or XTN_HAMMING */
	int32_t verbose;               /* -V: stage timings on stdout */
	int32_t device;                /* HIP device ordinal */
	int64_t max_records_per_chunk; /* 0: size key-space chunks from free HBM;
	                                  > 0: cap deletion records per chunk
	                                  (forces the multi-chunk path) */
	int64_t output_chunk_pairs;    /* NN candidate pairs verified per output
	                                  chunk (0: default); each chunk is one
	                                  callback, copied on a side stream */
	int32_t profile;               /* 1: time every kernel launch with HIP
	                                  events (see xtn_kernel_profile) */
	int32_t bucket_capacity;       /* records one workgroup groups in LDS
	                                  (0 or above 3072: 3072); larger
	                                  buckets take the global sort path */
	int32_t no_graph;              /* 1: never capture or replay a run as a
	                                  HIP graph (0: a single-chunk run is
	                                  captured once it repeats and replayed
	                                  while it keeps its sizes) */
	int32_t coarse_separate;       /* 1: the records' first ordering pass (by
	                                  their top 8 hash bits) runs as a pass
	                                  of its own; 0: the generator does it
	                                  when one chunk uses bucket regions */
	int64_t max_pairs_per_round;   /* 0: candidate pairs per round sized from
	                                  free HBM (at most 2^30 - 1); > 0: cap
	                                  them. A search with more candidates runs
	                                  in rounds over ranges of the smaller id
	                                  (xtn.cu:385-470), outputs in order */
} xtn_config;

typedef struct xtn_handle xtn_handle;

/* counters of the last run (the reference prints the same quantities in
 * verbose mode: xtn.cu:336, 378, 430, 468, 472-473) */
typedef struct {
	int64_t n_sequences;   /* rows searched (unique rows in overlap mode) */
	int64_t n_variants;    /* deletion records generated (V) */
	int64_t n_chunks;      /* key-space chunks */
	int64_t n_runs;        /* key runs holding >= 2 records */
	int64_t n_candidates;  /* candidate pairs i != j emitted from runs (with
	                          multiplicity; the copies a sequence's repeated
	                          variants would add are not there when the
	                          generator's coarse pass left them out) */
	int64_t n_unique;      /* unique candidate pairs verified (U) */
	int64_t n_output;      /* pairs (NN) or repertoire entries (overlap) out */
	double ms_total;       /* device pipeline wall time */
	double ms_generate, ms_sort, ms_pairs, ms_verify, ms_output;
	int64_t n_rounds;      /* candidate rounds (ranges of the smaller id) */
	int64_t tail_checked;  /* 1: the candidate tail ran with the host checking
	                          each step (after a device-side wait hit its
	                          limit on this handle) */
	int64_t n_launches;    /* kernel launches of the last call */
	int64_t n_syncs;       /* host waits on the device of the last call */
	int64_t graph_replays; /* runs issued as one captured graph launch */
	int64_t tail_passes;   /* 8-bit LSD passes of the last candidate sort: fewer
	                          than the pair bits need when the select ranks
	                          the low bits of each group in LDS, all of them
	                          after a group outgrew the select's window on
	                          this handle */
	int64_t coarse_in_generator; /* 1: the generator ordered the records by their
	                                top 8 hash bits itself; 0: a separate pass did
	                                (always after a sub-range of the generator's
	                                pass overflowed on this handle, until new
	                                sequences) */
	int64_t n_records;     /* deletion records written: n_variants less the
	                          repeated variants of a sequence (deleting any
	                          residue of a run of equal ones gives the same
	                          string), which the generator's coarse pass
	                          leaves out; n_variants on the separate pass */
} xtn_stats;

/* fills *cfg with defaults: distance 1, Levenshtein, device 0 */
void xtn_config_init(xtn_config* cfg);

/* replaces the implicit CUDA context set-up of main() (cli.cu:227-277) */
int xtn_open(xtn_handle** out, const xtn_config* cfg);
void xtn_close(xtn_handle* h);
int xtn_set_config(xtn_handle* h, const xtn_config* cfg);

/* replaces SeqArray::toDevice (codec.cu:211-218): n sequences, chars
 * concatenated without separators, offsets[n+1] (offsets[0] = 0). Residues
 * must be ACDEFGHIKLMNPQRSTVWY (codec.cu:25-38); anything else -> XTN_ERROR.
 * The library keeps a device-resident copy until the next call or close. */
int xtn_set_sequences(xtn_handle* h, const char* chars, const uint32_t* offsets, int32_t n);

/* replaces xtn_perform (xtn.cu:267). info == NULL: nearest-neighbour mode,
 * callback invoked once per output chunk in ascending (x, y) order.
 * info != NULL (n entries, original_index ignored and taken as the row):
 * repertoire-overlap mode, callback invoked exactly once. */
int xtn_perform(xtn_handle* h, const xtn_seq_info* info, xtn_callback callback, void* user);

/* the NN device pipeline without the host hand-off: device-resident input to
 * device-resident sorted (x, y, distance) triples. *out_len receives the
 * output length. Used for device-side timing. */
int xtn_run_nn_resident(xtn_handle* h, int64_t* out_len);
/* copies up to cap triples of the last xtn_run_nn_resident to host memory */
int xtn_copy_nn_output(xtn_handle* h, xtn_int2* pairs, int8_t* distances, int64_t cap);

/* Many independent NN searches over the rows set by xtn_set_sequences (no
 * reference counterpart; the reference runs one input per process). Query set
 * s is the row range [set_offsets[s], set_offsets[s+1]); pairs are only ever
 * reported inside a set. batched == 1 runs every set through one pipeline
 * pass (keys salted with the set, candidates across sets dropped at
 * verification); batched == 0 runs one pass per set, as separate reference
 * invocations would. The callback receives each set with output once, in set
 * order, with set-local row indexes in ascending (x, y) order; callback == NULL
 * leaves the results on the device (timing). */
typedef void (*xtn_set_callback)(void* user, int32_t set, const xtn_output* out);
int xtn_perform_sets(xtn_handle* h, const uint32_t* set_offsets, int32_t n_sets, int32_t batched,
                     xtn_set_callback callback, void* user);

int xtn_get_stats(xtn_handle* h, xtn_stats* out);

/* ---- engine parameters and their autotuner (no reference counterpart: the
 * reference's sizes are constants, xtn_inner.cu:14, xtn.cu:36-46) */
typedef struct {
	int32_t bucket_capacity;  /* records one workgroup groups in LDS: 2048, 3072 (default) or
	                             4096 (62 KB of LDS, two workgroups per CU) */
	int32_t bucket_target;    /* mean records per bucket (sets the bucket bits); default 2200 */
	int32_t sort_threads;     /* threads of a candidate-sort tile: 256, 512 (default) or 1024 */
	int32_t sort_tile;        /* keys of a candidate-sort tile: 4096 (default; with 256 or 512
	                             threads) or 8192 (with 512 or 1024 threads) */
	int32_t select_threads;   /* threads of a select tile, 8 candidates each: 256 (default) or 512 */
	int32_t from_cache;       /* xtn_autotune: 1 when read from the cache file */
	double ms;                /* xtn_autotune: run time with these parameters */
	double ms_default;        /* xtn_autotune: run time with the defaults */
	int32_t candidates;       /* xtn_autotune: settings timed */
	int32_t sort_low_bits;    /* candidate bits below which the sort stops, ranked inside each
	                             group by the select: 0 (default: full sort), 8, 16 or 24 */
	int32_t coarse_separate;  /* 1: the records' first ordering pass runs as a pass of its own
	                             (as xtn_config.coarse_separate); 0 (default): inside the generator */
	int32_t reserved;
} xtn_tuning;

int xtn_get_tuning(xtn_handle* h, xtn_tuning* out);
/* XTN_ERR_ARG for values outside the lists above (bucket_target must lie in
 * [bucket_capacity / 4, bucket_capacity * 0.9]) */
int xtn_set_tuning(xtn_handle* h, const xtn_tuning* t);
/* Chooses the parameters for the sequences set on the handle and its
 * distance and measure: each candidate setting is run `runs` times (after
 * two warm-up runs) as xtn_run_nn_resident, its output checked against the
 * default setting's; the search goes coordinate by coordinate (bucket
 * capacity and target, then the sort tile, then the select tile, then the
 * sort's low bits, then the generator's coarse pass on or off) and keeps
 * the fastest, the default unless another beats it by more than 1% when
 * both are timed again. With cache_path, the choice is read from that JSON
 * file when an entry for the workload (sequence count, distance, measure,
 * variants per sequence) is there, else written to it. The handle keeps the
 * choice; *out (nullable) receives it. */
int xtn_autotune(xtn_handle* h, const char* cache_path, int32_t runs, xtn_tuning* out);

/* ---- served form (no reference counterpart: query sets against a resident index).
 * A server keeps the sorted deletion index of n index sequences resident in
 * HBM and answers query sets against it: every (query row, index row) pair
 * within the distance. `slots` worker streams run admitted queries
 * concurrently; a query is admitted, in arrival order, while the estimated
 * HBM of the queries in flight stays within hbm_budget_bytes. */
typedef struct {
	int32_t distance;          /* >= 1 */
	int32_t measure;           /* XTN_LEVENSHTEIN or XTN_HAMMING */
	int32_t device;
	int32_t slots;             /* concurrent queries (0: 4) */
	int64_t hbm_budget_bytes;  /* admission budget (0: a quarter of free HBM) */
	int32_t max_query_rows;    /* largest query set accepted (0: 1 << 20) */
	int32_t reserved;
} xtn_server_config;

typedef struct xtn_server xtn_server;

void xtn_server_config_init(xtn_server_config* cfg);
/* builds the index (deletion variants of the index sequences, sorted by key) */
int xtn_server_open(xtn_server** out, const xtn_server_config* cfg, const char* chars,
                    const uint32_t* offsets, int32_t n);
/* copies the query set and queues it; returns at once with a ticket */
int xtn_server_submit(xtn_server* s, const char* chars, const uint32_t* offsets, int32_t n,
                      uint64_t* ticket);
/* waits for a ticket; one callback with index_pairs = (query row, index row)
 * ascending and pairwise_distances; *latency_ms = submit to completion */
int xtn_server_wait(xtn_server* s, uint64_t ticket, xtn_callback callback, void* user,
                    double* latency_ms);
void xtn_server_close(xtn_server* s);
const char* xtn_server_last_error(xtn_server* s);

/* ---- search without device allocation (no reference counterpart).
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. The same
 * search as xtn_run_nn_resident on the caller's device buffers: d_chars and
 * d_offsets[n + 1] as xtn_set_sequences takes them (residues are not
 * validated here), d_pairs / d_dist of out_cap entries receive the sorted
 * (x, y, distance) triples, *out_len their count; stream is a hipStream_t
 * (NULL: the default stream), waited on before the call returns. Every
 * workspace is a slice of one static device array of 52 MB per GPU, held
 * until the process ends; the call makes no allocation and creates no stream
 * or event. Limits: n <= 65536 sequences, at most 2^20 candidate pairs and
 * 2^20 output pairs; beyond them XTN_ERR_CAPACITY. Calls on one GPU share its
 * array: they are serialised inside the library, and calls from several
 * processes use separate arrays. */
int xtn_nn_noalloc(const char* d_chars, const uint32_t* d_offsets, int32_t n, int32_t distance, int32_t measure,
                   xtn_int2* d_pairs, int8_t* d_dist, int64_t out_cap, int64_t* out_len, void* stream);

/* profile == 1 only: launches, summed HIP-event duration (ms) and summed
 * algorithmic HBM bytes of one kernel over the runs since the last reset. */
int xtn_kernel_profile(xtn_handle* h, const char* kernel, int64_t* launches, double* ms,
                       double* bytes);
int xtn_kernel_profile_reset(xtn_handle* h);

const char* xtn_last_error(xtn_handle* h);
const char* xtn_status_string(int status);
/* library version string ("2.0.0-mi355x") */
const char* xtn_version(void);

#ifdef __cplusplus
}
#endif
#endif
