"""ctypes bindings for libxtn.so (the MI355X XTNeighbor engine) and
libxtn_synth.so (seeded synthetic CDR3 workloads).

The Engine class mirrors the reference entry point
``xtn_perform(XTNArgs, SeqArray*, SeqInfo*, callback)``
(xtneighbor_streaming/src/xtn.cu:267): sequences are handed over once
(``SeqArray::toDevice``, codec.cu:211-218), then ``nn()`` runs nearest-neighbour
mode and ``overlap()`` repertoire-overlap mode. Nothing here falls back to a CPU
path: when the library is missing or cannot reach a GPU the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libxtn.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libxtn_synth.so")
CLI_PATH = os.path.join(HERE, "bin", "xt_neighbor")
CLI1_PATH = os.path.join(HERE, "bin", "xtneighbor")  # the first, non-streaming command line

LEVENSHTEIN = 0
HAMMING = 1

STATUS = {
    0: "XTN_OK", 1: "XTN_ERROR", 2: "XTN_EXIT", 3: "XTN_ERR_ARG", 4: "XTN_ERR_HIP",
    5: "XTN_ERR_NOMEM", 6: "XTN_ERR_CAPACITY", 7: "XTN_ERR_STATE",
}

# every symbol include/xtn.h declares
EXPORTED = [
    "xtn_config_init", "xtn_open", "xtn_close", "xtn_set_config", "xtn_set_sequences",
    "xtn_perform", "xtn_run_nn_resident", "xtn_copy_nn_output", "xtn_get_stats",
    "xtn_kernel_profile", "xtn_kernel_profile_reset", "xtn_last_error", "xtn_status_string",
    "xtn_version", "xtn_perform_sets", "xtn_server_config_init", "xtn_server_open",
    "xtn_server_submit", "xtn_server_wait", "xtn_server_close", "xtn_server_last_error",
    "xtn_nn_noalloc", "xtn_get_tuning", "xtn_set_tuning", "xtn_autotune",
]


class XtnError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Int2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32)]


class SeqInfo(ctypes.Structure):
    _fields_ = [("frequency", ctypes.c_int32), ("repertoire", ctypes.c_int32),
                ("original_index", ctypes.c_int32)]


class Output(ctypes.Structure):
    _fields_ = [("index_pairs", ctypes.POINTER(Int2)),
                ("pairwise_distances", ctypes.POINTER(ctypes.c_int8)),
                ("pairwise_frequencies", ctypes.POINTER(ctypes.c_uint64)),
                ("len", ctypes.c_int64)]


class Config(ctypes.Structure):
    _fields_ = [("distance", ctypes.c_int32), ("measure", ctypes.c_int32),
                ("verbose", ctypes.c_int32), ("device", ctypes.c_int32),
                ("max_records_per_chunk", ctypes.c_int64),
                ("output_chunk_pairs", ctypes.c_int64),
                ("profile", ctypes.c_int32), ("bucket_capacity", ctypes.c_int32),
                ("no_graph", ctypes.c_int32), ("coarse_separate", ctypes.c_int32),
                ("max_pairs_per_round", ctypes.c_int64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("distance", ctypes.c_int32), ("measure", ctypes.c_int32),
                ("device", ctypes.c_int32), ("slots", ctypes.c_int32),
                ("hbm_budget_bytes", ctypes.c_int64), ("max_query_rows", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("n_sequences", ctypes.c_int64), ("n_variants", ctypes.c_int64),
                ("n_chunks", ctypes.c_int64), ("n_runs", ctypes.c_int64),
                ("n_candidates", ctypes.c_int64), ("n_unique", ctypes.c_int64),
                ("n_output", ctypes.c_int64), ("ms_total", ctypes.c_double),
                ("ms_generate", ctypes.c_double), ("ms_sort", ctypes.c_double),
                ("ms_pairs", ctypes.c_double), ("ms_verify", ctypes.c_double),
                ("ms_output", ctypes.c_double), ("n_rounds", ctypes.c_int64),
                ("tail_checked", ctypes.c_int64), ("n_launches", ctypes.c_int64),
                ("n_syncs", ctypes.c_int64), ("graph_replays", ctypes.c_int64),
                ("tail_passes", ctypes.c_int64), ("coarse_in_generator", ctypes.c_int64),
                ("n_records", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    _fields_ = [("bucket_capacity", ctypes.c_int32), ("bucket_target", ctypes.c_int32),
                ("sort_threads", ctypes.c_int32), ("sort_tile", ctypes.c_int32),
                ("select_threads", ctypes.c_int32), ("from_cache", ctypes.c_int32),
                ("ms", ctypes.c_double), ("ms_default", ctypes.c_double),
                ("candidates", ctypes.c_int32), ("sort_low_bits", ctypes.c_int32),
                ("coarse_separate", ctypes.c_int32), ("reserved", ctypes.c_int32)]


CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Output))
SET_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Output))

_lib = None
_synth = None


_HIP = None


def _hip_runtime():
    """the HIP runtime library (for pinned host buffers), or None"""
    global _HIP
    if _HIP is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7"):
            try:
                _HIP = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _HIP is not None:
            _HIP.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
            _HIP.hipHostFree.argtypes = [ctypes.c_void_p]
    return _HIP


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads libxtn.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C xt-neighbor_amd`")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.xtn_config_init.argtypes = [ctypes.POINTER(Config)]
    lib.xtn_config_init.restype = None
    lib.xtn_open.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Config)]
    lib.xtn_close.argtypes = [vp]
    lib.xtn_close.restype = None
    lib.xtn_set_config.argtypes = [vp, ctypes.POINTER(Config)]
    lib.xtn_set_sequences.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32]
    lib.xtn_perform.argtypes = [vp, ctypes.POINTER(SeqInfo), CALLBACK, vp]
    lib.xtn_run_nn_resident.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.xtn_copy_nn_output.argtypes = [vp, ctypes.POINTER(Int2), ctypes.POINTER(ctypes.c_int8), ctypes.c_int64]
    lib.xtn_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    lib.xtn_perform_sets.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32, ctypes.c_int32,
                                     SET_CALLBACK, vp]
    lib.xtn_kernel_profile.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.xtn_kernel_profile_reset.argtypes = [vp]
    lib.xtn_last_error.argtypes = [vp]
    lib.xtn_last_error.restype = ctypes.c_char_p
    lib.xtn_status_string.argtypes = [ctypes.c_int]
    lib.xtn_status_string.restype = ctypes.c_char_p
    lib.xtn_version.restype = ctypes.c_char_p
    lib.xtn_server_config_init.argtypes = [ctypes.POINTER(ServerConfig)]
    lib.xtn_server_config_init.restype = None
    lib.xtn_server_open.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(ServerConfig), ctypes.c_char_p,
                                    ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32]
    lib.xtn_server_open.restype = ctypes.c_int
    lib.xtn_server_submit.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_uint64)]
    lib.xtn_server_submit.restype = ctypes.c_int
    lib.xtn_server_wait.argtypes = [vp, ctypes.c_uint64, CALLBACK, vp, ctypes.POINTER(ctypes.c_double)]
    lib.xtn_server_wait.restype = ctypes.c_int
    lib.xtn_server_close.argtypes = [vp]
    lib.xtn_server_close.restype = None
    lib.xtn_server_last_error.argtypes = [vp]
    lib.xtn_server_last_error.restype = ctypes.c_char_p
    lib.xtn_nn_noalloc.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp,
                                   ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), vp]
    lib.xtn_nn_noalloc.restype = ctypes.c_int
    lib.xtn_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.xtn_get_tuning.restype = ctypes.c_int
    lib.xtn_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    lib.xtn_set_tuning.restype = ctypes.c_int
    lib.xtn_autotune.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(Tuning)]
    lib.xtn_autotune.restype = ctypes.c_int
    for name in ("xtn_open", "xtn_set_config", "xtn_set_sequences", "xtn_perform",
                 "xtn_run_nn_resident", "xtn_copy_nn_output", "xtn_get_stats", "xtn_perform_sets",
                 "xtn_kernel_profile", "xtn_kernel_profile_reset"):
        getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def load_synth(path: str = SYNTH_PATH) -> ctypes.CDLL:
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C xt-neighbor_amd`")
    s = ctypes.CDLL(path)
    c_char_pp = ctypes.c_char_p
    u32p = ctypes.POINTER(ctypes.c_uint32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    s.xs_generate_nn.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                 c_char_pp, ctypes.c_int64, u32p]
    s.xs_generate_nn.restype = ctypes.c_int64
    s.xs_overlap_sizes.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, i32p]
    s.xs_overlap_sizes.restype = ctypes.c_int64
    s.xs_generate_overlap.argtypes = [ctypes.c_uint64, ctypes.c_int32, i32p, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_double, ctypes.c_int32, c_char_pp, ctypes.c_int64,
                                      u32p, i32p]
    s.xs_generate_overlap.restype = ctypes.c_int64
    s.xs_generate_queries.argtypes = [ctypes.c_uint64, c_char_pp, u32p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_double, c_char_pp, ctypes.c_int64, u32p]
    s.xs_generate_queries.restype = ctypes.c_int64
    s.xs_write_csv.argtypes = [ctypes.c_char_p, c_char_pp, u32p, ctypes.c_int32, i32p]
    s.xs_write_csv.restype = ctypes.c_int32
    _synth = s
    return s


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _i32p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def encode(seqs: Iterable[str]) -> tuple[np.ndarray, np.ndarray]:
    """list of strings -> (uint8 chars, uint32 offsets[n+1])"""
    seqs = list(seqs)
    lens = np.fromiter((len(s) for s in seqs), dtype=np.int64, count=len(seqs))
    offsets = np.zeros(len(seqs) + 1, dtype=np.uint32)
    np.cumsum(lens, out=offsets[1:])
    chars = np.frombuffer("".join(seqs).encode("ascii"), dtype=np.uint8).copy()
    return chars, offsets


def decode(chars: np.ndarray, offsets: np.ndarray) -> list[str]:
    b = chars.tobytes()
    return [b[offsets[i]:offsets[i + 1]].decode("ascii") for i in range(len(offsets) - 1)]


def synth_nn(seed: int, n: int, mu: float = 15.0, sigma: float = 2.0, lmin: int = 8, lmax: int = 22,
             planted: float = 0.2, max_edits: int = 1) -> tuple[np.ndarray, np.ndarray]:
    """seeded unique synthetic CDR3 set (chars, offsets)"""
    s = load_synth()
    cap = int(n * (lmax + 2) + 64)
    chars = np.zeros(cap, dtype=np.uint8)
    offsets = np.zeros(n + 1, dtype=np.uint32)
    total = s.xs_generate_nn(seed, n, mu, sigma, lmin, lmax, planted, max_edits,
                             chars.ctypes.data_as(ctypes.c_char_p), cap, _u32p(offsets))
    if total < 0:
        raise RuntimeError(f"xs_generate_nn failed ({total})")
    return chars[:total].copy(), offsets


def synth_overlap(seed: int, n_reps: int, size_min: int, size_max: int, shared: float = 0.1,
                  mu: float = 15.0, sigma: float = 2.0, lmin: int = 8, lmax: int = 22,
                  zipf_s: float = 2.0, zipf_max: int = 100000):
    """repertoires (chars, offsets, freq, rep_sizes)"""
    s = load_synth()
    sizes = np.zeros(n_reps, dtype=np.int32)
    n = s.xs_overlap_sizes(seed, n_reps, size_min, size_max, _i32p(sizes))
    cap = int(n * (lmax + 2) + 64)
    chars = np.zeros(cap, dtype=np.uint8)
    offsets = np.zeros(n + 1, dtype=np.uint32)
    freq = np.zeros(n, dtype=np.int32)
    total = s.xs_generate_overlap(seed, n_reps, _i32p(sizes), shared, mu, sigma, lmin, lmax, zipf_s,
                                  zipf_max, chars.ctypes.data_as(ctypes.c_char_p), cap, _u32p(offsets),
                                  _i32p(freq))
    if total < 0:
        raise RuntimeError(f"xs_generate_overlap failed ({total})")
    return chars[:total].copy(), offsets, freq, sizes


def synth_queries(seed: int, idx_chars: np.ndarray, idx_offsets: np.ndarray, n: int,
                  mu: float = 15.0, sigma: float = 2.0, lmin: int = 8, lmax: int = 22,
                  planted: float = 0.3):
    s = load_synth()
    cap = int(n * (lmax + 2) + 64)
    chars = np.zeros(cap, dtype=np.uint8)
    offsets = np.zeros(n + 1, dtype=np.uint32)
    total = s.xs_generate_queries(seed, idx_chars.ctypes.data_as(ctypes.c_char_p), _u32p(idx_offsets),
                                  len(idx_offsets) - 1, n, mu, sigma, lmin, lmax, planted,
                                  chars.ctypes.data_as(ctypes.c_char_p), cap, _u32p(offsets))
    if total < 0:
        raise RuntimeError(f"xs_generate_queries failed ({total})")
    return chars[:total].copy(), offsets


def write_csv(path: str, chars: np.ndarray, offsets: np.ndarray, freq: np.ndarray | None = None) -> None:
    """the sequences (and frequencies) as the CLI's input CSV"""
    chars = np.ascontiguousarray(chars, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    fp = None
    if freq is not None:
        freq = np.ascontiguousarray(freq, dtype=np.int32)
        fp = _i32p(freq)
    rc = load_synth().xs_write_csv(path.encode(), chars.ctypes.data_as(ctypes.c_char_p), _u32p(offsets),
                                   len(offsets) - 1, fp)
    if rc != 0:
        raise OSError(f"cannot write {path}")


def reps_from_sizes(sizes: Sequence[int]) -> np.ndarray:
    """repertoire index per row, as parse_info assigns it (cli.cu:178-182)"""
    return np.repeat(np.arange(len(sizes), dtype=np.int32), np.asarray(sizes, dtype=np.int64))


class Engine:
    """One device handle (xtn_open). Not thread safe; use one per thread."""

    def __init__(self, distance: int = 1, measure: int | str = LEVENSHTEIN, device: int = 0,
                 max_records_per_chunk: int = 0, output_chunk_pairs: int = 0,
                 profile: bool = False, verbose: bool = False, bucket_capacity: int = 0,
                 max_pairs_per_round: int = 0, no_graph: bool = False):
        self.lib = load_library()
        self.cfg = Config()
        self.lib.xtn_config_init(ctypes.byref(self.cfg))
        self.cfg.distance = distance
        self.cfg.measure = {"leven": LEVENSHTEIN, "hamming": HAMMING}.get(measure, measure)
        self.cfg.device = device
        self.cfg.max_records_per_chunk = max_records_per_chunk
        self.cfg.output_chunk_pairs = output_chunk_pairs
        self.cfg.profile = 1 if profile else 0
        self.cfg.verbose = 1 if verbose else 0
        self.cfg.bucket_capacity = bucket_capacity
        self.cfg.max_pairs_per_round = max_pairs_per_round
        self.cfg.no_graph = 1 if no_graph else 0
        self.h = ctypes.c_void_p()
        st = self.lib.xtn_open(ctypes.byref(self.h), ctypes.byref(self.cfg))
        if st != 0:
            raise XtnError(st, "xtn_open failed (no GPU or invalid configuration)")
        self._keep = None
        self._pin = ctypes.c_void_p()
        self._pin_bytes = 0

    def _staging(self, nbytes: int):
        """a pinned host buffer of at least nbytes, kept for later calls: the
        library's uploads from it skip the runtime's pageable staging copy
        (None when the runtime cannot allocate one)"""
        if nbytes > self._pin_bytes:
            self._free_staging()
            hip = _hip_runtime()
            p = ctypes.c_void_p()
            if hip is None or hip.hipHostMalloc(ctypes.byref(p), nbytes, 0) != 0 or not p.value:
                return None
            self._pin, self._pin_bytes = p, nbytes
        return self._pin

    def _free_staging(self):
        if self._pin.value:
            _hip_runtime().hipHostFree(self._pin)
        self._pin, self._pin_bytes = ctypes.c_void_p(), 0

    def _check(self, st: int):
        if st != 0:
            raise XtnError(st, self.lib.xtn_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.xtn_close(self.h)
            self.h = ctypes.c_void_p()
        self._free_staging()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, **kw):
        for k, v in kw.items():
            if k == "measure":
                v = {"leven": LEVENSHTEIN, "hamming": HAMMING}.get(v, v)
            if k in ("profile", "no_graph"):
                v = 1 if v else 0
            setattr(self.cfg, k, v)
        self._check(self.lib.xtn_set_config(self.h, ctypes.byref(self.cfg)))

    def set_sequences(self, seqs=None, chars: np.ndarray | None = None, offsets: np.ndarray | None = None):
        if seqs is not None:
            chars, offsets = encode(seqs)
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        self._keep = (chars, offsets)
        n = len(offsets) - 1
        self._check(self.lib.xtn_set_sequences(self.h, chars.ctypes.data_as(ctypes.c_char_p),
                                               _u32p(offsets), n))

    def nn(self) -> tuple[np.ndarray, np.ndarray, int]:
        """(pairs int32[n,2], distances int8[n], number of callbacks)"""
        parts_p, parts_d = [], []

        def cb(_user, out_p):
            o = out_p.contents
            n = o.len
            if n:
                p = np.ctypeslib.as_array(ctypes.cast(o.index_pairs, ctypes.POINTER(ctypes.c_int32)),
                                          shape=(n * 2,)).reshape(n, 2).copy()
                d = np.ctypeslib.as_array(o.pairwise_distances, shape=(n,)).copy()
            else:
                p = np.zeros((0, 2), np.int32)
                d = np.zeros(0, np.int8)
            parts_p.append(p)
            parts_d.append(d)

        c = CALLBACK(cb)
        self._check(self.lib.xtn_perform(self.h, None, c, None))
        if not parts_p:
            return np.zeros((0, 2), np.int32), np.zeros(0, np.int8), 0
        return np.concatenate(parts_p), np.concatenate(parts_d), len(parts_p)

    def nn_resident(self) -> int:
        n = ctypes.c_int64()
        self._check(self.lib.xtn_run_nn_resident(self.h, ctypes.byref(n)))
        return n.value

    def copy_nn_output(self, n: int) -> tuple[np.ndarray, np.ndarray]:
        pairs = np.zeros((n, 2), dtype=np.int32)
        dist = np.zeros(n, dtype=np.int8)
        if n:
            self._check(self.lib.xtn_copy_nn_output(
                self.h, pairs.ctypes.data_as(ctypes.POINTER(Int2)),
                dist.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), n))
        return pairs, dist

    def overlap(self, freq: Sequence[int], rep: Sequence[int]) -> np.ndarray:
        """uint64[n,3] rows (A, B, overlap) sorted by (A, B)"""
        freq = np.asarray(freq, dtype=np.int32)
        rep = np.asarray(rep, dtype=np.int32)
        n = len(freq)
        dt = np.dtype([("frequency", "<i4"), ("repertoire", "<i4"), ("original_index", "<i4")])
        m = max(n, 1)
        pin = self._staging(m * dt.itemsize)  # rows built in pinned memory when possible
        if pin is not None:
            rows = np.ctypeslib.as_array(ctypes.cast(pin, ctypes.POINTER(ctypes.c_uint8)),
                                         shape=(m * dt.itemsize,)).view(dt)
            rows[n:] = 0
        else:
            rows = np.zeros(m, dtype=dt)
        rows["frequency"][:n] = freq
        rows["repertoire"][:n] = rep
        rows["original_index"][:n] = np.arange(n, dtype=np.int32)
        info = rows.ctypes.data_as(ctypes.POINTER(SeqInfo))
        parts = []
        calls = [0]

        def cb(_user, out_p):
            o = out_p.contents
            m = o.len
            calls[0] += 1
            arr = np.zeros((m, 3), dtype=np.uint64)
            if m:
                p = np.ctypeslib.as_array(ctypes.cast(o.index_pairs, ctypes.POINTER(ctypes.c_int32)),
                                          shape=(m * 2,)).reshape(m, 2)
                f = np.ctypeslib.as_array(o.pairwise_frequencies, shape=(m,))
                arr[:, 0] = p[:, 0]
                arr[:, 1] = p[:, 1]
                arr[:, 2] = f
            parts.append(arr)

        c = CALLBACK(cb)
        self._check(self.lib.xtn_perform(self.h, info, c, None))
        self.overlap_calls = calls[0]
        return np.concatenate(parts) if parts else np.zeros((0, 3), np.uint64)

    def nn_sets(self, set_offsets, batched: bool = True, collect: bool = True) -> dict:
        """{set: int64[n,3] (x, y, dist) set-local} for every set with output"""
        so = np.ascontiguousarray(set_offsets, dtype=np.uint32)
        res = {}

        def cb(_user, k, out_p):
            o = out_p.contents
            n = o.len
            p = np.ctypeslib.as_array(ctypes.cast(o.index_pairs, ctypes.POINTER(ctypes.c_int32)),
                                      shape=(n * 2,)).reshape(n, 2).astype(np.int64)
            d = np.ctypeslib.as_array(o.pairwise_distances, shape=(n,)).astype(np.int64)
            res[int(k)] = np.concatenate([p, d[:, None]], axis=1)

        c = SET_CALLBACK(cb) if collect else ctypes.cast(None, SET_CALLBACK)
        self._check(self.lib.xtn_perform_sets(self.h, _u32p(so), len(so) - 1, 1 if batched else 0, c, None))
        return res

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self.lib.xtn_get_tuning(self.h, ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def set_tuning(self, bucket_capacity: int = 3072, bucket_target: int = 2200, sort_threads: int = 512,
                   sort_tile: int = 4096, select_threads: int = 256, sort_low_bits: int = 0,
                   coarse_separate: int = 0):
        t = Tuning(bucket_capacity=bucket_capacity, bucket_target=bucket_target, sort_threads=sort_threads,
                   sort_tile=sort_tile, select_threads=select_threads, sort_low_bits=sort_low_bits,
                   coarse_separate=coarse_separate)
        self._check(self.lib.xtn_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, cache_path: str | None = None, runs: int = 5) -> dict:
        """chooses and keeps the engine parameters for the current sequences
        (xtn_autotune); returns them with the measured times"""
        t = Tuning()
        self._check(self.lib.xtn_autotune(self.h, cache_path.encode() if cache_path else None, runs,
                                          ctypes.byref(t)))
        return {name: getattr(t, name) for name, _ in Tuning._fields_}

    def stats(self) -> dict:
        s = Stats()
        self._check(self.lib.xtn_get_stats(self.h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in Stats._fields_}

    def kernel_profile(self, name: str) -> tuple[int, float, float]:
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        b = ctypes.c_double()
        self._check(self.lib.xtn_kernel_profile(self.h, name.encode(), ctypes.byref(n),
                                                ctypes.byref(ms), ctypes.byref(b)))
        return n.value, ms.value, b.value

    def kernel_profile_reset(self):
        self._check(self.lib.xtn_kernel_profile_reset(self.h))


class Server:
    """A resident deletion index over `index` sequences answering query sets
    (xtn_server_*). submit() returns a ticket at once; wait(ticket) returns
    (pairs int32[n,2] = (query row, index row), distances int8[n], latency ms)."""

    def __init__(self, chars: np.ndarray, offsets: np.ndarray, distance: int = 1,
                 measure: int | str = LEVENSHTEIN, slots: int = 4, hbm_budget_bytes: int = 0,
                 max_query_rows: int = 0, device: int = 0):
        self.lib = load_library()
        cfg = ServerConfig()
        self.lib.xtn_server_config_init(ctypes.byref(cfg))
        cfg.distance = distance
        cfg.measure = {"leven": LEVENSHTEIN, "hamming": HAMMING}.get(measure, measure)
        cfg.slots = slots
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_query_rows = max_query_rows
        cfg.device = device
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        self.s = ctypes.c_void_p()
        st = self.lib.xtn_server_open(ctypes.byref(self.s), ctypes.byref(cfg),
                                      chars.ctypes.data_as(ctypes.c_char_p), _u32p(offsets), len(offsets) - 1)
        if st != 0:
            raise XtnError(st, "xtn_server_open failed")

    def submit(self, chars: np.ndarray, offsets: np.ndarray) -> int:
        chars = np.ascontiguousarray(chars, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
        t = ctypes.c_uint64()
        st = self.lib.xtn_server_submit(self.s, chars.ctypes.data_as(ctypes.c_char_p), _u32p(offsets),
                                        len(offsets) - 1, ctypes.byref(t))
        if st != 0:
            raise XtnError(st, self.lib.xtn_server_last_error(self.s).decode())
        return t.value

    def wait(self, ticket: int):
        got = {}

        def cb(_user, out_p):
            o = out_p.contents
            n = o.len
            if n:
                got["p"] = np.ctypeslib.as_array(ctypes.cast(o.index_pairs, ctypes.POINTER(ctypes.c_int32)),
                                                 shape=(n * 2,)).reshape(n, 2).copy()
                got["d"] = np.ctypeslib.as_array(o.pairwise_distances, shape=(n,)).copy()

        c = CALLBACK(cb)
        lat = ctypes.c_double()
        st = self.lib.xtn_server_wait(self.s, ticket, c, None, ctypes.byref(lat))
        if st != 0:
            raise XtnError(st, self.lib.xtn_server_last_error(self.s).decode())
        return got.get("p", np.zeros((0, 2), np.int32)), got.get("d", np.zeros(0, np.int8)), lat.value

    def close(self):
        if self.s:
            self.lib.xtn_server_close(self.s)
            self.s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
