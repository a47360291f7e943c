"""ctypes access to the CPU oracle (oracle/lib/libxtn_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product (xt-neighbor_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libxtn_oracle.so")

_lib = None


class Stats(ctypes.Structure):
    _fields_ = [("n_variants", ctypes.c_int64), ("n_candidates", ctypes.c_int64),
                ("n_unique", ctypes.c_int64), ("n_output", ctypes.c_int64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    i32p = ctypes.POINTER(ctypes.c_int32)
    L.xo_char_encode.argtypes = [ctypes.c_char]
    L.xo_char_encode.restype = ctypes.c_int
    L.xo_str_encode.argtypes = [ctypes.c_char_p, u32p]
    L.xo_str_decode.argtypes = [u32p, ctypes.c_char_p]
    L.xo_len_decode.argtypes = [u32p]
    L.xo_len_decode.restype = ctypes.c_int
    L.xo_remove_char.argtypes = [u32p, ctypes.c_int, u32p]
    L.xo_combination_len.argtypes = [ctypes.c_int, ctypes.c_int]
    L.xo_combination_len.restype = ctypes.c_int
    L.xo_expand_keys.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int]
    L.xo_expand_keys.restype = ctypes.c_int
    L.xo_stage1.argtypes = [u32p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u32p), ctypes.POINTER(i32p)]
    L.xo_stage1.restype = ctypes.c_int64
    L.xo_sort_stage.argtypes = [u32p, i32p, ctypes.c_int64]
    L.xo_gen_pairs.argtypes = [u32p, i32p, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                               ctypes.POINTER(i32p)]
    L.xo_gen_pairs.restype = ctypes.c_int64
    L.xo_distance.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.xo_distance.restype = ctypes.c_int
    L.xo_nn_port.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, ctypes.POINTER(i32p), ctypes.POINTER(Stats)]
    L.xo_nn_port.restype = ctypes.c_int64
    L.xo_nn_port_mt.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(i32p), ctypes.POINTER(Stats)]
    L.xo_nn_port_mt.restype = ctypes.c_int64
    L.xo_nn_brute.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(i32p)]
    L.xo_nn_brute.restype = ctypes.c_int64
    L.xo_neighbors_of.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(i32p)]
    L.xo_neighbors_of.restype = ctypes.c_int64
    L.xo_overlap_terms.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, i32p, i32p, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint64))]
    L.xo_overlap_terms.restype = ctypes.c_int64
    L.xo_overlap_port_mt.argtypes = [ctypes.c_char_p, u32p, ctypes.c_int, i32p, i32p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint64))]
    L.xo_overlap_port_mt.restype = ctypes.c_int64
    L.xo_free.argtypes = [ctypes.c_void_p]
    _lib = L
    return L


def _words(w) -> np.ndarray:
    return np.asarray(w, dtype=np.uint32)


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def char_encode(c: str) -> int:
    return lib().xo_char_encode(c.encode())


def str_encode(s: str) -> np.ndarray:
    out = np.zeros(3, dtype=np.uint32)
    lib().xo_str_encode(s.encode(), _p(out))
    return out


def str_decode(w) -> str:
    w = _words(w)
    buf = ctypes.create_string_buffer(19)
    lib().xo_str_decode(_p(w), buf)
    return buf.value.decode()


def len_decode(w) -> int:
    w = _words(w)
    return lib().xo_len_decode(_p(w))


def remove_char(w, pos: int) -> np.ndarray:
    w = _words(w)
    out = np.zeros(3, dtype=np.uint32)
    lib().xo_remove_char(_p(w), pos, _p(out))
    return out


def combination_len(length: int, d: int) -> int:
    return lib().xo_combination_len(length, d)


def expand_keys(s: str, d: int) -> list[str]:
    w = str_encode(s)
    n = combination_len(len_decode(w), d)
    out = np.zeros(3 * n, dtype=np.uint32)
    lib().xo_expand_keys(_p(w), d, _p(out), n)
    return [str_decode(out[3 * i:3 * i + 3]) for i in range(n)]


def _take_u32(ptr, n):
    arr = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
    lib().xo_free(ctypes.cast(ptr, ctypes.c_void_p))
    return arr


def _take_i32(ptr, n):
    arr = np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.int32)
    lib().xo_free(ctypes.cast(ptr, ctypes.c_void_p))
    return arr


def stage1(seqs: list[str], d: int) -> tuple[list[str], list[int]]:
    words = np.concatenate([str_encode(s) for s in seqs]).astype(np.uint32)
    kp = ctypes.POINTER(ctypes.c_uint32)()
    vp = ctypes.POINTER(ctypes.c_int32)()
    v = lib().xo_stage1(_p(words), len(seqs), d, ctypes.byref(kp), ctypes.byref(vp))
    keys = _take_u32(kp, 3 * v)
    vals = _take_i32(vp, v)
    return [str_decode(keys[3 * i:3 * i + 3]) for i in range(v)], vals.tolist()


def sort_stage(keys: list[str], values: list[int]) -> tuple[list[str], list[int]]:
    k = np.concatenate([str_encode(s) for s in keys]).astype(np.uint32)
    v = np.asarray(values, dtype=np.int32)
    lib().xo_sort_stage(_p(k), v.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(values))
    return [str_decode(k[3 * i:3 * i + 3]) for i in range(len(values))], v.tolist()


def gen_pairs(keys: list[str], values: list[int], lowerbound: int):
    k = np.concatenate([str_encode(s) for s in keys]).astype(np.uint32)
    v = np.asarray(values, dtype=np.int32)
    n = ctypes.c_int64(len(values))
    pp = ctypes.POINTER(ctypes.c_int32)()
    np_ = lib().xo_gen_pairs(_p(k), v.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(n),
                             lowerbound, ctypes.byref(pp))
    pairs = _take_i32(pp, 2 * np_).reshape(-1, 2)
    kept = n.value
    return ([tuple(p) for p in pairs.tolist()],
            [str_decode(k[3 * i:3 * i + 3]) for i in range(kept)], v[:kept].tolist())


def distance(a: str, b: str, measure: int = 0) -> int:
    return lib().xo_distance(a.encode(), len(a), b.encode(), len(b), measure)


def _seq_args(chars: np.ndarray, offsets: np.ndarray):
    chars = np.ascontiguousarray(chars, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    return chars, offsets


def nn_port(chars, offsets, d: int, measure: int = 0, rounds: int = 1):
    """(triples int32[n,3], stats dict)"""
    chars, offsets = _seq_args(chars, offsets)
    out = ctypes.POINTER(ctypes.c_int32)()
    st = Stats()
    n = lib().xo_nn_port(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1, d,
                         measure, rounds, ctypes.byref(out), ctypes.byref(st))
    tri = _take_i32(out, 3 * n).reshape(-1, 3)
    return tri, {name: getattr(st, name) for name, _ in Stats._fields_}


def nn_port_mt(chars, offsets, d: int, measure: int = 0, threads: int = 1):
    """nn_port on `threads` OpenMP threads: (triples int32[n,3], stats dict)"""
    chars, offsets = _seq_args(chars, offsets)
    out = ctypes.POINTER(ctypes.c_int32)()
    st = Stats()
    n = lib().xo_nn_port_mt(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1, d,
                            measure, threads, ctypes.byref(out), ctypes.byref(st))
    tri = _take_i32(out, 3 * n).reshape(-1, 3)
    return tri, {name: getattr(st, name) for name, _ in Stats._fields_}


def nn_brute(chars, offsets, d: int, measure: int = 0) -> np.ndarray:
    chars, offsets = _seq_args(chars, offsets)
    out = ctypes.POINTER(ctypes.c_int32)()
    n = lib().xo_nn_brute(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1, d,
                          measure, ctypes.byref(out))
    return _take_i32(out, 3 * n).reshape(-1, 3)


def overlap(chars, offsets, freq, rep, d: int, measure: int = 0, which: int = 3) -> np.ndarray:
    """uint64[m,3] (A, B, sum) sorted by (A, B); which: 1 exact-match term,
    2 neighbour term, 3 both (the reference's output)"""
    chars, offsets = _seq_args(chars, offsets)
    f = np.ascontiguousarray(freq, dtype=np.int32)
    r = np.ascontiguousarray(rep, dtype=np.int32)
    out = ctypes.POINTER(ctypes.c_uint64)()
    m = lib().xo_overlap_terms(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1,
                               f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), d, measure, which,
                               ctypes.byref(out))
    arr = np.ctypeslib.as_array(out, shape=(3 * m,)).copy() if m else np.zeros(0, np.uint64)
    lib().xo_free(ctypes.cast(out, ctypes.c_void_p))
    return arr.reshape(-1, 3)


def overlap_mt(chars, offsets, freq, rep, d: int, measure: int = 0, threads: int = 1) -> np.ndarray:
    """overlap() with the neighbour search on `threads` OpenMP threads"""
    chars, offsets = _seq_args(chars, offsets)
    f = np.ascontiguousarray(freq, dtype=np.int32)
    r = np.ascontiguousarray(rep, dtype=np.int32)
    out = ctypes.POINTER(ctypes.c_uint64)()
    m = lib().xo_overlap_port_mt(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1,
                                 f.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                 r.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), d, measure, threads,
                                 ctypes.byref(out))
    arr = np.ctypeslib.as_array(out, shape=(3 * m,)).copy() if m else np.zeros(0, np.uint64)
    lib().xo_free(ctypes.cast(out, ctypes.c_void_p))
    return arr.reshape(-1, 3)


def neighbors_of(chars, offsets, row: int, d: int, measure: int = 0) -> np.ndarray:
    """int64[m,2] (j, dist): every other row within distance d of `row`"""
    chars, offsets = _seq_args(chars, offsets)
    out = ctypes.POINTER(ctypes.c_int32)()
    m = lib().xo_neighbors_of(chars.ctypes.data_as(ctypes.c_char_p), _p(offsets), len(offsets) - 1,
                              row, d, measure, ctypes.byref(out))
    return _take_i32(out, 2 * m).reshape(-1, 2).astype(np.int64)
