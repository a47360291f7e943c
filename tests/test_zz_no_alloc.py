"""The search without device allocation (xtn_nn_noalloc) against the
resident search (xtn_run_nn_resident) on the same inputs: bit-exact triples,
two calls in a row, the second on a smaller input (the static scratch keeps
the first call's contents), both measures; and its size limits."""
import ctypes

import numpy as np
import pytest

import xtn

pytestmark = pytest.mark.gpu


class DevBuf:
    """device memory owned by the test (the caller of the entry point)"""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(nbytes, 256)) == 0
        self.nbytes = nbytes

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0

    def download(self, dtype, count):
        out = np.zeros(count, dtype=dtype)
        if count:
            assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def hip_runtime():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def noalloc(lib, hip, chars, offsets, d, measure, out_cap):
    n = len(offsets) - 1
    dc, do = DevBuf(hip, len(chars)), DevBuf(hip, 4 * (n + 1))
    dp, dd = DevBuf(hip, 8 * out_cap), DevBuf(hip, out_cap)
    try:
        dc.upload(chars)
        do.upload(offsets.astype(np.uint32))
        out_len = ctypes.c_int64(-1)
        st = lib.xtn_nn_noalloc(dc.p, do.p, n, d, measure, dp.p, dd.p, out_cap, ctypes.byref(out_len), None)
        if st != 0:
            return st, None
        m = out_len.value
        pairs = dp.download(np.int32, 2 * m).reshape(-1, 2).astype(np.int64)
        dist = dd.download(np.int8, m).astype(np.int64)
        return 0, np.concatenate([pairs, dist[:, None]], axis=1)
    finally:
        for b in (dc, do, dp, dd):
            b.free()


def resident(chars, offsets, d, measure):
    with xtn.Engine(distance=d, measure=measure) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        n = eng.nn_resident()
        p, dist = eng.copy_nn_output(n)
    return np.concatenate([p.astype(np.int64), dist.astype(np.int64)[:, None]], axis=1)


@pytest.mark.parametrize("d,measure", [(2, 0), (1, 1)])
def test_noalloc_matches_resident_twice(d, measure):
    lib = xtn.load_library()
    hip = hip_runtime()
    big = xtn.synth_nn(61, 8000, planted=0.25, max_edits=2)
    small = xtn.synth_nn(62, 1500, planted=0.3, max_edits=2)
    for chars, offsets in (big, small):
        want = resident(chars, offsets, d, measure)
        assert len(want) > 0
        st, got = noalloc(lib, hip, chars, offsets, d, measure, 1 << 20)
        assert st == 0
        assert np.array_equal(got, want)


def test_noalloc_limits():
    lib = xtn.load_library()
    hip = hip_runtime()
    chars, offsets = xtn.synth_nn(63, 2000, planted=0.3, max_edits=1)
    want = resident(chars, offsets, 1, 0)
    assert len(want) > 2
    # an output buffer too small for the result
    st, _ = noalloc(lib, hip, chars, offsets, 1, 0, len(want) - 1)
    assert st == 6
    # too many sequences for the scratch
    big_c, big_o = xtn.synth_nn(64, 70000, planted=0.1, max_edits=1)
    st, _ = noalloc(lib, hip, big_c, big_o, 1, 0, 1 << 20)
    assert st == 6
    # and it still works after the rejected calls
    st, got = noalloc(lib, hip, chars, offsets, 1, 0, 1 << 20)
    assert st == 0 and np.array_equal(got, want)
