"""GPU parity of the batched form: many query sets of uneven size in one
pipeline pass equal one pass per set and the per-set brute force; no pair
ever crosses two sets, even when the sets hold identical strings."""
import numpy as np
import pytest

import oracle
import xtn
from helpers import clustered_seqs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = xtn.Engine(distance=1)
    yield e
    e.close()


def make_sets(rng, sizes, alphabet="ACDEFGHIK", shared_pool=None):
    seqs, offs = [], [0]
    for sz in sizes:
        if shared_pool is not None:
            part = [shared_pool[i] for i in rng.integers(0, len(shared_pool), sz)]
        else:
            part = clustered_seqs(rng, sz, 3, 24, alphabet, edits=2) if sz else []
        seqs += part[:sz]
        offs.append(len(seqs))
    return seqs, np.asarray(offs, dtype=np.uint32)


def per_set_brute(chars, offsets, set_offsets, d, measure):
    out = {}
    for k in range(len(set_offsets) - 1):
        a, b = int(set_offsets[k]), int(set_offsets[k + 1])
        sub_off = (offsets[a:b + 1] - offsets[a]).astype(np.uint32)
        sub_chars = chars[offsets[a]:offsets[b]]
        tri = oracle.nn_brute(sub_chars, sub_off, d, measure).astype(np.int64)
        if len(tri):
            out[k] = tri
    return out


def same(a, b):
    assert sorted(a) == sorted(b)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("d,measure", [(1, 0), (2, 0), (1, 1)])
def test_batched_equals_plain_and_brute(engine, seed, d, measure):
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(0, 300, 25)] + [0, 1, 2]
    seqs, so = make_sets(rng, sizes)
    chars, offsets = xtn.encode(seqs)
    engine.configure(distance=d, measure=measure)
    engine.set_sequences(chars=chars, offsets=offsets)
    batched = engine.nn_sets(so, batched=True)
    plain = engine.nn_sets(so, batched=False)
    same(batched, plain)
    same(batched, per_set_brute(chars, offsets, so, d, measure))


def test_identical_strings_in_different_sets(engine):
    rng = np.random.default_rng(3)
    pool = clustered_seqs(rng, 60, 4, 16, "ACDEF", edits=1)
    seqs, so = make_sets(rng, [50, 50, 50, 50], shared_pool=pool)
    chars, offsets = xtn.encode(seqs)
    engine.configure(distance=1, measure=0)
    engine.set_sequences(chars=chars, offsets=offsets)
    got = engine.nn_sets(so, batched=True)
    same(got, per_set_brute(chars, offsets, so, 1, 0))
    # the whole input as one search does find the cross-set pairs
    p, dist, _ = engine.nn()
    assert len(p) > sum(len(v) for v in got.values())


def test_sets_over_a_row_subrange(engine):
    chars, offsets = xtn.synth_nn(12, 5000, planted=0.3)
    so = np.array([1000, 1700, 1700, 3100, 4999], dtype=np.uint32)
    engine.configure(distance=1, measure=0)
    engine.set_sequences(chars=chars, offsets=offsets)
    got = engine.nn_sets(so, batched=True)
    same(got, engine.nn_sets(so, batched=False))
    same(got, per_set_brute(chars, offsets, so, 1, 0))


def _device_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_close_releases_device_memory():
    # every workspace of a handle (sort, bucket, set and output buffers) is
    # freed by xtn_close: repeated open / run / close cycles leave the free
    # device memory where it was
    rng = np.random.default_rng(11)
    seqs, offs = make_sets(rng, [int(x) for x in rng.integers(50, 400, 40)])
    chars, offsets = xtn.encode(seqs)

    def cycle():
        with xtn.Engine(distance=1) as e:
            e.set_sequences(chars=chars, offsets=offsets)
            e.nn()
            e.nn_sets(offs, batched=True)

    cycle()
    free0 = _device_free_bytes()
    for _ in range(3):
        cycle()
    assert free0 - _device_free_bytes() < (16 << 20)
