"""The autotuner and the engine parameters it chooses, on the GPU: every
bucket capacity (2048, 3072, 4096 records in LDS), bucket target, sort tile
width, select tile and number of low pair bits left to the select's in-group
ranking gives exactly the oracle's output; xtn_autotune keeps a setting it
has checked against the defaults and reads it back from its cache."""
import json
import os

import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def triples(p, d):
    return np.concatenate([p.astype(np.int64), d.astype(np.int64)[:, None]], axis=1)


@pytest.mark.parametrize("cap,target,nt,tile,sel,low", [(2048, 1500, 256, 4096, 256, 16),
                                                        (4096, 3400, 512, 8192, 512, 24),
                                                        (4096, 1100, 1024, 8192, 256, 0),
                                                        (3072, 2600, 512, 4096, 512, 8),
                                                        (3072, 2200, 512, 4096, 256, 24)])
def test_settings_match_oracle(cap, target, nt, tile, sel, low):
    chars, offsets = xtn.synth_nn(51, 150_000, planted=0.25, max_edits=2)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    idb = int(len(offsets) - 2).bit_length()
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        eng.set_tuning(bucket_capacity=cap, bucket_target=target, sort_threads=nt, sort_tile=tile,
                       select_threads=sel, sort_low_bits=low)
        for _ in range(3):  # eager, captured, replayed
            n = eng.nn_resident()
            assert np.array_equal(triples(*eng.copy_nn_output(n)), want)
            assert eng.stats()["tail_passes"] == max(1, (2 * idb - low + 7) // 8)
        assert eng.tuning()["bucket_capacity"] == cap
        assert eng.tuning()["sort_low_bits"] == low


def test_tuned_separate_pass_matches_oracle():
    # the tuned switch sends the records' first ordering pass out of the
    # generator as the config's does, on one chunk and on chunks cut by
    # bucket size, and back in when it is cleared
    chars, offsets = xtn.synth_nn(53, 300_000, planted=0.25, max_edits=2)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for cap, target, sep in ((3072, 2200, 1), (2048, 512, 1), (2048, 512, 0), (3072, 2200, 0)):
            eng.set_tuning(bucket_capacity=cap, bucket_target=target, coarse_separate=sep)
            for _ in range(2):
                n = eng.nn_resident()
                assert np.array_equal(triples(*eng.copy_nn_output(n)), want), (target, sep)
                assert eng.stats()["coarse_in_generator"] == 1 - sep
            assert eng.tuning()["coarse_separate"] == sep


def test_invalid_settings_rejected():
    with xtn.Engine(distance=1) as eng:
        for kw in (dict(bucket_capacity=1000), dict(sort_threads=128), dict(bucket_capacity=2048, bucket_target=2000),
                   dict(bucket_capacity=4096, bucket_target=500), dict(sort_threads=256, sort_tile=8192),
                   dict(sort_threads=1024, sort_tile=4096), dict(select_threads=1024), dict(sort_low_bits=12),
                   dict(coarse_separate=2)):
            with pytest.raises(xtn.XtnError) as e:
                eng.set_tuning(**kw)
            assert e.value.status == 3


def test_autotune_and_cache(tmp_path):
    chars, offsets = xtn.synth_nn(52, 120_000, planted=0.25, max_edits=2)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    cache = str(tmp_path / "tuning.json")
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        t = eng.autotune(cache, runs=2)
        assert t["from_cache"] == 0 and t["candidates"] >= 16
        assert t["ms"] <= t["ms_default"]
        n = eng.nn_resident()
        assert np.array_equal(triples(*eng.copy_nn_output(n)), want)
    with open(cache) as f:
        doc = json.load(f)
    assert len(doc["entries"]) == 1 and doc["entries"][0]["bucket_capacity"] == t["bucket_capacity"]
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        t2 = eng.autotune(cache, runs=2)
        assert t2["from_cache"] == 1
        keys = ("bucket_capacity", "bucket_target", "sort_threads", "sort_tile", "select_threads", "sort_low_bits",
                "coarse_separate")
        assert [t2[k] for k in keys] == [t[k] for k in keys]
        n = eng.nn_resident()
        assert np.array_equal(triples(*eng.copy_nn_output(n)), want)
        # another workload gets an entry of its own
        eng.configure(distance=1)
        eng.autotune(cache, runs=1)
    with open(cache) as f:
        assert len(json.load(f)["entries"]) == 2


@pytest.mark.parametrize("low", [16, 24])
def test_partial_sort_in_batched_sets_and_rounds(low):
    # the grouped select over query sets (cross-set candidates dropped) and
    # over rounds of the smaller id, against the fully sorted tail
    chars, offsets = xtn.synth_nn(53, 80_000, planted=0.3, max_edits=2)
    n = len(offsets) - 1
    so = np.array([0, 5000, 5000, 30000, 61000, n], dtype=np.uint32)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        base = eng.nn_sets(so, batched=True)
        eng.set_tuning(sort_low_bits=low)
        for _ in range(3):  # eager, captured, replayed
            got = eng.nn_sets(so, batched=True)
            assert got.keys() == base.keys()
            for k in base:
                assert np.array_equal(got[k], base[k]), k
        n_cand = None
        for cap in (0, 1 << 15):
            eng.configure(max_pairs_per_round=cap)
            m = eng.nn_resident()
            assert np.array_equal(triples(*eng.copy_nn_output(m)), want)
            st = eng.stats()
            n_cand = st["n_candidates"] if n_cand is None else n_cand
            assert st["n_candidates"] == n_cand
        assert st["n_rounds"] > 1


def test_chunks_cut_by_bucket_target():
    # 37M records over a 512-record bucket target exceed the 65536 buckets a
    # chunk may hold, so the run splits into chunks by bucket size: the
    # records are generated once and each chunk reads its part; a chunk size
    # set by the caller generates every part on its own. All match the oracle,
    # over one round and over several.
    chars, offsets = xtn.synth_nn(54, 300_000, planted=0.25, max_edits=2)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        eng.set_tuning(bucket_capacity=2048, bucket_target=512)
        # generated once through the coarse pass (one round and several), once
        # with the separate pass, and per part
        for chunk, pairs, sep in ((0, 0, 0), (0, 1 << 16, 0), (0, 0, 1), (12_000_000, 0, 0)):
            eng.configure(max_records_per_chunk=chunk, max_pairs_per_round=pairs, coarse_separate=sep)
            n = eng.nn_resident()
            assert np.array_equal(triples(*eng.copy_nn_output(n)), want), (chunk, pairs, sep)
            st = eng.stats()
            assert st["n_chunks"] > 1
            assert (st["n_rounds"] > 1) == (pairs > 0)
            assert st["coarse_in_generator"] == (1 if chunk == 0 and sep == 0 else 0)
