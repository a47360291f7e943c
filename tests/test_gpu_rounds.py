"""Bounded candidate memory and the tail's fallback, on the GPU.

Rounds: with a cap on the candidate pairs one round may hold, a search whose
candidates exceed it runs in rounds over ranges of the smaller id (the
reference's lower-bound rounds, xtneighbor_streaming/src/xtn.cu:385-470) and
still returns exactly the oracle's triples, in order. The input is heavily
duplicated (a thousand copies of a few strings: every copy shares every
deletion variant, so one key group holds a thousand records and the
candidates grow with its square), which also sends groups past the LDS
bucket capacity onto the global run path.

Window overflow: with the partial sort chosen (sort_low_bits 16), the sort of
the candidate tail stops short of the low pair bits and the select ranks each
group of equal high bits in LDS; a group too large for the select's window
raises a flag and the run is redone with the full sort. The duplicated input
above has such groups.

Fallback: a tiny spin limit on the tail's bounded look-back waits (set from
the host through XTN_LOOKBACK_SPIN_LIMIT) makes some wait give up and raise
the error flag; the run is then redone with the host-checked tail, and the
output is unchanged. Every kernel still exits normally.
"""
import os

import numpy as np
import pytest

import oracle
import xtn
from helpers import clustered_seqs, random_seqs

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def triples(p, d):
    return np.concatenate([p.astype(np.int64), d.astype(np.int64)[:, None]], axis=1)


def port(chars, offsets, d, measure):
    tri, st = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)
    return tri.astype(np.int64), st


def duplicated_input(seed, copies, n_random):
    rng = np.random.default_rng(seed)
    base = ["CASSLGQAYEQYF", "CASSPDRGNTEAFF", "CSARDGTGNGYTF"]
    seqs = []
    for s in base:
        seqs += [s] * copies
    seqs += random_seqs(rng, n_random, 8, 20)
    seqs += clustered_seqs(rng, n_random // 2, 10, 18, edits=1)
    order = rng.permutation(len(seqs))
    return xtn.encode([seqs[i] for i in order])


def test_rounds_bounded_candidates_match_oracle():
    chars, offsets = duplicated_input(11, 1000, 20000)
    want, wst = port(chars, offsets, 1, 0)
    assert len(want) > 1_500_000  # the copies pair with each other at distance 0
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        eng.set_tuning(sort_low_bits=16)
        n = eng.nn_resident()
        one = triples(*eng.copy_nn_output(n))
        st1 = eng.stats()
        assert np.array_equal(one, want)
        # the copies' candidates of one row outgrow the select's window of
        # a partly sorted tail: the run was redone fully sorted
        idb = int(len(offsets) - 2).bit_length()
        assert st1["tail_passes"] == (2 * idb + 7) // 8
        cap = max(st1["n_candidates"] // 5, 1 << 16)
        eng.configure(max_pairs_per_round=cap)
        n = eng.nn_resident()
        got = triples(*eng.copy_nn_output(n))
        st = eng.stats()
        assert st["n_rounds"] >= 5
        assert st["n_candidates"] == st1["n_candidates"]
        assert np.array_equal(got, want)
        # the host hand-off over the same rounds
        p, d, _ = eng.nn()
        assert np.array_equal(triples(p, d), want)


def test_rounds_with_forced_chunks():
    chars, offsets = duplicated_input(12, 400, 30000)
    want, _ = port(chars, offsets, 2, 0)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        eng.configure(max_records_per_chunk=400_000, max_pairs_per_round=1 << 20)
        n = eng.nn_resident()
        got = triples(*eng.copy_nn_output(n))
        st = eng.stats()
        assert st["n_chunks"] > 1 and st["n_rounds"] > 1
        assert np.array_equal(got, want)


def test_lookback_limit_falls_back_to_checked_tail(monkeypatch):
    chars, offsets = xtn.synth_nn(31, 300_000, planted=0.25, max_edits=2)
    want, _ = port(chars, offsets, 2, 0)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        n = eng.nn_resident()
        assert np.array_equal(triples(*eng.copy_nn_output(n)), want)
        assert eng.stats()["tail_checked"] == 0
    monkeypatch.setenv("XTN_LOOKBACK_SPIN_LIMIT", "0")
    fell_back = False
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for _ in range(3):
            n = eng.nn_resident()
            assert np.array_equal(triples(*eng.copy_nn_output(n)), want)
            fell_back = fell_back or eng.stats()["tail_checked"] == 1
    assert fell_back
