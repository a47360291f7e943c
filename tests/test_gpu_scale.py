"""Exact parity at the benchmark sizes: the whole output of the HIP path is
compared array for array with the oracle's OpenMP restatement of the
reference algorithm on the same seeded inputs.

  1M CDR3, Levenshtein d=2 (one chunk, and the key space forced into chunks)
  5M CDR3 (lengths 6..30), d=1, Levenshtein and Hamming, forced chunks
  64 repertoires of 10k..500k rows, Zipf(2) frequencies, Levenshtein d=1
  a 10M-row served index, 8 query sets of 1k..100k rows, Levenshtein d=1

Size-independent properties (strict ordering, the closed-form variant count,
the relation between the two measures at d=1) are checked beside equality.
"""
import math
import os

import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def check_triples(tri, n, d):
    assert np.all(tri[:, 0] < tri[:, 1])
    assert np.all((tri[:, 0] >= 0) & (tri[:, 1] < n))
    assert np.all((tri[:, 2] >= 0) & (tri[:, 2] <= d))
    key = tri[:, 0] * n + tri[:, 1]
    assert np.all(np.diff(key) > 0)  # strictly ascending (i, j), no duplicates


def variant_total(offsets, d):
    lens = np.minimum(np.diff(offsets.astype(np.int64)), 18)
    per = {L: sum(math.comb(L, k) for k in range(min(d, L) + 1)) for L in range(19)}
    return int(sum(per[int(L)] for L in lens))


def run(eng, d, measure, chunk=0):
    eng.configure(distance=d, measure=measure, max_records_per_chunk=chunk)
    n = eng.nn_resident()
    p, dist = eng.copy_nn_output(n)
    return np.concatenate([p.astype(np.int64), dist.astype(np.int64)[:, None]], axis=1)


def port(chars, offsets, d, measure):
    tri, st = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)
    return tri.astype(np.int64), st


def test_config2_1M_d2():
    """1,000,000 synthetic CDR3, 25% planted 1-2 edit relatives, Levenshtein d=2"""
    chars, offsets = xtn.synth_nn(2, 1_000_000, planted=0.25, max_edits=2)
    n = len(offsets) - 1
    want, wst = port(chars, offsets, 2, 0)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        tri = run(eng, 2, 0)
        st = eng.stats()
        assert st["n_variants"] == variant_total(offsets, 2) == wst["n_variants"]
        check_triples(tri, n, 2)
        assert np.array_equal(tri, want)
        forced = run(eng, 2, 0, chunk=st["n_variants"] // 5)
        assert eng.stats()["n_chunks"] >= 8
        assert np.array_equal(forced, want)


def test_config3_5M_d1_both_measures():
    """5,000,000 CDR3 (lengths ~N(15,2.5) in [6,30]), d=1, Hamming and
    Levenshtein, the key space forced into several chunks"""
    chars, offsets = xtn.synth_nn(3, 5_000_000, mu=15.0, sigma=2.5, lmin=6, lmax=30,
                                  planted=0.25, max_edits=1)
    n = len(offsets) - 1
    lens = np.diff(offsets.astype(np.int64))
    want_lev, _ = port(chars, offsets, 1, 0)
    want_ham, _ = port(chars, offsets, 1, 1)
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        lev = run(eng, 1, 0)
        check_triples(lev, n, 1)
        assert np.array_equal(lev, want_lev)
        lev_chunked = run(eng, 1, 0, chunk=20_000_000)
        assert eng.stats()["n_chunks"] > 1
        assert np.array_equal(lev_chunked, want_lev)
        ham = run(eng, 1, 1, chunk=20_000_000)
        assert eng.stats()["n_chunks"] > 1
        assert np.array_equal(ham, want_ham)
        ham_one = run(eng, 1, 1)
        assert np.array_equal(ham_one, want_ham)
    # at d = 1 two strings of equal length are within Levenshtein 1 exactly
    # when they differ by one substitution, i.e. Hamming 1
    same_len = want_lev[lens[want_lev[:, 0]] == lens[want_lev[:, 1]]]
    assert np.array_equal(want_ham, same_len)


def test_config4_64_repertoire_overlap():
    """64 repertoires of log-uniform size in [10k, 500k], Zipf(2) frequencies
    capped at 1e5, 10% shared or 1-edit-shared sequences, Levenshtein d=1:
    the whole R x R matrix against the oracle"""
    chars, offsets, freq, sizes = xtn.synth_overlap(4, 64, 10_000, 500_000, shared=0.1)
    rep = xtn.reps_from_sizes(sizes)
    assert len(sizes) == 64 and len(freq) > 1_000_000
    want = oracle.overlap_mt(chars, offsets, freq, rep, 1, 0, THREADS)
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        got = eng.overlap(freq, rep)
        assert eng.overlap_calls == 1
    assert np.array_equal(got, want)
    assert np.all(got[:, 0] <= got[:, 1])


def test_config5_10M_served_index():
    """a 10,000,000-row resident index (Levenshtein d=1) and 8 query sets of
    log-uniform size in [1k, 100k] with 30% planted relatives of index rows:
    every query's (query row, index row, distance) triples, alone and with all
    eight in flight, against the cross pairs of one oracle NN search over the
    index and all the queries"""
    ichars, ioffs = xtn.synth_nn(5, 10_000_000, 15.0, 2.0, 8, 22, 0.25, 1)
    ni = len(ioffs) - 1
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(100000), 8)).astype(int)
    queries = [xtn.synth_queries(5000 + k, ichars, ioffs, int(sz), planted=0.3)
               for k, sz in enumerate(sizes)]
    # oracle: one NN over index + every query; query k's answer is the pairs
    # with one row in the index and the other in query k
    parts_c = [ichars] + [q[0] for q in queries]
    parts_o = [ioffs]
    base_row = [ni]
    base_char = int(ioffs[-1])
    for qc, qo in queries:
        parts_o.append(qo[1:].astype(np.int64) + base_char)
        base_char += int(qo[-1])
        base_row.append(base_row[-1] + len(qo) - 1)
    all_c = np.concatenate(parts_c)
    all_o = np.concatenate([p.astype(np.int64) for p in parts_o]).astype(np.uint32)
    tri, _ = port(all_c, all_o, 1, 0)
    tri = tri[(tri[:, 0] < ni) & (tri[:, 1] >= ni)]
    want = []
    for k in range(len(queries)):
        lo, hi = base_row[k], base_row[k + 1]
        sel = tri[(tri[:, 1] >= lo) & (tri[:, 1] < hi)]
        w = np.stack([sel[:, 1] - lo, sel[:, 0], sel[:, 2]], axis=1)
        want.append(w[np.lexsort((w[:, 1], w[:, 0]))])
    assert sum(len(w) for w in want) > 0

    def triples(p, d):
        return np.concatenate([p.astype(np.int64), d.astype(np.int64)[:, None]], axis=1)

    with xtn.Server(ichars, ioffs, distance=1, slots=4, max_query_rows=1 << 17) as srv:
        for k, (qc, qo) in enumerate(queries):
            p, d, _ = srv.wait(srv.submit(qc, qo))
            assert np.array_equal(triples(p, d), want[k]), k
        tickets = [srv.submit(qc, qo) for qc, qo in queries]
        for k, t in enumerate(tickets):
            p, d, _ = srv.wait(t)
            assert np.array_equal(triples(p, d), want[k]), k
