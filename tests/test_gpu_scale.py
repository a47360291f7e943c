"""Parity at the benchmark sizes through size-independent properties:
sortedness and validity of every triple, oracle distances on a sample of
output pairs, completeness of sampled rows against a one-vs-all brute force
over the whole input, chunking invariance, exact relations between the
measures, and the closed-form variant count."""
import math

import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu


def check_triples(tri, n, d):
    assert np.all(tri[:, 0] < tri[:, 1])
    assert np.all((tri[:, 0] >= 0) & (tri[:, 1] < n))
    assert np.all((tri[:, 2] >= 0) & (tri[:, 2] <= d))
    key = tri[:, 0] * n + tri[:, 1]
    assert np.all(np.diff(key) > 0)  # strictly ascending (i, j), no duplicates


def sample_distances(tri, chars, offsets, measure, k, seed):
    seqs_at = lambda r: chars[offsets[r]:offsets[r + 1]].tobytes().decode()
    rng = np.random.default_rng(seed)
    for q in rng.choice(len(tri), size=min(k, len(tri)), replace=False):
        i, j, dd = (int(x) for x in tri[q])
        assert oracle.distance(seqs_at(i), seqs_at(j), measure) == dd


def sample_completeness(tri, chars, offsets, d, measure, rows):
    for r in rows:
        mine = tri[(tri[:, 0] == r) | (tri[:, 1] == r)]
        other = np.where(mine[:, 0] == r, mine[:, 1], mine[:, 0])
        got = sorted(zip(other.tolist(), mine[:, 2].tolist()))
        want = [tuple(x) for x in oracle.neighbors_of(chars, offsets, int(r), d, measure).tolist()]
        assert got == want, r


def variant_total(offsets, d):
    lens = np.minimum(np.diff(offsets.astype(np.int64)), 18)
    per = {L: sum(math.comb(L, k) for k in range(min(d, L) + 1)) for L in range(19)}
    return int(sum(per[int(L)] for L in lens))


def run(eng, d, measure, chunk=0):
    eng.configure(distance=d, measure=measure, max_records_per_chunk=chunk)
    n = eng.nn_resident()
    p, dist = eng.copy_nn_output(n)
    return np.concatenate([p.astype(np.int64), dist.astype(np.int64)[:, None]], axis=1)


def test_config2_1M_d2():
    """1,000,000 synthetic CDR3, 25% planted 1-2 edit relatives, Levenshtein d=2"""
    chars, offsets = xtn.synth_nn(2, 1_000_000, planted=0.25, max_edits=2)
    n = len(offsets) - 1
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        tri = run(eng, 2, 0)
        st = eng.stats()
        assert st["n_variants"] == variant_total(offsets, 2)
        check_triples(tri, n, 2)
        sample_distances(tri, chars, offsets, 0, 3000, 1)
        rng = np.random.default_rng(7)
        rows = list(rng.choice(n, 6, replace=False)) + list(tri[rng.choice(len(tri), 6), 0])
        sample_completeness(tri, chars, offsets, 2, 0, rows)
        forced = run(eng, 2, 0, chunk=st["n_variants"] // 5)
        assert eng.stats()["n_chunks"] >= 8
        assert np.array_equal(tri, forced)


def test_config3_5M_d1_both_measures():
    """5,000,000 CDR3 (lengths ~N(15,2.5) in [6,30]), d=1, Hamming and
    Levenshtein, the key space forced into several chunks"""
    chars, offsets = xtn.synth_nn(3, 5_000_000, mu=15.0, sigma=2.5, lmin=6, lmax=30,
                                  planted=0.25, max_edits=1)
    n = len(offsets) - 1
    lens = np.diff(offsets.astype(np.int64))
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        lev = run(eng, 1, 0)
        check_triples(lev, n, 1)
        lev_chunked = run(eng, 1, 0, chunk=20_000_000)
        assert eng.stats()["n_chunks"] > 1
        assert np.array_equal(lev, lev_chunked)
        ham = run(eng, 1, 1, chunk=20_000_000)
        check_triples(ham, n, 1)
        # at d = 1 two strings of equal length are within Levenshtein 1 exactly
        # when they differ by one substitution, i.e. Hamming 1
        same_len = lev[lens[lev[:, 0]] == lens[lev[:, 1]]]
        assert np.array_equal(ham, same_len)
        sample_distances(lev, chars, offsets, 0, 2000, 2)
        rng = np.random.default_rng(9)
        rows = list(rng.choice(n, 3, replace=False)) + list(lev[rng.choice(len(lev), 3), 0])
        sample_completeness(lev, chars, offsets, 1, 0, rows)
        sample_completeness(ham, chars, offsets, 1, 1, rows[:3])
