"""Pins the CPU oracle: reference known-answer vectors, independent brute
force, closed forms and identities. No GPU needed."""
import functools
import math

import numpy as np
import pytest

import oracle
import xtn
from conftest import read_golden
from helpers import AA20, clustered_seqs, random_seqs


def test_char_encode(golden):
    for c, v in golden["char_encode"]["cases"]:
        assert oracle.char_encode(c) == v
    # the 20 residues are exactly the accepted letters
    accepted = [chr(c) for c in range(32, 127) if oracle.char_encode(chr(c)) != -1]
    assert "".join(accepted) == AA20


def test_str_encode_decode(golden):
    g = golden["str_encode"]
    assert oracle.str_encode(g["input"]).tolist() == g["words"]
    assert oracle.str_decode(oracle.str_encode(g["roundtrip"])) == g["roundtrip"]
    assert oracle.len_decode(oracle.str_encode(golden["len_decode"]["input"])) == golden["len_decode"]["len"]
    # truncation at 18 residues and field terminators
    assert oracle.str_decode(oracle.str_encode("ACDEFGHIKLMNPQRSTVWY")) == "ACDEFGHIKLMNPQRSTV"
    for term in "\n,\t":
        assert oracle.str_decode(oracle.str_encode("CAS" + term + "QY")) == "CAS"


def test_remove_char(golden):
    g = golden["remove_char"]
    w = oracle.str_encode(g["input"])
    for pos, exp in enumerate(g["expected"]):
        assert oracle.str_decode(oracle.remove_char(w, pos)) == exp
    w2 = oracle.str_encode(g["input2"])
    for pos, exp in enumerate(g["expected2"]):
        assert oracle.str_decode(oracle.remove_char(w2, pos)) == exp


def test_generate_combination_order(golden):
    for case in golden["generate_combination"]["cases"]:
        assert oracle.expand_keys(case["input"], case["distance"]) == case["expected"]


@pytest.mark.parametrize("d", [1, 2, 3, 5])
def test_variants_closed_form(d):
    rng = np.random.default_rng(d)
    for s in random_seqs(rng, 40, 1, 25):
        L = min(len(s), 18)
        want = sum(math.comb(L, k) for k in range(0, min(d, L) + 1))
        assert oracle.combination_len(L, d) == want
        got = oracle.expand_keys(s, d)
        assert len(got) == want
        # the multiset of variants is every subsequence of the prefix with
        # at most d deletions, one per deletion set
        import itertools
        p = s[:18]
        exp = []
        for k in range(0, min(d, L) + 1):
            for dels in itertools.combinations(range(L), k):
                exp.append("".join(c for i, c in enumerate(p) if i not in dels))
        assert sorted(got) == sorted(exp)


def test_stage1_vector(golden):
    g = golden["stream1"]
    keys, vals = oracle.stage1(g["seqs"], g["distance"])
    assert keys == g["keys"]
    assert vals == g["values"]


def test_stage2_resort_vector(golden):
    g = golden["stream2"]
    keys, vals = oracle.sort_stage(g["keys_in"], g["values_in"])
    assert keys == g["keys"]
    assert vals == g["values"]


def test_stage3_pairs_vector(golden):
    g = golden["stream3"]
    pairs, kk, kv = oracle.gen_pairs(g["keys_in"], g["values_in"], g["lowerbound"])
    assert [list(p) for p in pairs] == g["pairs"]
    assert kk == g["kept_keys"]
    assert kv == g["kept_values"]


def _lev_recursive(a, b):
    @functools.lru_cache(maxsize=None)
    def f(i, j):
        if i == len(a):
            return len(b) - j
        if j == len(b):
            return len(a) - i
        return min(f(i + 1, j) + 1, f(i, j + 1) + 1, f(i + 1, j + 1) + (a[i] != b[j]))
    return f(0, 0)


def test_distance_against_recursion_and_identities():
    rng = np.random.default_rng(7)
    seqs = random_seqs(rng, 60, 0, 9, alphabet="ACD") + random_seqs(rng, 20, 15, 30)
    for a in seqs[:40]:
        for b in seqs[40:]:
            d = oracle.distance(a, b, 0)
            assert d == _lev_recursive(a, b)
            assert d == oracle.distance(b, a, 0)
            assert d >= abs(len(a) - len(b))
            h = oracle.distance(a, b, 1)
            if len(a) != len(b):
                assert h == 77
            else:
                assert h == sum(x != y for x, y in zip(a, b))
                assert d <= h
        assert oracle.distance(a, a, 0) == 0


def test_stage4_vector(golden):
    """test_stream4.cu: all pairs of 7 sequences (two of 22 residues)"""
    g = golden["stream4"]
    chars, offsets = xtn.encode(g["seqs"])
    got = oracle.nn_brute(chars, offsets, g["distance"], 0).tolist()
    assert got == g["expected"]
    got_port = oracle.nn_port(chars, offsets, g["distance"], 0)[0].tolist()
    assert got_port == g["expected"]


def _lines(triples):
    return "".join(f"{a} {b} {c}\n" for a, b, c in triples.tolist()).encode()


@pytest.mark.parametrize("d,measure,fixture", [(1, 0, "expected_dummy_d1.txt"),
                                               (1, 1, "expected_dummy_d1.txt"),
                                               (2, 0, "expected_dummy_d2.txt")])
def test_dummy_fixture(d, measure, fixture):
    seqs = read_golden("dummy_input.txt").decode().split("\n")[1:]
    seqs = [s for s in seqs if s]
    chars, offsets = xtn.encode(seqs)
    assert _lines(oracle.nn_brute(chars, offsets, d, measure)) == read_golden(fixture)
    assert _lines(oracle.nn_port(chars, offsets, d, measure)[0]) == read_golden(fixture)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("d", [1, 2, 3])
def test_port_equals_brute(seed, d):
    rng = np.random.default_rng(100 * seed + d)
    alphabet = "ACDE" if seed % 2 else AA20
    seqs = clustered_seqs(rng, 300, 1, 30, alphabet, edits=d + 1)
    chars, offsets = xtn.encode(seqs)
    for measure in (0, 1):
        brute = oracle.nn_brute(chars, offsets, d, measure)
        port, st = oracle.nn_port(chars, offsets, d, measure)
        assert np.array_equal(port, brute)
        assert st["n_output"] == len(brute)
    rounds3, _ = oracle.nn_port(chars, offsets, d, 0, rounds=3)
    assert np.array_equal(rounds3, oracle.nn_brute(chars, offsets, d, 0))


def test_overlap_init_vector(golden):
    g = golden["overlap_init"]
    chars, offsets = xtn.encode(g["seqs"])
    got = oracle.overlap(chars, offsets, g["frequency"], g["repertoire"], 1, 0, which=1)
    assert got.tolist() == g["expected"]


def test_overlap_stage4_vector(golden):
    g = golden["stream4_overlap"]
    chars, offsets = xtn.encode(g["seqs"])
    got = oracle.overlap(chars, offsets, g["frequency"], g["repertoire"], g["distance"], 0, which=2)
    assert got.tolist() == g["expected"]


def test_overlap_airr_fixture():
    # dummy_input_airr.tsv: cdr3_aa column 6, duplicate_count column 2
    rows = read_golden("dummy_input_airr.tsv").decode().split("\n")[1:]
    rows = [r.split("\t") for r in rows if r]
    seqs = [r[6] for r in rows]
    freq = [int(r[2]) for r in rows]
    rep = xtn.reps_from_sizes([2, 3])
    chars, offsets = xtn.encode(seqs)
    got = oracle.overlap(chars, offsets, freq, rep, 1, 0)
    text = "".join(f"{a} {b} {c}\n" for a, b, c in got.tolist()).encode()
    assert text == read_golden("expected_airr_overlap_d1.txt")


@pytest.mark.parametrize("seed", range(4))
def test_overlap_closed_form(seed):
    """upper triangle of R^T (A + I) R on input without repeats inside a
    repertoire (repeats across repertoires allowed)"""
    rng = np.random.default_rng(seed)
    pool = clustered_seqs(rng, 120, 3, 20, "ACDEFG", edits=1)
    pool = list(dict.fromkeys(pool))
    R = 4
    sizes = [int(rng.integers(5, 30)) for _ in range(R)]
    seqs, rep = [], []
    for r, sz in enumerate(sizes):
        pick = rng.choice(len(pool), size=min(sz, len(pool)), replace=False)
        seqs += [pool[i] for i in pick]
        rep += [r] * len(pick)
    freq = rng.integers(1, 1000, len(seqs)).astype(np.int64)
    chars, offsets = xtn.encode(seqs)
    got = oracle.overlap(chars, offsets, freq, rep, 1, 0)

    uniq = sorted(set(seqs))
    ui = {s: i for i, s in enumerate(uniq)}
    Rm = np.zeros((len(uniq), R), dtype=object)
    for s, r, f in zip(seqs, rep, freq):
        Rm[ui[s], r] += int(f)
    A = np.zeros((len(uniq), len(uniq)), dtype=object)
    for i in range(len(uniq)):
        for j in range(len(uniq)):
            if i != j and _lev_recursive(uniq[i], uniq[j]) <= 1:
                A[i, j] = 1
    M = Rm.T.dot(A + np.identity(len(uniq), dtype=object)).dot(Rm)
    exp = [[a, b, int(M[a, b])] for a in range(R) for b in range(a, R) if M[a, b] != 0]
    assert got.tolist() == exp


def test_neighbors_of_matches_brute():
    rng = np.random.default_rng(42)
    seqs = clustered_seqs(rng, 250, 2, 25, "ACDEF", edits=2)
    chars, offsets = xtn.encode(seqs)
    for d, measure in [(1, 0), (2, 0), (1, 1)]:
        tri = oracle.nn_brute(chars, offsets, d, measure)
        for row in (0, 7, 120, 249):
            want = sorted([(int(b), int(c)) for a, b, c in tri if a == row] +
                          [(int(a), int(c)) for a, b, c in tri if b == row])
            assert oracle.neighbors_of(chars, offsets, row, d, measure).tolist() == [list(x) for x in want]


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_port_mt_equals_port(threads):
    rng = np.random.default_rng(threads)
    seqs = clustered_seqs(rng, 1500, 1, 24, "ACDEFG", edits=2) + ["CASS"] * 7
    chars, offsets = xtn.encode(seqs)
    for d, measure in [(1, 0), (2, 0), (1, 1)]:
        want, wst = oracle.nn_port(chars, offsets, d, measure)
        got, gst = oracle.nn_port_mt(chars, offsets, d, measure, threads)
        assert np.array_equal(got, want)
        assert gst == wst
    empty, _ = oracle.nn_port_mt(*xtn.encode(["CASS"]), 1, 0, threads)
    assert len(empty) == 0


@pytest.mark.parametrize("threads", [2, 8])
def test_overlap_mt_equals_overlap(threads):
    """the OpenMP overlap (the checker of the 64-repertoire workload) against
    the single-thread restatement, with duplicates inside and across
    repertoires and negative frequencies (sums that wrap mod 2^64)"""
    rng = np.random.default_rng(40 + threads)
    seqs = clustered_seqs(rng, 1200, 3, 22, "ACDEFGH", edits=1)
    chars, offsets = xtn.encode(seqs)
    n = len(seqs)
    sizes = [300, 200, 400, 300]
    rep = xtn.reps_from_sizes(sizes)
    freq = rng.integers(-2**31 + 1, 2**31 - 1, n).astype(np.int32)
    freq[freq == 0] = 1
    for d, measure in [(1, 0), (2, 0), (1, 1)]:
        want = oracle.overlap(chars, offsets, freq, rep, d, measure)
        got = oracle.overlap_mt(chars, offsets, freq, rep, d, measure, threads)
        assert np.array_equal(got, want)
