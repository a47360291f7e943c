"""The records' first ordering pass (by their top 8 hash bits) done by the
generator itself, on the GPU: it gives exactly the separate pass's output and
the oracle's, for both measures and several distances, while writing fewer
records (a sequence's repeated variants left out); a sub-range that
overflows (many copies of one string in one generator tile) sends the run back
to the separate pass, which the handle keeps until new sequences; and a bucket
that outgrows its region after the generator's pass is regrouped exactly from
the compacted sub-ranges."""
import os

import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def triples(p, d):
    return np.concatenate([p.astype(np.int64), d.astype(np.int64)[:, None]], axis=1)


def run(eng, **cfg):
    eng.configure(**cfg)
    n = eng.nn_resident()
    return triples(*eng.copy_nn_output(n)), eng.stats()


@pytest.mark.parametrize("d,measure,n", [(1, 0, 45_000), (1, 1, 45_000), (2, 0, 60_000), (2, 1, 60_000),
                                         (3, 0, 40_000)])
def test_generator_pass_equals_separate_pass(d, measure, n):
    chars, offsets = xtn.synth_nn(61 + d, n, planted=0.3, max_edits=d)
    want = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=d) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for sep in (0, 1, 0):
            for _ in range(3):  # eager, captured, replayed
                got, st = run(eng, distance=d, measure=measure, coarse_separate=sep)
                assert np.array_equal(got, want), (sep, st)
                assert st["coarse_in_generator"] == 1 - sep
                # repeated variants of a sequence are left out by the
                # generator's pass (tiles with hash tables: d >= 2)
                if sep or d == 1:
                    assert st["n_records"] == st["n_variants"]
                else:
                    assert 0 < st["n_records"] < st["n_variants"]


def test_overflowing_sub_range_falls_back():
    # 455 copies of an 8-letter string fill the first generator tile (9
    # records each at d=1): each of its 9 variants puts 455 records into one
    # sub-range whose capacity is ~165 against a mean of ~44
    base = xtn.decode(*xtn.synth_nn(64, 45_000, planted=0.3, max_edits=1))
    seqs = ["CDEFGHIK"] * 455 + base
    chars, offsets = xtn.encode(seqs)
    want = oracle.nn_port_mt(chars, offsets, 1, 0, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=1) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for _ in range(3):
            got, st = run(eng, distance=1, measure=0, coarse_separate=0)
            assert np.array_equal(got, want)
            assert st["coarse_in_generator"] == 0
        # new sequences without the copies: the generator's pass again
        chars2, offsets2 = xtn.encode(base)
        want2 = oracle.nn_port_mt(chars2, offsets2, 1, 0, THREADS)[0].astype(np.int64)
        eng.set_sequences(chars=chars2, offsets=offsets2)
        got, st = run(eng, distance=1, measure=0, coarse_separate=0)
        assert np.array_equal(got, want2)
        assert st["coarse_in_generator"] == 1


@pytest.mark.parametrize("cap", [75, 1900])
def test_region_overflow_after_generator_pass(cap):
    # buckets above `cap` records outgrow their regions: the chunk is
    # regrouped on the exact sort from the compacted sub-ranges
    chars, offsets = xtn.synth_nn(65, 60_000, planted=0.3, max_edits=2)
    want = oracle.nn_port_mt(chars, offsets, 2, 0, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        got, st = run(eng, distance=2, measure=0, bucket_capacity=cap, coarse_separate=0)
        assert np.array_equal(got, want)
        assert st["coarse_in_generator"] == 1


@pytest.mark.parametrize("seed,n,d,measure,dup", [(71, 60_000, 1, 0, 0.0), (72, 150_000, 1, 1, 0.2),
                                                  (73, 80_000, 2, 0, 0.3), (74, 50_000, 2, 1, 0.0),
                                                  (75, 25_000, 3, 1, 0.1)])
def test_generator_pass_sweep(seed, n, d, measure, dup):
    # short and long strings (lengths 4..30), planted relatives, and a share
    # of rows repeated a few times each, shuffled: uneven sub-range fills
    chars, offsets = xtn.synth_nn(seed, n, 14.0, 4.0, 4, 30, planted=0.3, max_edits=d)
    seqs = xtn.decode(chars, offsets)
    rng = np.random.default_rng(seed)
    extra = [seqs[i] for i in rng.integers(0, len(seqs), int(dup * len(seqs)))]
    seqs = seqs + extra
    seqs = [seqs[i] for i in rng.permutation(len(seqs))]
    chars, offsets = xtn.encode(seqs)
    want = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=d) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        for _ in range(2):
            got, st = run(eng, distance=d, measure=measure, coarse_separate=0)
            assert np.array_equal(got, want)
            assert st["coarse_in_generator"] == 1


def runs_input(d, measure):
    # strings made of runs (homopolymers, doubled residues) and of
    # alternations (ABAB...), whose repeats the run test does not catch,
    # mixed with ordinary rows and their planted relatives
    rng = np.random.default_rng(80 + d + 10 * measure)
    aa = "ACDEFGHIKLMNPQRSTVWY"
    seqs = []
    for _ in range(6000):
        kind = rng.integers(0, 3)
        n = int(rng.integers(4, 21))
        if kind == 0:
            seqs.append("".join(aa[int(rng.integers(0, 3))] * int(rng.integers(1, 5)) for _ in range(n // 2 + 1))[:n])
        elif kind == 1:
            a, b = aa[int(rng.integers(0, 20))], aa[int(rng.integers(0, 20))]
            seqs.append("".join(a if i % 2 == 0 else b for i in range(n)))
        else:
            seqs.append("".join(c * 2 for c in rng.choice(list(aa), n // 2 + 1))[:n])
    base = xtn.decode(*xtn.synth_nn(90 + d, 20_000, planted=0.3, max_edits=d))
    seqs = seqs + base
    seqs = [seqs[i] for i in rng.permutation(len(seqs))]
    return xtn.encode(seqs)


@pytest.mark.parametrize("d,measure", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_runs_of_equal_residues(d, measure):
    # the generator's pass leaves out most of these rows' deletion sets and
    # the output is the oracle's
    chars, offsets = runs_input(d, measure)
    want = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=d) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        got, st = run(eng, distance=d, measure=measure, coarse_separate=0)
        assert np.array_equal(got, want)
        assert st["coarse_in_generator"] == 1
        assert st["n_records"] < st["n_variants"]


@pytest.mark.parametrize("d,measure", [(2, 0), (2, 1), (3, 0), (3, 1)])
def test_runs_of_equal_residues_separate_pass(d, measure):
    # every deletion set written (165 equal records of one id for one variant
    # here): buckets outgrow the LDS and their long runs go through the
    # flattened pair kernel, whose pairs were once filtered by the wrong range
    chars, offsets = runs_input(d, measure)
    want = oracle.nn_port_mt(chars, offsets, d, measure, THREADS)[0].astype(np.int64)
    with xtn.Engine(distance=d) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        got, st = run(eng, distance=d, measure=measure, coarse_separate=1)
        assert st["coarse_in_generator"] == 0
        assert st["n_records"] == st["n_variants"]
        assert np.array_equal(got, want)


@pytest.fixture(scope="module")
def runs_d3():
    chars, offsets = runs_input(3, 0)
    return chars, offsets, oracle.nn_port_mt(chars, offsets, 3, 0, THREADS)[0].astype(np.int64)


@pytest.mark.parametrize("name,tuning,cfg", [
    ("chunks cut by memory", {}, dict(max_records_per_chunk=-3)),
    ("rounds, separate pass", {}, dict(max_pairs_per_round=8 << 20, coarse_separate=1)),
    ("small buckets, separate pass", dict(bucket_capacity=2048, bucket_target=1500), dict(coarse_separate=1)),
    ("small buckets", dict(bucket_capacity=2048, bucket_target=1500), {}),
])
def test_runs_of_equal_residues_fallback_paths(runs_d3, name, tuning, cfg):
    # the same rows through the paths that take over when buckets, rounds or
    # memory overflow (long runs of equal keys reach the flattened pair kernel)
    chars, offsets, want = runs_d3
    with xtn.Engine(distance=3) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        eng.nn_resident()
        v = eng.stats()["n_variants"]
        eng.set_tuning(**tuning)
        c = dict(cfg)
        if c.get("max_records_per_chunk") == -3:
            c["max_records_per_chunk"] = v // 3
        for _ in range(2):
            got, st = run(eng, distance=3, measure=0, **c)
            assert np.array_equal(got, want), (name, st)
        if "max_records_per_chunk" in c:
            assert st["n_chunks"] > 1
        if "max_pairs_per_round" in c:
            assert st["n_rounds"] > 1
