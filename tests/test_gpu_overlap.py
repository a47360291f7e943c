"""GPU parity of repertoire-overlap mode against the oracle's restatement of
the reference (itself pinned by test_overlap_init.cu, test_stream4_overlap.cu,
the AIRR fixture and the closed form R^T (A + I) R in test_oracle.py)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import xtn
from conftest import GOLDEN, read_golden
from helpers import clustered_seqs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = xtn.Engine(distance=1)
    yield e
    e.close()


def run(engine, seqs=None, chars=None, offsets=None, freq=None, rep=None, d=1, measure=0, chunk=0):
    engine.configure(distance=d, measure=measure, max_records_per_chunk=chunk)
    engine.set_sequences(seqs, chars, offsets)
    out = engine.overlap(freq, rep)
    assert engine.overlap_calls == 1  # exactly one callback (xtn.cu:476-480)
    return out


def test_bad_repertoire_indexes(engine):
    # checked on the device: the first row with a negative repertoire index,
    # and more repertoires than the dense R x R matrix holds
    seqs = ["CASSLGF", "CASSLGW", "CASSPGF", "CATSLGF"]
    engine.configure(distance=1, measure=0, max_records_per_chunk=0)
    engine.set_sequences(seqs)
    with pytest.raises(xtn.XtnError) as e:
        engine.overlap([1, 2, 3, 4], [0, 1, -2, -1])
    assert e.value.status == 3 and "row 2" in str(e.value)
    with pytest.raises(xtn.XtnError) as e:
        engine.overlap([1, 1, 1, 1], [0, 46340, 1, 2])
    assert e.value.status == 6 and "46341 repertoires" in str(e.value)
    out = engine.overlap([1, 1, 1, 1], [0, 1999, 1, 2])  # the handle still works
    want = oracle.overlap(*xtn.encode(seqs), [1, 1, 1, 1], [0, 1999, 1, 2], 1, 0)
    assert out.tolist() == want.tolist()


def test_reference_vectors(engine, golden):
    for key in ("overlap_init", "stream4_overlap"):
        g = golden[key]
        chars, offsets = xtn.encode(g["seqs"])
        got = run(engine, chars=chars, offsets=offsets, freq=g["frequency"], rep=g["repertoire"], d=1)
        want = oracle.overlap(chars, offsets, g["frequency"], g["repertoire"], 1, 0)
        assert got.tolist() == want.tolist(), key


def test_airr_fixture(engine):
    rows = [r.split("\t") for r in read_golden("dummy_input_airr.tsv").decode().split("\n")[1:] if r]
    seqs, freq = [r[6] for r in rows], [int(r[2]) for r in rows]
    got = run(engine, seqs, freq=freq, rep=xtn.reps_from_sizes([2, 3]), d=1)
    text = "".join(f"{a} {b} {c}\n" for a, b, c in got.tolist()).encode()
    assert text == read_golden("expected_airr_overlap_d1.txt")


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("d,measure", [(1, 0), (2, 0), (1, 1)])
def test_random_repertoires(engine, seed, d, measure):
    """includes strings repeated inside one repertoire (the reference's
    exact-match term counts such a pair once, kernel.cu:576-583)"""
    rng = np.random.default_rng(seed * 7 + d)
    pool = clustered_seqs(rng, 400, 3, 24, "ACDEFGHIK", edits=2)
    R = [3, 9, 70][seed % 3]  # 70 > 64 exercises the global-memory matrix
    seqs, rep = [], []
    for r in range(R):
        k = int(rng.integers(1, 60))
        seqs += [pool[i] for i in rng.integers(0, len(pool), k)]
        rep += [r] * k
    freq = rng.integers(-5, 10000, len(seqs)).astype(np.int32)
    freq[freq == 0] = 1
    chars, offsets = xtn.encode(seqs)
    got = run(engine, chars=chars, offsets=offsets, freq=freq, rep=rep, d=d, measure=measure)
    want = oracle.overlap(chars, offsets, freq, rep, d, measure)
    assert np.array_equal(got, want)


def test_overlap_many_rows_hash_collisions(engine):
    """~300k unique rows: 32-bit hash runs holding different strings occur,
    exercising the exact string grouping"""
    chars, offsets, freq, sizes = xtn.synth_overlap(5, 8, 20000, 60000, shared=0.1, zipf_max=1000)
    rep = xtn.reps_from_sizes(sizes)
    got = run(engine, chars=chars, offsets=offsets, freq=freq, rep=rep, d=1)
    want = oracle.overlap(chars, offsets, freq, rep, 1, 0)
    assert np.array_equal(got, want)
    # the multi-chunk NN path inside overlap mode
    got2 = run(engine, chars=chars, offsets=offsets, freq=freq, rep=rep, d=1, chunk=1 << 20)
    assert engine.stats()["n_chunks"] > 1
    assert np.array_equal(got2, want)


def test_overlap_edge_cases(engine):
    assert run(engine, [], freq=[], rep=[]).shape == (0, 3)
    got = run(engine, ["CASS"], freq=[7], rep=[0])
    assert got.tolist() == [[0, 0, 49]]
    # a frequency product that wraps modulo 2^64 (size_t in the reference)
    got = run(engine, ["CASS", "CASS"], freq=[-3, 5], rep=[0, 1])
    want = oracle.overlap(*xtn.encode(["CASS", "CASS"]), [-3, 5], [0, 1], 1, 0)
    assert np.array_equal(got, want)
    assert got[1, 2] == np.uint64(2 ** 64 - 15)


def test_cli_overlap_airr(tmp_path):
    env = dict(os.environ, LC_ALL="C")
    out = tmp_path / "ov.txt"
    r = subprocess.run([xtn.CLI_PATH, "-a", "-i", os.path.join(GOLDEN, "dummy_input_airr.tsv"), "-n", "5",
                        "-I", os.path.join(GOLDEN, "dummy_info.csv"), "-N", "2", "-o", str(out)],
                       capture_output=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out.read_bytes() == read_golden("expected_airr_overlap_d1.txt")
    # the same data as a two-column CSV
    rows = [r.split("\t") for r in read_golden("dummy_input_airr.tsv").decode().split("\n")[1:] if r]
    csv = tmp_path / "seqs.csv"
    csv.write_text("cdr3,count\n" + "".join(f"{r[6]},{r[2]}\n" for r in rows))
    out2 = tmp_path / "ov2.txt"
    r = subprocess.run([xtn.CLI_PATH, "-i", str(csv), "-n", "5", "-I", os.path.join(GOLDEN, "dummy_info.csv"),
                        "-N", "2", "-o", str(out2)], capture_output=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    assert out2.read_bytes() == read_golden("expected_airr_overlap_d1.txt")
