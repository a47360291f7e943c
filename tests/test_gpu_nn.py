"""GPU parity of nearest-neighbour mode: the HIP path through the C ABI
against the oracle (reference algorithm restated) and the brute-force
definition, bit-exact on (i, j, distance) triples."""
import os
import subprocess

import numpy as np
import pytest

import oracle
import xtn
from conftest import GOLDEN, read_golden
from helpers import AA20, clustered_seqs, random_seqs, triples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    e = xtn.Engine(distance=1)
    yield e
    e.close()


def run(engine, seqs=None, chars=None, offsets=None, d=1, measure=0, **kw):
    engine.configure(distance=d, measure=measure, max_records_per_chunk=kw.get("chunk", 0),
                     output_chunk_pairs=kw.get("out_chunk", 0), bucket_capacity=kw.get("cap", 0))
    engine.set_sequences(seqs, chars, offsets)
    p, dist, ncb = engine.nn()
    return triples(p, dist), ncb


def test_dummy_fixture(engine):
    seqs = [s for s in read_golden("dummy_input.txt").decode().split("\n")[1:] if s]
    for d, measure, fx in [(1, 0, "expected_dummy_d1.txt"), (1, 1, "expected_dummy_d1.txt"),
                           (2, 0, "expected_dummy_d2.txt")]:
        got, _ = run(engine, seqs, d=d, measure=measure)
        text = "".join(f"{a} {b} {c}\n" for a, b, c in got.tolist()).encode()
        assert text == read_golden(fx)


def test_stage4_vector(engine, golden):
    g = golden["stream4"]
    got, _ = run(engine, g["seqs"], d=g["distance"])
    assert got.tolist() == g["expected"]


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("d", [1, 2, 3])
def test_random_vs_brute(engine, seed, d):
    rng = np.random.default_rng(1000 * seed + d)
    alphabet = ["ACDE", AA20, "AC", AA20][seed % 4]
    seqs = clustered_seqs(rng, 600, 1, 32, alphabet, edits=d + 1)
    chars, offsets = xtn.encode(seqs)
    for measure in (0, 1):
        got, _ = run(engine, chars=chars, offsets=offsets, d=d, measure=measure)
        want = oracle.nn_brute(chars, offsets, d, measure)
        assert np.array_equal(got, want), (seed, d, measure)


def test_large_groups_use_flattened_pairs(engine):
    # a tiny alphabet makes long runs of equal keys (> 32 records)
    rng = np.random.default_rng(5)
    seqs = random_seqs(rng, 1500, 2, 6, alphabet="AC")
    chars, offsets = xtn.encode(seqs)
    got, _ = run(engine, chars=chars, offsets=offsets, d=2)
    assert np.array_equal(got, oracle.nn_brute(chars, offsets, 2, 0))


@pytest.mark.parametrize("d,base_len,ins", [(2, 10, 2), (3, 6, 3)])
def test_hash_paths_agree_across_tiles(engine, d, base_len, ins):
    # at d = 2 and 3 the generator keys the variants of a tile of at most 56
    # sequences from per-sequence tables and builds those of a fuller tile
    # slot by slot: here the originals (length 10 at d = 2: ~73 sequences per
    # tile; length 6 at d = 3: ~97) take the second path and their insertion
    # relatives (length 12: ~52 per tile; length 9: ~31) the first, and every
    # pair between the two halves must still share a key
    rng = np.random.default_rng(40 + d)
    base = random_seqs(rng, 3000, base_len, base_len)
    rel = []
    for s in base[:1500]:
        t = list(s)
        for _ in range(ins):
            t.insert(int(rng.integers(0, len(t) + 1)), AA20[int(rng.integers(20))])
        rel.append("".join(t))
    chars, offsets = xtn.encode(base + rel)
    got, _ = run(engine, chars=chars, offsets=offsets, d=d)
    assert np.array_equal(got, oracle.nn_port(chars, offsets, d, 0)[0])
    assert sum(1 for i, j, _ in got.tolist() if i < 3000 <= j) >= 1500


@pytest.mark.parametrize("d", [1, 2])
def test_short_strings_fill_tiles(engine, d):
    # strings of 2-4 residues have 3-11 variants: a 4096-record generator
    # tile then spans more than 448 sequences, the form that reads each
    # record's packed prefix from global memory (mixed here with fuller
    # tiles of longer strings)
    rng = np.random.default_rng(70 + d)
    seqs = random_seqs(rng, 2500, 2, 4) + random_seqs(rng, 1500, 8, 14, alphabet="ACDE")
    chars, offsets = xtn.encode(seqs)
    got, _ = run(engine, chars=chars, offsets=offsets, d=d)
    assert np.array_equal(got, oracle.nn_brute(chars, offsets, d, 0))


@pytest.mark.parametrize("copies", [3, 4, 7])
def test_runs_across_tile_boundaries(engine, copies):
    # every string repeated: thousands of runs of 3..10 records, many of
    # them straddling the 4096-record tiles of the run detector
    base = xtn.decode(*xtn.synth_nn(6, 6000 // copies, planted=0.3))
    seqs = [s for s in base for _ in range(copies)]
    chars, offsets = xtn.encode(seqs)
    for d in (1, 2):
        got, _ = run(engine, chars=chars, offsets=offsets, d=d)
        port, _ = oracle.nn_port(chars, offsets, d, 0)
        assert np.array_equal(got, port), d


@pytest.mark.parametrize("n,cap", [(3000, 1000), (3000, 1430), (30000, 75), (60000, 1900)])
def test_bucket_overflow_fallback(engine, n, cap):
    # buckets above `cap` records take the global sort/run/pair path; with
    # more than 8 bucket bits (30000 and 60000 sequences) a bucket above `cap`
    # also outgrows its region, so the chunk is redone on the exact sort
    # (60000/1900: a few buckets of ~1780 records on average overflow)
    chars, offsets = xtn.synth_nn(8, n, planted=0.3, max_edits=2)
    base, _ = run(engine, chars=chars, offsets=offsets, d=2)
    forced, _ = run(engine, chars=chars, offsets=offsets, d=2, cap=cap)
    assert np.array_equal(base, forced)
    assert np.array_equal(base, oracle.nn_port(chars, offsets, 2, 0)[0])


@pytest.mark.parametrize("one_row_heavy", [False, True])
def test_many_duplicate_candidates(engine, one_row_heavy):
    # more than 2^19 candidates, most of them repeats of one pair: 20
    # sequences 60 times each, shuffled; or 400 copies of one sequence at the
    # front, so that row 0 alone holds ~6000 candidates
    rng = np.random.default_rng(int(one_row_heavy))
    extra = clustered_seqs(rng, 3000, 10, 20, AA20, edits=1)
    if one_row_heavy:
        seqs = ["CASSLGQGAYEQYF"] * 400 + extra
    else:
        seqs = [s for s in random_seqs(rng, 20, 12, 18) for _ in range(60)] + extra
        seqs = [seqs[i] for i in rng.permutation(len(seqs))]
    chars, offsets = xtn.encode(seqs)
    got, _ = run(engine, chars=chars, offsets=offsets, d=1)
    assert engine.stats()["n_candidates"] > (1 << 19)
    assert np.array_equal(got, oracle.nn_brute(chars, offsets, 1, 0))


@pytest.mark.parametrize("chunk", [1 << 16, 70000])
def test_multi_chunk_equals_single(engine, chunk):
    chars, offsets = xtn.synth_nn(3, 20000, planted=0.25, max_edits=2)
    single, _ = run(engine, chars=chars, offsets=offsets, d=2)
    multi, _ = run(engine, chars=chars, offsets=offsets, d=2, chunk=chunk)
    assert engine.stats()["n_chunks"] > 1
    assert np.array_equal(single, multi)
    port, _ = oracle.nn_port(chars, offsets, 2, 0)
    assert np.array_equal(single, port)


def test_output_chunks_are_ascending(engine):
    chars, offsets = xtn.synth_nn(4, 20000, planted=0.3)
    full, ncb1 = run(engine, chars=chars, offsets=offsets, d=1)
    parts, ncb = run(engine, chars=chars, offsets=offsets, d=1, out_chunk=1000)
    assert ncb1 == 1 and ncb == (len(full) + 999) // 1000
    assert np.array_equal(full, parts)


def test_config1_10k(engine):
    """10,000 synthetic CDR3, Levenshtein d=1 (the CPU-runnable config)"""
    chars, offsets = xtn.synth_nn(1, 10000, planted=0.2, max_edits=1)
    got, _ = run(engine, chars=chars, offsets=offsets, d=1)
    port, st = oracle.nn_port(chars, offsets, 1, 0)
    assert np.array_equal(got, port)
    assert np.array_equal(got, oracle.nn_brute(chars, offsets, 1, 0))
    s = engine.stats()
    assert s["n_variants"] == st["n_variants"]
    assert s["n_output"] == len(port)


@pytest.mark.parametrize("measure", [0, 1])
def test_100k_d2_vs_port(engine, measure):
    chars, offsets = xtn.synth_nn(2, 100000, planted=0.25, max_edits=2)
    got, _ = run(engine, chars=chars, offsets=offsets, d=2, measure=measure)
    port, _ = oracle.nn_port(chars, offsets, 2, measure)
    assert np.array_equal(got, port)


@pytest.mark.parametrize("measure", [0, 1])
def test_100k_d3_vs_port(engine, measure):
    # the table form of distance 3 (three deleted slots, G^-2 term) over
    # thousands of generator tiles
    chars, offsets = xtn.synth_nn(5, 100000, planted=0.25, max_edits=3)
    got, _ = run(engine, chars=chars, offsets=offsets, d=3, measure=measure)
    port, _ = oracle.nn_port_mt(chars, offsets, 3, measure, min(16, os.cpu_count() or 1))
    assert np.array_equal(got, port.astype(np.int64))


@pytest.mark.parametrize("measure", [0, 1])
def test_20k_d4_vs_port(engine, measure):
    # distance 4: variants unranked in colex order, keyed by the word hash
    chars, offsets = xtn.synth_nn(6, 20000, planted=0.3, max_edits=4)
    got, _ = run(engine, chars=chars, offsets=offsets, d=4, measure=measure)
    port, _ = oracle.nn_port_mt(chars, offsets, 4, measure, min(16, os.cpu_count() or 1))
    assert np.array_equal(got, port.astype(np.int64))


def test_shuffle_invariance(engine):
    """01_correctness.ipynb cell 18: permute the input, map indexes back"""
    chars, offsets = xtn.synth_nn(9, 50000, planted=0.25, max_edits=2)
    seqs = xtn.decode(chars, offsets)
    base, _ = run(engine, chars=chars, offsets=offsets, d=2)
    perm = np.random.default_rng(0).permutation(len(seqs))
    got, _ = run(engine, [seqs[i] for i in perm], d=2)
    i, j = perm[got[:, 0]], perm[got[:, 1]]
    remapped = np.stack([np.minimum(i, j), np.maximum(i, j), got[:, 2]], axis=1)
    remapped = remapped[np.lexsort((remapped[:, 1], remapped[:, 0]))]
    assert np.array_equal(remapped, base)


def test_edge_cases(engine):
    # empty input, one sequence, duplicates, single residues, long strings
    assert run(engine, [], d=1)[0].shape == (0, 3)
    assert run(engine, ["CASS"], d=1)[0].shape == (0, 3)
    got, _ = run(engine, ["CASS", "CASS", "CASS"], d=1)
    assert got.tolist() == [[0, 1, 0], [0, 2, 0], [1, 2, 0]]
    got, _ = run(engine, ["A", "C", "AC", "D"], d=1)
    assert np.array_equal(got, oracle.nn_brute(*xtn.encode(["A", "C", "AC", "D"]), 1, 0))
    rng = np.random.default_rng(11)
    longs = clustered_seqs(rng, 300, 17, 120, "ACDEF", edits=2)
    chars, offsets = xtn.encode(longs)
    for d in (1, 2, 4):
        got, _ = run(engine, chars=chars, offsets=offsets, d=d)
        assert np.array_equal(got, oracle.nn_brute(chars, offsets, d, 0)), d
    # distances above the register band templates (generic path)
    got, _ = run(engine, chars=chars[:offsets[80]], offsets=offsets[:81], d=10)
    assert np.array_equal(got, oracle.nn_brute(chars[:offsets[80]], offsets[:81], 10, 0))


def test_invalid_residue_rejected(engine):
    with pytest.raises(xtn.XtnError) as e:
        engine.set_sequences(["CASS", "CAXS"])
    assert e.value.status == 1
    # the residue check runs in 64 KiB blocks: a bad byte in the second block,
    # each excluded letter and bytes either side of A..Y
    res = np.frombuffer(b"ACDEFGHIKLMNPQRSTVWY", np.uint8)
    chars = res[np.arange(70_000) % 20].copy()
    offsets = np.array([0, 40_000, 70_000], np.uint32)
    for bad in b"BJOUX@Z[az\x00\xff":
        c = chars.copy()
        c[65_536] = bad
        with pytest.raises(xtn.XtnError) as e:
            engine.set_sequences(chars=c, offsets=offsets)
        assert e.value.status == 1, bad
    engine.set_sequences(chars=chars, offsets=offsets)
    assert engine.nn()[0].shape == (0, 2)


def test_rejected_input_keeps_previous_sequences(engine):
    # a rejected xtn_set_sequences (bad residue past the first MiB, bad
    # offsets) reaches nothing on the device: the sequences set before are
    # still searched, with the same output
    # (the module's engine keeps the configuration of the test before)
    engine.configure(distance=1, measure=0, max_records_per_chunk=0, output_chunk_pairs=0, bucket_capacity=0)
    chars, offsets = xtn.synth_nn(61, 150_000, planted=0.3, max_edits=1)
    engine.set_sequences(chars=chars, offsets=offsets)
    p0, d0, _ = engine.nn()
    assert len(d0) > 0
    other_c, other_o = xtn.synth_nn(62, 120_000, planted=0.3, max_edits=1)
    bad_c = other_c.copy()
    bad_c[(1 << 20) + 5] = ord("X")
    with pytest.raises(xtn.XtnError) as e:
        engine.set_sequences(chars=bad_c, offsets=other_o)
    assert e.value.status == 1
    bad_o = other_o.copy()
    bad_o[1000] = bad_o[1001] + 1
    with pytest.raises(xtn.XtnError) as e:
        engine.set_sequences(chars=other_c, offsets=bad_o)
    assert e.value.status == 3
    p1, d1, _ = engine.nn()
    assert np.array_equal(p1, p0) and np.array_equal(d1, d0)


def test_cli_end_to_end(tmp_path):
    env = dict(os.environ, LC_ALL="C")
    for d, measure, fx in [("1", "leven", "expected_dummy_d1.txt"), ("1", "hamming", "expected_dummy_d1.txt"),
                           ("2", "leven", "expected_dummy_d2.txt")]:
        out = tmp_path / f"o{d}{measure}.txt"
        r = subprocess.run([xtn.CLI_PATH, "-i", os.path.join(GOLDEN, "dummy_input.txt"), "-n", "5",
                            "-d", d, "-m", measure, "-o", str(out)], capture_output=True, env=env, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout == b"" and r.stderr == b""
        assert out.read_bytes() == read_golden(fx)
    # no -o: nothing written, exit 0; -V ends with the total line
    r = subprocess.run([xtn.CLI_PATH, "-i", os.path.join(GOLDEN, "dummy_input.txt"), "-n", "5", "-V"],
                       capture_output=True, env=env, timeout=120)
    assert r.returncode == 0
    assert r.stdout.startswith(b"XTNArgs{\n\tdistance: 1\n")
    assert r.stdout.endswith(b"total output length: 5\n")


def test_cli1_outputs_and_check(tmp_path):
    """the non-streaming xtneighbor command line on the reference's sample
    inputs and a random set (lines longer than 18 residues are cut there):
    the output file against the oracle's brute force over the 18-residue
    forms, and -c's own brute-force check"""
    env = dict(os.environ, LC_ALL="C")
    rng = np.random.default_rng(17)
    rnd = tmp_path / "rnd.txt"
    rows = random_seqs(rng, 3000, 8, 26, "ACDEFGHIK")
    rnd.write_text("".join(s + "\n" for s in rows))
    cases = [(os.path.join(GOLDEN, "sample_input1.txt"), 4), (os.path.join(GOLDEN, "sample_input2.txt"), 3),
             (str(rnd), 3000)]
    for path, n in cases:
        lines = [l for l in open(path).read().split("\n") if l not in ("", " ")]
        cut = [l[:18] for l in lines]
        chars, offsets = xtn.encode(cut)
        for d in (1, 2):
            out = tmp_path / "o.txt"
            r = subprocess.run([xtn.CLI1_PATH, "-p", path, "-n", str(n), "-d", str(d), "-o", str(out), "-c", "-V"],
                               capture_output=True, env=env, timeout=300)
            assert r.returncode == 0, r.stderr
            assert b"output is verrified to be correct\n" in r.stdout
            want = oracle.nn_brute(chars, offsets, d, 0)
            text = "".join(f"{a} {b} {c}\n" for a, b, c in want.tolist())
            assert out.read_text() == text
            assert r.stdout.startswith(b"XTNArgs{\n\tdistance: ")
            assert r.stdout.endswith(f"Success! Number of triplet: {len(want)}\n".encode())
