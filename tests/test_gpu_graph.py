"""Graph capture and replay of a single-chunk run, on the GPU.

A run whose sizes and buffers match the previous one is captured as one HIP
graph and replayed from then on. Every replay must give exactly the oracle's
output, also with other (eager) work and host synchronisations between the
replays, after a configuration change and back, and in the batched form.
"""
import os

import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def triples(p, d):
    return np.concatenate([p.astype(np.int64), d.astype(np.int64)[:, None]], axis=1)


def port(chars, offsets, d, measure=0):
    return oracle.nn_port_mt(chars, offsets, d, measure, THREADS)[0].astype(np.int64)


def test_replays_match_oracle_with_eager_work_between():
    chars, offsets = xtn.synth_nn(41, 200_000, planted=0.25, max_edits=2)
    want2 = port(chars, offsets, 2)
    want1 = port(chars, offsets, 1)
    other_c, other_o = xtn.synth_nn(42, 30_000, planted=0.3, max_edits=1)
    want_other = port(other_c, other_o, 1)
    replays = []
    with xtn.Engine(distance=2) as eng, xtn.Engine(distance=1) as side:
        eng.set_sequences(chars=chars, offsets=offsets)
        side.set_sequences(chars=other_c, offsets=other_o)
        for k in range(8):
            n = eng.nn_resident()
            assert np.array_equal(triples(*eng.copy_nn_output(n)), want2), k
            st = eng.stats()
            replays.append(st["graph_replays"])
            assert st["n_syncs"] >= 1
            # eager work of another handle and a host sync between replays
            m = side.nn_resident()
            assert np.array_equal(triples(*side.copy_nn_output(m)), want_other)
            if k == 4:
                # a different run on the same handle, then back
                eng.configure(distance=1)
                n = eng.nn_resident()
                assert np.array_equal(triples(*eng.copy_nn_output(n)), want1)
                eng.configure(distance=2)
    assert replays[0] == 0  # the first run is eager
    assert sum(replays) >= 5  # captured on a repeat, replayed after
    assert replays[-1] == 1


def test_no_graph_and_batched_replays():
    chars, offsets = xtn.synth_nn(43, 60_000, planted=0.25, max_edits=2)
    n = len(offsets) - 1
    so = np.array([0, 1000, 1000, 15000, 40000, n], dtype=np.uint32)
    with xtn.Engine(distance=2, no_graph=True) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        base = eng.nn_sets(so, batched=True)
        for _ in range(3):
            eng.nn_resident()
            assert eng.stats()["graph_replays"] == 0
    with xtn.Engine(distance=2) as eng:
        eng.set_sequences(chars=chars, offsets=offsets)
        seen = 0
        for _ in range(5):
            got = eng.nn_sets(so, batched=True)
            seen += eng.stats()["graph_replays"]
            assert got.keys() == base.keys()
            for k in base:
                assert np.array_equal(got[k], base[k]), k
        assert seen >= 3
    # the batched sets against the oracle run set by set
    for k in range(len(so) - 1):
        a, b = int(so[k]), int(so[k + 1])
        sub_o = (offsets[a:b + 1] - offsets[a]).astype(np.uint32)
        sub_c = chars[int(offsets[a]):int(offsets[b])]
        want = port(sub_c, sub_o, 2) if b - a > 1 else np.zeros((0, 3), np.int64)
        got = base.get(k, np.zeros((0, 3), np.int64))
        assert np.array_equal(got, want), k
