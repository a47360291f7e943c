"""GPU parity of the served form: query sets against a resident index give
exactly the cross pairs (query row, index row) of a nearest-neighbour search
over index + query (the completeness lemma makes the index lookup exact),
and every result is the same alone and inside a concurrent mix."""
import numpy as np
import pytest

import oracle
import xtn

pytestmark = pytest.mark.gpu


def cross_pairs(ichars, ioffs, qchars, qoffs, d, measure):
    """oracle: NN over index + query, pairs with one row on each side"""
    ni = len(ioffs) - 1
    chars = np.concatenate([ichars, qchars])
    offs = np.concatenate([ioffs, qoffs[1:] + ioffs[-1]]).astype(np.uint32)
    tri, _ = oracle.nn_port(chars, offs, d, measure)
    cross = tri[(tri[:, 0] < ni) & (tri[:, 1] >= ni)]
    out = np.stack([cross[:, 1] - ni, cross[:, 0], cross[:, 2]], axis=1).astype(np.int64)
    return out[np.lexsort((out[:, 1], out[:, 0]))]


def as_triples(p, dist):
    return np.concatenate([p.astype(np.int64), dist.astype(np.int64)[:, None]], axis=1)


@pytest.fixture(scope="module")
def index():
    return xtn.synth_nn(21, 20000, planted=0.25, max_edits=1)


@pytest.mark.parametrize("d,measure", [(1, 0), (2, 0), (1, 1)])
def test_queries_match_oracle(index, d, measure):
    ichars, ioffs = index
    with xtn.Server(ichars, ioffs, distance=d, measure=measure, slots=3) as srv:
        for seed, n in [(1, 500), (2, 3000), (3, 1)]:
            qchars, qoffs = xtn.synth_queries(seed, ichars, ioffs, n, planted=0.3)
            p, dist, lat = srv.wait(srv.submit(qchars, qoffs))
            assert lat > 0
            assert np.array_equal(as_triples(p, dist), cross_pairs(ichars, ioffs, qchars, qoffs, d, measure))


def test_same_alone_and_in_mix(index):
    ichars, ioffs = index
    queries = [xtn.synth_queries(100 + k, ichars, ioffs, int(n), planted=0.3)
               for k, n in enumerate([2000, 150, 5000, 800, 2500, 40, 4000, 1200])]
    with xtn.Server(ichars, ioffs, distance=1, slots=4) as srv:
        alone = []
        for qc, qo in queries:
            p, dist, _ = srv.wait(srv.submit(qc, qo))
            alone.append(as_triples(p, dist))
        # all in flight together, submitted in a shuffled order
        order = np.random.default_rng(0).permutation(len(queries))
        tickets = {int(k): srv.submit(*queries[k]) for k in order}
        for k, t in tickets.items():
            p, dist, _ = srv.wait(t)
            assert np.array_equal(as_triples(p, dist), alone[k]), k
    for k, (qc, qo) in enumerate(queries):
        assert np.array_equal(alone[k], cross_pairs(ichars, ioffs, qc, qo, 1, 0)), k


def test_tight_budget_still_serves_everything(index):
    ichars, ioffs = index
    qs = [xtn.synth_queries(200 + k, ichars, ioffs, 1000, planted=0.3) for k in range(6)]
    with xtn.Server(ichars, ioffs, distance=1, slots=4, hbm_budget_bytes=1) as srv:
        tickets = [srv.submit(qc, qo) for qc, qo in qs]
        for (qc, qo), t in zip(qs, tickets):
            p, dist, _ = srv.wait(t)
            assert np.array_equal(as_triples(p, dist), cross_pairs(ichars, ioffs, qc, qo, 1, 0))


def test_empty_and_invalid_queries(index):
    ichars, ioffs = index
    with xtn.Server(ichars, ioffs, distance=1, slots=2) as srv:
        p, dist, _ = srv.wait(srv.submit(np.zeros(0, np.uint8), np.zeros(1, np.uint32)))
        assert p.shape == (0, 2)
        bad_c, bad_o = xtn.encode(["CASSX"])
        with pytest.raises(xtn.XtnError) as e:
            srv.submit(bad_c, bad_o)
        assert e.value.status == 1


def test_ticket_collected_once_and_null_chars(index):
    """a ticket is collected by one wait: a second wait on it, a wait on a
    ticket never issued and a second concurrent waiter all get XTN_ERR_ARG
    instead of blocking; NULL residues with a non-empty query are rejected"""
    import ctypes
    import threading

    ichars, ioffs = index
    with xtn.Server(ichars, ioffs, distance=1, slots=2) as srv:
        qc, qo = xtn.synth_queries(7, ichars, ioffs, 300, planted=0.3)
        t = srv.submit(qc, qo)
        p, dist, _ = srv.wait(t)
        assert np.array_equal(as_triples(p, dist), cross_pairs(ichars, ioffs, qc, qo, 1, 0))
        for bad in (t, t + 1000, 0):
            with pytest.raises(xtn.XtnError) as e:
                srv.wait(bad)
            assert e.value.status == 3
        # two threads waiting on one ticket: one result, one error
        t2 = srv.submit(qc, qo)
        out = []

        def w():
            try:
                out.append(("ok", srv.wait(t2)))
            except xtn.XtnError as err:
                out.append(("err", err.status))

        ths = [threading.Thread(target=w) for _ in range(2)]
        for th in ths:
            th.start()
        for th in ths:
            th.join(timeout=60)
        assert sorted(k for k, _ in out) == ["err", "ok"]
        assert [v for k, v in out if k == "err"] == [3]
        tk = ctypes.c_uint64()
        st = srv.lib.xtn_server_submit(srv.s, None, xtn._u32p(np.ascontiguousarray(qo, np.uint32)),
                                       len(qo) - 1, ctypes.byref(tk))
        assert st == 3
        assert b"invalid" in srv.lib.xtn_server_last_error(srv.s)
