"""Served form (continuous batching) through the C ABI: a request's result is
the same alone or in any mix of other requests, equals the batch path and
the oracle, and cancel / poll / replay behave as declared."""

import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def setup(golden_dir):
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    path = os.path.join(golden_dir, "c1_hnsw_l2.glass")
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    return X, Q, path, s


def test_alone_equals_in_mix_equals_batch_and_oracle(setup):
    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=1024, max_k=64, max_inflight=2)
    print("server alone vs mix", flush=True)
    a_ids, a_d = srv.wait(srv.submit(Q[7], k=10, ef=64))
    rng = np.random.default_rng(synth.EFK_SEED)
    tickets = []
    row = 0
    for i in range(60):
        nq = int(rng.integers(1, 9))
        ef = int(rng.choice([16, 32, 64, 128, 256]))
        k = int(rng.integers(1, 51))
        rows = np.arange(row, row + nq) % len(Q)
        tickets.append((srv.submit(Q[rows], k=k, ef=ef), rows, ef, k))
        row += nq
        if i == 30:
            tickets.append((srv.submit(Q[7], k=10, ef=64), np.array([7]), 64, 10))
    o = O.Searcher(O.Graph(path), X, "L2", "SQ4U", "FP32")
    for t, rows, ef, k in tickets:
        ids, dis = srv.wait(t)
        bi, bd = s.batch_search_efk(Q[rows], ef, k)
        np.testing.assert_array_equal(ids, np.stack(bi))
        np.testing.assert_array_equal(dis, np.stack(bd))
        if len(rows) == 1 and rows[0] == 7 and ef == 64 and k == 10:
            np.testing.assert_array_equal(ids, a_ids)
            np.testing.assert_array_equal(dis, a_d)
        o.set_ef(ef)
        oi, od = o.batch_search(Q[rows[:1]], k)
        np.testing.assert_array_equal(ids[:1], oi)
        np.testing.assert_array_equal(dis[:1], od)
    st = srv.stats()
    assert st["requests"] == len(tickets) + 1
    srv.close()


def test_poll_cancel_and_errors(setup):
    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=1024, max_k=16, max_inflight=1)
    big = srv.submit(Q[:1000], k=10, ef=256)
    small = srv.submit(Q[3], k=5, ef=32)
    try:
        srv.cancel(small)
        cancelled = True
    except pg.GlassError:
        cancelled = False  # the dispatcher had already taken it
    ids, _ = srv.wait(big)
    assert ids.shape == (1000, 10)
    if cancelled:
        with pytest.raises(pg.GlassError):
            srv.cancel(small)  # the ticket is gone
    else:
        r = srv.wait(small)
        assert r[0].shape == (1, 5)
    t = srv.submit(Q[4], k=5, ef=32)
    res = None
    for _ in range(200000):
        res = srv.poll(t)
        if res is not None:
            break
    assert res is not None
    with pytest.raises(pg.GlassError):
        srv.submit(Q[:2], k=17, ef=32)  # k above max_k
    with pytest.raises(pg.GlassError):
        srv.submit(Q[:2], k=5, ef=40000)  # a pool larger than a CU's LDS
    srv.close()


def test_replay_poisson_matches_batch(setup):
    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=2048, max_k=10, max_inflight=2)
    rng = np.random.default_rng(synth.ARRIVAL_SEED)
    n = 2000
    gaps = rng.exponential(1e9 / 200_000, n)  # 200k requests/s offered
    arr = np.cumsum(gaps).astype(np.int64)
    rows = rng.integers(0, len(Q), n)
    print("server replay", flush=True)
    lat, ids, dis = srv.replay(Q[rows], 1, 48, 10, arr)
    s.set_ef(48)
    bi, bd = s.batch_search(Q[rows], 10)
    np.testing.assert_array_equal(ids.reshape(n, 10), bi)
    np.testing.assert_array_equal(dis.reshape(n, 10), bd)
    assert np.all(lat > 0)
    st = srv.stats()
    assert st["queries"] == n and st["batches"] > 1
    srv.close()


def test_large_pool_request_does_not_stop_the_server(setup):
    """A request whose pool needs fewer waves per workgroup to fit LDS runs
    like any other, and the requests after it are served."""
    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=1024, max_k=16, max_inflight=2)
    print("server ef=3500 then ef=64", flush=True)
    t1 = srv.submit(Q[:4], k=10, ef=3500)
    t2 = srv.submit(Q[4:6], k=10, ef=64)
    i1, d1 = srv.wait(t1)
    i2, d2 = srv.wait(t2)
    for (ids, dis), rows, ef in [((i1, d1), slice(0, 4), 3500), ((i2, d2), slice(4, 6), 64)]:
        bi, bd = s.batch_search_efk(Q[rows], ef, 10)
        np.testing.assert_array_equal(ids, np.stack(bi))
        np.testing.assert_array_equal(dis, np.stack(bd))
    i3, _ = srv.wait(srv.submit(Q[9], k=10, ef=32))
    assert i3.shape == (1, 10)
    srv.close()


@pytest.mark.parametrize("quant", ["SQ8", "SQ8U"])
def test_served_768d_ip_requests(tmp_path, quant):
    """The serving workload's shape at reduced N: 768-d unit rows, NSG R=48,
    inner product, per-dim SQ8 (float) or SQ8U (integer) codes + FP32
    re-rank; requests of U[1, 256] queries with per-request ef in {32, 64,
    128, 256} and k = 10 arriving as a Poisson process.  Each request gets
    the same ids and distances alone, in the mix, through
    batch_search_efk and from the oracle."""
    X = synth.unit_rows(3000, 768, synth.BASE_SEED)
    Q = synth.unit_rows(1500, 768, synth.QUERY_SEED)
    path = str(tmp_path / "g768.glass")
    pg.build_graph("NSG", X, path, "IP", "FP32", 48, 80)
    s = pg.Searcher(pg.Graph(path), X, "IP", quant, "FP32")
    rng = np.random.default_rng(synth.ARRIVAL_SEED)
    nreq = 24
    nq = rng.integers(1, 257, nreq).astype(np.int32)
    ef = rng.choice([32, 64, 128, 256], nreq).astype(np.int32)
    starts = np.concatenate([[0], np.cumsum(nq)])
    rows = np.arange(starts[-1]) % len(Q)
    arr = np.cumsum(rng.exponential(1e9 / 20000, nreq)).astype(np.int64)
    srv = pg.Server(s, max_batch=1024, max_k=10, max_inflight=3)
    alone = {}
    for j in (0, 7, 15):
        r = rows[starts[j]:starts[j + 1]]
        alone[j] = srv.wait(srv.submit(Q[r], k=10, ef=int(ef[j])))
    print(f"served 768-d IP {quant}+FP32: {nreq} requests, {starts[-1]} queries", flush=True)
    lat, ids, dis = srv.replay(Q[rows], nq, ef, 10, arr)
    srv.close()
    o = O.Searcher(O.Graph(path), X, "IP", quant, "FP32")
    for j in range(nreq):
        r = rows[starts[j]:starts[j + 1]]
        mi = ids[starts[j] * 10:starts[j + 1] * 10].reshape(-1, 10)
        md = dis[starts[j] * 10:starts[j + 1] * 10].reshape(-1, 10)
        bi, bd = s.batch_search_efk(Q[r], int(ef[j]), 10)
        np.testing.assert_array_equal(mi, np.stack(bi), err_msg=f"request {j}")
        np.testing.assert_array_equal(md, np.stack(bd), err_msg=f"request {j}")
        if j in alone:
            np.testing.assert_array_equal(mi, alone[j][0])
            np.testing.assert_array_equal(md, alone[j][1])
        o.set_ef(int(ef[j]))
        oi, od = o.batch_search(Q[r], 10)
        np.testing.assert_array_equal(mi, oi, err_msg=f"request {j} vs oracle")
        np.testing.assert_array_equal(md, od, err_msg=f"request {j} vs oracle")
    assert np.all(lat > 0)


def test_request_larger_than_a_batch_is_split(setup):
    """A request of more queries than max_batch runs over several batches and
    completes once all of them have; its results equal the batch path's,
    alone and among other requests (submit and replay)."""
    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=256, max_k=16, max_inflight=3)
    inf = srv.info()
    assert 1 <= inf["max_inflight"] <= 3 and inf["max_batch"] == 256
    print("server split request", flush=True)
    t_small = srv.submit(Q[5:8], k=10, ef=48)
    t_big = srv.submit(Q[:1000], k=10, ef=64)
    t_after = srv.submit(Q[9], k=5, ef=32)
    ib, db = srv.wait(t_big)
    bi, bd = s.batch_search_efk(Q[:1000], 64, 10)
    np.testing.assert_array_equal(ib, np.stack(bi))
    np.testing.assert_array_equal(db, np.stack(bd))
    np.testing.assert_array_equal(srv.wait(t_small)[0], np.stack(s.batch_search_efk(Q[5:8], 48, 10)[0]))
    np.testing.assert_array_equal(srv.wait(t_after)[0], np.stack(s.batch_search_efk(Q[9:10], 32, 5)[0]))
    nq = np.array([700, 3, 900, 1], np.int32)
    ef = np.array([64, 32, 128, 16], np.int32)
    Q2 = np.concatenate([Q, Q])[:1604]
    lat, ids, dis = srv.replay(Q2, nq, ef, 10, np.zeros(4, np.int64))
    start = 0
    for n_, e_ in zip(nq, ef):
        bi, bd = s.batch_search_efk(Q2[start:start + n_], int(e_), 10)
        np.testing.assert_array_equal(ids[start * 10:(start + n_) * 10].reshape(-1, 10), np.stack(bi))
        np.testing.assert_array_equal(dis[start * 10:(start + n_) * 10].reshape(-1, 10), np.stack(bd))
        start += n_
    st = srv.stats()
    assert st["requests"] == 7 and st["queries"] == 1004 + 1604
    srv.close()


def test_concurrent_clients_and_replay(setup):
    """Six client threads submit, wait and poll at once while a seventh
    replays: requests that lie whole in one batch (completed by that batch's
    worker without the server lock) and split ones (settled under it) all
    equal the batch path, and every request is counted once."""
    import threading

    X, Q, path, s = setup
    srv = pg.Server(s, max_batch=256, max_k=16, max_inflight=4)
    rng = np.random.default_rng(synth.EFK_SEED + 7)
    plans = []
    row = 0
    for t in range(6):
        reqs = []
        for i in range(120):
            nq = int(rng.choice([1, 1, 1, 2, 5, 300]))
            ef = int(rng.choice([16, 64, 128]))
            k = int(rng.choice([1, 10, 16]))
            rows = np.arange(row, row + nq) % len(Q)
            row += nq
            reqs.append((rows, ef, k))
        plans.append(reqs)
    rep_nq = rng.integers(1, 9, 400).astype(np.int32)
    rep_ef = rng.choice([32, 64], 400).astype(np.int32)
    rep_rows = np.arange(int(rep_nq.sum())) % len(Q)
    # expected results: the batch path with per-query ef and k
    allreq = [r for p in plans for r in p]
    Qa = np.concatenate([Q[rows] for rows, _, _ in allreq])
    efa = np.concatenate([np.full(len(rows), ef, np.int32) for rows, ef, _ in allreq])
    ka = np.concatenate([np.full(len(rows), k, np.int32) for rows, _, k in allreq])
    bi, bd = s.batch_search_efk(Qa, efa, ka)
    ri, rd = s.batch_search_efk(Q[rep_rows], np.repeat(rep_ef, rep_nq), 10)
    got = [None] * 6
    errors = []
    print("server: 6 client threads + replay", flush=True)

    def client(t):
        try:
            tickets = [srv.submit(Q[rows], k=k, ef=ef) for rows, ef, k in plans[t]]
            out = [None] * len(tickets)
            for j in range(len(tickets) - 1, -1, -2):
                out[j] = srv.wait(tickets[j])
            todo = [j for j in range(len(tickets)) if out[j] is None]
            while todo:
                left = []
                for j in todo:
                    r = srv.poll(tickets[j])
                    if r is None:
                        left.append(j)
                    else:
                        out[j] = r
                todo = left
            got[t] = out
        except Exception as e:  # re-raised below
            errors.append(e)

    rep = {}

    def replayer():
        try:
            rep["r"] = srv.replay(Q[rep_rows], rep_nq, rep_ef, 10, np.zeros(len(rep_nq), np.int64))
        except Exception as e:
            errors.append(e)

    th = [threading.Thread(target=client, args=(t,)) for t in range(6)] + [threading.Thread(target=replayer)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=120)
    assert not any(x.is_alive() for x in th), "a client thread did not finish"
    if errors:
        raise errors[0]
    q0 = 0
    for t in range(6):
        for (rows, ef, k), (ids, dis) in zip(plans[t], got[t]):
            n = len(rows)
            np.testing.assert_array_equal(ids, np.stack(bi[q0:q0 + n]))
            np.testing.assert_array_equal(dis, np.stack(bd[q0:q0 + n]))
            q0 += n
    _, rids, rdis = rep["r"]
    np.testing.assert_array_equal(rids.reshape(-1, 10), np.stack(ri))
    np.testing.assert_array_equal(rdis.reshape(-1, 10), np.stack(rd))
    st = srv.stats()
    assert st["requests"] == len(allreq) + len(rep_nq)
    assert st["queries"] == len(Qa) + len(rep_rows)
    srv.close()
