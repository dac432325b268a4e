"""HIP path vs oracle on the cases the golden graphs do not reach: the
hand-worked HNSW toy graph, ef >= N brute force (both pool-merge paths),
per-query ef/k, the visited-set overflow path, every launch shape, short
pools (k larger than the reachable set) and the device-pointer entry point.
All through the C ABI; ids and distances bit-exact.
"""

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from oracle import toy_graphs as T
from pyglass_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


def ring_graph(n, k, seed):
    rng = np.random.default_rng(seed)
    adj = np.full((n, k), -1, np.int32)
    for i in range(n):
        nb = [(i + 1) % n] + list(rng.choice(n, k - 1, replace=False))
        nb = [v for v in dict.fromkeys(nb) if v != i][:k]
        adj[i, :len(nb)] = nb
    return adj


def mixture(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3, 3, (8, d)).astype(np.float32)
    return (c[rng.integers(0, 8, n)] + rng.standard_normal((n, d)).astype(np.float32)).astype(np.float32)


@pytest.mark.parametrize("batch", [False, True])
def test_toy_hnsw_descent(tmp_path, batch):
    path = str(tmp_path / "toy.glass")
    X, q = T.toy_hnsw(path)
    print(f"toy hnsw batch={batch}", flush=True)
    s = pg.Searcher(pg.Graph(path), X, "L2", "FP32")
    s.set_ef(3)
    if batch:
        ids, dis = s.batch_search(q[None], 3)
        ids, dis = ids[0], dis[0]
        assert s.get_last_search_dist_cmps() == T.TOY_DIST_CMPS
    else:
        ids, dis = s.search(q, 3)
    assert list(ids) == T.TOY_RESULT_IDS
    assert list(dis) == T.TOY_RESULT_DISTS


def test_short_pool_pads_like_oracle(tmp_path):
    # k = 5 but only 4 nodes are reachable from the entry: the tail is -1/+inf
    path = str(tmp_path / "toy.glass")
    X, q = T.toy_hnsw(path)
    s = pg.Searcher(pg.Graph(path), X, "L2", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "FP32")
    s.set_ef(3)
    o.set_ef(3)
    I, D = s.batch_search(q[None], 5)
    Io, Do = o.batch_search(q[None], 5)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    assert I[0, -1] == -1 and np.isinf(D[0, -1])


@pytest.mark.parametrize("n", [240, 600])
@pytest.mark.parametrize("quant,metric,refine", [("SQ4U", "L2", "FP32"), ("SQ8U", "IP", ""), ("SQ8", "L2", ""),
                                                 ("SQ6", "L2", ""), ("SQ4", "IP", ""), ("PQ8", "L2", "FP32"),
                                                 ("BF16", "IP", "")])
def test_ef_at_least_n_matches_oracle(n, quant, metric, refine):
    # PQ8 at d = 64: 16 sub-spaces (the 16-block of the LUT sum); with n = 240
    # its k-means has fewer sample rows than centroids (k-means++ repeats picks)
    # n=240 keeps the pool in registers (capacity <= 256); n=600 takes the LDS merge
    X = mixture(n, 64, 31)
    Q = mixture(40, 64, 32)
    adj = ring_graph(n, 12, 33)
    s = pg.Searcher(pg.Graph.from_arrays(adj, [3]), X, metric, quant, refine)
    o = O.Searcher(O.Graph(adj=adj, eps=[3]), X, metric, quant, refine)
    print(f"ef>=N n={n} {quant} {metric} {refine or 'none'}", flush=True)
    for ef, k in [(n, 10), (n + 7, 1), (n // 2, 25)]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q, k)
        Io, Do = o.batch_search(Q, k)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
        for i in range(0, 40, 13):
            a, b = s.search(Q[i], k)
            c, d = o.search(Q[i], k)
            np.testing.assert_array_equal(a, c)
            np.testing.assert_array_equal(b, d)


@pytest.fixture(scope="module")
def c1(golden_dir):
    import os

    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    path = os.path.join(golden_dir, "c1_hnsw_l2.glass")
    return X, Q, path


@pytest.mark.parametrize("quant,refine", [("SQ8U", ""), ("SQ4U", "FP32"), ("E5M2", ""), ("PQ8", "FP32"), ("SQ6", ""),
                                         ("BF16", "")])
def test_per_query_ef_k_matches_oracle(c1, quant, refine):
    X, Q, path = c1
    nq = 300
    rng = np.random.default_rng(synth.EFK_SEED)
    ef = np.rint(np.exp(rng.uniform(np.log(16), np.log(512), nq))).astype(np.int32)
    k = rng.integers(1, 101, nq).astype(np.int32)
    s = pg.Searcher(pg.Graph(path), X, "L2", quant, refine)
    o = O.Searcher(O.Graph(path), X, "L2", quant, refine)
    print(f"per-query ef/k {quant} {refine or 'none'}", flush=True)
    ids, dis, cnt = s.batch_search_efk(Q[:nq], ef, k, with_counters=True)
    for i in range(nq):
        o.set_ef(int(ef[i]))
        Io, Do = o.batch_search(Q[i:i + 1], int(k[i]))
        np.testing.assert_array_equal(ids[i], Io[0], err_msg=f"query {i} ef={ef[i]} k={k[i]}")
        np.testing.assert_array_equal(dis[i], Do[0], err_msg=f"query {i}")
        assert cnt[i, 0] == o.last_counts[0][0]
        assert cnt[i, 1] == o.last_counts[1][0]


@pytest.mark.parametrize("quant", ["SQ8U", "PQ8"])
@pytest.mark.parametrize("entries", [256, 512])
def test_visited_overflow_to_global_bitset(c1, entries, quant):
    # a small LDS table forces most queries past half occupancy onto the
    # slot's global bitset part-way through the search; PQ8's lookup table
    # sits in front of that bitset in the same slot scratch
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", quant)
    o = O.Searcher(O.Graph(path), X, "L2", quant)
    s.set_visited_capacity(entries)
    print(f"visited overflow {quant} entries={entries}", flush=True)
    for ef in [64, 200]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q[:200], 10)
        Io, Do = o.batch_search(Q[:200], 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()


@pytest.mark.parametrize("team", [4, 8, 16])
@pytest.mark.parametrize("wpb", [1, 2, 4])
def test_every_launch_shape_gives_the_same_result(c1, team, wpb):
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ8U", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "SQ8U", "FP32")
    s.set_launch(team, wpb, 0)
    s.set_ef(48)
    o.set_ef(48)
    print(f"launch team={team} wpb={wpb}", flush=True)
    I, D = s.batch_search(Q[:256], 10)
    Io, Do = o.batch_search(Q[:256], 10)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    # a grid smaller than the batch: the persistent waves loop over the queue
    s.set_launch(team, wpb, 3)
    I2, D2 = s.batch_search(Q[:256], 10)
    np.testing.assert_array_equal(I2, Io)


def test_device_pointer_entry_point_matches_host_batch(c1):
    import torch

    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    s.set_ef(64)
    I, D = s.batch_search(Q[:500], 10)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(Q[:500].copy()).to(dev)
    ids = torch.full((500, 10), -7, dtype=torch.int32, device=dev)
    dis = torch.zeros((500, 10), dtype=torch.float32, device=dev)
    cnt = torch.zeros((500, 3), dtype=torch.int32, device=dev)
    st = torch.cuda.Stream(device=dev)
    print("device-pointer batch", flush=True)
    s.search_batch_dev(q.data_ptr(), 500, 10, ids.data_ptr(), dis.data_ptr(), cnt.data_ptr(), stream=st.cuda_stream)
    st.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy(), I)
    np.testing.assert_array_equal(dis.cpu().numpy(), D)
    assert cnt.cpu().numpy()[:, 0].mean() == s.get_last_search_dist_cmps()


def test_device_batches_on_two_streams_back_to_back(c1):
    """Two device-pointer batches on different streams with no host sync
    between them share the searcher's work counter and visited bitsets; the
    second launch is ordered after the first, so both match the host path."""
    import torch

    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    s.set_ef(64)
    I1, D1 = s.batch_search(Q[:400], 10)
    s.set_ef(128)
    I2, D2 = s.batch_search(Q[400:1000], 10)
    dev = torch.device("cuda:0")
    q = torch.from_numpy(Q[:1000].copy()).to(dev)
    out = [(torch.full((n, 10), -7, dtype=torch.int32, device=dev), torch.zeros((n, 10), dtype=torch.float32,
                                                                                 device=dev)) for n in (400, 600)]
    st = [torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)]
    torch.cuda.synchronize()
    print("two streams back to back", flush=True)
    s.set_graph_capture(False)
    for rep in range(3):
        s.set_ef(64)
        s.search_batch_dev(q.data_ptr(), 400, 10, out[0][0].data_ptr(), out[0][1].data_ptr(), 0,
                           stream=st[0].cuda_stream)
        s.set_ef(128)
        s.search_batch_dev(q[400:].data_ptr(), 600, 10, out[1][0].data_ptr(), out[1][1].data_ptr(), 0,
                           stream=st[1].cuda_stream)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(out[0][0].cpu().numpy(), I1)
        np.testing.assert_array_equal(out[0][1].cpu().numpy(), D1)
        np.testing.assert_array_equal(out[1][0].cpu().numpy(), I2)
        np.testing.assert_array_equal(out[1][1].cpu().numpy(), D2)


def test_more_streams_than_launch_states(c1):
    """Batches cycling over six streams with no host sync between them: the
    searcher keeps four per-stream launch states, so the fifth and sixth
    streams take over the least recently used states, each after waiting for
    that state's last launch.  Every batch matches the host path."""
    import torch

    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    parts = [(0, 150, 32), (150, 350, 64), (350, 500, 16), (500, 700, 128), (700, 850, 48), (850, 1000, 96)]
    want = []
    for a, b, ef in parts:
        s.set_ef(ef)
        want.append(s.batch_search(Q[a:b], 10))
    dev = torch.device("cuda:0")
    q = torch.from_numpy(Q[:1000].copy()).to(dev)
    out = [(torch.full((b - a, 10), -7, dtype=torch.int32, device=dev),
            torch.zeros((b - a, 10), dtype=torch.float32, device=dev)) for a, b, _ in parts]
    st = [torch.cuda.Stream(device=dev) for _ in parts]
    torch.cuda.synchronize()
    s.set_graph_capture(False)
    for rep in range(3):
        print(f"six streams, round {rep}", flush=True)
        for j, (a, b, ef) in enumerate(parts):
            out[j][0].fill_(-7)
        torch.cuda.synchronize()
        for j, (a, b, ef) in enumerate(parts):
            s.set_ef(ef)
            s.search_batch_dev(q[a:].data_ptr(), b - a, 10, out[j][0].data_ptr(), out[j][1].data_ptr(), 0,
                               stream=st[(j + rep) % len(st)].cuda_stream)
        torch.cuda.synchronize()
        for j in range(len(parts)):
            np.testing.assert_array_equal(out[j][0].cpu().numpy(), want[j][0])
            np.testing.assert_array_equal(out[j][1].cpu().numpy(), want[j][1])


@pytest.mark.parametrize("ef", [1500, 3500])
def test_large_pools_fit_lds_by_fewer_waves(c1, ef):
    """Pools whose LDS at the tuned shape exceeds a CU run with fewer waves
    per workgroup (and, if needed, a smaller visited table)."""
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ8U", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "SQ8U", "FP32")
    s.set_ef(ef)
    o.set_ef(ef)
    print(f"large pool ef={ef}", flush=True)
    I, D = s.batch_search(Q[:64], 10)
    Io, Do = o.batch_search(Q[:64], 10)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    ki = s.kernel_info()
    assert ki["lds_per_block"] <= 160 * 1024


@pytest.mark.parametrize("kind,R", [("HNSW", 128), ("NSG", 100)])
@pytest.mark.parametrize("quant,refine", [("SQ8U", ""), ("SQ4U", "FP32"), ("FP32", "")])
def test_degree_above_64_matches_oracle(tmp_path, kind, R, quant, refine):
    """Rows of more than 64 neighbours (HNSW R=128: 128 at level 0, 64 in the
    upper layers; NSG R=100) are read in chunks of 64 entries."""
    X = mixture(4000, 64, 3)
    Q = mixture(128, 64, 4)
    path = str(tmp_path / f"{kind}{R}.glass")
    pg.build_graph(kind, X, path, "L2", "FP32", R, 200 if kind == "HNSW" else R + 32)
    G = pg.Graph(path)
    assert G.info()["K"] == R
    s = pg.Searcher(G, X, "L2", quant, refine)
    o = O.Searcher(O.Graph(path), X, "L2", quant, refine)
    print(f"degree {R} {kind} {quant}+{refine or 'none'}", flush=True)
    for ef in [16, 100]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q, 10)
        Io, Do = o.batch_search(Q, 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
        for i in range(8):
            si, sd = s.search(Q[i], 10)
            oi, od = o.search(Q[i], 10)
            np.testing.assert_array_equal(si, oi)
            np.testing.assert_array_equal(sd, od)


def test_pool_beyond_4096_entries(c1):
    """Pools are bounded by a CU's LDS, not by a fixed count: ef = 6000 and
    a re-rank depth of 4,500 entries against the oracle; a pool too large
    for LDS is an error, not a truncation."""
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "SQ4U", "FP32")
    print("pool ef=6000", flush=True)
    s.set_ef(6000)
    o.set_ef(6000)
    I, D = s.batch_search(Q[:16], 10)
    Io, Do = o.batch_search(Q[:16], 10)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    s.set_ef(64)
    o.set_ef(64)
    print("pool reorder_k=4500", flush=True)
    I, D = s.batch_search(Q[:8], 3000)
    Io, Do = o.batch_search(Q[:8], 3000)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    s.set_ef(40000)
    with pytest.raises(pg.GlassError):
        s.batch_search(Q[:2], 10)


def test_empty_batch_and_bad_arguments(c1):
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ8U")
    I, D = s.batch_search(Q[:0], 10)
    assert I.shape == (0, 10)
    with pytest.raises(pg.GlassError):
        s.batch_search(Q[:4], 0)
    with pytest.raises(pg.GlassError):
        s.set_visited_capacity(300)


@pytest.mark.parametrize("quant,metric", [("SQ4U", "L2"), ("SQ2U", "L2"), ("SQ8U", "L2"), ("SQ8U", "IP")])
@pytest.mark.parametrize("team", [8, 16])
@pytest.mark.parametrize("gather", [0, 1])
def test_neighbour_gather_modes_match_oracle(c1, quant, metric, team, gather):
    # gather=1 loads every neighbour's code row before the visited test (when
    # one lane group per neighbour fits the wave); results must not change
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, metric, quant, "FP32" if quant != "SQ8U" else "")
    o = O.Searcher(O.Graph(path), X, metric, quant, "FP32" if quant != "SQ8U" else "")
    s.set_launch(team, 1, 0)
    s.set_gather(gather)
    print(f"gather={gather} team={team} {quant} {metric}", flush=True)
    for ef, k in [(24, 10), (96, 10), (300, 50)]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q[:200], k)
        Io, Do = o.batch_search(Q[:200], k)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()


@pytest.mark.parametrize("quant,refine", [("SQ4U", "FP32"), ("PQ8", "FP32"), ("SQ6", ""), ("FP16", "")])
def test_graph_replay_matches_eager(c1, quant, refine):
    # search_batch_dev captures (counter-reset kernel + search kernel) after
    # the first eager call and replays the graph while the arguments repeat;
    # every replay is checked against the eager result, with the outputs
    # cleared before it and eager searches + synchronisations in between.
    # PQ8 rebuilds each query's lookup table in the slot scratch inside the
    # replayed launch.
    import torch

    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", quant, refine)
    dev = torch.device("cuda:0")
    st = torch.cuda.Stream(device=dev)
    q = torch.from_numpy(Q[:500].copy()).to(dev)
    ids = torch.empty((500, 10), dtype=torch.int32, device=dev)
    dis = torch.empty((500, 10), dtype=torch.float32, device=dev)
    for ef in [64, 32]:
        s.set_ef(ef)
        I_ref, D_ref = s.batch_search(Q[:500], 10)
        for rep in range(5):
            print(f"graph replay {quant} ef={ef} rep={rep}", flush=True)
            with torch.cuda.stream(st):
                ids.fill_(-7)
                dis.fill_(0.0)
            s.search_batch_dev(q.data_ptr(), 500, 10, ids.data_ptr(), dis.data_ptr(), stream=st.cuda_stream)
            st.synchronize()
            np.testing.assert_array_equal(ids.cpu().numpy(), I_ref)
            np.testing.assert_array_equal(dis.cpu().numpy(), D_ref)
            # the program's own eager work between replays
            Ie, _ = s.batch_search(Q[500:600], 10)
            torch.cuda.synchronize()
    s.set_graph_capture(False)
    with torch.cuda.stream(st):
        ids.fill_(-7)
    s.search_batch_dev(q.data_ptr(), 500, 10, ids.data_ptr(), dis.data_ptr(), stream=st.cuda_stream)
    st.synchronize()
    np.testing.assert_array_equal(ids.cpu().numpy(), I_ref)
