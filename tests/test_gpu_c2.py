"""The headline workload's data (uint8-valued Gaussian mixture, HNSW R=32
L=200, L2, SQ4U + FP32 re-rank) through the C ABI, bit-exact:
  * against the compiled reference's own ids and distances on a 20,000-row
    slice with the reference-built graph (tests/golden/make_golden_c2.py),
    re-rank factors 1.5 (the reference default), 4 and 8, ef 16-256, k 1 and
    10, Search and SearchBatch semantics;
  * against the oracle on a 100,000-row slice with this build's HNSW graph,
    per-query dist_cmps included;
  * at the full 1,000,000 x 128 size: an oracle spot check of a sample of
    the 10,000-query batch, and the size-independent property that a query's
    result does not depend on the batch it runs in.
Reference: graph_searcher.hpp:131-159, refiner.hpp:45-110,
sq4u_quant.hpp:13-80, neighbor.hpp:361-396.
"""

import hashlib
import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu

FACTORS = [1.5, 4.0, 8.0]
EFS = [16, 48, 64, 128, 256]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def golden(golden_dir):
    g = np.load(os.path.join(golden_dir, "search_c2.npz"))
    X = synth.uint8_mixture(20000, 128, synth.BASE_SEED)
    Q = synth.uint8_mixture(100, 128, synth.QUERY_SEED)
    assert _sha(X) == str(g["data_sha"]) and _sha(Q) == str(g["query_sha"])
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, "c2_mix20k_hnsw.glass")), X, "L2", "SQ4U", "FP32")
    return g, X, Q, s


@pytest.mark.parametrize("factor", FACTORS)
def test_mixture_matches_reference(golden, factor):
    g, X, Q, s = golden
    s.set_refine_factor(factor)
    assert s.info()["refine_factor"] == factor
    for ef in EFS:
        s.set_ef(ef)
        for k in [1, 10]:
            key = f"{factor}/{ef}/{k}"
            print(f"c2 golden {key}", flush=True)
            ids, dis = s.batch_search(Q, k)
            np.testing.assert_array_equal(ids, g[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(dis, g[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(g[f"{key}/batch_dist_cmps"])
            for i in range(len(g[f"{key}/single_ids"])):
                si, sd = s.search(Q[i], k)
                np.testing.assert_array_equal(si, g[f"{key}/single_ids"][i], err_msg=f"{key} single {i}")
                np.testing.assert_array_equal(sd, g[f"{key}/single_dists"][i], err_msg=f"{key} single {i}")
    s.set_refine_factor(0)


@pytest.fixture(scope="module")
def slice100k(tmp_path_factory):
    X = synth.uint8_mixture(100000, 128, synth.BASE_SEED)
    Q = synth.uint8_mixture(300, 128, synth.QUERY_SEED)
    path = str(tmp_path_factory.mktemp("c2") / "mix100k.glass")
    print("c2 100k: building HNSW R=32 L=200", flush=True)
    pg.build_graph("HNSW", X, path, "L2", "FP32", 32, 200)
    return X, Q, path


@pytest.mark.parametrize("factor", FACTORS)
def test_mixture_100k_matches_oracle(slice100k, factor):
    X, Q, path = slice100k
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    s.set_refine_factor(factor)
    o = O.Searcher(O.Graph(path), X, "L2", "SQ4U", "FP32", refine_factor=factor)
    for ef in EFS:
        for k in [1, 10]:
            print(f"c2 100k factor={factor} ef={ef} k={k}", flush=True)
            efv = np.full(len(Q), ef, np.int32)
            kv = np.full(len(Q), k, np.int32)
            ids, dis, cnt = s.batch_search_efk(Q, efv, kv, with_counters=True)
            o.set_ef(ef)
            Io, Do = o.batch_search(Q, k)
            np.testing.assert_array_equal(np.stack(ids), Io)
            np.testing.assert_array_equal(np.stack(dis), Do)
            np.testing.assert_array_equal(cnt[:, 0], o.last_counts[0])
            np.testing.assert_array_equal(cnt[:, 1], o.last_counts[1])
            s.set_ef(ef)
            for i in range(16):
                si, sd = s.search(Q[i], k)
                oi, od = o.search(Q[i], k)
                np.testing.assert_array_equal(si, oi)
                np.testing.assert_array_equal(sd, od)


def test_mixture_full_size(tmp_path):
    """1M x 128 rows, 10,000 queries, factor 8, ef 64 (the headline operating
    point): a sample checked against the oracle, and every query's result
    equal whether it runs in the full batch, in a small batch or alone."""
    X = synth.uint8_mixture(1_000_000, 128, synth.BASE_SEED)
    Q = synth.uint8_mixture(10_000, 128, synth.QUERY_SEED)
    path = str(tmp_path / "mix1m.glass")
    print("c2 1M: building HNSW R=32 L=200", flush=True)
    pg.build_graph("HNSW", X, path, "L2", "FP32", 32, 200)
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    s.set_refine_factor(8.0)
    s.set_ef(64)
    print("c2 1M: batch of 10000", flush=True)
    ids, dis = s.batch_search(Q, 10)
    rng = np.random.default_rng(7)
    sample = np.sort(rng.choice(len(Q), 64, replace=False))
    o = O.Searcher(O.Graph(path), X, "L2", "SQ4U", "FP32", refine_factor=8.0)
    o.set_ef(64)
    Io, Do = o.batch_search(Q[sample], 10)
    np.testing.assert_array_equal(ids[sample], Io)
    np.testing.assert_array_equal(dis[sample], Do)
    Is, Ds = s.batch_search(Q[sample], 10)
    np.testing.assert_array_equal(Is, ids[sample])
    np.testing.assert_array_equal(Ds, dis[sample])
    for i in sample[:4]:
        bi, bd = s.batch_search(Q[i:i + 1], 10)
        np.testing.assert_array_equal(bi[0], ids[i])
        np.testing.assert_array_equal(bd[0], dis[i])
    # a small LDS table and ef = 2500: queries set more bitset words than the
    # slot's dirty list holds and clear the whole bitset at their end; the
    # same queries again on the same slots match the oracle
    s.set_visited_capacity(256)
    s.set_launch(0, 0, 2)
    o.set_ef(2500)
    s.set_ef(2500)
    print("c1 1M: ef=2500 with a 256-entry table", flush=True)
    Io, Do = o.batch_search(Q[:6], 10)
    for _ in range(2):
        Ib, Db = s.batch_search(Q[:6], 10)
        np.testing.assert_array_equal(Ib, Io)
        np.testing.assert_array_equal(Db, Do)
    s.set_visited_capacity(0)
    s.set_launch(0, 0, 0)
    # recall sanity at the operating point (exact kNN of the sample in float64)
    xs = X.astype(np.float64)
    hit = 0
    for j, i in enumerate(sample[:16]):
        d = ((xs - Q[i].astype(np.float64)) ** 2).sum(1)
        gt = np.lexsort((np.arange(len(d)), d))[:10]
        hit += len(np.intersect1d(gt, ids[i]))
    assert hit / 160 > 0.85
