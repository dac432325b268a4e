"""Several waves on one query (set_query_waves / the tuned choice for small
batches) against the oracle: the leading wave keeps the pool, the visited
set and the merges, every wave computes a share of each expansion's
distances, and ids, distances and counters stay bit-exact
(SearchImpl2, graph_searcher.hpp:196-224).  Also the visited bitsets' dirty
lists: a query clears only the words it set, so bitsets are clean for the
next query on the same wave slot."""

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
def c1(golden_dir):
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    return X, Q, os.path.join(golden_dir, "c1_hnsw_l2.glass")


@pytest.mark.parametrize("waves", [2, 4])
@pytest.mark.parametrize("quant,refine", [("SQ4U", "FP32"), ("SQ8U", ""), ("E5M2", ""), ("PQ8", "FP32"),
                                          ("SQ6", ""), ("FP32", ""), ("SQ2U", "SQ8U")])
def test_waves_per_query_match_oracle(c1, quant, refine, waves):
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", quant, refine)
    o = O.Searcher(O.Graph(path), X, "L2", quant, refine)
    s.set_query_waves(waves)
    print(f"query waves={waves} {quant}+{refine or 'none'}", flush=True)
    for ef in [16, 100, 300]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q[:200], 10)
        assert s.get_query_waves()[0] == waves
        Io, Do = o.batch_search(Q[:200], 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
        for i in range(4):
            si, sd = s.search(Q[i], 10)
            oi, od = o.search(Q[i], 10)
            np.testing.assert_array_equal(si, oi)
            np.testing.assert_array_equal(sd, od)
    ef, k = synth.per_query_ef_k(120)
    ids, dis, cnt = s.batch_search_efk(Q[:120], ef, k, with_counters=True)
    for i in range(120):
        o.set_ef(int(ef[i]))
        Io, Do = o.batch_search(Q[i:i + 1], int(k[i]))
        np.testing.assert_array_equal(ids[i], Io[0])
        np.testing.assert_array_equal(dis[i], Do[0])
        assert cnt[i, 0] == o.last_counts[0][0]


def test_waves_per_query_on_long_rows(tmp_path):
    """960-d SQ8U rows on an NSG R=64 graph and 768-d IP SQ8 + FP32, the
    shapes the shared query is for."""
    X = synth.gamma_rows(2000, 960, synth.BASE_SEED)
    Q = synth.gamma_rows(60, 960, synth.QUERY_SEED)
    path = str(tmp_path / "g960.glass")
    pg.build_graph("NSG", X, path, "L2", "FP32", 64, 96)
    for waves in [2, 4]:
        s = pg.Searcher(pg.Graph(path), X, "L2", "SQ8U")
        o = O.Searcher(O.Graph(path), X, "L2", "SQ8U")
        s.set_query_waves(waves)
        print(f"960-d query waves={waves}", flush=True)
        ef, k = synth.per_query_ef_k(len(Q))
        ids, dis, cnt = s.batch_search_efk(Q, ef, k, with_counters=True)
        for i in range(len(Q)):
            o.set_ef(int(ef[i]))
            Io, Do = o.batch_search(Q[i:i + 1], int(k[i]))
            np.testing.assert_array_equal(ids[i], Io[0])
            np.testing.assert_array_equal(dis[i], Do[0])
    X = synth.unit_rows(2000, 768, synth.BASE_SEED)
    Q = synth.unit_rows(64, 768, synth.QUERY_SEED)
    path = str(tmp_path / "g768.glass")
    pg.build_graph("NSG", X, path, "IP", "FP32", 48, 80)
    s = pg.Searcher(pg.Graph(path), X, "IP", "SQ8", "FP32")
    o = O.Searcher(O.Graph(path), X, "IP", "SQ8", "FP32")
    s.set_query_waves(4)
    for ef in [32, 256]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q, 10)
        Io, Do = o.batch_search(Q, 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)


def test_tuned_waves_for_small_batches(c1):
    """optimize() records a choice per ef bucket; a batch that leaves most
    wave slots idle uses it, a large one runs one wave per query, and both
    match the oracle."""
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ8U", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "SQ8U", "FP32")
    s.optimize()
    s.set_ef(64)
    o.set_ef(64)
    tuned = s.get_query_waves()[1]
    assert tuned in (1, 2, 4)
    print(f"tuned waves for small batches: {tuned}", flush=True)
    I, D = s.batch_search(Q[:50], 10)
    assert s.get_query_waves()[0] == tuned
    Io, Do = o.batch_search(Q[:50], 10)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
    Qbig = np.tile(Q, (12, 1))
    I, D = s.batch_search(Qbig, 10)
    assert s.get_query_waves()[0] == 1
    np.testing.assert_array_equal(I[:1000], o.batch_search(Q, 10)[0])


@pytest.mark.parametrize("waves", [1, 4])
def test_overflowed_bitsets_are_clean_for_the_next_query(c1, waves):
    """A 256-entry LDS table sends nearly every query to its slot's global
    bitset; a query must leave that bitset clean, so a second batch on the
    same slots, and a batch of a different ef, still match the oracle."""
    X, Q, path = c1
    s = pg.Searcher(pg.Graph(path), X, "L2", "SQ4U", "FP32")
    o = O.Searcher(O.Graph(path), X, "L2", "SQ4U", "FP32")
    s.set_visited_capacity(256)
    s.set_query_waves(waves)
    s.set_launch(0, 0, 4)  # few workgroups: every slot runs many queries in a row
    for ef in [200, 200, 40, 600]:
        s.set_ef(ef)
        o.set_ef(ef)
        print(f"bitset reuse waves={waves} ef={ef}", flush=True)
        I, D = s.batch_search(Q[:300], 10)
        Io, Do = o.batch_search(Q[:300], 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
