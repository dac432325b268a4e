"""The C++ drop-in: include/glass_hip_searcher.hpp (HipSearcher, the
glass::GraphSearcherBase a maintainer adds to the reference) compiled against
the reference's own headers (oracle/hip_searcher_harness.cpp, built by
oracle/Makefile into oracle/_ref/libhipsearcher.so), driven only through the
GraphSearcherBase virtuals (glass/searcher/searcher_base.hpp:7-19) on a graph
the reference's Graph::load read.  Results equal the Python mirror's and the
oracle's, bit for bit."""

import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O
from pyglass_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libhipsearcher.so")


def _load():
    if not os.path.exists(LIB):
        pytest.skip("oracle/_ref/libhipsearcher.so is built from the reference sources (make -C oracle ref)")
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.hs_new.argtypes = [C.c_char_p, vp, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p]
    L.hs_new.restype = vp
    L.hs_free.argtypes = [vp]
    L.hs_set_ef.argtypes = [vp, C.c_int]
    L.hs_get_ef.argtypes = [vp]
    L.hs_optimize.argtypes = [vp]
    L.hs_search.argtypes = [vp, vp, C.c_int, vp, vp]
    L.hs_search_batch.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp]
    L.hs_dist_cmps.argtypes = [vp]
    L.hs_dist_cmps.restype = C.c_double
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_harness_exports_the_searcher_calls():
    L = _load()
    for name in ["hs_new", "hs_free", "hs_set_ef", "hs_get_ef", "hs_optimize", "hs_search", "hs_search_batch",
                 "hs_dist_cmps"]:
        assert hasattr(L, name)


@pytest.mark.gpu
@pytest.mark.parametrize("graph,metric,quant,refine", [("c1_hnsw_l2", "L2", "SQ4U", "FP32"),
                                                      ("nsg_l2", "L2", "SQ8U", ""),
                                                      ("ip_hnsw", "IP", "SQ8", "FP32")])
def test_hip_searcher_through_graph_searcher_base(golden_dir, graph, metric, quant, refine):
    import pyglass_amd as pg

    if pg.device_count() < 1:
        pytest.skip("no HIP device")
    L = _load()
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(200, 128, synth.QUERY_SEED)
    if graph == "nsg_l2":
        X = X[:5000].copy()
    if graph == "ip_hnsw":
        X = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
        Q = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    path = os.path.join(golden_dir, f"{graph}.glass")
    print(f"HipSearcher {graph} {quant}+{refine or 'none'}", flush=True)
    h = L.hs_new(path.encode(), _p(X), X.shape[0], X.shape[1], metric.encode(), quant.encode(), refine.encode())
    assert h
    try:
        o = O.Searcher(O.Graph(path), X, metric, quant, refine)
        s = pg.Searcher(pg.Graph(path), X, metric, quant, refine)
        assert L.hs_optimize(h) == 0
        for ef in [16, 64]:
            assert L.hs_set_ef(h, ef) == 0 and L.hs_get_ef(h) == ef
            o.set_ef(ef)
            s.set_ef(ef)
            ids = np.zeros((len(Q), 10), np.int32)
            dis = np.zeros((len(Q), 10), np.float32)
            assert L.hs_search_batch(h, _p(Q), len(Q), 10, _p(ids), _p(dis)) == 0
            Io, Do = o.batch_search(Q, 10)
            np.testing.assert_array_equal(ids, Io)
            np.testing.assert_array_equal(dis, Do)
            assert L.hs_dist_cmps(h) == o.get_last_search_dist_cmps()
            Ip, Dp = s.batch_search(Q, 10)
            np.testing.assert_array_equal(ids, Ip)
            for i in range(4):
                si = np.zeros(10, np.int32)
                sd = np.zeros(10, np.float32)
                assert L.hs_search(h, _p(Q[i]), 10, _p(si), _p(sd)) == 0
                oi, od = o.search(Q[i], 10)
                np.testing.assert_array_equal(si, oi)
                np.testing.assert_array_equal(sd, od)
        # errors surface as exceptions inside the drop-in, a status here
        assert L.hs_search_batch(h, _p(Q), len(Q), 0, _p(ids), _p(dis)) == -1
    finally:
        L.hs_free(h)
