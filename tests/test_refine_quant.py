"""SQ8U and SQ8 as re-rank quantizers (make_refiner,
glass/searcher/graph_searcher.hpp:240-254; Refiner,
glass/searcher/refiner.hpp:45-110) against the compiled reference's vectors
(tests/golden/make_golden_refine.py): the oracle on the CPU, the HIP path on
the GPU."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from pyglass_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = [("SQ4U", "SQ8U"), ("SQ4U", "SQ8"), ("SQ2U", "SQ8U"), ("SQ8U", "SQ8U"), ("E5M2", "SQ8")]
GRAPHS = ["c1_hnsw_l2", "ip_hnsw"]


@pytest.fixture(scope="module")
def data():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)[:48]
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP")}, np.load(os.path.join(GOLD, "search_refine.npz"))


def _check(searcher, key0, Q, g):
    for ef in [16, 64]:
        searcher.set_ef(ef)
        for k in [1, 10]:
            key = f"{key0}/{ef}/{k}"
            ids, dis = searcher.batch_search(Q, k)
            np.testing.assert_array_equal(ids, g[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(dis, g[f"{key}/batch_dists"], err_msg=key)
            assert searcher.get_last_search_dist_cmps() == float(g[f"{key}/batch_dist_cmps"]), key
            for i in range(len(Q)):
                si, sd = searcher.search(Q[i], k)
                np.testing.assert_array_equal(si, g[f"{key}/single_ids"][i], err_msg=f"{key} single {i}")
                np.testing.assert_array_equal(sd, g[f"{key}/single_dists"][i], err_msg=f"{key} single {i}")


@pytest.mark.parametrize("graph", GRAPHS)
@pytest.mark.parametrize("quant,refine", CASES)
def test_oracle_refine_quantizers(data, graph, quant, refine):
    ds, g = data
    X, Q, metric = ds[graph]
    o = O.Searcher(O.Graph(os.path.join(GOLD, f"{graph}.glass")), X, metric, quant, refine)
    _check(o, f"{graph}/{quant}/{refine}", Q, g)


@pytest.mark.gpu
@pytest.mark.parametrize("graph", GRAPHS)
@pytest.mark.parametrize("quant,refine", CASES)
def test_hip_refine_quantizers(data, graph, quant, refine):
    import pyglass_amd as pg

    if pg.device_count() < 1:
        pytest.skip("no HIP device")
    ds, g = data
    X, Q, metric = ds[graph]
    print(f"refine {graph} {quant}+{refine}", flush=True)
    s = pg.Searcher(pg.Graph(os.path.join(GOLD, f"{graph}.glass")), X, metric, quant, refine)
    assert s.info()["rerank_row_bytes"] == (128 if refine == "SQ8U" else 128)
    _check(s, f"{graph}/{quant}/{refine}", Q, g)
