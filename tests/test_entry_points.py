"""Entry-point lists that repeat ids or hold more than 64 of them
(Graph::initialize_search, glass/graph.hpp:95-104: every entry point is
inserted in order -- a repeated one twice -- and marked visited), with and
without the FP32 Refiner, against the compiled reference's vectors
(tests/golden/make_golden_eps.py): the oracle on the CPU, the HIP path
through the C ABI on the GPU."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from pyglass_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = [("SQ8U", ""), ("SQ8U", "FP32"), ("SQ4U", "FP32")]
GRAPHS = ["eps_dup", "eps_many"]


@pytest.fixture(scope="module")
def data():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)[:5000].copy()
    Q = synth.gaussian(32, 128, synth.QUERY_SEED + 3)
    return X, Q, np.load(os.path.join(GOLD, "search_eps.npz"))


def _check(searcher, name, quant, refine, Q, g, single=True):
    for ef in [16, 64]:
        searcher.set_ef(ef)
        for k in [1, 10]:
            key = f"{name}/{quant}/{refine or 'none'}/{ef}/{k}"
            ids, dis = searcher.batch_search(Q, k)
            np.testing.assert_array_equal(ids, g[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(dis, g[f"{key}/batch_dists"], err_msg=key)
            assert searcher.get_last_search_dist_cmps() == float(g[f"{key}/batch_dist_cmps"]), key
            if single:
                for i in range(len(Q)):
                    si, sd = searcher.search(Q[i], k)
                    np.testing.assert_array_equal(si, g[f"{key}/single_ids"][i], err_msg=f"{key} single {i}")
                    np.testing.assert_array_equal(sd, g[f"{key}/single_dists"][i], err_msg=f"{key} single {i}")


def test_golden_graphs_have_the_entry_lists(data):
    import pyglass_amd as pg

    _, eps = pg.Graph(os.path.join(GOLD, "eps_dup.glass")).arrays()
    assert len(eps) == 5 and eps[0] == eps[1] == eps[3]
    _, eps = pg.Graph(os.path.join(GOLD, "eps_many.glass")).arrays()
    assert len(eps) == 100 and len(set(eps.tolist())) == 90


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("quant,refine", CASES)
def test_oracle_entry_points(data, name, quant, refine):
    X, Q, g = data
    o = O.Searcher(O.Graph(os.path.join(GOLD, f"{name}.glass")), X, "L2", quant, refine)
    _check(o, name, quant, refine, Q, g)


@pytest.mark.gpu
@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("quant,refine", CASES)
def test_hip_entry_points(data, name, quant, refine):
    import pyglass_amd as pg

    if pg.device_count() < 1:
        pytest.skip("no HIP device")
    X, Q, g = data
    print(f"entry points {name} {quant}+{refine or 'none'}", flush=True)
    s = pg.Searcher(pg.Graph(os.path.join(GOLD, f"{name}.glass")), X, "L2", quant, refine)
    _check(s, name, quant, refine, Q, g)
