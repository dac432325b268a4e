"""The other workload shapes as GPU parity cases, at reduced N (graphs built
here by the CPU builders), ids and distances bit-exact against the oracle:
  * 960-d Gamma rows, NSG R=64, SQ8U / L2, per-query ef (log-uniform 16-512)
    and k (1-100) with SearchBatch semantics per query;
  * 768-d unit rows, NSG R=48, IP with SQ8 (per-dim float) and SQ8U (integer)
    codes + FP32 re-rank, ef in {32, 64, 128, 256}, k = 10;
  * 128-d 4096-center mixture, HNSW R=32 L=100, SQ2U + FP32 re-rank at the
    reference's factor 3.0 (reorder_k = 30) and at 4.0 (reorder_k = 40).
"""

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


def _pair(tmp_path, name, X, kind, R, L, metric, quant, refine):
    path = str(tmp_path / f"{name}.glass")
    pg.build_graph(kind, X, path, metric, "FP32", R, L)
    return pg.Searcher(pg.Graph(path), X, metric, quant, refine), O.Searcher(O.Graph(path), X, metric, quant, refine)


def test_960d_nsg_sq8u_per_query_ef_k(tmp_path):
    X = synth.gamma_rows(2000, 960, synth.BASE_SEED)
    Q = synth.gamma_rows(60, 960, synth.QUERY_SEED)
    print("960-d NSG SQ8U per-query ef/k", flush=True)
    s, o = _pair(tmp_path, "g960", X, "NSG", 64, 96, "L2", "SQ8U", "")
    ef, k = synth.per_query_ef_k(len(Q))
    ids, dis, cnt = s.batch_search_efk(Q, ef, k, with_counters=True)
    for i in range(len(Q)):
        o.set_ef(int(ef[i]))
        Io, Do = o.batch_search(Q[i:i + 1], int(k[i]))
        np.testing.assert_array_equal(ids[i], Io[0], err_msg=f"query {i} ef={ef[i]} k={k[i]}")
        np.testing.assert_array_equal(dis[i], Do[0], err_msg=f"query {i}")
        assert cnt[i, 0] == o.last_counts[0][0]


@pytest.mark.parametrize("quant", ["SQ8", "SQ8U"])
def test_768d_nsg_ip_rerank(tmp_path, quant):
    X = synth.unit_rows(2000, 768, synth.BASE_SEED)
    Q = synth.unit_rows(100, 768, synth.QUERY_SEED)
    print(f"768-d NSG IP {quant}+FP32", flush=True)
    s, o = _pair(tmp_path, "g768", X, "NSG", 48, 80, "IP", quant, "FP32")
    for ef in [32, 64, 128, 256]:
        s.set_ef(ef)
        o.set_ef(ef)
        I, D = s.batch_search(Q, 10)
        Io, Do = o.batch_search(Q, 10)
        np.testing.assert_array_equal(I, Io)
        np.testing.assert_array_equal(D, Do)
        assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()


@pytest.mark.parametrize("factor", [None, 4.0])
def test_128d_hnsw_sq2u_rerank(tmp_path, factor):
    X = synth.normal_mixture(20000, 128, synth.BASE_SEED)
    Q = synth.normal_mixture(200, 128, synth.QUERY_SEED)
    print(f"128-d HNSW SQ2U+FP32 factor={factor}", flush=True)
    s, o = _pair(tmp_path, "g128", X, "HNSW", 32, 100, "L2", "SQ2U", "FP32")
    if factor:
        s.set_refine_factor(factor)
        o = O.Searcher(O.Graph(str(tmp_path / "g128.glass")), X, "L2", "SQ2U", "FP32", refine_factor=factor)
    assert s.info()["refine_factor"] == pytest.approx(factor or 3.0)
    s.set_ef(128)
    o.set_ef(128)
    I, D = s.batch_search(Q, 10)
    Io, Do = o.batch_search(Q, 10)
    np.testing.assert_array_equal(I, Io)
    np.testing.assert_array_equal(D, Do)
