"""PQ8 (ProductQuant, 4 dims x 256 centroids per sub-space): the searcher's
codebook (trained on the host as glass::Clustering does) and its device codes
against the oracle, HIP search vs the oracle through the C ABI, HIP results
vs the compiled reference's stored ones (tests/golden/search_pq8.npz), and the
served path, whose batches bring their own per-slot lookup tables.

Bar: codebook, codes, ids, distances and dist_cmps bit-exact (the LUT is
built and summed in the reference's float order).
"""

import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu


def _datasets():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    X100 = synth.gaussian(3000, 100, synth.BASE_SEED)
    Q100 = synth.gaussian(200, 100, synth.QUERY_SEED)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP"), "nsg_l2": (X[:5000].copy(), Q, "L2"),
            "g100_hnsw_l2": (X100, Q100, "L2")}


@pytest.fixture(scope="module")
def dsets():
    return _datasets()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("gname", ["c1_hnsw_l2", "g100_hnsw_l2"])
@pytest.mark.parametrize("metric", ["L2", "IP"])
def test_codebook_and_codes_match_oracle(dsets, golden_dir, metric, gname):
    # nsq = 32 (one 32-block of the LUT sum) and 25 (a 16-block and a tail of 9)
    X, _, _ = dsets[gname]
    print(f"codes PQ8 {metric} {gname}", flush=True)
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), X, metric, "PQ8")
    oq = O.Quant("PQ8", metric, X.shape[1]).fit(X)
    np.testing.assert_array_equal(s.calibration().reshape(-1, 256, 4), oq.codebook())
    np.testing.assert_array_equal(s.codes(), oq.codes)


CASES = [(g, "") for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"]] + [("c1_hnsw_l2", "FP32"),
                                                                                    ("g100_hnsw_l2", "FP16"),
                                                                                    ("ip_hnsw", "FP32"),
                                                                                    ("nsg_l2", "FP16")]


@pytest.mark.parametrize("gname,refine", CASES)
def test_search_matches_oracle(dsets, golden_dir, gname, refine):
    XX, QQ, metric = dsets[gname]
    path = os.path.join(golden_dir, f"{gname}.glass")
    s = pg.Searcher(pg.Graph(path), XX, metric, "PQ8", refine)
    o = O.Searcher(O.Graph(path), XX, metric, "PQ8", refine)
    nq = 200
    for ef in [16, 40, 96, 200]:
        for k in [1, 10, 50]:
            print(f"{gname} PQ8 {refine or 'none'} ef={ef} k={k}", flush=True)
            s.set_ef(ef)
            o.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            Io, Do = o.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, Io)
            np.testing.assert_array_equal(D, Do)
            assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
            for i in range(0, nq, 25):
                a, b = s.search(QQ[i], k)
                c, d = o.search(QQ[i], k)
                np.testing.assert_array_equal(a, c)
                np.testing.assert_array_equal(b, d)


@pytest.mark.parametrize("gname,refine", [(g, "none") for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"]]
                         + [("c1_hnsw_l2", "FP32")])
def test_reference_ids_directly(dsets, golden_dir, gname, refine):
    """HIP results equal the compiled reference's stored ones (not only the oracle's)."""
    z = np.load(os.path.join(golden_dir, "search_pq8.npz"))
    XX, QQ, metric = dsets[gname]
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), XX, metric, "PQ8",
                    "" if refine == "none" else refine)
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/PQ8/{refine}/{ef}/{k}"
            print(key, flush=True)
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:64], k)
            np.testing.assert_array_equal(I, z[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, z[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(z[f"{key}/batch_dist_cmps"]), key
            for i in range(0, 64, 16):
                a, b = s.search(QQ[i], k)
                np.testing.assert_array_equal(a, z[f"{key}/single_ids"][i], err_msg=key)
                np.testing.assert_array_equal(b, z[f"{key}/single_dists"][i], err_msg=key)


def test_served_equals_batch(dsets, golden_dir):
    """Server batches use their own slots (and so their own lookup tables)."""
    XX, QQ, metric = dsets["c1_hnsw_l2"]
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, "c1_hnsw_l2.glass")), XX, metric, "PQ8")
    srv = pg.Server(s, max_batch=512, max_k=64, max_inflight=2)
    try:
        tickets = []
        for t in range(8):
            rows = np.arange(t * 60, t * 60 + 60)
            ef, k = [(16, 1), (40, 10), (96, 50), (64, 10)][t % 4]
            tickets.append((srv.submit(QQ[rows], k=k, ef=ef), rows, ef, k))
        for tk, rows, ef, k in tickets:
            print(f"served rows {rows[0]} ef={ef} k={k}", flush=True)
            ids, dis = srv.wait(tk)
            bi, bd = s.batch_search_efk(QQ[rows], ef, k)
            np.testing.assert_array_equal(ids, bi)
            np.testing.assert_array_equal(dis, bd)
    finally:
        srv.close()
