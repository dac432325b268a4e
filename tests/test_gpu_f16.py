"""FP16 / BF16 quantizers and the FP16 re-rank: HIP search vs the oracle
through the C ABI, on the golden graphs, and HIP ids vs the compiled
reference's stored ids (tests/golden/search_f16.npz).

Bar: ids, distances and dist_cmps bit-exact (both kernels widen the 16-bit
codes exactly and run the fp32 kernel's accumulation order).
"""

import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu


def _datasets():
    # the 100-d fixture was written with its input followed by zeros in
    # memory (tests/golden/make_golden_f16.py): the reference's fp16 encoder
    # then pads the last row with zeros, as the device encoder does
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    X100 = synth.gaussian(3000, 100, synth.BASE_SEED)
    Q100 = synth.gaussian(200, 100, synth.QUERY_SEED)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP"), "g100_hnsw_l2": (X100, Q100, "L2")}


@pytest.fixture(scope="module")
def dsets():
    return _datasets()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.mark.parametrize("gname", ["c1_hnsw_l2", "g100_hnsw_l2"])
@pytest.mark.parametrize("metric", ["L2", "IP"])
@pytest.mark.parametrize("quant", ["FP16", "BF16"])
def test_device_codes_match_oracle(dsets, golden_dir, quant, metric, gname):
    # 100-d rows: the fp16 padding holds the next row's first values
    X, _, _ = dsets[gname]
    print(f"codes {quant} {metric} {gname}", flush=True)
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), X, metric, quant)
    oq = O.Quant(quant, metric, X.shape[1]).fit(X)
    np.testing.assert_array_equal(s.codes(), oq.codes)


CASES = [(g, q, "") for g in ["c1_hnsw_l2", "ip_hnsw", "g100_hnsw_l2"] for q in ["FP16", "BF16"]]
CASES += [(g, "SQ4U", "FP16") for g in ["c1_hnsw_l2", "ip_hnsw", "g100_hnsw_l2"]]
CASES += [("c1_hnsw_l2", "SQ8U", "FP16"), ("g100_hnsw_l2", "SQ8", "FP16")]


@pytest.mark.parametrize("gname,quant,refine", CASES)
def test_search_matches_oracle(dsets, golden_dir, gname, quant, refine):
    XX, QQ, metric = dsets[gname]
    path = os.path.join(golden_dir, f"{gname}.glass")
    s = pg.Searcher(pg.Graph(path), XX, metric, quant, refine)
    o = O.Searcher(O.Graph(path), XX, metric, quant, refine)
    nq = 200
    for ef in [16, 40, 96, 200]:
        for k in [1, 10, 50]:
            print(f"{gname} {quant} {refine or 'none'} ef={ef} k={k}", flush=True)
            s.set_ef(ef)
            o.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            Io, Do = o.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, Io)
            np.testing.assert_array_equal(D, Do)
            assert s.get_last_search_dist_cmps() == o.get_last_search_dist_cmps()
            for i in range(0, nq, 50):
                a, b = s.search(QQ[i], k)
                c, d = o.search(QQ[i], k)
                np.testing.assert_array_equal(a, c)
                np.testing.assert_array_equal(b, d)


@pytest.mark.parametrize("gname,quant,refine", [("c1_hnsw_l2", "FP16", "none"), ("ip_hnsw", "BF16", "none"),
                                                ("g100_hnsw_l2", "FP16", "none"), ("c1_hnsw_l2", "SQ4U", "FP16")])
def test_reference_ids_directly(dsets, golden_dir, gname, quant, refine):
    """HIP ids equal the compiled reference's stored ids (not only the oracle's)."""
    z = np.load(os.path.join(golden_dir, "search_f16.npz"))
    XX, QQ, metric = dsets[gname]
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), XX, metric, quant,
                    "" if refine == "none" else refine)
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/{quant}/{refine}/{ef}/{k}"
            print(key, flush=True)
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:64], k)
            np.testing.assert_array_equal(I, z[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, z[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(z[f"{key}/batch_dist_cmps"]), key
