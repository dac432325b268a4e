"""Per-dim scalar quantizers SQ6 and SQ4: device calibration and codes, HIP
search vs the oracle through the C ABI, and HIP ids vs the compiled
reference's stored ids (tests/golden/search_sq6.npz, search_sq4.npz).

Bar: calibration, codes, ids, distances and dist_cmps bit-exact.  The device
kernels follow the reference's summation orders (SQ6: 8 accumulators and the
per-context rounding of the decode; SQ4: one serial sum over element pairs),
so distances are equal, not merely close.
"""

import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu

PERDIM = ["SQ6", "SQ4"]


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


@pytest.mark.parametrize("quant", PERDIM)
@pytest.mark.parametrize("gname", ["c1_hnsw_l2", "g100_hnsw_l2"])
@pytest.mark.parametrize("metric", ["L2", "IP"])
def test_device_codes_match_oracle(dsets, golden_dir, metric, gname, quant):
    # SQ6: 128-d rows are 96 bytes (dword aligned), 100-d rows 78 bytes (2-byte aligned)
    X, _, _ = dsets[gname]
    print(f"codes {quant} {metric} {gname}", flush=True)
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), X, metric, quant)
    oq = O.Quant(quant, metric, X.shape[1]).fit(X)
    da = oq.dim_align
    np.testing.assert_array_equal(s.calibration()[:2 * da], oq.params()[:2 * da])
    np.testing.assert_array_equal(s.codes(), oq.codes)


CASES = [(q, g, "") for q in PERDIM for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"]]
CASES += [(q, "c1_hnsw_l2", "FP32") for q in PERDIM]


@pytest.mark.parametrize("quant,gname,refine", CASES)
def test_search_matches_oracle(dsets, golden_dir, quant, gname, refine):
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
            for i in range(0, nq, 25):
                a, b = s.search(QQ[i], k)
                c, d = o.search(QQ[i], k)
                np.testing.assert_array_equal(a, c)
                np.testing.assert_array_equal(b, d)


@pytest.mark.parametrize("quant", PERDIM)
@pytest.mark.parametrize("gname", ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"])
def test_reference_ids_directly(dsets, golden_dir, gname, quant):
    """HIP results equal the compiled reference's stored ones (not only the oracle's)."""
    z = np.load(os.path.join(golden_dir, f"search_{quant.lower()}.npz"))
    XX, QQ, metric = dsets[gname]
    s = pg.Searcher(pg.Graph(os.path.join(golden_dir, f"{gname}.glass")), XX, metric, quant)
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/{quant}/none/{ef}/{k}"
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
