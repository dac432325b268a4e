"""glass.Clustering and the quant_* trainers through the C ABI: k-means with
its distance sweeps on the GPU against the compiled reference's stored
centroids, predictions and transforms (tests/golden/clustering.npz) and the
oracle; trainer parameters against the oracle's calibration.

Bar: centroids, predictions, transformed rows and parameters bit-exact.
"""

import hashlib
import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O
from pyglass_amd import synth

pytestmark = pytest.mark.gpu

DIMS = [4, 7, 12, 16, 30, 100, 128]


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


@pytest.fixture(scope="module")
def zcl(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "clustering.npz")))


@pytest.mark.parametrize("init", ["random", "kmeans++"])
@pytest.mark.parametrize("d", DIMS)
def test_clustering_matches_reference(zcl, d, init):
    print(f"clustering d={d} {init}", flush=True)
    X = synth.gaussian(1500, d, synth.BASE_SEED)
    P = synth.gaussian(200, d, synth.QUERY_SEED)
    c = pg.Clustering(64, 10, init)
    c.fit(X)
    key = f"{d}/{init}"
    np.testing.assert_array_equal(c.centroids(), zcl[f"{key}/centroids"])
    np.testing.assert_array_equal(c.predict(P), zcl[f"{key}/predict"])
    t = P.copy()
    assert c.transform(t, inplace=True) is None
    np.testing.assert_array_equal(t[:4], zcl[f"{key}/transform_head"])
    assert hashlib.sha256(t.tobytes()).hexdigest() == str(zcl[f"{key}/transform_sha"])
    assert c.transform(P) is None  # the binding returns None and leaves data alone without inplace


def test_clustering_larger_than_one_sample_cap():
    """More rows than the 100,000-row sample: the fit samples, predict covers
    every row; equal to the oracle."""
    X = synth.gaussian(120000, 24, synth.BASE_SEED)
    print("clustering 120000 x 24 k=256", flush=True)
    c = pg.Clustering(256, 3, "kmeans++")
    c.fit(X)
    o = O.Clustering(256, 3, "kmeans++").fit(X)
    np.testing.assert_array_equal(c.centroids(), o.centroids)
    np.testing.assert_array_equal(c.predict(X[:5000]), o.predict(X[:5000]))


def test_clustering_errors():
    with pytest.raises(pg.GlassError):
        pg.Clustering(8, 10, "other")
    c = pg.Clustering(8, 10, "random")
    with pytest.raises(pg.GlassError):
        c.predict(np.zeros((4, 3), np.float32))  # not fitted
    with pytest.raises(pg.GlassError):
        c.fit(np.zeros((5, 3), np.float32))  # random seeding needs k of the n - 1 sampled rows
    c.fit(synth.gaussian(100, 3, synth.BASE_SEED))
    with pytest.raises(pg.GlassError):
        c.predict(np.zeros((4, 5), np.float32))  # wrong dim


@pytest.mark.parametrize("name,quant", [("quant_sq8", "SQ8"), ("quant_sq8u", "SQ8U"), ("quant_fp32", "FP32"),
                                        ("quant_fp16", "FP16"), ("quant_bf16", "BF16"), ("quant_e5m2", "E5M2")])
def test_quant_trainers(name, quant):
    X = synth.gaussian(3000, 100, synth.BASE_SEED)
    print(f"trainer {name}", flush=True)
    t = getattr(pg, name)(100)
    t.train(X)
    oq = O.Quant(quant, "L2", 100).fit(X)
    ref = oq.params()
    if quant == "SQ8":
        np.testing.assert_array_equal(t.params(), ref[: 2 * oq.dim_align])
    elif quant == "SQ8U":
        np.testing.assert_array_equal(t.params(), ref[:2])
    else:
        assert t.params().size == 0
