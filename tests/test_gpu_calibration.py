"""Device calibration (the calibrators' order statistics found by a radix
select on the GPU) against the oracle's sort-based order statistics, through
the C ABI, on inputs chosen for the selection's edge cases: many ties and
signed zeros, negative and subnormal values, heavy tails, and inputs so small
that the 10% sample holds only a couple of items.

Bar: calibration parameters and codes bit-exact.
"""

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if pg.device_count() < 1:
        pytest.skip("no HIP device")


def _data(kind: str, n: int, d: int) -> np.ndarray:
    rng = np.random.default_rng(7)
    if kind == "ties":  # integers in [-3, 3]: long runs of equal keys, +0 and -0
        X = rng.integers(-3, 4, (n, d)).astype(np.float32)
        X[X == 0] = np.where(rng.random((X == 0).sum()) < 0.5, -0.0, 0.0)
        return X
    if kind == "heavy":  # Cauchy tails: order statistics far from the bulk
        return rng.standard_cauchy((n, d)).astype(np.float32)
    if kind == "subnormal":
        return (rng.standard_normal((n, d)) * 1e-39).astype(np.float32)
    return (rng.standard_normal((n, d)) * 3 + 1).astype(np.float32)


def _graph(n: int) -> pg.Graph:
    adj = np.full((n, 4), -1, np.int32)
    for i in range(n):
        adj[i, : min(4, n - 1)] = [(i + j + 1) % n for j in range(min(4, n - 1))]
    return pg.Graph.from_arrays(adj, eps=[0])


def _params(quant, metric, cal, oq):
    ref = oq.params()
    if quant in ("SQ8", "SQ6", "SQ4"):
        return cal, ref[: 2 * oq.dim_align]
    if quant == "SQ8U" and metric == "IP":
        return cal, ref[:1]
    return cal, ref[:2]


CASES = [(q, m) for q in ["SQ8U", "SQ4U", "SQ2U", "SQ8", "SQ6", "SQ4"] for m in ["L2", "IP"]]


@pytest.mark.parametrize("kind", ["ties", "heavy", "subnormal", "shifted"])
@pytest.mark.parametrize("quant,metric", CASES)
def test_device_calibration_matches_oracle(quant, metric, kind):
    n, d = 1500, 40
    X = _data(kind, n, d)
    print(f"calibration {quant} {metric} {kind}", flush=True)
    s = pg.Searcher(_graph(n), X, metric, quant)
    oq = O.Quant(quant, metric, d).fit(X)
    got, want = _params(quant, metric, s.calibration(), oq)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(s.codes(), oq.codes)


@pytest.mark.parametrize("quant", ["SQ4U", "SQ2U", "SQ6", "SQ4"])
@pytest.mark.parametrize("n,d", [(3, 8), (12, 16), (101, 3)])
def test_device_calibration_tiny_inputs(quant, n, d):
    # SQ4U/SQ2U sample int(n*d*0.1f) items: 2, 19 and 30 here
    X = _data("shifted", n, d)
    print(f"tiny {quant} n={n} d={d}", flush=True)
    s = pg.Searcher(_graph(n), X, "L2", quant)
    oq = O.Quant(quant, "L2", d).fit(X)
    got, want = _params(quant, "L2", s.calibration(), oq)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(s.codes(), oq.codes)
