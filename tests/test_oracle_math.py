"""Pins the oracle by what the mathematics fixes, independently of the
golden vectors: closed-form integer distances, the code layouts and the
calibrators' order statistics restated in numpy, a model of LinearPool,
brute force on small graphs (ef >= N), a hand-worked HNSW descent and
fp16 rounding.  Each test is built so that a plausible slip in the oracle
(a swapped nibble, lower_bound for upper_bound, a non-strict comparison, a
different summation order) fails it.
"""

from __future__ import annotations

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from oracle import oracle as O
from oracle import toy_graphs as T

MUL = {"SQ8U": 127, "SQ4U": 15, "SQ2U": 3}
BITS = {"SQ8U": 8, "SQ4U": 4, "SQ2U": 2}


def unpack(codes: np.ndarray, kind: str, metric: str, dim_align: int) -> np.ndarray:
    """Code layouts of sq8u_quant.hpp:38-42, sq4u_quant.hpp:40-50 and
    sq2u_quant.hpp:38-52: dim j sits in bits [b*(j % (8/b)), ...) of byte
    j // (8/b); SQ8U codes are int8 (signed for the symmetric IP calibrator)."""
    codes = np.asarray(codes, np.uint8)
    if kind == "SQ8U":
        v = codes.view(np.int8).astype(np.int64)
        return v[..., :dim_align]
    b = BITS[kind]
    per = 8 // b
    shifts = np.arange(per) * b
    fields = (codes[..., :, None].astype(np.int64) >> shifts) & ((1 << b) - 1)
    return fields.reshape(codes.shape[:-1] + (-1,))[..., :dim_align]


def mixture(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3, 3, (8, d)).astype(np.float32)
    return (c[rng.integers(0, 8, n)] + rng.standard_normal((n, d)).astype(np.float32)).astype(np.float32)


# ---------------------------------------------------------------- calibrators


def order_stat_minmax(flat: np.ndarray, ratio: float):
    """find_minmax (quant/utils.hpp:31-40): a MaxHeap of top = int64(n*ratio)+1
    smallest values leaves the top-th smallest on top; the same for -x gives
    the top-th largest.  The count is computed in float as in the source."""
    n = flat.size
    top = int(np.float32(n) * np.float32(ratio)) + 1
    lo = np.partition(flat, top - 1)[top - 1]
    hi = np.partition(flat, n - top)[n - top]
    return np.float32(lo), np.float32(hi)


@pytest.mark.parametrize("kind,ratio,train_ratio", [("SQ8U", 0.0, None), ("SQ4U", 0.01, 0.1), ("SQ2U", 0.05, 0.1)])
def test_affine_calibration_is_the_order_statistic(kind, ratio, train_ratio):
    X = mixture(3000, 37, 11)
    q = O.Quant(kind, "L2", 37)
    q.train(X)
    flat = X.ravel()
    if train_ratio is not None:
        # (int64_t)n * dim * train_ratio evaluated in float, truncated
        nitems = int(np.float32(flat.size) * np.float32(train_ratio))
        flat = flat[:nitems]
    lo, hi = order_stat_minmax(flat, ratio)
    p = q.params()
    assert p[0] == lo
    assert p[1] == np.float32(hi - lo)


def test_symmetric_calibration_is_absmax():
    X = mixture(2000, 40, 12)
    q = O.Quant("SQ8U", "IP", 40)
    q.train(X)
    assert q.params()[0] == np.float32(np.abs(X).max())


def test_sq8_per_dim_calibration_bins_flat_items_by_i_over_n():
    # find_minmax_perdim (quant/utils.hpp:52-70) sends flat item i to bin i/n,
    # i.e. bin b covers items [b*n, (b+1)*n) of the row-major array, not column
    # b; padded bins stay 0.
    n, d = 500, 20
    X = mixture(n, d, 13)
    q = O.Quant("SQ8", "L2", d)
    q.train(X)
    da = q.dim_align
    p = q.params()
    bins = X.ravel().reshape(d, n)
    np.testing.assert_array_equal(p[:d], bins.min(axis=1))
    np.testing.assert_array_equal(p[da:da + d], (bins.max(axis=1) - bins.min(axis=1)).astype(np.float32))
    assert not np.any(p[d:da]) and not np.any(p[da + d:2 * da])
    # the per-column reading gives different parameters on this data
    assert not np.array_equal(p[:d], X.min(axis=0))


# ---------------------------------------------------------------- encoders


@pytest.mark.parametrize("kind,metric", [("SQ8U", "L2"), ("SQ8U", "IP"), ("SQ4U", "L2"), ("SQ2U", "L2")])
def test_codes_follow_the_rounding_formula_and_layout(kind, metric):
    # code = round_half_away(clamp((x - min) / dif, 0, 1) * mul)
    # (calibrator.hpp:28-32), symmetric: round(clamp(x / max, -1, 1) * mul).
    d = 45
    X = mixture(400, d, 14)
    q = O.Quant(kind, metric, d).fit(X)
    p = q.params()
    if metric == "IP" and kind == "SQ8U":
        v = np.clip(X.astype(np.float64) / np.float64(p[0]), -1, 1) * MUL[kind]
    else:
        v = np.clip((X.astype(np.float64) - np.float64(p[0])) / np.float64(p[1]), 0, 1) * MUL[kind]
    want = np.sign(v) * np.floor(np.abs(v) + 0.5)
    got = unpack(q.codes, kind, metric, q.dim_align)
    near_half = np.abs(np.abs(v) - np.floor(np.abs(v)) - 0.5) < 1e-4
    assert np.all((got[:, :d] == want) | near_half)
    assert near_half.mean() < 0.01
    # padding dims are 0 (zeroed storage, memory.hpp:27-29)
    assert not np.any(got[:, d:])


def test_e5m2_is_the_high_byte_of_fp16():
    X = mixture(50, 24, 15) * 37.0
    q = O.Quant("E5M2", "L2", 24).fit(X)
    want = (X.astype(np.float16).view(np.uint16) >> 8).astype(np.uint8)
    np.testing.assert_array_equal(q.codes[:, :24], want)


def test_fp16_conversion_rounds_to_nearest_even():
    rng = np.random.default_rng(16)
    xs = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(100.0),
        rng.standard_normal(1000).astype(np.float32) * np.float32(1e-5),   # fp16 subnormals
        np.array([65504.0, 65519.0, 65520.0, 1e6, -1e6, 0.0, -0.0, 2.0**-24, 2.0**-25, 3 * 2.0**-26], np.float32),
        # exact ties between two fp16 values: 1 + (2k+1) * 2^-11
        (1.0 + (2 * np.arange(64) + 1) * 2.0**-11).astype(np.float32),
    ])
    got = np.array([O.f32_to_f16(x) for x in xs], np.uint16)
    with np.errstate(over="ignore"):
        np.testing.assert_array_equal(got, xs.astype(np.float16).view(np.uint16))
    back = np.array([O.f16_to_f32(h) for h in got[:200]], np.float32)
    np.testing.assert_array_equal(back, got[:200].view(np.float16).astype(np.float32))


# ---------------------------------------------------------------- distances


@pytest.mark.parametrize("kind,metric", [("SQ8U", "L2"), ("SQ8U", "IP"), ("SQ4U", "L2"), ("SQ4U", "IP"),
                                         ("SQ2U", "L2"), ("SQ2U", "IP")])
def test_integer_distances_are_the_closed_form(kind, metric):
    # helpa's scalar references (ref/l2_ref.hpp, ref/dot_ref.hpp): L2 sum (x-y)^2
    # over dim_align; SQ8U IP -sum x*y; SQ4U/SQ2U use the L2 kernel for both
    # metrics (sq4u_quant.hpp:64, graph_searcher.hpp:357).
    d = 77
    X = mixture(300, d, 17)
    Q = mixture(6, d, 18)
    q = O.Quant(kind, metric, d).fit(X)
    base = unpack(q.codes, kind, metric, q.dim_align)
    for x in Q:
        qc = unpack(q.prepare_query(x)[None, :q.code_size], kind, metric, q.dim_align)[0]
        if kind == "SQ8U" and metric == "IP":
            want = -(base * qc).sum(axis=1)
        else:
            want = ((base - qc) ** 2).sum(axis=1)
        got = q.distances(x, range(len(X)))
        np.testing.assert_array_equal(got, want.astype(np.float64))


def test_fp32_distance_summation_order_matters_and_oracle_is_exact_on_integers():
    # integer-valued rows: every order gives the exact sum
    rng = np.random.default_rng(19)
    X = rng.integers(-20, 20, (50, 33)).astype(np.float32)
    x = rng.integers(-20, 20, 33).astype(np.float32)
    q = O.Quant("FP32", "L2", 33).fit(X)
    np.testing.assert_array_equal(q.distances(x, range(50)), ((X - x).astype(np.float64) ** 2).sum(axis=1))
    # generic floats: a plain sequential float32 sum differs from the oracle's
    # 16-lane FMA + tree on some rows, so the golden comparison of the float
    # distances is sensitive to the order (a planted reorder would fail it)
    Xf = mixture(400, 128, 20)
    xf = mixture(1, 128, 21)[0]
    qf = O.Quant("FP32", "L2", 128).fit(Xf)
    orc = qf.distances(xf, range(400)).astype(np.float32)
    seq = np.zeros(400, np.float32)
    for j in range(128):
        t = (xf[j] - Xf[:, j]).astype(np.float32)
        seq = (seq + t * t).astype(np.float32)
    assert np.any(orc != seq)
    np.testing.assert_allclose(orc, seq, rtol=1e-5)


# ---------------------------------------------------------------- LinearPool


class PoolModel:
    """LinearPool restated as a set: the pool is the `cap` smallest admitted
    entries by (dist, insertion sequence); pop takes the first unchecked one
    while its index is below ef (neighbor.hpp:320-433 read as a set rule)."""

    def __init__(self, cap, ef):
        self.cap, self.ef = cap, ef
        self.items = []  # (dist, seq, id)
        self.checked = set()
        self.seq = 0

    def insert(self, u, d):
        self.items.append((d, self.seq, u))
        self.seq += 1
        self.items.sort()
        del self.items[self.cap:]

    def cur(self):
        for i, (_, _, u) in enumerate(self.items):
            if u not in self.checked:
                return i
        return len(self.items)

    def pop(self):
        c = self.cur()
        if c < len(self.items) and c < self.ef:
            u = self.items[c][2]
            self.checked.add(u)
            return u
        return -1


@settings(max_examples=400, deadline=None)
@given(st.integers(1, 12), st.data())
def test_linear_pool_matches_set_model(cap, data):
    ef = data.draw(st.integers(1, cap))
    n = data.draw(st.integers(0, 60))
    kinds = data.draw(st.lists(st.booleans(), min_size=n, max_size=n))
    dvals = data.draw(st.lists(st.integers(0, 6), min_size=n, max_size=n))
    ops, dists, nid = [], [], 0
    for is_pop, dv in zip(kinds, dvals):
        if is_pop:
            ops.append(-1)
            dists.append(0.0)
        else:
            ops.append(nid)
            dists.append(float(dv))
            nid += 1
    popped, ids, dd, cur = O.pool_run(cap, ef, ops, dists)
    m = PoolModel(cap, ef)
    for i, (o, d) in enumerate(zip(ops, dists)):
        if o >= 0:
            m.insert(o, d)
            assert popped[i] == -2
        else:
            assert popped[i] == m.pop(), f"op {i}"
    assert [int(u) & 0x7FFFFFFF for u in ids] == [u for _, _, u in m.items]
    assert list(dd) == [d for d, _, _ in m.items]
    assert [bool(int(u) & 0x80000000) for u in ids] == [u in m.checked for _, _, u in m.items]
    assert cur == m.cur()


def test_linear_pool_ties_go_after_equal_keys():
    # upper_bound insertion (neighbor.hpp:361-371): a later equal key lands after
    popped, ids, dd, _ = O.pool_run(4, 4, [0, 1, 2, 3], [1.0, 1.0, 0.5, 1.0])
    assert list(ids) == [2, 0, 1, 3]


# ---------------------------------------------------------------- search


def ring_graph(n, k, seed):
    """Random out-edges plus a ring (i -> i+1) so every node is reachable."""
    rng = np.random.default_rng(seed)
    adj = np.full((n, k), -1, np.int32)
    for i in range(n):
        nb = [(i + 1) % n] + list(rng.choice(n, k - 1, replace=False))
        nb = [v for v in dict.fromkeys(nb) if v != i][:k]
        adj[i, :len(nb)] = nb
    return adj


@pytest.mark.parametrize("kind,metric", [("SQ8U", "L2"), ("SQ4U", "L2"), ("SQ2U", "IP"), ("E5M2", "IP"),
                                         ("SQ8", "L2"), ("FP32", "L2")])
@pytest.mark.parametrize("batch", [False, True])
def test_search_with_ef_at_least_n_is_exact_knn(kind, metric, batch):
    # ef >= N: nothing is ever dropped, every reachable node is evaluated and
    # the result is the k smallest (dist, insertion order) pairs.
    n, d, k = 240, 32, 10
    X = mixture(n, d, 22)
    Q = mixture(8, d, 23)
    g = O.Graph(adj=ring_graph(n, 8, 24), eps=[5])
    s = O.Searcher(g, X, metric, kind)
    s.set_ef(n)
    for x in Q:
        if batch:
            ids, dis = s.batch_search(x[None], k)
            ids, dis = ids[0], dis[0]
            assert s.get_last_search_dist_cmps() == n
        else:
            ids, dis = s.search(x, k)
        full = s.q.distances(x, range(n), ctx="batch" if batch else "search")
        want = np.sort(full)[:k].astype(np.float32)
        np.testing.assert_array_equal(dis, want)
        for j in range(k):
            if np.sum(full == full[ids[j]]) == 1:
                assert ids[j] == np.argsort(full, kind="stable")[j]
            assert full[ids[j]] == np.float64(dis[j]) or np.float32(full[ids[j]]) == dis[j]


def test_refined_search_with_ef_at_least_n_reranks_the_quantized_top():
    n, d, k = 240, 32, 10
    X = mixture(n, d, 25)
    Q = mixture(6, d, 26)
    g = O.Graph(adj=ring_graph(n, 8, 27), eps=[0])
    s = O.Searcher(g, X, "L2", "SQ4U", "FP32")
    assert s.factor == np.float32(1.5)  # get_refine_factor(SQ4U), graph_searcher.hpp:227-238
    s.set_ef(n)
    rk = int(k * 1.5)
    for x in Q:
        ids, dis = s.batch_search(x[None], k)
        qd = s.q.distances(x, range(n), ctx="batch")
        order = np.argsort(qd, kind="stable")
        if qd[order[rk - 1]] == qd[order[rk]]:
            continue  # the reorder_k boundary is a tie; insertion order decides
        cand = order[:rk]
        exact = ((X[cand].astype(np.float64) - x) ** 2).sum(axis=1)
        want = cand[np.lexsort((cand, exact))][:k]
        np.testing.assert_array_equal(ids[0], want)
        np.testing.assert_allclose(dis[0], np.sort(exact)[:k], rtol=1e-5)


@pytest.mark.parametrize("batch", [False, True])
def test_hnsw_descent_on_hand_worked_toy_graph(tmp_path, batch):
    path = str(tmp_path / "toy.glass")
    X, q = T.toy_hnsw(path)
    s = O.Searcher(O.Graph(path), X, "L2", "FP32")
    s.set_ef(3)
    if batch:
        ids, dis = s.batch_search(q[None], 3)
        ids, dis = ids[0], dis[0]
        assert s.get_last_search_dist_cmps() == T.TOY_DIST_CMPS
    else:
        ids, dis = s.search(q, 3)
    assert list(ids) == T.TOY_RESULT_IDS
    assert list(dis) == T.TOY_RESULT_DISTS


def test_toy_graph_hand_derivation_matches_compiled_reference(tmp_path):
    from oracle import refwrap as R

    if not R.available():
        pytest.skip("compiled reference (oracle/_ref) not built here")
    path = str(tmp_path / "toy.glass")
    X, q = T.toy_hnsw(path)
    r = R.RefSearcher(path, X, "L2", "FP32", "")
    r.set_ef(3)
    ids, dis = r.batch_search(q[None], 3)
    assert list(ids[0]) == T.TOY_RESULT_IDS
    assert list(dis[0]) == T.TOY_RESULT_DISTS
    assert r.get_last_search_dist_cmps() == T.TOY_DIST_CMPS
    ids, dis = r.search(q, 3)
    assert list(ids) == T.TOY_RESULT_IDS
