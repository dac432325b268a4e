"""CPU graph builders (glass.Index counterparts): structural invariants of
the produced graphs and search quality through the oracle searcher.  The
reference builders are nondeterministic (OpenMP, shared RNG), so graphs are
checked by properties, not against reference files:
  * every row holds at most R distinct ids in [0, N), -1 padded, no self loop;
  * NSG / SSG: every node is reachable from the entry (tree_grow,
    glass/nsg/nsg.hpp:314-409);
  * HNSW: the entry has the top level, upper lists only name nodes of at
    least that level and hold at most R/2 ids (glass/hnsw/hnsw.hpp:42-65);
  * exact-kNN recall@10 of an oracle FP32 search at ef=256 is high (the
    32-d mixture has a high intrinsic dimension: at ef=64 HNSW reaches ~0.85
    and NSG ~0.72 on it).
"""

import os

import numpy as np
import pytest

import pyglass_amd as pg
from oracle import oracle as O


def mixture(n, d, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3, 3, (16, d)).astype(np.float32)
    return (c[rng.integers(0, 16, n)] + rng.standard_normal((n, d)).astype(np.float32)).astype(np.float32)


def check_rows(adj, R, n):
    assert adj.shape == (n, R)
    for i, row in enumerate(adj):
        ids = row[row >= 0]
        assert np.all(row[len(ids):] == -1), f"row {i} not -1 padded"
        assert len(set(ids.tolist())) == len(ids), f"row {i} has duplicates"
        assert np.all(ids < n) and i not in ids


def reachable(adj, ep):
    seen = np.zeros(len(adj), bool)
    seen[ep] = True
    stack = [ep]
    while stack:
        u = stack.pop()
        for v in adj[u]:
            if v < 0:
                break
            if not seen[v]:
                seen[v] = True
                stack.append(v)
    return seen


def recall_at_10(path, X, Q):
    s = O.Searcher(O.Graph(path), X, "L2", "FP32")
    s.set_ef(256)
    ids, _ = s.batch_search(Q, 10)
    d = ((Q[:, None, :].astype(np.float64) - X[None].astype(np.float64)) ** 2).sum(-1)
    gt = np.argsort(d, axis=1)[:, :10]
    return np.mean([len(np.intersect1d(a, b)) / 10 for a, b in zip(ids, gt)])


@pytest.mark.parametrize("kind", ["NSG", "SSG"])
def test_nsg_graph_invariants_and_recall(tmp_path, kind):
    X = mixture(4000, 32, 41)
    Q = mixture(100, 32, 42)
    g = pg.Index(kind, "L2", "FP32", 24, 64).build(X)
    info = g.info()
    assert info["N"] == 4000 and info["K"] == 24 and info["nep"] == 1 and not info["has_upper"]
    adj, eps = g.arrays()
    check_rows(adj, 24, 4000)
    assert reachable(adj, eps[0]).all()
    path = str(tmp_path / "g.glass")
    g.save(path)
    assert recall_at_10(path, X, Q) >= 0.9


def test_hnsw_graph_invariants_and_recall(tmp_path):
    X = mixture(4000, 32, 43)
    Q = mixture(100, 32, 44)
    path = str(tmp_path / "h.glass")
    pg.build_graph("HNSW", X, path, "L2", "FP32", 24, 100)
    g = pg.Graph(path)
    adj, eps = g.arrays()
    check_rows(adj, 24, 4000)
    # upper layers straight from the file (HNSWInitializer.hpp:64-86)
    raw = np.fromfile(path, np.int32)
    nep = raw[0]
    off = 1 + nep + 2 + 4000 * 24
    n, kup, ep = raw[off:off + 3]
    assert (n, kup) == (4000, 12)
    off += 3
    levels = np.zeros(n, np.int64)
    lists = []
    for i in range(n):
        cnt = raw[off]
        levels[i] = cnt // kup
        lists.append(raw[off + 1:off + 1 + cnt].reshape(-1, kup) if cnt else np.zeros((0, kup), np.int32))
        off += 1 + cnt
    assert off == len(raw)
    assert levels[ep] == levels.max() and eps[0] == ep
    for i in range(n):
        for lv in range(1, levels[i] + 1):
            row = lists[i][lv - 1]
            ids = row[row >= 0]
            assert np.all(row[len(ids):] == -1)
            assert np.all(levels[ids] >= lv)
    assert recall_at_10(path, X, Q) >= 0.95


def test_builder_rejects_bad_arguments():
    X = mixture(100, 8, 45)
    with pytest.raises(pg.GlassError):
        pg.Index("KNN", "L2").build(X)
    with pytest.raises(pg.GlassError):
        pg.Index("HNSW", "COSINE").build(X)
    with pytest.raises(pg.GlassError):
        pg.Index("NSG", "L2", "FP32", 160, 64).build(X)  # degree above the 128 the searcher supports
