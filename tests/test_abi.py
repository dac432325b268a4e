"""The C-ABI library loads and exports every symbol include/glass_hip.h
declares; host-only entry points (graph I/O, argument checking) work without
a GPU."""

import os
import re

import numpy as np
import pytest

from pyglass_amd import _lib
import pyglass_amd as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "glass_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(glass_hip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_version_and_device_count():
    assert "gfx950" in pg.version()
    assert pg.device_count() >= 0


def test_graph_roundtrip(tmp_path, golden_dir):
    src = os.path.join(golden_dir, "c1_hnsw_l2.glass")
    g = pg.Graph(src)
    inf = g.info()
    assert inf == {"N": 10000, "K": 32, "nep": 1, "has_upper": True}
    out = tmp_path / "copy.glass"
    g.save(str(out))
    assert open(src, "rb").read() == open(out, "rb").read()


def test_graph_from_arrays_and_nsg_format(tmp_path):
    rng = np.random.default_rng(0)
    n, k = 50, 8
    adj = rng.integers(0, n, size=(n, k)).astype(np.int32)
    adj[:, 6:] = -1
    g = pg.Graph.from_arrays(adj, [3])
    a2, e2 = g.arrays()
    np.testing.assert_array_equal(a2, adj)
    np.testing.assert_array_equal(e2, [3])
    # nsg format: K, ep, then (cnt, ids) per node (glass/graph.hpp:220-245)
    p = tmp_path / "g.nsg"
    with open(p, "wb") as f:
        np.array([k, 3], np.int32).tofile(f)
        for i in range(n):
            row = adj[i][adj[i] >= 0]
            np.array([len(row)], np.int32).tofile(f)
            row.astype(np.int32).tofile(f)
    g2 = pg.Graph(str(p), "nsg")
    a3, e3 = g2.arrays()
    np.testing.assert_array_equal(a3, adj)
    np.testing.assert_array_equal(e3, [3])


def _write_hnswlib(path, adj, ep, dim=4, max_level=0, delete_mark_row=None):
    """hnswlib's saveIndex layout as glass/graph.hpp:184-218 reads it."""
    n, m0 = adj.shape
    spe = 4 * (1 + m0) + 4 * dim + 8
    with open(path, "wb") as f:
        np.array([0, n, n, spe, 4 * (1 + m0) + 4 * dim, 4 * (1 + m0)], np.uint64).tofile(f)
        np.array([max_level, ep], np.int32).tofile(f)
        np.array([m0 // 2, m0, m0 // 2], np.uint64).tofile(f)
        np.array([1.0 / np.log(m0 // 2)], np.float64).tofile(f)
        np.array([100], np.uint64).tofile(f)
        for i in range(n):
            row = adj[i][adj[i] >= 0]
            hdr = len(row) | ((1 << 16) if i == delete_mark_row else 0)
            np.array([hdr], np.uint32).tofile(f)
            np.concatenate([row, np.zeros(m0 - len(row), np.int32)]).astype(np.int32).tofile(f)
            np.zeros(dim, np.float32).tofile(f)
            np.array([i], np.uint64).tofile(f)
        np.zeros(n, np.int32).tofile(f)  # per-element upper-level list sizes


def test_hnswlib_format(tmp_path):
    rng = np.random.default_rng(1)
    n, m0 = 40, 6
    adj = np.full((n, m0), -1, np.int32)
    for i in range(n):
        c = int(rng.integers(1, m0 + 1))
        adj[i, :c] = rng.choice([j for j in range(n) if j != i], c, replace=False)
    p = tmp_path / "g.hnswlib"
    _write_hnswlib(p, adj, 7, delete_mark_row=5)
    a, e = pg.Graph(str(p), "hnswlib").arrays()
    np.testing.assert_array_equal(a, adj)
    np.testing.assert_array_equal(e, [7])
    _write_hnswlib(p, adj, 7, max_level=2)
    with pytest.raises(pg.GlassError):
        pg.Graph(str(p), "hnswlib")


def test_errors_are_raised_not_crashes(tmp_path):
    with pytest.raises(pg.GlassError):
        pg.Graph(str(tmp_path / "missing.glass"))
    bad = tmp_path / "bad.glass"
    bad.write_bytes(b"\x01\x00")
    with pytest.raises(pg.GlassError):
        pg.Graph(str(bad))
    adj = np.full((4, 2), -1, np.int32)
    adj[0, 0] = 7  # out of range
    g = pg.Graph.from_arrays(adj, [0])
    with pytest.raises(pg.GlassError, match="outside"):
        pg.Searcher(g, np.zeros((4, 8), np.float32), "L2", "SQ8U")
    g = pg.Graph.from_arrays(np.full((4, 2), -1, np.int32), [0])
    with pytest.raises(pg.GlassError, match="metric"):
        pg.Searcher(g, np.zeros((4, 8), np.float32), "COS", "SQ8U")
    g = pg.Graph.from_arrays(np.full((4, 2), -1, np.int32), [0])
    with pytest.raises(pg.GlassError, match="quantizer"):
        pg.Searcher(g, np.zeros((4, 8), np.float32), "L2", "SQ8P")


def test_clustering_argument_errors_without_a_gpu():
    # checked before any device work
    with pytest.raises(pg.GlassError, match="init"):
        pg.Clustering(8, 10, "other")
    c = pg.Clustering(8, 10, "random")
    with pytest.raises(pg.GlassError, match="not fitted"):
        c.predict(np.zeros((4, 3), np.float32))
    with pytest.raises(pg.GlassError, match="too few"):
        c.fit(np.zeros((5, 3), np.float32))  # random seeding draws k of the n - 1 sampled rows
    big = pg.Clustering(100001, 1, "random")
    with pytest.raises(pg.GlassError, match="too few"):
        big.fit(np.zeros((200000, 1), np.float32))  # the sample is capped at 100,000 rows
    with pytest.raises(pg.GlassError, match="dim"):
        pg.quant_sq8(0)


def test_host_allocation_failure_returns_a_status():
    # (2^31 - 1) x 2^20 centroid floats (8 PiB) cannot be allocated on the host.
    # The entry point catches std::bad_alloc and returns GLASS_HIP_ENOMEM; an
    # exception escaping through extern "C" would end the process instead.
    c = pg.Clustering(2**31 - 1, 1, "kmeans++")
    with pytest.raises(pg.GlassError, match=r"error -3: host allocation failed"):
        c.fit(np.zeros((2, 1 << 20), np.float32))

def test_diskann_format_matches_reference(golden_dir):
    """Graph(path, "diskann") reads a ragged diskann file as the reference's
    Graph::load_diskann does (glass/graph.hpp:155-182): its .glass save of the
    same file (tests/golden/make_golden_formats.py) has the same adjacency,
    -1 padding and entry point."""
    g = pg.Graph(os.path.join(golden_dir, "diskann_1k.bin"), "diskann")
    r = pg.Graph(os.path.join(golden_dir, "diskann_1k_ref.glass"))
    assert g.info() == r.info() == {"N": 1000, "K": 24, "nep": 1, "has_upper": False}
    (ga, ge), (ra, re_) = g.arrays(), r.arrays()
    np.testing.assert_array_equal(ga, ra)
    np.testing.assert_array_equal(ge, re_)
    assert ge[0] == 417 and (ga == -1).any() and (ga[:, 0] >= 0).any()
