"""The oracle reproduces the compiled reference's golden vectors.

Golden vectors (tests/golden/*.npz, *.glass) were written by
tests/golden/make_golden.py through the reference headers compiled with the
reference's own flags.  This pins calibration, encoders, every distance
kernel, LinearPool/expansion semantics, the HNSW initializer and the refiner
of oracle/glass_oracle.c.
"""

import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from pyglass_amd import synth

QUANTS = ["SQ8U", "SQ4U", "SQ2U", "E5M2", "SQ8", "FP32"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def qz(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "quant.npz")))


@pytest.fixture(scope="module")
def sz(golden_dir):
    return np.load(os.path.join(golden_dir, "search.npz"))


@pytest.fixture(scope="module")
def c1():
    return synth.gaussian(10000, 128, synth.BASE_SEED), synth.gaussian(1000, 128, synth.QUERY_SEED)


def test_inputs_regenerate_identically(qz, c1):
    X, Q = c1
    assert sha(X) == str(qz["data_sha"])
    assert sha(Q) == str(qz["query_sha"])


@pytest.mark.parametrize("metric", ["L2", "IP"])
@pytest.mark.parametrize("quant", QUANTS)
def test_quantizer_matches_reference(qz, c1, quant, metric):
    X, Q = c1
    key = f"{quant}_{metric}"
    oq = O.Quant(quant, metric, X.shape[1]).fit(X)
    assert oq.code_size == int(qz[f"{key}/code_size"])
    assert oq.dim_align == int(qz[f"{key}/dim_align"])
    ref_params = qz[f"{key}/params"]
    np.testing.assert_array_equal(oq.params(), ref_params)
    np.testing.assert_array_equal(oq.codes[:8], qz[f"{key}/codes_head"])
    assert sha(oq.codes) == str(qz[f"{key}/codes_sha"])
    ids = qz["ids"]
    got = np.stack([oq.distances(Q[i], ids) for i in range(16)])
    np.testing.assert_array_equal(got, qz[f"{key}/dist"])


def _datasets():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP"), "nsg_l2": (X[:5000].copy(), Q, "L2")}


CASES = []
for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2"]:
    for q in QUANTS:
        for r in ["none"] + (["FP32"] if q in ("SQ8U", "SQ4U", "SQ2U", "SQ8") else []):
            CASES.append((g, q, r))


@pytest.fixture(scope="module")
def dsets():
    return _datasets()


@pytest.mark.parametrize("gname,quant,refine", CASES)
def test_search_matches_reference(sz, dsets, golden_dir, gname, quant, refine):
    XX, QQ, metric = dsets[gname]
    assert sha(XX) == str(sz[f"{gname}/data_sha"])
    g = O.Graph(os.path.join(golden_dir, f"{gname}.glass"))
    s = O.Searcher(g, XX, metric, quant, "" if refine == "none" else refine)
    nq = 64
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/{quant}/{refine}/{ef}/{k}"
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, sz[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, sz[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(sz[f"{key}/batch_dist_cmps"]), key
            si = np.stack([s.search(QQ[i], k)[0] for i in range(nq)])
            sd = np.stack([s.search(QQ[i], k)[1] for i in range(nq)])
            np.testing.assert_array_equal(si, sz[f"{key}/single_ids"], err_msg=key)
            np.testing.assert_array_equal(sd, sz[f"{key}/single_dists"], err_msg=key)


# FP16 / BF16 (tests/golden/make_golden_f16.py)

@pytest.fixture(scope="module")
def qz16(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "quant_f16.npz")))


@pytest.fixture(scope="module")
def sz16(golden_dir):
    return np.load(os.path.join(golden_dir, "search_f16.npz"))


def _datasets_f16():
    d = _datasets()
    X100 = synth.gaussian(3000, 100, synth.BASE_SEED)
    Q100 = synth.gaussian(200, 100, synth.QUERY_SEED)
    return {"c1_hnsw_l2": d["c1_hnsw_l2"], "ip_hnsw": d["ip_hnsw"], "g100_hnsw_l2": (X100, Q100, "L2")}


@pytest.fixture(scope="module")
def dsets16():
    return _datasets_f16()


@pytest.mark.parametrize("d", [128, 100])
@pytest.mark.parametrize("metric", ["L2", "IP"])
@pytest.mark.parametrize("quant", ["FP16", "BF16"])
def test_f16_quantizer_matches_reference(qz16, dsets16, quant, metric, d):
    # d = 100: the fp16 padding of every row holds the floats that follow it
    # in memory (the next row's first values), as the reference's block
    # converter leaves it; past the last row the fixture's input is zeros
    X, Q, _ = dsets16["c1_hnsw_l2" if d == 128 else "g100_hnsw_l2"]
    assert sha(X) == str(qz16[f"{d}/data_sha"])
    key = f"{quant}_{metric}_{d}"
    oq = O.Quant(quant, metric, d).fit(X)
    assert oq.code_size == int(qz16[f"{key}/code_size"])
    assert oq.dim_align == int(qz16[f"{key}/dim_align"])
    np.testing.assert_array_equal(oq.codes[:8], qz16[f"{key}/codes_head"])
    assert sha(oq.codes) == str(qz16[f"{key}/codes_sha"])
    ids = qz16[f"{d}/ids"]
    got = np.stack([oq.distances(Q[i], ids) for i in range(16)])
    np.testing.assert_array_equal(got, qz16[f"{key}/dist"])


def test_f16_padding_holds_the_next_row():
    # the layout rule itself, on numbers small enough to read: dim 3, padded
    # to 16; row i's padding j (3 <= j < 16) is fp16(data[i*3 + j]) while
    # that lies inside the input, else 0; bf16 padding is always 0
    X = np.arange(1, 1 + 6 * 3, dtype=np.float32).reshape(6, 3)
    q16 = O.Quant("FP16", "L2", 3).fit(X)
    h = q16.codes.view(np.uint16)
    flat = X.reshape(-1)
    for i in range(6):
        for j in range(16):
            at = i * 3 + j
            want = O.f32_to_f16(float(flat[at])) if at < flat.size else 0
            assert h[i, j] == want, (i, j)
    qb = O.Quant("BF16", "L2", 3).fit(X)
    hb = qb.codes.view(np.uint16)
    assert (hb[:, 3:] == 0).all()
    np.testing.assert_array_equal(hb[:, :3], (X.view(np.uint32) + 0x8000) >> 16)


CASES16 = [(g, q, "none") for g in ["c1_hnsw_l2", "ip_hnsw", "g100_hnsw_l2"] for q in ["FP16", "BF16"]]
CASES16 += [(g, "SQ4U", "FP16") for g in ["c1_hnsw_l2", "ip_hnsw"]]


@pytest.mark.parametrize("gname,quant,refine", CASES16)
def test_f16_search_matches_reference(sz16, dsets16, golden_dir, gname, quant, refine):
    XX, QQ, metric = dsets16[gname]
    assert sha(XX) == str(sz16[f"{gname}/data_sha"])
    g = O.Graph(os.path.join(golden_dir, f"{gname}.glass"))
    s = O.Searcher(g, XX, metric, quant, "" if refine == "none" else refine)
    nq = 64
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/{quant}/{refine}/{ef}/{k}"
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, sz16[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, sz16[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(sz16[f"{key}/batch_dist_cmps"]), key
            si = np.stack([s.search(QQ[i], k)[0] for i in range(nq)])
            sd = np.stack([s.search(QQ[i], k)[1] for i in range(nq)])
            np.testing.assert_array_equal(si, sz16[f"{key}/single_ids"], err_msg=key)
            np.testing.assert_array_equal(sd, sz16[f"{key}/single_dists"], err_msg=key)


# Per-dim scalar quantizers SQ6 and SQ4 (tests/golden/make_golden_perdim.py)

PERDIM = ["SQ6", "SQ4"]


@pytest.fixture(scope="module")
def qzp(golden_dir):
    return {q: dict(np.load(os.path.join(golden_dir, f"quant_{q.lower()}.npz"))) for q in PERDIM}


@pytest.fixture(scope="module")
def szp(golden_dir):
    return {q: np.load(os.path.join(golden_dir, f"search_{q.lower()}.npz")) for q in PERDIM}


def _unpack(quant: str, codes: np.ndarray, da: int) -> np.ndarray:
    """Components of packed rows.  SQ6: 4 per 3 bytes, six bits each from the
    least significant end; SQ4: dim j in the low (even j) or high nibble of
    byte j/2."""
    if quant == "SQ4":
        b = codes.astype(np.uint32)
        c = np.stack([b & 15, b >> 4], axis=-1)
    else:
        b = codes.astype(np.uint32).reshape(codes.shape[0], -1, 3)
        w = b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16)
        c = np.stack([(w >> (6 * s)) & 63 for s in range(4)], axis=-1)
    return c.reshape(codes.shape[0], -1)[:, :da]


@pytest.mark.parametrize("quant", PERDIM)
@pytest.mark.parametrize("d", [128, 100])
@pytest.mark.parametrize("metric", ["L2", "IP"])
def test_perdim_quantizer_matches_reference(qzp, dsets16, quant, metric, d):
    z = qzp[quant]
    X, Q, _ = dsets16["c1_hnsw_l2" if d == 128 else "g100_hnsw_l2"]
    assert sha(X) == str(z[f"{d}/data_sha"])
    key = f"{quant}_{metric}_{d}"
    oq = O.Quant(quant, metric, d).fit(X)
    assert oq.code_size == int(z[f"{key}/code_size"])
    assert oq.dim_align == int(z[f"{key}/dim_align"])
    np.testing.assert_array_equal(oq.params(), z[f"{key}/params"])
    np.testing.assert_array_equal(oq.codes[:8], z[f"{key}/codes_head"])
    assert sha(oq.codes) == str(z[f"{key}/codes_sha"])
    ids = z[f"{d}/ids"]
    got = np.stack([oq.distances(Q[i], ids) for i in range(16)])
    np.testing.assert_array_equal(got, z[f"{key}/dist"])


@pytest.mark.parametrize("quant,levels", [("SQ6", 63), ("SQ4", 15)])
def test_perdim_codes_follow_the_quantile_rule(dsets16, quant, levels):
    """Calibration and packing restated in numpy: chunk j of the flattened
    input (items [j*n, (j+1)*n)) gives min = its top-th smallest value and
    max its top-th largest, top = int(n * 0.01f) + 1; code =
    round(clamp((x - min) / (max - min), 0, 1) * levels), half away from zero."""
    X, _, _ = dsets16["g100_hnsw_l2"]
    n, d = X.shape
    oq = O.Quant(quant, "L2", d).fit(X)
    flat = X.reshape(-1)
    top = int(np.float32(n) * np.float32(0.01)) + 1
    mins = np.zeros(d, np.float32)
    difs = np.zeros(d, np.float32)
    for j in range(d):
        s = np.sort(flat[j * n:(j + 1) * n])
        mins[j], difs[j] = s[top - 1], np.float32(s[n - top] - s[top - 1])
    da = oq.dim_align
    np.testing.assert_array_equal(oq.params()[:d], mins)
    np.testing.assert_array_equal(oq.params()[da:da + d], difs)
    v = np.clip((X - mins) / difs, np.float32(0), np.float32(1)) * np.float32(levels)
    want = np.floor(v).astype(np.int64) + (v - np.floor(v) >= 0.5)
    got = _unpack(quant, oq.codes, da)
    np.testing.assert_array_equal(got[:, :d], want)
    assert (got[:, d:] == 0).all()


@pytest.mark.parametrize("quant", PERDIM)
@pytest.mark.parametrize("metric", ["L2", "IP"])
def test_perdim_distance_is_the_decoded_closed_form(dsets16, quant, metric):
    """float32 kernel vs the float64 sum over decoded values.  SQ6 decodes
    y = (c + 0.5) / 63 * dif + min.  SQ4 decodes with (c + 0.5) / 16 although
    it encodes with 15 levels, and its L2 is that of 16x against 16y (a factor
    256).  A wrong component order, shift or decode offset moves the value far
    outside the tolerance."""
    X, Q, _ = dsets16["g100_hnsw_l2"]
    d = X.shape[1]
    oq = O.Quant(quant, metric, d).fit(X)
    da = oq.dim_align
    p = oq.params().astype(np.float64)
    ids = np.arange(0, X.shape[0], 37)
    div, l2s = (63.0, 1.0) if quant == "SQ6" else (16.0, 256.0)
    y = (_unpack(quant, oq.codes[ids], da) + 0.5) / div * p[da:2 * da] + p[:da]
    for qi in range(8):
        x = np.zeros(da)
        x[:d] = Q[qi]
        want = l2s * ((x - y) ** 2).sum(1) if metric == "L2" else -(x * y).sum(1)
        got = oq.distances(Q[qi], ids)
        np.testing.assert_allclose(got, want, rtol=2e-5, atol=2e-4 * l2s)


def test_sq6_contexts_round_differently(dsets16):
    """Search's expansions and SearchBatch evaluate SQ6 with different
    roundings; the golden Search and SearchBatch distances below tell them
    apart only if the two forms disagree somewhere."""
    X, Q, _ = dsets16["c1_hnsw_l2"]
    oq = O.Quant("SQ6", "L2", 128).fit(X)
    ids = np.arange(0, 2000, 3)
    a = np.concatenate([oq.distances(Q[i], ids, "search") for i in range(4)])
    b = np.concatenate([oq.distances(Q[i], ids, "batch") for i in range(4)])
    assert (a != b).sum() > 0
    np.testing.assert_allclose(a, b, rtol=1e-5)


CASESP = [(q, g, "none") for q in PERDIM for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"]]
CASESP += [(q, "c1_hnsw_l2", "FP32") for q in PERDIM]


@pytest.fixture(scope="module")
def dsetsp(dsets, dsets16):
    return {**dsets, "g100_hnsw_l2": dsets16["g100_hnsw_l2"]}


@pytest.mark.parametrize("quant,gname,refine", CASESP)
def test_perdim_search_matches_reference(szp, dsetsp, golden_dir, quant, gname, refine):
    """Ids, distances and dist-cmp counts of Search and SearchBatch; an FP32
    re-rank is a no-op for SQ6 and SQ4 (get_refine_factor 1.0)."""
    z = szp[quant]
    XX, QQ, metric = dsetsp[gname]
    assert sha(XX) == str(z[f"{gname}/data_sha"])
    g = O.Graph(os.path.join(golden_dir, f"{gname}.glass"))
    s = O.Searcher(g, XX, metric, quant, "" if refine == "none" else refine)
    nq = 64
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/{quant}/{refine}/{ef}/{k}"
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, z[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, z[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(z[f"{key}/batch_dist_cmps"]), key
            si = np.stack([s.search(QQ[i], k)[0] for i in range(nq)])
            sd = np.stack([s.search(QQ[i], k)[1] for i in range(nq)])
            np.testing.assert_array_equal(si, z[f"{key}/single_ids"], err_msg=key)
            np.testing.assert_array_equal(sd, z[f"{key}/single_dists"], err_msg=key)


# PQ8 (tests/golden/make_golden_pq.py)

@pytest.fixture(scope="module")
def qpq(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "quant_pq8.npz")))


@pytest.fixture(scope="module")
def spq(golden_dir):
    return np.load(os.path.join(golden_dir, "search_pq8.npz"))


def test_mt19937_known_answer():
    """std::mt19937 default-seeded: first output 3499211612, and the 10000th
    4123659995 (the value the C++ standard requires, [rand.predef])."""
    out = O.mt19937_outputs(10000)
    assert out[0] == 3499211612
    assert out[-1] == 4123659995


def test_kmeans_fit4_invariants():
    """Lloyd fixed point on separable data: 4 well separated blobs, k = 4;
    every centroid ends at the (float) mean of its blob, each blob found
    once, whatever the seeding picked."""
    rng = np.random.default_rng(3)
    centers = np.array([[0, 0, 0, 0], [10, 0, 0, 0], [0, 10, 0, 0], [0, 0, 10, 10]], np.float32)
    X = np.concatenate([c + rng.standard_normal((50, 4)).astype(np.float32) * 0.1 for c in centers])
    cb = O.kmeans_fit4(X, k=4, epochs=10)
    # the fit samples n - 1 rows: compare against the blob means within tolerance
    order = [int(np.argmin(((cb - c) ** 2).sum(1))) for c in centers]
    assert sorted(order) == [0, 1, 2, 3]
    for c, j in zip(centers, order):
        np.testing.assert_allclose(cb[j], c, atol=0.1)


def test_pq_lut_sum_orders():
    """The LUT sum against the float64 sum (tolerance) on all block shapes
    (nsq = 8: tail only; 25: 16-block + tail; 32: one 32-block; 57: 32 + 16 +
    9), and against the serial float sum, which differs from the blocked one
    for some rows (so the blocked order is what the golden data pins)."""
    rng = np.random.default_rng(5)
    differs = 0
    for nsq in [8, 25, 32, 57]:
        lut = rng.standard_normal((nsq, 256)).astype(np.float32) * 100
        for _ in range(50):
            code = rng.integers(0, 256, nsq).astype(np.uint8)
            vals = lut[np.arange(nsq), code]
            got = O.pq_lut_sum(lut, code)
            np.testing.assert_allclose(got, vals.astype(np.float64).sum(), rtol=1e-5, atol=1e-3)
            serial = np.float32(0)
            for v in vals:
                serial = np.float32(serial + v)
            differs += got != serial
    assert differs > 0


@pytest.mark.parametrize("d", [128, 100])
@pytest.mark.parametrize("metric", ["L2", "IP"])
def test_pq8_quantizer_matches_reference(qpq, dsets16, metric, d):
    X, Q, _ = dsets16["c1_hnsw_l2" if d == 128 else "g100_hnsw_l2"]
    assert sha(X) == str(qpq[f"{d}/data_sha"])
    key = f"PQ8_{metric}_{d}"
    oq = O.Quant("PQ8", metric, d).fit(X)
    cb = oq.codebook()
    np.testing.assert_array_equal(cb[0], qpq[f"{key}/codebook_head"])
    assert sha(cb) == str(qpq[f"{key}/codebook_sha"])
    np.testing.assert_array_equal(oq.codes[:8], qpq[f"{key}/codes_head"])
    assert sha(oq.codes) == str(qpq[f"{key}/codes_sha"])
    lut = np.stack([oq.lut(Q[i]) for i in range(4)])
    np.testing.assert_array_equal(lut[0, 0], qpq[f"{key}/lut_head"])
    assert sha(lut) == str(qpq[f"{key}/lut_sha"])
    ids = qpq[f"{d}/ids"]
    got = np.stack([oq.distances(Q[i], ids) for i in range(16)])
    np.testing.assert_array_equal(got, qpq[f"{key}/dist"])


def test_pq8_lut_is_the_subvector_distance(dsets16):
    """LUT entry (sq, k) vs the float64 L2 / -dot of the query's sub-vector sq
    and centroid k; codes are the nearest centroid of each sub-vector."""
    X, Q, _ = dsets16["g100_hnsw_l2"]
    for metric in ["L2", "IP"]:
        oq = O.Quant("PQ8", metric, 100).fit(X)
        cb = oq.codebook().astype(np.float64)
        x = Q[0].astype(np.float64).reshape(25, 1, 4)
        want = ((x - cb) ** 2).sum(-1) if metric == "L2" else -(x * cb).sum(-1)
        np.testing.assert_allclose(oq.lut(Q[0]), want, rtol=1e-5, atol=1e-5)
    xr = X[:50].astype(np.float64).reshape(50, 25, 1, 4)
    near = ((xr - cb[None]) ** 2).sum(-1).argmin(-1)
    assert (near == oq.codes[:50]).mean() > 0.99


CASESPQ = [(g, "none") for g in ["c1_hnsw_l2", "ip_hnsw", "nsg_l2", "g100_hnsw_l2"]] + [("c1_hnsw_l2", "FP32")]


@pytest.mark.parametrize("gname,refine", CASESPQ)
def test_pq8_search_matches_reference(spq, dsetsp, golden_dir, gname, refine):
    """Ids, distances and dist-cmp counts of Search and SearchBatch; PQ8 re-ranks
    int(1.5 k) candidates with FP32 rows (get_refine_factor)."""
    XX, QQ, metric = dsetsp[gname]
    assert sha(XX) == str(spq[f"{gname}/data_sha"])
    g = O.Graph(os.path.join(golden_dir, f"{gname}.glass"))
    s = O.Searcher(g, XX, metric, "PQ8", "" if refine == "none" else refine)
    nq = 64
    for ef in [16, 40, 96]:
        for k in [1, 10, 50]:
            key = f"{gname}/PQ8/{refine}/{ef}/{k}"
            s.set_ef(ef)
            I, D = s.batch_search(QQ[:nq], k)
            np.testing.assert_array_equal(I, spq[f"{key}/batch_ids"], err_msg=key)
            np.testing.assert_array_equal(D, spq[f"{key}/batch_dists"], err_msg=key)
            assert s.get_last_search_dist_cmps() == float(spq[f"{key}/batch_dist_cmps"]), key
            si = np.stack([s.search(QQ[i], k)[0] for i in range(nq)])
            sd = np.stack([s.search(QQ[i], k)[1] for i in range(nq)])
            np.testing.assert_array_equal(si, spq[f"{key}/single_ids"], err_msg=key)
            np.testing.assert_array_equal(sd, spq[f"{key}/single_dists"], err_msg=key)


# Clustering (tests/golden/make_golden_clustering.py)

CL_DIMS = [4, 7, 12, 16, 30, 100, 128]


@pytest.fixture(scope="module")
def zcl(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "clustering.npz")))


@pytest.mark.parametrize("init", ["random", "kmeans++"])
@pytest.mark.parametrize("d", CL_DIMS)
def test_clustering_matches_reference(zcl, d, init):
    X = synth.gaussian(1500, d, synth.BASE_SEED)
    P = synth.gaussian(200, d, synth.QUERY_SEED)
    assert sha(X) == str(zcl[f"{d}/data_sha"])
    c = O.Clustering(64, 10, init).fit(X)
    key = f"{d}/{init}"
    np.testing.assert_array_equal(c.centroids, zcl[f"{key}/centroids"])
    np.testing.assert_array_equal(c.predict(P), zcl[f"{key}/predict"])
    t = c.transform(P)
    np.testing.assert_array_equal(t[:4], zcl[f"{key}/transform_head"])
    assert sha(t) == str(zcl[f"{key}/transform_sha"])


@pytest.mark.parametrize("d", [7, 30, 128])
def test_clustering_is_lloyd_fixed_point_like(d):
    """After the epochs each centroid is near the float64 mean of the sample
    rows nearest to it (one more Lloyd step moves it little), and predict
    agrees with a float64 brute-force nearest-centroid search."""
    X = synth.gaussian(1500, d, synth.BASE_SEED)
    c = O.Clustering(16, 20, "kmeans++").fit(X)
    ids = c.predict(X)
    d64 = ((X.astype(np.float64)[:, None, :] - c.centroids.astype(np.float64)[None]) ** 2).sum(-1)
    assert (ids == d64.argmin(1)).mean() > 0.999
    for j in range(16):
        rows = X[ids == j].astype(np.float64)
        if len(rows) > 20:
            assert np.abs(rows.mean(0) - c.centroids[j]).max() < 0.2
