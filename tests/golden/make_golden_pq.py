"""Writes the PQ8 golden vectors under tests/golden/ from the compiled
reference (glass/quant/product_quant.hpp trained by glass::Clustering,
glass/algorithms/clustering.hpp).

Run in the build container only (it needs oracle/_ref/libglassref.so, built
from /root/reference by `make -C oracle ref`):

    python tests/golden/make_golden_pq.py

Outputs (inputs are regenerated from seeds; their SHA-256 is stored):
  quant_pq8.npz   per (L2|IP, dim 128|100): digest of the codebook and its
                  first sub-space, digest of the codes and the first 8 rows,
                  digest of 4 query LUTs and the first LUT's first sub-space,
                  16 x 97 query->base distances of a bare computer
  search_pq8.npz  per (graph, refine, ef, k): Search and SearchBatch ids /
                  distances over 64 queries and the batch dist-cmp average
"""

from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refwrap as R  # noqa: E402
from pyglass_amd import synth  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def datasets():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    X100 = synth.gaussian(3000, 100, synth.BASE_SEED)
    Q100 = synth.gaussian(200, 100, synth.QUERY_SEED)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP"), "nsg_l2": (X[:5000].copy(), Q, "L2"),
            "g100_hnsw_l2": (X100, Q100, "L2")}


def main():
    R.set_num_threads(8)
    ds = datasets()

    out_q = {}
    for gname, d in [("c1_hnsw_l2", 128), ("g100_hnsw_l2", 100)]:
        X, Q, _ = ds[gname]
        n = X.shape[0]
        ids = np.arange(3, n - 1, (n - 4) // 97 + 1, dtype=np.int32)[:97]
        out_q[f"{d}/ids"] = ids
        out_q[f"{d}/data_sha"] = np.array(sha(X))
        out_q[f"{d}/query_sha"] = np.array(sha(Q))
        for metric in ["L2", "IP"]:
            cb, codes, lut, dist = R.pq_probe(metric, X, Q[:16], ids)
            key = f"PQ8_{metric}_{d}"
            out_q[f"{key}/codebook_sha"] = np.array(sha(cb))
            out_q[f"{key}/codebook_head"] = cb[0].copy()
            out_q[f"{key}/codes_sha"] = np.array(sha(codes))
            out_q[f"{key}/codes_head"] = codes[:8].copy()
            out_q[f"{key}/lut_sha"] = np.array(sha(lut[:4]))
            out_q[f"{key}/lut_head"] = lut[0, 0].copy()
            out_q[f"{key}/dist"] = dist
    np.savez_compressed(os.path.join(HERE, "quant_pq8.npz"), **out_q)

    out_s = {}
    nq = 64
    cases = [(g, "") for g in ds] + [("c1_hnsw_l2", "FP32")]
    for gname, refine in cases:
        XX, QQ, metric = ds[gname]
        out_s[f"{gname}/data_sha"] = np.array(sha(XX))
        s = R.RefSearcher(os.path.join(HERE, f"{gname}.glass"), XX, metric, "PQ8", refine)
        for ef in [16, 40, 96]:
            for k in [1, 10, 50]:
                s.set_ef(ef)
                key = f"{gname}/PQ8/{refine or 'none'}/{ef}/{k}"
                I, D = s.batch_search(QQ[:nq], k)
                out_s[f"{key}/batch_ids"] = I
                out_s[f"{key}/batch_dists"] = D
                out_s[f"{key}/batch_dist_cmps"] = np.float64(s.get_last_search_dist_cmps())
                si, sd = zip(*[s.search(QQ[i], k) for i in range(nq)])
                out_s[f"{key}/single_ids"] = np.stack(si)
                out_s[f"{key}/single_dists"] = np.stack(sd)
        del s
        print("PQ8", gname, refine or "none", flush=True)
    np.savez_compressed(os.path.join(HERE, "search_pq8.npz"), **out_s)


if __name__ == "__main__":
    main()
