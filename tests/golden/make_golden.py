"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in the build container only (it needs oracle/_ref/libglassref.so, built
from /root/reference by `make -C oracle ref`):

    python tests/golden/make_golden.py

Outputs (inputs are regenerated from seeds; their SHA-256 is stored so a
test can prove the regeneration is identical):
  c1_hnsw_l2.glass   HNSW R=32 L=50 over 10,000 x 128 N(0,1) (seed 1234), L2
  ip_hnsw.glass      HNSW R=32 L=50 over 5,000 unit-normalised rows, IP
  nsg_l2.glass       NSG R=32 over 5,000 x 128 N(0,1), L2 (fixed entry point)
  quant.npz          per (quantizer, metric): calibration, codes digest,
                     first 8 code rows, 16 x 97 query->base distances
  search.npz         per (graph, quantizer, refine, ef, k, Search|SearchBatch):
                     ids, distances and the batch dist-cmp average over 64 queries
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

QUANTS = ["SQ8U", "SQ4U", "SQ2U", "E5M2", "SQ8", "FP32"]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def datasets():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)
    Xn = X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)
    Qn = Q / np.linalg.norm(Q, axis=1, keepdims=True)
    return {
        "c1_hnsw_l2": (X, Q, "L2"),
        "ip_hnsw": (Xn.astype(np.float32), Qn.astype(np.float32), "IP"),
        "nsg_l2": (X[:5000].copy(), Q, "L2"),
    }


def main():
    R.set_num_threads(8)
    ds = datasets()
    R.build_graph("HNSW", ds["c1_hnsw_l2"][0], os.path.join(HERE, "c1_hnsw_l2.glass"), "L2", "FP32", 32, 50)
    R.build_graph("HNSW", ds["ip_hnsw"][0], os.path.join(HERE, "ip_hnsw.glass"), "IP", "FP32", 32, 50)
    R.build_graph("NSG", ds["nsg_l2"][0], os.path.join(HERE, "nsg_l2.glass"), "L2", "FP32", 32, 0)

    out_q = {}
    X, Q, _ = ds["c1_hnsw_l2"]
    ids = np.arange(3, 10000, 103, dtype=np.int32)  # 97 ids
    for metric in ["L2", "IP"]:
        for quant in QUANTS:
            codes, params, dist, cs, da = R.quant_probe(quant, metric, X, Q[:16], ids)
            key = f"{quant}_{metric}"
            out_q[f"{key}/params"] = params
            out_q[f"{key}/codes_sha"] = np.array(sha(codes))
            out_q[f"{key}/codes_head"] = codes[:8].copy()
            out_q[f"{key}/code_size"] = np.int32(cs)
            out_q[f"{key}/dim_align"] = np.int32(da)
            out_q[f"{key}/dist"] = dist
    out_q["ids"] = ids
    out_q["data_sha"] = np.array(sha(X))
    out_q["query_sha"] = np.array(sha(Q))
    np.savez_compressed(os.path.join(HERE, "quant.npz"), **out_q)

    out_s = {}
    nq = 64
    for gname, (XX, QQ, metric) in ds.items():
        out_s[f"{gname}/data_sha"] = np.array(sha(XX))
        gpath = os.path.join(HERE, f"{gname}.glass")
        for quant in QUANTS:
            refines = [""] + (["FP32"] if quant in ("SQ8U", "SQ4U", "SQ2U", "SQ8") else [])
            for refine in refines:
                s = R.RefSearcher(gpath, XX, metric, quant, refine)
                for ef in [16, 40, 96]:
                    for k in [1, 10, 50]:
                        s.set_ef(ef)
                        key = f"{gname}/{quant}/{refine or 'none'}/{ef}/{k}"
                        I, D = s.batch_search(QQ[:nq], k)
                        out_s[f"{key}/batch_ids"] = I
                        out_s[f"{key}/batch_dists"] = D
                        out_s[f"{key}/batch_dist_cmps"] = np.float64(s.get_last_search_dist_cmps())
                        si, sd = zip(*[s.search(QQ[i], k) for i in range(nq)])
                        out_s[f"{key}/single_ids"] = np.stack(si)
                        out_s[f"{key}/single_dists"] = np.stack(sd)
                del s
                print(gname, quant, refine, flush=True)
    np.savez_compressed(os.path.join(HERE, "search.npz"), **out_s)


if __name__ == "__main__":
    main()
