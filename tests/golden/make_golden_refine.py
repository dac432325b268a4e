"""Writes golden vectors for the SQ8U and SQ8 re-rank quantizers from the
compiled reference.

Run in the build container only (it needs oracle/_ref/libglassref.so):

    python tests/golden/make_golden_refine.py

make_refiner (glass/searcher/graph_searcher.hpp:240-254) wraps the searcher
in Refiner<SQ8QuantizerUniform> for refine_quant "SQ8U" and
Refiner<SQ8Quantizer> for "SQ8"; the re-rank distances are then the refine
quantizer's (an integer for SQ8U, per-dim float for SQ8), ordered by
(dist, id) (glass/searcher/refiner.hpp:92-110).
Output: search_refine.npz, per (graph, search quantizer, refine quantizer,
ef, k): SearchBatch and Search ids / distances over 48 queries and the batch
dist-cmp average, on the golden graphs c1_hnsw_l2 (L2) and ip_hnsw (IP).
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refwrap as R  # noqa: E402
from pyglass_amd import synth  # noqa: E402

CASES = [("SQ4U", "SQ8U"), ("SQ4U", "SQ8"), ("SQ2U", "SQ8U"), ("SQ8U", "SQ8U"), ("E5M2", "SQ8")]


def data():
    X = synth.gaussian(10000, 128, synth.BASE_SEED)
    Q = synth.gaussian(1000, 128, synth.QUERY_SEED)[:48]
    Xn = (X[:5000] / np.linalg.norm(X[:5000], axis=1, keepdims=True)).astype(np.float32)
    Qn = (Q / np.linalg.norm(Q, axis=1, keepdims=True)).astype(np.float32)
    return {"c1_hnsw_l2": (X, Q, "L2"), "ip_hnsw": (Xn, Qn, "IP")}


def main():
    R.set_num_threads(8)
    out = {}
    for gname, (X, Q, metric) in data().items():
        gpath = os.path.join(HERE, f"{gname}.glass")
        for quant, refine in CASES:
            s = R.RefSearcher(gpath, X, metric, quant, refine)
            for ef in [16, 64]:
                s.set_ef(ef)
                for k in [1, 10]:
                    key = f"{gname}/{quant}/{refine}/{ef}/{k}"
                    ids, dis = s.batch_search(Q, k)
                    out[f"{key}/batch_ids"] = ids
                    out[f"{key}/batch_dists"] = dis
                    out[f"{key}/batch_dist_cmps"] = np.float64(s.get_last_search_dist_cmps())
                    si, sd = zip(*[s.search(Q[i], k) for i in range(len(Q))])
                    out[f"{key}/single_ids"] = np.stack(si)
                    out[f"{key}/single_dists"] = np.stack(sd)
            del s
    np.savez_compressed(os.path.join(HERE, "search_refine.npz"), **out)


if __name__ == "__main__":
    main()
