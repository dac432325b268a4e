"""Writes golden vectors on the headline workload's data (uint8-valued
Gaussian mixture, SQ4U + FP32 re-rank) from the compiled reference.

Run in the build container only (it needs oracle/_ref/libglassref.so, built
from /root/reference by `make -C oracle ref`):

    python tests/golden/make_golden_c2.py

The rows are integers in [0, 255] quantized to 16 levels, so SQ4U distances
tie far more often than on N(0,1) rows: this pins the LinearPool tie order
(new keys after equal old keys, glass/neighbor.hpp:361-396) and the
Refiner's (dist, id) order (glass/searcher/refiner.hpp:92-110) on the data
the headline number is measured on.

Outputs (inputs are regenerated from seeds; their SHA-256 is stored so a
test can prove the regeneration is identical):
  c2_mix20k_hnsw.glass  HNSW R=32 L=200 built by the reference
                        (glass/hnsw/hnsw.hpp:27-69) over 20,000 x 128 rows of
                        the mixture (seed 1234)
  c2_mix10k_nsg.glass   NSG R=32 L=64 built by the reference
                        (glass/nsg/nsg.hpp:56-104) over 10,000 x 128 rows
                        (seed 4321), for the builder comparison
  search_c2.npz         per (refine factor, ef, k): SearchBatch ids,
                        distances and dist-cmp average over 100 queries, and
                        Search ids / distances over the first 32
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

N_BASE = 20000
N_NSG = 10000
NSG_SEED = 4321
N_QUERY = 100
N_SINGLE = 32
FACTORS = [None, 4.0, 8.0]  # None: get_refine_factor("SQ4U") = 1.5 (graph_searcher.hpp:227-238)
EFS = [16, 48, 64, 128, 256]
KS = [1, 10]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def data():
    X = synth.uint8_mixture(N_BASE, 128, synth.BASE_SEED)
    Q = synth.uint8_mixture(N_QUERY, 128, synth.QUERY_SEED)
    Xn = synth.uint8_mixture(N_NSG, 128, NSG_SEED)
    return X, Q, Xn


def main():
    R.set_num_threads(8)
    X, Q, Xn = data()
    gpath = os.path.join(HERE, "c2_mix20k_hnsw.glass")
    R.build_graph("HNSW", X, gpath, "L2", "FP32", 32, 200)
    R.build_graph("NSG", Xn, os.path.join(HERE, "c2_mix10k_nsg.glass"), "L2", "FP32", 32, 64)
    out = {"data_sha": np.array(sha(X)), "query_sha": np.array(sha(Q)), "nsg_data_sha": np.array(sha(Xn))}
    for f in FACTORS:
        s = R.RefSearcher(gpath, X, "L2", "SQ4U", "FP32", refine_factor=f)
        for ef in EFS:
            s.set_ef(ef)
            for k in KS:
                key = f"{f or 1.5}/{ef}/{k}"
                ids, dis = s.batch_search(Q, k)
                out[f"{key}/batch_ids"] = ids
                out[f"{key}/batch_dists"] = dis
                out[f"{key}/batch_dist_cmps"] = np.float64(s.get_last_search_dist_cmps())
                si, sd = zip(*[s.search(Q[i], k) for i in range(N_SINGLE)])
                out[f"{key}/single_ids"] = np.stack(si)
                out[f"{key}/single_dists"] = np.stack(sd)
        print("factor", f or 1.5, flush=True)
        del s
    np.savez_compressed(os.path.join(HERE, "search_c2.npz"), **out)


if __name__ == "__main__":
    main()
