"""Writes the Clustering golden vectors under tests/golden/ from the compiled
reference (glass::Clustering, glass/algorithms/clustering.hpp, driven as
python/bindings.cc:155-203 drives it).

Run in the build container only (it needs oracle/_ref/libglassref.so, built
from /root/reference by `make -C oracle ref`):

    python tests/golden/make_golden_clustering.py

Output clustering.npz, per (dim, init): the 64 fitted centroids (1,500 rows
N(0,1), seed 1234, 10 epochs), predict() of 200 other rows (seed 4321) and
the digest and first rows of transform(inplace=True) of those rows.  The dims
cover every branch of helpa::l2_fp32_fp32: 4 and 7 (4-wide step and scalar
tail), 12 (8-wide and 4-wide), 16 (body only), 30 (body, 8, 4, 2), 100 and
128.
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

DIMS = [4, 7, 12, 16, 30, 100, 128]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    R.set_num_threads(8)
    out = {}
    for d in DIMS:
        X = synth.gaussian(1500, d, synth.BASE_SEED)
        P = synth.gaussian(200, d, synth.QUERY_SEED)
        out[f"{d}/data_sha"] = np.array(sha(X))
        for init in ["random", "kmeans++"]:
            cent, ids, tr = R.clustering_probe(X, 64, 10, init, P)
            key = f"{d}/{init}"
            out[f"{key}/centroids"] = cent
            out[f"{key}/predict"] = ids
            out[f"{key}/transform_sha"] = np.array(sha(tr))
            out[f"{key}/transform_head"] = tr[:4].copy()
        print(d, flush=True)
    np.savez_compressed(os.path.join(HERE, "clustering.npz"), **out)


if __name__ == "__main__":
    main()
