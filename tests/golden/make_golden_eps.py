"""Writes golden vectors for graphs whose entry-point list repeats ids or
holds more than 64 of them, from the compiled reference.

Run in the build container only (it needs oracle/_ref/libglassref.so):

    python tests/golden/make_golden_eps.py

Graph::initialize_search (glass/graph.hpp:95-104) inserts every entry point
into the pool in order, a repeated one twice, and marks it visited; the
Refiner then re-ranks both copies (glass/searcher/refiner.hpp:92-110).  The
graphs are nsg_l2.glass (NSG R=32 over 5,000 x 128 N(0,1) rows) with its
entry list replaced:
  eps_dup.glass   [ep, ep, 17, ep, 4242]
  eps_many.glass  100 entry points: 90 seeded random ids, then 10 of them again
Output: search_eps.npz with SearchBatch and Search ids / distances for
SQ8U, SQ8U + FP32 and SQ4U + FP32 at ef 16 / 64, k 1 / 10 over 32 queries.
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

CASES = [("SQ8U", ""), ("SQ8U", "FP32"), ("SQ4U", "FP32")]


def read_glass(path):
    raw = np.fromfile(path, np.uint8)
    nep = int(raw[:4].view(np.int32)[0])
    eps = raw[4:4 + 4 * nep].view(np.int32).copy()
    off = 4 + 4 * nep
    n, k = raw[off:off + 8].view(np.int32)
    adj = raw[off + 8:off + 8 + 4 * int(n) * int(k)].view(np.int32).reshape(int(n), int(k)).copy()
    return eps, adj


def write_glass(path, eps, adj):
    eps = np.asarray(eps, np.int32)
    with open(path, "wb") as f:
        f.write(np.int32(len(eps)).tobytes())
        f.write(eps.tobytes())
        f.write(np.array(adj.shape, np.int32).tobytes())
        f.write(np.ascontiguousarray(adj, np.int32).tobytes())


def entry_lists(ep0, n):
    rng = np.random.default_rng(99)
    many = rng.choice(n, 90, replace=False).astype(np.int32)
    return {"eps_dup": np.array([ep0, ep0, 17, ep0, 4242], np.int32),
            "eps_many": np.concatenate([many, many[:10]])}


def main():
    R.set_num_threads(8)
    X = synth.gaussian(10000, 128, synth.BASE_SEED)[:5000].copy()
    Q = synth.gaussian(32, 128, synth.QUERY_SEED + 3)
    eps0, adj = read_glass(os.path.join(HERE, "nsg_l2.glass"))
    out = {}
    for name, eps in entry_lists(int(eps0[0]), X.shape[0]).items():
        path = os.path.join(HERE, f"{name}.glass")
        write_glass(path, eps, adj)
        for quant, refine in CASES:
            s = R.RefSearcher(path, X, "L2", quant, refine)
            for ef in [16, 64]:
                s.set_ef(ef)
                for k in [1, 10]:
                    key = f"{name}/{quant}/{refine or 'none'}/{ef}/{k}"
                    ids, dis = s.batch_search(Q, k)
                    out[f"{key}/batch_ids"] = ids
                    out[f"{key}/batch_dists"] = dis
                    out[f"{key}/batch_dist_cmps"] = np.float64(s.get_last_search_dist_cmps())
                    si, sd = zip(*[s.search(Q[i], k) for i in range(len(Q))])
                    out[f"{key}/single_ids"] = np.stack(si)
                    out[f"{key}/single_dists"] = np.stack(sd)
            del s
    np.savez_compressed(os.path.join(HERE, "search_eps.npz"), **out)


if __name__ == "__main__":
    main()
