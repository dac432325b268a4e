"""Writes a diskann-format graph and the compiled reference's reading of it.

Run in the build container only (it needs oracle/_ref/libglassref.so):

    python tests/golden/make_golden_formats.py

Graph::load_diskann (glass/graph.hpp:155-182) reads a u64 file size, the
maximum degree K, the entry point, a u64 it skips, then per node a count and
that many ids.  The file here has 1,000 nodes of 0-24 neighbours (seeded);
the reference loads it and saves it as .glass (graph.hpp:106-119):
  diskann_1k.bin        the diskann file
  diskann_1k_ref.glass  the reference's .glass of it
"""

from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import refwrap as R  # noqa: E402


def main():
    rng = np.random.default_rng(77)
    n, K, ep = 1000, 24, 417
    body = bytearray()
    for i in range(n):
        c = int(rng.integers(0, K + 1))
        ids = rng.choice(n, c, replace=False).astype(np.int32)
        body += np.int32(c).tobytes() + ids.tobytes()
    size = 24 + len(body)
    head = np.uint64(size).tobytes() + np.int32(K).tobytes() + np.int32(ep).tobytes() + np.uint64(0).tobytes()
    src = os.path.join(HERE, "diskann_1k.bin")
    with open(src, "wb") as f:
        f.write(head + bytes(body))
    assert R.graph_convert(src, "diskann", os.path.join(HERE, "diskann_1k_ref.glass")) == n


if __name__ == "__main__":
    main()
