"""Test infrastructure: small hand-built graphs for the oracle and the HIP
path, written in the .glass layout of glass/graph.hpp:106-119 (header
nep, eps, N, K, adjacency) followed by the HNSW upper layers of
glass/hnsw/HNSWInitializer.hpp:64-86 (N, K_up, ep, then per node a count of
levels[i]*K_up ints and the lists, level l at offset (l-1)*K_up).
"""

from __future__ import annotations

import numpy as np


def write_glass(path, adj, eps, upper=None):
    """upper = (K_up, ep, {node: [level-1 list, level-2 list, ...]}) or None."""
    adj = np.asarray(adj, np.int32)
    n, k = adj.shape
    parts = [np.array([len(eps)], np.int32), np.asarray(eps, np.int32), np.array([n, k], np.int32), adj.ravel()]
    if upper is not None:
        kup, ep, lists = upper
        parts.append(np.array([n, kup, ep], np.int32))
        for i in range(n):
            lv = lists.get(i, [])
            flat = np.full(len(lv) * kup, -1, np.int32)
            for li, lst in enumerate(lv):
                flat[li * kup:li * kup + len(lst)] = lst
            parts.append(np.array([len(flat)], np.int32))
            parts.append(flat)
    with open(path, "wb") as f:
        for p in parts:
            f.write(np.ascontiguousarray(p, np.int32).tobytes())


# A 3-level toy HNSW on a line.  Node i sits at POS[i] on the first axis (the
# other axes are 0) and the query sits at 10, so L2 distances are the exact
# integers (POS[i] - 10)^2:
#   node  0    1   2   3   4    5   6   7
#   pos   0    4   7  12  30    9  10  11
#   dist 100  36   9   4  400   1   0   1
POS = [0.0, 4.0, 7.0, 12.0, 30.0, 9.0, 10.0, 11.0]
QUERY_POS = 10.0
TOY_KUP = 4
TOY_EP = 0
# lists[node] = [level-1 list, level-2 list]
# (every node named in a level-l list has level >= l, as in any HNSW)
TOY_UPPER = {
    0: [[1], [1, 2, 4]],
    1: [[], []],
    2: [[], [3, 0]],
    3: [[7, 5], [2]],
    5: [[]],
    6: [[7, 5]],
    7: [[6]],
    4: [[], []],
}
# Greedy descent (HNSWInitializer.hpp:41-62) worked by hand:
#   start u=0 d=100                                   1 evaluation
#   level 2, pass 1 over list(0): 1 (36<100 -> u=1), 2 (9<36 -> u=2), 4 (400)
#                                                     3 evaluations
#   level 2, pass 2 over list(2): 3 (4<9 -> u=3), 0 (100)      2
#   level 2, pass 3 over list(3): 2 (9)              1, no change
#   level 1, pass 1 over list(3): 7 (1<4 -> u=7), 5 (1, not < 1)  2
#   level 1, pass 2 over list(7): 6 (0<1 -> u=6)               1
#   level 1, pass 3 over list(6): 7 (1), 5 (1)       2, no change
# -> entry u=6, d=0 after 12 evaluations.  Nodes evaluated on the upper
# levels are not marked visited, so the base-layer expansion of node 6
# (neighbours 7, 5, 3) evaluates them again: 15 evaluations in total.  With
# k = ef = 3 the pool ends as 6 (0), 7 (1), 5 (1): the tie keeps insertion
# order and node 3 (4) is rejected by the full pool.
TOY_ENTRY = 6
TOY_UPPER_EVALS = 12
TOY_BASE_ADJ = {6: [7, 5, 3]}
TOY_RESULT_IDS = [6, 7, 5]
TOY_RESULT_DISTS = [0.0, 1.0, 1.0]
TOY_DIST_CMPS = 15


def toy_hnsw(path, dim=16):
    """Writes the toy graph; returns (data[8, dim], query[dim])."""
    adj = np.full((len(POS), 4), -1, np.int32)
    for u, row in TOY_BASE_ADJ.items():
        adj[u, :len(row)] = row
    write_glass(path, adj, [TOY_EP], (TOY_KUP, TOY_EP, TOY_UPPER))
    data = np.zeros((len(POS), dim), np.float32)
    data[:, 0] = POS
    q = np.zeros(dim, np.float32)
    q[0] = QUERY_POS
    return data, q
