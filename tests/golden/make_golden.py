"""Writes the golden fixtures used by tests/test_oracle.py.

Run in the build container (it reads the reference's dataset file and uses
networkx 3.4.2's FP64 Brandes as the independent library routine):

    python tests/golden/make_golden.py /path/to/ck-bc

Fixtures (inputs + expected outputs, nothing executable):
  facebook.npz  edges of dataset/graph/facebook_combined.txt (the reference's
                CK dataset), networkx BC over sources 0..999 (the CK run's
                -N 1000 sweep, bc2d.c:1017-1018) and exact BC over all sources
                (both unnormalised, unordered pairs, as the reference's result
                file halves its ordered sums, bc2d.c:1065)
  rmat10.npz    R-MAT scale 10 edge factor 16 seed 20240101 generated by this
                engine, with networkx exact BC
"""
import os
import sys

import networkx as nx
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ck-bc_amd"))


def nx_bc(n, edges, sources=None):
    G = nx.Graph()
    G.add_nodes_from(range(n))
    G.add_edges_from((int(a), int(b)) for a, b in edges if a != b)
    if sources is None:
        d = nx.betweenness_centrality(G, normalized=False)
    else:
        d = nx.betweenness_centrality_subset(G, sources=list(sources), targets=list(range(n)), normalized=False)
    return np.array([d[i] for i in range(n)], np.float64)


def main(ref_root):
    path = os.path.join(ref_root, "dataset", "graph", "facebook_combined.txt")
    e = np.loadtxt(path, dtype=np.int64, comments="#")
    n = 4039
    bc1000 = nx_bc(n, e, range(1000))
    bcall = nx_bc(n, e)
    np.savez_compressed(os.path.join(HERE, "facebook.npz"), n=n, edges=e.astype(np.int16),
                        bc_src1000=bc1000, bc_all=bcall)
    import bcmi
    g = bcmi.Graph.rmat(10, 16, seed=20240101)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    keep = src < g.adj
    edges = np.stack([src[keep], g.adj[keep]], 1)
    np.savez_compressed(os.path.join(HERE, "rmat10.npz"), n=g.n, edges=edges.astype(np.int16),
                        bc_all=nx_bc(g.n, edges))
    print("wrote fixtures")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
