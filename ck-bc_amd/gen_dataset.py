"""Writes the synthetic graphs of the benchmark configurations as CK dataset
files: a binary CSR (what `ck-bc-hip -f` reads fastest) or a text edge list
in the reference's input format (one `u v` pair per line, each undirected
edge once).  Seeds and shapes are fixed by the dataset entries under
dataset/*/.cm/meta.json, which call this script to create missing files.

    python gen_dataset.py rmat --scale 22 --edgefactor 16 --seed 20250101 --out rmat-s22-ef16.csr
    python gen_dataset.py chain --blocks 32768 --seed 20250101 --out chain-32768.csr
    python gen_dataset.py rmat --scale 18 --pendant 0.3 --seed 20250101 --out rmat-s18-trees.txt --text
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bcmi  # noqa: E402


def build(a):
    if a.kind == "rmat":
        g = bcmi.Graph.rmat(a.scale, a.edgefactor, seed=a.seed)
    else:
        g = bcmi.Graph.block_chain(a.blocks, a.block, a.degree, seed=a.seed)
    if a.pendant > 0:
        g = g.with_pendant_trees(a.pendant, seed=a.seed + 2)
    return g


def write_text(g, path, chunk=1 << 22):
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.off))
    keep = src < g.adj                       # each undirected edge once
    u, v = src[keep], g.adj[keep].astype(np.int64)
    with open(path, "w") as f:
        f.write(f"# Nodes: {g.n} Edges: {len(u)}\n")
        for i in range(0, len(u), chunk):
            f.write("\n".join(f"{x} {y}" for x, y in zip(u[i:i + chunk], v[i:i + chunk])))
            f.write("\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["rmat", "chain"])
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--blocks", type=int, default=32768)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--degree", type=float, default=8.0)
    ap.add_argument("--pendant", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--out", required=True)
    ap.add_argument("--text", action="store_true", help="text edge list instead of binary CSR")
    a = ap.parse_args(argv)
    g = build(a)
    tmp = a.out + ".part"
    if a.text:
        write_text(g, tmp)
    else:
        g.save(tmp)
    os.replace(tmp, a.out)
    print(f"{a.out}: n={g.n} arcs={g.m}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
