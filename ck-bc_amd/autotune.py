"""Autotuner for the engine's launch and algorithm parameters.

Coordinate descent on measured throughput (GTEPS of a sample of batches):
block shape, grid, record size (edges per work item), the top-down ->
bottom-up switch threshold (batched and plain), the frontier floor below
which bottom-up is not used, the frontier size below which path counts are
pushed from the frontier, the levels per captured graph, and the batch width.  (The
small-grid size for sparse levels is left at the engine default, which a
sweep on bridged-block chains set; R-MAT graphs do not depend on it.)  Results go to a JSON table keyed by a
graph class ("rmat-s22-ef16", "file:<name>", ...) that bench.py and callers
apply with `apply(engine, table, key)`.

    python autotune.py --scale 22 --edgefactor 16 --out tuned.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bcmi  # noqa: E402

SPACE = {
    "block": [64, 128, 256],
    "grid": [1024, 2048, 4096, 8192, 16384],
    "grid_back": [0, 4096, 8192, 16384, 32768],
    "alpha_batched": [125, 250, 500, 1000, 2000, 4000, 1 << 60],
    "beta_batched": [4, 8, 16, 32, 64, 128, 512],
    "gamma_push": [0, 4, 8, 16, 32],
    "fuse_bu": [1, 2],
    "pack_bu": [0, 1],
    "gamma_back": [0, 4, 8, 16],
    "steps": [8, 10, 12, 16],
    "width": [32, 48, 64],
}
PLAIN_SPACE = {"alpha_plain": [100, 300, 1000, 3000, 10000, 1 << 60], "grid_plain": [2048, 4096, 8192]}
CHUNKS = [256, 512, 1024]
DEFAULT_PATH = os.path.join(HERE, "tuned.json")


def graph_key(kind, **kw):
    if kind == "rmat":
        return f"rmat-s{kw['scale']}-ef{kw['edgefactor']}"
    return f"{kind}"


def measure(e, src, mode, reps=1):
    best = 0.0
    for _ in range(reps):
        e.reset()
        t = time.perf_counter()
        st = e.run(src, mode=mode)
        t = time.perf_counter() - t
        best = max(best, st["te"] / t / 1e9)
    return best


def tune(g, src_batched, src_plain, log=print):
    params = {"chunk": 256, "block": 256, "grid": 4096, "grid_plain": 4096, "alpha_batched": 1000, "beta_batched": 16, "gamma_push": 8,
              "fuse_bu": 2, "pack_bu": 1, "gamma_back": 0, "steps": 16, "width": 64, "grid_back": 0,
              "alpha_plain": 1000}
    best_overall = (-1.0, None)
    for ch in CHUNKS:
        e = bcmi.Engine(0, chunk=ch)
        e.upload(g)
        cur = dict(params, chunk=ch)
        for k in ("block", "grid", "grid_back", "alpha_batched", "beta_batched", "gamma_push", "fuse_bu", "pack_bu",
                  "gamma_back", "steps", "width"):
            if k != "chunk":
                e.set_param(k, cur[k])
        e.run(src_batched[:64])  # capture
        score = measure(e, src_batched, bcmi.MODE_BATCHED)
        for k, vals in SPACE.items():
            best_v, best_s = cur[k], score
            for v in vals:
                if v == cur[k]:
                    continue
                e.set_param(k, v)
                if k == "block":  # keep the thread count per kernel
                    e.set_param("grid", max(1, cur["grid"] * cur["block"] // v))
                s = measure(e, src_batched, bcmi.MODE_BATCHED)
                log(f"  chunk {ch} {k}={v}: {s:.1f} GTEPS")
                if s > best_s:
                    best_v, best_s = v, s
                e.set_param(k, cur[k])
                if k == "block":
                    e.set_param("grid", cur["grid"])
            cur[k] = best_v
            if k == "block":
                cur["grid"] = max(1, params["grid"] * params["block"] // best_v)
            e.set_param(k, best_v)
            if k == "block":
                e.set_param("grid", cur["grid"])
            score = best_s
        log(f"chunk {ch}: best {score:.1f} GTEPS with {cur}")
        if score > best_overall[0]:
            # plain sweep parameters on this engine (coordinate descent)
            ps = measure(e, src_plain, bcmi.MODE_PLAIN)
            for pk, vals in PLAIN_SPACE.items():
                pv = cur.get(pk, vals[0])
                for v in vals:
                    e.set_param(pk, v)
                    s = measure(e, src_plain, bcmi.MODE_PLAIN)
                    log(f"  plain {pk}={v}: {s:.2f} GTEPS")
                    if s > ps:
                        pv, ps = v, s
                e.set_param(pk, pv)
                cur[pk] = pv
            cur["plain_gteps"] = round(ps, 3)
            cur["batched_gteps"] = round(score, 2)
            best_overall = (score, dict(cur))
        e.close()
    return best_overall[1]


# engine parameters a table entry may set (the rest of an entry are notes)
APPLIED = ("chunk", "block", "grid", "grid_back", "grid_plain", "alpha_batched", "beta_batched", "gamma_push",
           "fuse_bu", "pack_bu", "gamma_back", "steps", "alpha_plain", "width")


def apply(engine, table, key):
    """Sets the tuned parameters of `key` (or table['default']) on an engine
    whose graph is not uploaded yet.  Returns the parameters used."""
    p = table.get(key) or table.get("default")
    if not p:
        return None
    for k in APPLIED:
        if k in p:
            engine.set_param(k, int(p[k]))
    return p


def load(path=DEFAULT_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--batches", type=int, default=12, help="batches of 64 sources per measurement")
    ap.add_argument("--plain", type=int, default=16, help="sources per plain measurement")
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    g = bcmi.Graph.rmat(a.scale, a.edgefactor, seed=a.seed)
    src = g.sample_sources(16384, seed=a.seed + 99)
    best = tune(g, src[:64 * a.batches], src[:a.plain])
    key = graph_key("rmat", scale=a.scale, edgefactor=a.edgefactor)
    table = load(a.out)
    table[key] = best
    table.setdefault("default", best)
    table[key]["tuned_on"] = time.strftime("%Y-%m-%d")
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({key: best}))


if __name__ == "__main__":
    main()
