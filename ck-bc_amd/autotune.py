"""Autotuner for the engine's launch and algorithm parameters.

Coordinate descent on measured throughput (GTEPS) over the whole search
space: record size (edges per work item, which needs a fresh upload), block
shape, forward and backward grids, the top-down -> bottom-up switch
threshold and the bottom-up floor, the frontier size below which path counts
are pushed, bottom-up fusion and packing, pushed backward levels, the levels
per captured graph, the batch width, the path-count row width and the
batches run at once (each on its own workspace and stream; 0 is the engine's
choice); then the plain sweep's switch threshold and grid.  (The small-grid size for sparse levels is left at the
engine default, which sweeps on bridged-block chains set; R-MAT graphs do not
depend on it.)

A measurement runs batches of the workload itself: the engine's grouping of
all the workload's sources into batches (bcmi_batch_order), `--batches` of
them spread evenly over that schedule, as bench.py times them (on peeled
graphs, whose sources the engine maps to core roots, the first sources).
Each coordinate re-measures the incumbent first and moves only on a gain
above `--min-gain`, so run-to-run noise does not walk the parameters; the
descent repeats `--passes` times from the entry already in the table (or the
defaults).  Results go to a JSON table keyed by graph class
("rmat-s22-ef16", ...), with the command and sample that produced them, and
bench.py and callers apply them with `apply(engine, table, key)`.

    python autotune.py --scale 22 --edgefactor 16 --sources 16384 --batches 32
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
    "chunk": [256, 512, 1024],
    "block": [64, 128, 256],
    "grid": [1024, 2048, 4096, 8192, 16384],
    "grid_back": [0, 4096, 8192, 16384, 32768],
    "alpha_batched": [125, 250, 500, 1000, 2000, 4000, 1 << 60],
    "beta_batched": [4, 8, 16, 32, 64, 128, 512, 2048],
    "gamma_push": [0, 4, 8, 16, 32],
    "fuse_bu": [1, 2],
    "pack_bu": [0, 1],
    "gamma_back": [0, 4, 8, 16],
    "steps": [6, 8, 10, 12, 16],
    "width": [32, 48, 64],
    "narrow_sigma": [1, 2],
    "concurrency": [0, 1, 2, 3, 4, 6, 8],
}
PLAIN_SPACE = {"alpha_plain": [100, 300, 1000, 3000, 10000, 1 << 60], "grid_plain": [1024, 2048, 4096, 8192]}
DEFAULTS = {"chunk": 256, "block": 256, "grid": 4096, "grid_back": 0, "alpha_batched": 1000, "beta_batched": 16,
            "gamma_push": 8, "fuse_bu": 2, "pack_bu": 1, "gamma_back": 0, "steps": 16, "width": 64,
            "narrow_sigma": 1, "concurrency": 0, "alpha_plain": 1000, "grid_plain": 4096}
DEFAULT_PATH = os.path.join(HERE, "tuned.json")
# engine parameters a table entry may set (the rest of an entry are notes)
APPLIED = ("chunk", "block", "grid", "grid_back", "grid_plain", "alpha_batched", "beta_batched", "gamma_push",
           "fuse_bu", "pack_bu", "gamma_back", "steps", "alpha_plain", "width", "narrow_sigma", "concurrency")
NOTES = ("batched_gteps", "plain_gteps", "tuned_on", "tuned_by", "command", "sample")


def graph_key(kind, **kw):
    if kind == "rmat":
        return f"rmat-s{kw['scale']}-ef{kw['edgefactor']}"
    return f"{kind}"


class Tuner:
    """Measures batched and plain throughput of one graph under parameter sets."""

    def __init__(self, g, sources, batches, plain, heuristic=bcmi.H_NONE, log=print, min_time=0.3):
        self.g, self.sources, self.batches, self.plain, self.h, self.log = g, sources, batches, plain, heuristic, log
        self.e, self.chunk, self.width = None, None, None
        # a measurement repeats its run until it has timed at least min_time
        # seconds (small graphs: one run of a few milliseconds is mostly noise)
        self.min_time = min_time

    def _rate(self, e, src, mode):
        te = t_sum = 0.0
        while True:
            e.reset()
            t = time.perf_counter()
            st = e.run(src, mode=mode)
            t_sum += time.perf_counter() - t
            te += st["te"]
            if t_sum >= self.min_time:
                return te / t_sum / 1e9

    def _engine(self, p):
        """An engine holding the graph uploaded with p's record size, p applied."""
        if self.e is None or self.chunk != p["chunk"]:
            if self.e is not None:
                self.e.close()
            self.e = bcmi.Engine(0, chunk=p["chunk"])
            self.e.upload(self.g, heuristic=self.h)
            self.chunk, self.width = p["chunk"], None
        for k in APPLIED:
            if k != "chunk" and k in p:
                self.e.set_param(k, p[k])
        return self.e

    def _timed_sources(self, e, width):
        """`batches` batches of `width` spread evenly over the engine's grouping of
        all the sources (peeled graphs: the first sources), and a warmup batch."""
        if self.h != bcmi.H_NONE:
            e.set_param("batching", 4)
            k = self.batches * width
            return self.sources[k:k + width], self.sources[:k]
        e.set_param("batching", 4)
        order = e.batch_order(self.sources)
        nb = len(order) // width
        idx = sorted({int((i + 0.5) * nb / self.batches) for i in range(min(self.batches, nb))})
        warm = next(i for i in range(nb) if i not in set(idx)) if nb > len(idx) else idx[0]
        e.set_param("batching", 0)
        pick = lambda ix: np.concatenate([order[i * width:(i + 1) * width] for i in ix])
        return pick([warm]), pick(idx)

    def batched(self, p):
        e = self._engine(p)
        warm, timed = self._timed_sources(e, p["width"])
        e.run(warm, mode=bcmi.MODE_BATCHED)  # captures the level loops for p
        return self._rate(e, timed, bcmi.MODE_BATCHED)

    def plain_rate(self, p):
        e = self._engine(p)
        src = self.sources[:self.plain]
        e.run(src[:2], mode=bcmi.MODE_PLAIN)
        return self._rate(e, src, bcmi.MODE_PLAIN)

    def close(self):
        if self.e is not None:
            self.e.close()
            self.e = None


def descend(measure, start, space, passes, min_gain, log, tag):
    """Coordinate descent: each coordinate re-measures the incumbent, tries every
    other value and moves to the best one only if it beats the incumbent by
    more than min_gain (relative).  Returns the parameters and their last score."""
    cur = dict(start)
    score = 0.0
    for it in range(passes):
        moved = False
        for k, vals in space.items():
            base = measure(cur)
            top_v, top_s = cur[k], base
            for v in vals:
                if v == cur[k]:
                    continue
                s = measure(dict(cur, **{k: v}))
                log(f"  {tag} pass {it + 1} {k}={v}: {s:.2f} GTEPS (incumbent {k}={cur[k]}: {base:.2f})")
                if s > top_s:
                    top_v, top_s = v, s
            if top_v != cur[k] and top_s > base * (1.0 + min_gain):
                cur[k] = top_v
                moved = True
                score = top_s
            else:
                score = base
        log(f"{tag} pass {it + 1}: {score:.2f} GTEPS with {cur}")
        if not moved:
            break
    return cur, score


def tune(g, sources, batches=32, plain=16, start=None, space=None, plain_space=None, passes=2, min_gain=0.005,
         heuristic=bcmi.H_NONE, log=print, min_time=0.3):
    """Returns the tuned parameters (APPLIED keys) and the measured GTEPS."""
    space = SPACE if space is None else space
    plain_space = PLAIN_SPACE if plain_space is None else plain_space
    p = dict(DEFAULTS)
    if start:
        p.update({k: int(v) for k, v in start.items() if k in APPLIED})
    T = Tuner(g, sources, batches, plain, heuristic, log, min_time)
    try:
        p, score = descend(T.batched, p, space, passes, min_gain, log, "batched")
        p, pscore = descend(T.plain_rate, p, plain_space, passes, min_gain, log, "plain")
    finally:
        T.close()
    return p, score, pscore


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
    ap.add_argument("--pendant", type=float, default=0.0, help="add pendant trees (this fraction) and peel them")
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--sources", type=int, default=16384, help="sources of the workload (as bench.py samples them)")
    ap.add_argument("--batches", type=int, default=32, help="batches timed per measurement")
    ap.add_argument("--plain", type=int, default=16, help="sources per plain measurement")
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--min-gain", type=float, default=0.005)
    ap.add_argument("--min-time", type=float, default=0.3, help="seconds each measurement times at least (runs repeat)")
    ap.add_argument("--fresh", action="store_true", help="start from the engine defaults, not the table entry")
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    g = bcmi.Graph.rmat(a.scale, a.edgefactor, seed=a.seed)
    heur = bcmi.H_NONE
    if a.pendant > 0:
        g = g.with_pendant_trees(a.pendant, seed=a.seed + 2)
        heur = bcmi.H_PEEL
    src = g.sample_sources(a.sources, seed=a.seed + 1)  # the sources bench.py samples
    key = graph_key("rmat", scale=a.scale, edgefactor=a.edgefactor)
    table = load(a.out)
    start = None if a.fresh else table.get(key)
    t0 = time.time()
    p, score, pscore = tune(g, src, a.batches, a.plain, start, passes=a.passes, min_gain=a.min_gain,
                            heuristic=heur, log=log, min_time=a.min_time)
    entry = {k: int(p[k]) for k in APPLIED}
    entry.update({
        "batched_gteps": round(score, 2), "plain_gteps": round(pscore, 3), "tuned_on": time.strftime("%Y-%m-%d"),
        "tuned_by": "autotune.py",
        "command": "python ck-bc_amd/autotune.py " + " ".join(sys.argv[1:]),
        "sample": (f"R-MAT scale {a.scale} ef {a.edgefactor}" + (f" + {a.pendant:.0%} pendant trees, peeled" if heur else "")
                   + f", seed {a.seed}; {a.batches} batches of the {a.sources}-source workload per measurement, "
                   f"{a.plain} plain sweeps, each run repeated to at least {a.min_time:g} s; {a.passes} passes from "
                   + ("the defaults" if start is None else "the previous entry")
                   + f"; {time.time() - t0:.0f} s")})
    table[key] = entry
    if a.scale == 22 and a.edgefactor == 16 and not heur:
        table["default"] = dict(entry)
    with open(a.out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps({key: entry}))


if __name__ == "__main__":
    main()
