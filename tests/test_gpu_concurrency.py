"""Sweeps of one call run at once (parameter concurrency: batches, or plain
sources, handed out to several workspaces, each with its own stream): the BC
equals the oracle's and the one-at-a-time run's (the per-workspace arrays are
added at the end, so only the rounding of the sum over sources differs), the
work totals are the same, and the device timer sums the timed kernel's
launches over every workspace."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [bcmi.MODE_BATCHED, bcmi.MODE_PLAIN])
def test_concurrent_sweeps_match(mode):
    g = bcmi.Graph.rmat(12, 16, seed=31)
    src = g.sample_sources(640 if mode == bcmi.MODE_BATCHED else 96, seed=32)
    ref = O.brandes(g.off, g.adj, src, nthreads=8)
    runs = {}
    for c in (1, 2, 3, 4):
        e = bcmi.Engine(concurrency=c, ktime=2 if mode == bcmi.MODE_BATCHED else 7)
        e.upload(g)
        e.run(src[:64], mode=mode)  # first use: helper workspaces and captured graphs
        e.reset()
        st = e.run(src, mode=mode)
        ms, launches = e.kernel_time()
        runs[c] = (e.bc(), st, launches)
        e.close()
    for c, (bc, st, launches) in runs.items():
        assert bcmi.rel_err(bc, ref) <= 1e-9, c
        assert bcmi.rel_err(bc, runs[1][0]) <= 1e-12, c
        for k in ("sources_run", "batches", "sum_n", "sum_m", "sum_a", "te"):
            assert st[k] == runs[1][1][k], (c, k)
        assert launches == runs[1][2], c   # same batches, same levels, same launches


def test_concurrency_with_peeling_and_expansion():
    g = bcmi.Graph.rmat(11, 8, seed=41).with_pendant_trees(0.3, seed=42)
    src = g.sample_sources(400, seed=43)
    for expand in (False, True):
        ref = O.bc_pipeline(g.off, g.adj, src, heuristic=1, expand=expand, nthreads=8)
        for c in (1, 3):
            e = bcmi.Engine(concurrency=c)
            e.upload(g, heuristic=bcmi.H_PEEL)
            e.run(src, mode=bcmi.MODE_BATCHED, expand=expand)
            assert bcmi.rel_err(e.bc(), ref) <= 1e-9, (expand, c)
            e.close()
