"""Parity at the full sizes of the larger workloads, through checks whose
cost does not grow with the number of sources: the pruned set and weights
against the oracle, peeled vs unpeeled runs, plain vs batched sweeps, the
CPU oracle on a few sources, and the distance checksum (the BC sum of a
source equals the sum over its reached targets of d(s,t) - 1), which holds
whatever the size of the path counts.

  * R-MAT scale 18 + 30% pendant trees (n = 340,787), 8,192 sources;
  * R-MAT scale 22 (4.2M vertices, ~128M arcs), all 16,384 sources with the
    launch shape tuned for it;
  * R-MAT scale 25 (33.5M vertices, ~1.07G arcs);
  * a bridged-block chain of 32,768 blocks of 256 vertices (8.4M vertices,
    path counts far beyond 2^64, tens of thousands of levels).
"""
import os
import sys
import time

import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SEED = 20250101
NTH = min(16, os.cpu_count() or 1)


def say(msg):
    print(f"[configs] {msg}", file=sys.stderr, flush=True)


def rel_sum(a, b):
    return abs(float(a) - float(b)) / max(1.0, abs(float(b)))


def test_pendant_trees_scale18():
    t = time.time()
    g = bcmi.Graph.rmat(18, 16, seed=SEED).with_pendant_trees(0.3, seed=SEED + 2)
    src = g.sample_sources(8192, seed=SEED + 1)
    say(f"pendant18: n={g.n} arcs={g.m} built in {time.time() - t:.1f}s")
    e = bcmi.Engine()
    e.upload(g, heuristic=bcmi.H_PEEL)
    p = e.peel_info()
    o = O.peel(g.off, g.adj)
    assert np.array_equal(p["round"], o["round"])
    assert np.array_equal(p["w"], o["w"])
    rem = o["round"] > 0
    assert rem.sum() > 0.2 * g.n
    assert np.array_equal(p["parent"][rem], o["parent"][rem])
    e.reset()
    e.run(src, mode=bcmi.MODE_BATCHED)
    bc_peel = e.bc()
    say("pendant18: peeled batched run done")
    # the same sources without peeling
    u = bcmi.Engine()
    u.upload(g)
    u.reset()
    u.run(src, mode=bcmi.MODE_BATCHED)
    assert bcmi.rel_err(bc_peel, u.bc()) <= 1e-9
    u.close()
    # distance checksum over all 8,192 sources
    ds = O.depth_sums(g.off, g.adj, src, nthreads=NTH)
    assert rel_sum(bc_peel.sum(), ds.sum()) <= 1e-9
    say("pendant18: checksum done")
    # plain sweep with peeling vs the CPU pipeline on 256 sources
    sub = src[:256]
    e.reset()
    e.run(sub, mode=bcmi.MODE_PLAIN)
    ref = O.bc_pipeline(g.off, g.adj, sub, heuristic=1, nthreads=NTH)
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    e.close()


@pytest.fixture(scope="module")
def rmat22():
    """R-MAT scale 22 uploaded to an engine with the launch shape tuned for it."""
    import autotune
    t = time.time()
    g = bcmi.Graph.rmat(22, 16, seed=SEED)
    src = g.sample_sources(16384, seed=SEED + 1)
    say(f"rmat22: n={g.n} arcs={g.m} built in {time.time() - t:.1f}s")
    e = bcmi.Engine()
    assert autotune.apply(e, autotune.load(), autotune.graph_key("rmat", scale=22, edgefactor=16))
    e.upload(g)
    yield g, src, e
    e.close()


def test_rmat_scale22_all_sources_tuned(rmat22):
    """The batched workload as measured: all 16,384 sources in batches of 64
    with the tuned launch shape.  BC is linear in the sources, so the whole
    run equals the sum of two runs over its halves (batched differently);
    a 512-source run matches the oracle's distance checksum; two sources
    match the CPU oracle; plain and batched sweeps agree on 64 sources."""
    g, src, e = rmat22
    e.reset()
    st = e.run(src, mode=bcmi.MODE_BATCHED)
    assert st["batches"] == 256
    full = e.bc()
    half = np.zeros_like(full)
    for part in (src[:8192], src[8192:]):
        e.reset()
        e.run(part, mode=bcmi.MODE_BATCHED)
        half += e.bc()
    assert bcmi.rel_err(full, half) <= 1e-9
    say("rmat22: all sources = sum of the halves")
    sub = src[:512]
    e.reset()
    e.run(sub, mode=bcmi.MODE_BATCHED)
    ds = O.depth_sums(g.off, g.adj, sub, nthreads=NTH)
    assert rel_sum(e.bc().sum(), ds.sum()) <= 1e-9
    say("rmat22: 512-source checksum done")
    two = src[:2]
    e.reset()
    e.run(two, mode=bcmi.MODE_BATCHED)
    ref = O.brandes(g.off, g.adj, two, nthreads=2)
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    sixty4 = src[:64]
    e.reset()
    e.run(sixty4, mode=bcmi.MODE_BATCHED)
    bb = e.bc()
    e.reset()
    e.run(sixty4, mode=bcmi.MODE_PLAIN)
    assert bcmi.rel_err(bb, e.bc()) <= 1e-9
    say("rmat22: oracle and plain-vs-batched done")


def test_rmat_scale22_depth_sigma_bit_exact(rmat22):
    """Forward rule (cuda_kernels.cu:524-537) at full size: one batch of 64
    sources swept with the tuned launch shape and read back per source;
    depths and path counts of 8 of its lanes equal the oracle's bit for bit,
    their dependencies within 1e-12.  Two plain sweeps likewise."""
    g, src, e = rmat22
    batch = src[:64]
    d, sg, dl = e.msbfs(batch)
    say("rmat22: batch read back")
    for lane in range(0, 64, 9):
        od, osg, odl, _ = O.sssp(g.off, g.adj, int(batch[lane]))
        assert np.array_equal(d[lane], od), lane
        assert np.array_equal(sg[lane], osg), lane
        odl[batch[lane]] = 0.0
        assert bcmi.rel_err(dl[lane], odl) <= 1e-12, lane
    del d, sg, dl
    say("rmat22: 8 lanes of the batch bit-exact")
    for s in src[64:66]:
        d, sg, dl = e.sssp(int(s))
        od, osg, odl, _ = O.sssp(g.off, g.adj, int(s))
        assert np.array_equal(d, od) and np.array_equal(sg, osg)
        odl[s] = 0.0
        assert bcmi.rel_err(dl, odl) <= 1e-12
    say("rmat22: plain sweeps bit-exact")


def test_rmat_scale25():
    t = time.time()
    g = bcmi.Graph.rmat(25, 16, seed=SEED)
    src = g.sample_sources(4096, seed=SEED + 1)
    say(f"rmat25: n={g.n} arcs={g.m} built in {time.time() - t:.1f}s")
    assert g.m > 1_000_000_000
    e = bcmi.Engine()
    t = time.time()
    e.upload(g)
    say(f"rmat25: uploaded in {time.time() - t:.1f}s")
    sub = src[:64]
    e.reset()
    e.run(sub, mode=bcmi.MODE_BATCHED)
    bc = e.bc()
    ds = O.depth_sums(g.off, g.adj, sub, nthreads=NTH)
    assert rel_sum(bc.sum(), ds.sum()) <= 1e-9
    say("rmat25: batched 64 sources + checksum done")
    # batched vs plain vs the CPU oracle on two sources
    two = src[:2]
    e.reset()
    e.run(two, mode=bcmi.MODE_BATCHED)
    bb = e.bc()
    e.reset()
    e.run(two, mode=bcmi.MODE_PLAIN)
    bp = e.bc()
    ref = O.brandes(g.off, g.adj, two, nthreads=2)
    assert bcmi.rel_err(bb, ref) <= 1e-9
    assert bcmi.rel_err(bp, ref) <= 1e-9
    say("rmat25: two sources vs oracle done")
    e.close()


def test_chain_32768_blocks():
    """The 8.4M-vertex bridged-block chain (diameter past 30,000, path counts
    past 2^64 and past binary64's range): 256 sources through workgroup
    sweeps, their BC total against the oracle's distance checksum, and 16 of
    them vertex by vertex against the oracle's scaled-count Brandes."""
    t = time.time()
    g = bcmi.Graph.block_chain(32768, 256, 8.0, seed=SEED)
    src = g.sample_sources(1024, seed=SEED + 1)
    say(f"chain32768: n={g.n} arcs={g.m} built in {time.time() - t:.1f}s")
    e = bcmi.Engine(wg_mem=48 << 30)
    e.upload(g)
    sub = src[:256]
    e.reset()
    t = time.time()
    st = e.run(sub, mode=bcmi.MODE_BATCHED)
    say(f"chain32768: 256 sources in {time.time() - t:.1f}s, {st['max_levels']} levels")
    assert st["max_levels"] > 30000
    assert st["wg_sources"] == len(sub)                # deep graph: a workgroup per source
    assert st["scaled_batches"] == len(sub)            # path counts beyond 2^64
    bc = e.bc()
    assert np.all(np.isfinite(bc)) and np.all(bc >= 0)
    ds = O.depth_sums(g.off, g.adj, sub, nthreads=NTH)
    assert rel_sum(bc.sum(), ds.sum()) <= 1e-9
    few = sub[:16]
    e.reset()
    e.run(few, mode=bcmi.MODE_BATCHED)
    t = time.time()
    ref = O.brandes_scaled(g.off, g.adj, few, nthreads=NTH)
    say(f"chain32768: scaled-count oracle on 16 sources in {time.time() - t:.1f}s")
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    e.close()
