"""One workgroup per source (the engine for deep graphs): BC must match the
oracle in both sweep modes, with peeling in both source semantics, on
directed graphs, with path counts beyond 2^64 (scaled FP64, against the
big-integer oracle), when the level table has to grow, and with fewer
workgroups than sources; the exact mode must refuse a wrapped count and
leave BC untouched."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O
import graphs as G

pytestmark = pytest.mark.gpu


def run(g, src, mode=bcmi.MODE_BATCHED, heuristic=bcmi.H_NONE, expand=False, **params):
    e = bcmi.Engine(**params)
    e.upload(g, heuristic=heuristic)
    e.reset()
    st = e.run(src, mode=mode, expand=expand)
    bc = e.bc()
    e.close()
    return bc, st


@pytest.mark.parametrize("mode", [bcmi.MODE_BATCHED, bcmi.MODE_PLAIN])
def test_forced_on_shallow_graphs(mode):
    for g in (bcmi.Graph.rmat(11, 16, seed=3), bcmi.Graph.from_edges(*G.grid(12, 17)),
              bcmi.Graph.from_edges(*G.random_graph(2000, 9000, 4))):
        src = g.sample_sources(min(300, int((g.degrees() > 0).sum())), seed=2)
        ref = O.brandes(g.off, g.adj, src, nthreads=4)
        bc, st = run(g, src, mode=mode, strategy=2)
        assert st["wg_sources"] == len(src)
        assert bcmi.rel_err(bc, ref) <= 1e-9
        lv, st1 = run(g, src, mode=mode, strategy=1)
        assert st1["wg_sources"] == 0
        assert st["sum_m"] == st1["sum_m"] and st["sum_n"] == st1["sum_n"]
        assert bcmi.rel_err(bc, lv) <= 1e-12


def test_directed():
    e, n = G.random_graph(1500, 6000, 9)
    g = bcmi.Graph.from_edges(e, n, directed=True)
    src = np.arange(0, n, 5, dtype=np.int32)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    bc, _ = run(g, src, strategy=2)
    assert bcmi.rel_err(bc, ref) <= 1e-9


@pytest.mark.parametrize("expand", [False, True])
def test_peeling(expand):
    g = bcmi.Graph.block_chain(60, 64, 6.0, seed=5).with_pendant_trees(0.3, seed=6)
    src = g.sample_sources(250, seed=3)
    ref = O.bc_pipeline(g.off, g.adj, src, heuristic=1, expand=expand, nthreads=4)
    bc, st = run(g, src, heuristic=bcmi.H_PEEL, expand=expand, strategy=2)
    assert st["wg_sources"] > 0
    assert bcmi.rel_err(bc, ref) <= 1e-9


def test_deep_chain_auto_and_beyond_uint64():
    g = bcmi.Graph.block_chain(90, 128, 8.0, seed=6)
    src = np.array([0, 1, 5, g.n // 3, g.n // 2, g.n - 1], np.int32)
    ref = O.brandes_highprec(g.off, g.adj, src)
    bc, st = run(g, src)                                  # auto: deep graph, auto sigma
    assert st["wg_sources"] == len(src)
    assert st["scaled_batches"] >= 1
    assert bcmi.rel_err(bc, ref) <= 1e-9
    bc, st = run(g, src, sigma=1)
    assert st["scaled_batches"] == len(src)
    assert bcmi.rel_err(bc, ref) <= 1e-9
    e = bcmi.Engine(sigma=0)
    e.upload(g)
    e.reset()
    with pytest.raises(bcmi.BcmiError) as ex:
        e.run(src)
    assert ex.value.code == bcmi.ERR_OVERFLOW
    assert np.all(e.bc() == 0)
    e.close()


def test_level_table_growth_and_few_slots():
    # the hubs sit in a shallow R-MAT part; a separate long path needs more
    # levels than the initial table
    r = bcmi.Graph.rmat(10, 16, seed=8)
    pe, pn = G.path(5000)
    edges = np.concatenate([np.stack([np.repeat(np.arange(r.n), np.diff(r.off)), r.adj], 1),
                            pe + r.n])
    g = bcmi.Graph.from_edges(edges, r.n + pn)
    src = np.concatenate([r.sample_sources(40, seed=1), r.n + np.array([0, 1, 2500, 4999])]).astype(np.int32)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    bc, st = run(g, src, strategy=2, wg_slots=3)
    assert st["max_levels"] >= 4999
    assert bcmi.rel_err(bc, ref) <= 1e-9
    bc, _ = run(g, src, strategy=2, wg_mem=1 << 20)       # one slot
    assert bcmi.rel_err(bc, ref) <= 1e-9
