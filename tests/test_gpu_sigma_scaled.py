"""Path counts beyond 2^64 (long bridged-block chains): the batched sweep
holds them as FP64 scaled per level.  Checked against a big-integer /
50-digit decimal Brandes; the exact mode must refuse (and leave BC untouched);
on graphs without overflow the scaled mode must agree with the uint64 path."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_chain_beyond_uint64():
    g = bcmi.Graph.block_chain(90, 128, 8.0, seed=6)
    src = np.array([0, 1, 5, g.n // 3, g.n // 2, g.n - 1], np.int32)
    ref = O.brandes_highprec(g.off, g.adj, src)
    e = bcmi.Engine(strategy=1)       # the device-wide level loop (deep graphs default to
    e.upload(g)                       # a workgroup per source, tested on its own)
    e.set_param("sigma", 0)
    e.reset()
    with pytest.raises(bcmi.BcmiError) as ex:
        e.run(src, mode=bcmi.MODE_BATCHED)
    assert ex.value.code == bcmi.ERR_OVERFLOW
    assert np.all(e.bc() == 0)              # the invalid sweep added nothing
    e.set_param("sigma", 2)                 # auto: overflowing batches redone scaled
    e.reset()
    st = e.run(src, mode=bcmi.MODE_BATCHED)
    assert st["scaled_batches"] == 1
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    e.set_param("sigma", 1)
    e.reset()
    e.run(src, mode=bcmi.MODE_BATCHED)
    assert bcmi.rel_err(e.bc(), ref) <= 1e-9
    with pytest.raises(bcmi.BcmiError) as ex:   # the plain sweep is exact-only
        e.run(src[:1], mode=bcmi.MODE_PLAIN)
    assert ex.value.code == bcmi.ERR_OVERFLOW
    e.close()


@pytest.mark.parametrize("peel", [0, 1])
def test_scaled_equals_exact_without_overflow(peel):
    g = bcmi.Graph.rmat(11, 16, seed=31)
    if peel:
        g = g.with_pendant_trees(0.3, seed=2)
    src = g.sample_sources(200, seed=5)
    ref = O.bc_pipeline(g.off, g.adj, src, heuristic=peel, nthreads=4)
    e = bcmi.Engine()
    e.upload(g, heuristic=peel)
    for mode in (0, 1, 2):
        e.set_param("sigma", mode)
        e.reset()
        st = e.run(src, mode=bcmi.MODE_BATCHED)
        assert st["scaled_batches"] == (st["batches"] if mode == 1 else 0)
        assert bcmi.rel_err(e.bc(), ref) <= 1e-9, mode
    e.close()


def test_long_path_many_levels_scale_table_growth():
    """A component deeper than the hubs' (the scale table starts from their
    BFS depth, at least 4096 levels) forces the table to grow."""
    import graphs as G
    r = bcmi.Graph.rmat(10, 16, seed=8)
    pe, pn = G.path(6000)
    edges = np.concatenate([np.stack([np.repeat(np.arange(r.n), np.diff(r.off)), r.adj], 1), pe + r.n])
    g = bcmi.Graph.from_edges(edges, r.n + pn)
    src = np.array([r.n, r.n + pn - 1, r.n + 3000, 5], np.int32)
    ref = O.brandes(g.off, g.adj, src)
    e = bcmi.Engine(strategy=1)
    e.upload(g)
    e.set_param("sigma", 1)
    e.reset()
    st = e.run(src, mode=bcmi.MODE_BATCHED)
    assert st["max_levels"] > 4096
    assert bcmi.rel_err(e.bc(), ref) <= 1e-12
    e.close()
