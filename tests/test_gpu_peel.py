"""Degree-1 reduction on the device: the peeled set, parents, subtree weights
and rounds must equal the oracle's bit for bit; BC with peeling must equal the
oracle's (and plain Brandes) within 1e-9, in both source semantics."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O
import graphs as G

pytestmark = pytest.mark.gpu
BC_TOL = 1e-9


def _graphs():
    out = []
    for seed in range(4):
        e, n = G.random_forest_plus(300, seed)
        out.append(("forest%d" % seed, bcmi.Graph.from_edges(e, n)))
    e, n = G.path(33)
    out.append(("path33", bcmi.Graph.from_edges(e, n)))
    e, n = G.star(9)
    out.append(("star9", bcmi.Graph.from_edges(e, n)))
    out.append(("k2s", bcmi.Graph.from_edges([(0, 1), (2, 3), (5, 4), (7, 6), (6, 8)], 10)))
    core = bcmi.Graph.rmat(11, 8, seed=12)
    out.append(("rmat11+trees", core.with_pendant_trees(0.3, seed=5)))
    return out


GRAPHS = _graphs()


@pytest.mark.parametrize("name,g", GRAPHS, ids=[x[0] for x in GRAPHS])
def test_peel_matches_oracle(engine, name, g):
    engine.upload(g, heuristic=bcmi.H_PEEL)
    p = engine.peel_info()
    o = O.peel(g.off, g.adj)
    assert np.array_equal(p["round"], o["round"]), name
    assert np.array_equal(p["w"], o["w"]), name
    rem = o["round"] > 0
    assert np.array_equal(p["parent"][rem], o["parent"][rem]), name
    assert np.all(p["parent"][~rem] == -1)
    assert p["rounds"] == o["rounds"]
    lab, _ = O.components(g.off, g.adj)
    assert np.array_equal(engine.labels(), lab)


@pytest.mark.parametrize("mode", [bcmi.MODE_PLAIN, bcmi.MODE_BATCHED])
@pytest.mark.parametrize("name,g", GRAPHS, ids=[x[0] for x in GRAPHS])
def test_peeled_bc(engine, name, g, mode):
    engine.upload(g, heuristic=bcmi.H_PEEL)
    rng = np.random.default_rng(3)
    for S in (np.arange(g.n), rng.choice(g.n, size=max(1, g.n // 5), replace=False)):
        engine.reset()
        engine.run(S, mode=mode)
        ref = O.bc_pipeline(g.off, g.adj, S, heuristic=1, nthreads=4)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL, name
        plain = O.brandes(g.off, g.adj, S, nthreads=4)
        assert np.allclose(ref, plain, rtol=1e-11, atol=1e-8)
        engine.reset()
        engine.run(S, mode=mode, expand=True)
        ref = O.bc_pipeline(g.off, g.adj, S, heuristic=1, expand=True, nthreads=4)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL, name


def test_pendant_config_scale14(engine):
    """Degree-1 stress (smaller stand-in of R-MAT 18 + 30% pendant trees)."""
    g = bcmi.Graph.rmat(14, 16, seed=18).with_pendant_trees(0.3, seed=19)
    engine.upload(g, heuristic=bcmi.H_PEEL)
    p = engine.peel_info()
    o = O.peel(g.off, g.adj)
    assert np.array_equal(p["round"], o["round"]) and np.array_equal(p["w"], o["w"])
    src = g.sample_sources(512, seed=20)
    for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
        engine.reset()
        engine.run(src, mode=mode)
        ref = O.bc_pipeline(g.off, g.adj, src, heuristic=1, nthreads=8)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL
