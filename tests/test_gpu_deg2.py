"""GPU parity of the 2-degree heuristic (BCMI_SRC_DEG2; CLI -H 2 and -H 3):
plain runs whose degree-2 sources take their forward state from their two
neighbours' sweeps give the BC of ordinary plain runs and of the oracle, and
the same reached-vertex, arc and shortest-path-DAG-arc totals (the DAG arcs
of a derived sweep are counted in its backward pass)."""
import os

import numpy as np
import pytest

import bcmi
from oracle import oracle as O
import graphs as G
from test_deg2_rule import subdivided

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BC_TOL = 1e-9


def theta(k, tail):
    """Two hubs joined by k paths of length 2, plus a path of `tail` vertices
    hanging off the first hub: many degree-2 sources with the same neighbours."""
    e = [(0, 2 + i) for i in range(k)] + [(2 + i, 1) for i in range(k)]
    nxt = 2 + k
    prev = 0
    for _ in range(tail):
        e.append((prev, nxt))
        prev = nxt
        nxt += 1
    return np.array(e), nxt


CASES = {
    "cycle12": G.cycle(12),
    "grid6x6_sub": subdivided(*G.grid(6, 6)),
    "rand_sub": subdivided(*G.random_graph(80, 160, 7)),
    "theta": theta(9, 5),
    "rand300": G.random_graph(300, 420, 11),
}


def runs(engine, S, **kw):
    engine.reset()
    st0 = engine.run(S, mode=bcmi.MODE_PLAIN, **kw)
    bc0 = engine.bc()
    engine.reset()
    st1 = engine.run(S, mode=bcmi.MODE_PLAIN, deg2=True, **kw)
    return bc0, st0, engine.bc(), st1


def same_totals(st0, st1):
    for k in ("sources_run", "sum_n", "sum_m", "sum_a", "te"):
        assert st1[k] == st0[k], k


@pytest.mark.parametrize("name", sorted(CASES))
def test_deg2_all_sources(engine, name):
    e, n = CASES[name]
    g = bcmi.Graph.from_edges(e, n)
    engine.upload(g)
    S = np.arange(n, dtype=np.int32)
    bc0, st0, bc1, st1 = runs(engine, S)
    ref = O.brandes(g.off, g.adj, S, nthreads=4)
    assert bcmi.rel_err(bc1, ref) <= BC_TOL, name
    assert bcmi.rel_err(bc1, bc0) <= 1e-12, name
    assert st1["deg2_sources"] > 0, name
    assert st0["deg2_sources"] == 0
    same_totals(st0, st1)


def test_deg2_facebook_all_sources(engine):
    f = np.load(os.path.join(ROOT, "tests", "golden", "facebook.npz"))
    g = bcmi.Graph.from_edges(f["edges"], int(f["n"]))
    engine.upload(g)
    S = np.arange(g.n, dtype=np.int32)
    bc0, st0, bc1, st1 = runs(engine, S)
    assert bcmi.rel_err(bc1, f["bc_all"] * 2) <= BC_TOL   # networkx golden (unordered pairs)
    assert bcmi.rel_err(bc1, bc0) <= 1e-12
    assert st1["deg2_sources"] > 0
    same_totals(st0, st1)


def test_deg2_rmat12_all_sources_and_slots(engine):
    g = bcmi.Graph.rmat(12, 8, seed=1212)
    engine.upload(g)
    S = np.arange(g.n, dtype=np.int32)
    ref = O.brandes(g.off, g.adj, S, nthreads=8)
    derived = []
    try:
        for slots in (8, 2, 0):   # 0: as many slots as the memory budget holds (all anchors here)
            engine.set_param("deg2_slots", slots)
            bc0, st0, bc1, st1 = runs(engine, S)
            assert bcmi.rel_err(bc1, ref) <= BC_TOL, slots
            same_totals(st0, st1)
            derived.append(st1["deg2_sources"])
        engine.set_param("deg2_slots", 1)   # off
        _, _, bc1, st1 = runs(engine, S)
        assert st1["deg2_sources"] == 0
        assert bcmi.rel_err(bc1, ref) <= BC_TOL
    finally:
        engine.set_param("deg2_slots", 0)
    assert min(derived) > 0, derived
    # with every anchor in a slot, every degree-2 source the selection takes is derived
    deg = np.diff(g.off)
    assert derived[2] >= derived[0]
    assert derived[2] >= 0.9 * np.count_nonzero(deg == 2), derived


def test_degree1_sources_from_their_neighbour(engine):
    """Without peeling, a degree-1 source takes its forward state from its one
    neighbour's sweep: every leaf of a star and of a tree of paths."""
    e, n = G.star(9)
    g = bcmi.Graph.from_edges(e, n)
    engine.upload(g)
    S = np.arange(n, dtype=np.int32)
    bc0, st0, bc1, st1 = runs(engine, S)
    assert st1["deg2_sources"] == 9          # the leaves, from the centre's sweep
    assert bcmi.rel_err(bc1, O.brandes(g.off, g.adj, S)) <= BC_TOL
    same_totals(st0, st1)
    e, n = G.random_forest_plus(200, 7)
    g = bcmi.Graph.from_edges(e, n)
    engine.upload(g)
    S = np.arange(n, dtype=np.int32)
    bc0, st0, bc1, st1 = runs(engine, S)
    deg = np.diff(g.off)
    assert st1["deg2_sources"] >= np.count_nonzero(deg == 1) // 3
    assert bcmi.rel_err(bc1, O.brandes(g.off, g.adj, S, nthreads=4)) <= BC_TOL
    same_totals(st0, st1)


def test_deg2_duplicates_and_subsets(engine):
    e, n = subdivided(*G.grid(7, 5))
    g = bcmi.Graph.from_edges(e, n)
    engine.upload(g)
    rng = np.random.default_rng(5)
    for S in (np.concatenate([np.arange(n), np.arange(0, n, 3)]).astype(np.int32),
              rng.choice(n, size=n // 2, replace=False).astype(np.int32)):
        bc0, st0, bc1, st1 = runs(engine, S)
        ref = O.brandes(g.off, g.adj, S, nthreads=4)
        assert bcmi.rel_err(bc1, ref) <= BC_TOL
        same_totals(st0, st1)


@pytest.mark.parametrize("expand", [False, True])
def test_deg2_with_peeling(engine, expand):
    """-H 3: the degree-2 vertices of the peeled core, with subtree weights."""
    g = bcmi.Graph.rmat(11, 4, seed=77).with_pendant_trees(0.3, seed=78)
    engine.upload(g, heuristic=bcmi.H_PEEL)
    rng = np.random.default_rng(9)
    for S in (np.arange(g.n, dtype=np.int32), rng.choice(g.n, size=g.n // 3, replace=False).astype(np.int32)):
        bc0, st0, bc1, st1 = runs(engine, S, expand=expand)
        ref = O.bc_pipeline(g.off, g.adj, S, heuristic=1, expand=expand, nthreads=8)
        assert bcmi.rel_err(bc1, ref) <= BC_TOL
        assert bcmi.rel_err(bc1, bc0) <= 1e-12
        same_totals(st0, st1)
    assert st1["deg2_sources"] > 0 or expand


def test_deg2_overflow_reported(engine):
    """Path counts past 2^64 (exact counts asked for): derived sweeps report
    the overflow like ordinary ones, and the engine stays usable."""
    e, n = G.diamonds(70)
    g = bcmi.Graph.from_edges(e, n)
    # a deep graph: keep the one-workspace level loop, where the heuristic applies
    prm = {"sigma": 0, "strategy": 1, "streams": 1}
    old = {k: engine.get_param(k) for k in prm}
    try:
        for k, v in prm.items():
            engine.set_param(k, v)
        engine.upload(g)
        S = np.arange(n, dtype=np.int32)
        codes = []
        for deg2 in (False, True):
            engine.reset()
            with pytest.raises(bcmi.BcmiError) as ex:
                engine.run(S, mode=bcmi.MODE_PLAIN, deg2=deg2)
            codes.append(ex.value.code)
        assert codes == [bcmi.ERR_OVERFLOW, bcmi.ERR_OVERFLOW]
        # a chain whose counts fit: derived sweeps on the deep level loop
        e, n = G.diamonds(20)
        g = bcmi.Graph.from_edges(e, n)
        engine.upload(g)
        S = np.arange(n, dtype=np.int32)
        bc0, st0, bc1, st1 = runs(engine, S)
        assert st1["deg2_sources"] > 0
        assert bcmi.rel_err(bc1, O.brandes(g.off, g.adj, S)) <= BC_TOL
        same_totals(st0, st1)
    finally:
        for k, v in old.items():
            engine.set_param(k, v)


def test_deg2_ignored_when_batched(engine):
    e, n = subdivided(*G.grid(6, 6))
    g = bcmi.Graph.from_edges(e, n)
    engine.upload(g)
    S = np.arange(n, dtype=np.int32)
    engine.reset()
    st = engine.run(S, mode=bcmi.MODE_BATCHED, deg2=True)
    assert st["deg2_sources"] == 0
    assert bcmi.rel_err(engine.bc(), O.brandes(g.off, g.adj, S, nthreads=4)) <= BC_TOL
