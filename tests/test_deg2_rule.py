"""The 2-degree rule on the CPU: for a vertex v0 of degree 2 with neighbours
a and b, every t != v0 has d0(t) = 1 + min(da(t), db(t)) and
sigma0(t) = sum of sigma_x(t) over the sides x with dx(t) = d0(t) - 1.
Checked against the oracle's own forward sweep from v0 on graphs full of
degree-2 vertices; the planted variants (a side counted when it is not
minimal, the max instead of the min) must fail somewhere."""
import os

import numpy as np
import pytest

from oracle import oracle as O
import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def subdivided(edges, n):
    """Every edge u-v becomes u-x-v with a new middle vertex x."""
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(np.sort(e, axis=1), axis=0)
    mid = n + np.arange(e.shape[0])
    return np.concatenate([np.stack([e[:, 0], mid], 1), np.stack([mid, e[:, 1]], 1)]), n + e.shape[0]


def combine(da, sa, db, sb, v0, rule="min"):
    if db is None:  # degree 1: every path leaves through the one neighbour
        d0 = np.where(da >= 0, da + 1, -1).astype(np.int32)
        s0 = np.where(da >= 0, sa, 0).astype(np.uint64)
        d0[v0] = 0
        s0[v0] = 1
        return d0, s0
    ra, rb = da >= 0, db >= 0
    big = np.iinfo(np.int32).max
    xa = np.where(ra, da, big)
    xb = np.where(rb, db, big)
    m = np.minimum(xa, xb) if rule == "min" else np.maximum(xa, xb)
    reached = ra | rb
    d0 = np.where(reached, m + 1, -1).astype(np.int32)
    if rule == "both":
        s0 = np.where(ra, sa, 0) + np.where(rb, sb, 0)
    else:
        s0 = np.where(xa == m, sa, 0).astype(np.uint64) + np.where(xb == m, sb, 0).astype(np.uint64)
    s0 = np.where(reached, s0, 0).astype(np.uint64)
    d0[v0] = 0
    s0[v0] = 1
    return d0, s0


def degree2(off, adj):
    deg = np.diff(off)
    v = np.nonzero(deg == 2)[0]
    return [(int(x), int(adj[off[x]]), int(adj[off[x] + 1])) for x in v]


def graphs():
    fb = np.load(os.path.join(ROOT, "tests", "golden", "facebook.npz"))
    yield "facebook", O.csr_from_edges(fb["edges"], int(fb["n"]))
    yield "cycle11", O.csr_from_edges(*G.cycle(11))
    yield "grid5x6_sub", O.csr_from_edges(*subdivided(*G.grid(5, 6)))
    yield "rand_sub", O.csr_from_edges(*subdivided(*G.random_graph(60, 120, 7)))
    yield "diamonds", O.csr_from_edges(*G.diamonds(9))
    yield "rand", O.csr_from_edges(*G.random_graph(300, 420, 11))


@pytest.mark.parametrize("name,g", list(graphs()), ids=lambda x: x if isinstance(x, str) else "")
def test_degree1_rule_equals_forward_sweep(name, g):
    """Degree 1: d0(t) = 1 + da(t), sigma0(t) = sigma_a(t) for t != v0."""
    off, adj = g
    deg = np.diff(off)
    leaves = np.nonzero(deg == 1)[0][:100]
    for v0 in leaves:
        a = int(adj[off[v0]])
        d, sg, _, _ = O.sssp(off, adj, int(v0))
        da, sa, _, _ = O.sssp(off, adj, a)
        d0, s0 = combine(da, sa, None, None, int(v0))
        assert np.array_equal(d0, d) and np.array_equal(s0, sg), (name, v0)


@pytest.mark.parametrize("name,g", list(graphs()), ids=lambda x: x if isinstance(x, str) else "")
def test_rule_equals_forward_sweep(name, g):
    off, adj = g
    cases = degree2(off, adj)
    assert cases, name
    cache = {}

    def sweep(s):
        if s not in cache:
            d, sg, _, _ = O.sssp(off, adj, s)
            cache[s] = (d, sg)
        return cache[s]

    for v0, a, b in cases[:200]:
        d, sg = sweep(v0)
        d0, s0 = combine(*sweep(a), *sweep(b), v0)
        assert np.array_equal(d0, d), (name, v0)
        assert np.array_equal(s0, sg), (name, v0)


def test_planted_mistakes_fail():
    fails = {"max": 0, "both": 0}
    for name, (off, adj) in graphs():
        for v0, a, b in degree2(off, adj)[:50]:
            d, sg, _, _ = O.sssp(off, adj, v0)
            da, sa, _, _ = O.sssp(off, adj, a)
            db, sb, _, _ = O.sssp(off, adj, b)
            for rule in fails:
                d0, s0 = combine(da, sa, db, sb, v0, rule)
                if not (np.array_equal(d0, d) and np.array_equal(s0, sg)):
                    fails[rule] += 1
    assert all(v > 0 for v in fails.values()), fails
