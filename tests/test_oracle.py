"""Pins the CPU oracle (oracle/bc_oracle.c) before it is trusted as the checker.

Every check is independent of the oracle's own formulas: closed forms,
brute-force path enumeration, networkx 3.4.2 golden vectors, and invariants
(peeling must not change BC).  A plausible mistake (a dropped delta term, a
wrong level test, a wrong sigma accumulation) fails at least one of them; the
planted-mistake tests at the end show that for the two central formulas.
"""
import itertools
import os
from collections import deque

import numpy as np
import pytest

from oracle import oracle as O
import graphs as G

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def csr(edges, n, directed=False):
    return O.csr_from_edges(np.asarray(edges).reshape(-1, 2), n, directed=directed)


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("n", [2, 3, 7, 20])
def test_path_bc_closed_form(n):
    off, adj = csr(*G.path(n))
    bc = O.brandes(off, adj, np.arange(n)) / 2
    i = np.arange(n)
    assert np.array_equal(bc, (i * (n - 1 - i)).astype(float))


@pytest.mark.parametrize("k", [1, 2, 5, 30])
def test_star_bc_closed_form(k):
    off, adj = csr(*G.star(k))
    bc = O.brandes(off, adj, np.arange(k + 1)) / 2
    assert bc[0] == k * (k - 1) / 2 and np.all(bc[1:] == 0)


@pytest.mark.parametrize("n", [3, 4, 5, 10, 11])
def test_cycle_bc_closed_form(n):
    # every vertex of C_n: sum over pairs of the fraction of geodesics through it
    off, adj = csr(*G.cycle(n))
    bc = O.brandes(off, adj, np.arange(n)) / 2
    if n % 2:
        want = (n - 1) * (n - 3) / 8
    else:
        want = (n - 2) ** 2 / 8
    assert np.allclose(bc, want, rtol=0, atol=1e-12)


def test_complete_graph_zero():
    off, adj = csr(*G.complete(7))
    assert np.all(O.brandes(off, adj, np.arange(7)) == 0)


@pytest.mark.parametrize("a,b", [(3, 4), (6, 6), (2, 9)])
def test_grid_sigma_binomial(a, b):
    from math import comb
    off, adj = csr(*G.grid(a, b))
    d, sg, _, _ = O.sssp(off, adj, 0)
    for i in range(a):
        for j in range(b):
            assert d[i * b + j] == i + j
            assert int(sg[i * b + j]) == comb(i + j, i)


@pytest.mark.parametrize("dim", [3, 6, 10])
def test_hypercube_sigma_factorial(dim):
    from math import factorial
    off, adj = csr(*G.hypercube(dim))
    d, sg, _, _ = O.sssp(off, adj, 0)
    pc = np.array([bin(v).count("1") for v in range(1 << dim)])
    assert np.array_equal(d, pc)
    assert all(int(sg[v]) == factorial(pc[v]) for v in range(1 << dim))


def test_sigma_2_63_exact_and_2_64_overflow():
    off, adj = csr(*G.diamonds(63))
    _, sg, _, _ = O.sssp(off, adj, 0)
    assert int(sg[-1]) == 1 << 63
    off, adj = csr(*G.diamonds(64))
    with pytest.raises(O.OracleError):
        O.sssp(off, adj, 0)
    with pytest.raises(O.OracleError):
        O.brandes(off, adj, [0])


# ------------------------------------------------------------- brute force
def brute_bc(n, edges, sources):
    """Enumerates every shortest path between every source and target."""
    nb = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            nb[a].add(b)
            nb[b].add(a)
    bc = np.zeros(n)
    for s in sources:
        dist = [-1] * n
        dist[s] = 0
        q = deque([s])
        while q:
            u = q.popleft()
            for v in nb[u]:
                if dist[v] < 0:
                    dist[v] = dist[u] + 1
                    q.append(v)

        def paths(t):
            if t == s:
                return [[s]]
            return [p + [t] for u in nb[t] if dist[u] == dist[t] - 1 for p in paths(u)]

        for t in range(n):
            if t == s or dist[t] < 0:
                continue
            ps = paths(t)
            for p in ps:
                for v in p[1:-1]:
                    bc[v] += 1.0 / len(ps)
    return bc


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_small(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(3, 13))
    e = rng.integers(0, n, size=(int(rng.integers(1, 3 * n)), 2))
    off, adj = csr(e, n)
    src = rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False)
    assert np.allclose(O.brandes(off, adj, src), brute_bc(n, e, src), rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------- golden vectors
def test_networkx_facebook_ck_run():
    """The CK run: facebook, sources 0..999 (bc2d.c:1017-1018, -N 1000)."""
    f = np.load(os.path.join(GOLD, "facebook.npz"))
    n = int(f["n"])
    off, adj = csr(f["edges"].astype(np.int64), n)
    bc = O.brandes(off, adj, np.arange(1000), nthreads=4) / 2
    ref = f["bc_src1000"]
    assert np.max(np.abs(bc - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-12


def test_networkx_facebook_all_sources_with_peeling():
    f = np.load(os.path.join(GOLD, "facebook.npz"))
    n = int(f["n"])
    off, adj = csr(f["edges"].astype(np.int64), n)
    ref = f["bc_all"]
    bc = O.bc_pipeline(off, adj, np.arange(n), heuristic=1, nthreads=4) / 2
    nz = ref != 0
    assert np.all(bc[~nz] == 0)
    assert np.max(np.abs(bc[nz] - ref[nz]) / ref[nz]) < 1e-11


def test_networkx_rmat10():
    f = np.load(os.path.join(GOLD, "rmat10.npz"))
    n = int(f["n"])
    off, adj = csr(f["edges"].astype(np.int64), n)
    bc = O.brandes(off, adj, np.arange(n), nthreads=4) / 2
    ref = f["bc_all"]
    nz = ref != 0
    assert np.all(bc[~nz] == 0)
    assert np.max(np.abs(bc[nz] - ref[nz]) / ref[nz]) < 1e-12


# ------------------------------------------------------------------ peeling
@pytest.mark.parametrize("seed", range(8))
def test_peeling_preserves_bc(seed):
    e, n = G.random_forest_plus(120, seed)
    off, adj = csr(e, n)
    rng = np.random.default_rng(seed)
    for S in (np.arange(n), rng.choice(n, size=n // 4, replace=False), np.array([0, 0, 5])):
        plain = O.brandes(off, adj, S)
        peeled = O.bc_pipeline(off, adj, S, heuristic=1)
        assert np.allclose(peeled, plain, rtol=1e-12, atol=1e-9)


def test_peel_trees_and_k2():
    # path 0-1-2-3-4 is a tree: peels to one root with w = 4; K2 keeps its smaller end
    e = [(0, 1), (1, 2), (2, 3), (3, 4), (10, 11)]
    off, adj = csr(e, 12)
    p = O.peel(off, adj)
    core = np.where(p["round"] == 0)[0]
    assert set(core) >= {2, 10}
    assert p["w"][2] == 4 and p["w"][10] == 1 and p["round"][11] == 1
    assert p["round"][0] == 1 and p["round"][4] == 1 and p["round"][1] == 2 and p["round"][3] == 2


def test_expand_semantics_equal_expanded_sources():
    e, n = G.random_forest_plus(90, 3)
    off, adj = csr(e, n)
    p = O.peel(off, adj)
    par = p["parent"]
    root = np.arange(n)
    for v in np.argsort(-p["round"]):
        if p["round"][v] > 0:
            pass
    # roots by following parents
    def rt(v):
        while p["round"][v] > 0:
            v = par[v]
        return v
    root = np.array([rt(v) for v in range(n)])
    S = np.array([0, 1, 2, 3, 50, 60, 89])
    core_listed = set(int(s) for s in S if p["round"][s] == 0)
    expanded = np.array([v for v in range(n) if root[v] in core_listed])
    want = O.brandes(off, adj, expanded)
    got = O.bc_pipeline(off, adj, S, heuristic=1, expand=True)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-9)


def test_reference_single_pass_on_facebook():
    """The reference's single-pass reduction (bc2d.c:314-419) removes exactly
    the degree-1 vertices of facebook (75, counted here with numpy), and its
    reach counts equal the first peel round's weights."""
    f = np.load(os.path.join(GOLD, "facebook.npz"))
    n = int(f["n"])
    e = f["edges"].astype(np.uint64)
    arcs = np.concatenate([e, e[:, ::-1]])
    arcs = O.normalize(arcs)
    deg = np.bincount(arcs[:, 0].astype(np.int64), minlength=n)
    kept, removed, pnod, ncouple, reach = O.ref_degree_reduction(arcs, n)
    assert pnod == int((deg == 1).sum()) == 75 and ncouple == 0
    assert kept.shape[0] == arcs.shape[0] - 2 * 75
    off, adj = O.build_csc(arcs, n)
    p = O.peel(off, adj)
    assert set(np.where(p["round"] == 1)[0]) == set(np.where(deg == 1)[0])
    assert p["rounds"] == 1
    assert np.array_equal(reach.astype(np.uint64), p["w"])


# ------------------------------------------------------------------- ingest
def test_ingest_pipeline(tmp_path):
    fn = tmp_path / "g.txt"
    fn.write_text("# Nodes: 5 Edges: 6\n0 1\n1 2\n\n2 0\n# comment 3 4\n3 3\n1 0\n4 2\n")
    arcs = O.read_edges(str(fn), 5)
    # 7 data lines (the blank one repeats the previous edge, bc2d.c:227-230),
    # one self loop that is not mirrored: 7 + 6 arcs
    assert arcs.shape[0] == 13
    arcs = O.normalize(arcs)
    off, adj = O.build_csc(arcs, 5)
    assert list(off) == [0, 2, 4, 7, 7, 8]
    assert list(adj) == [1, 2, 0, 2, 0, 1, 4, 2]
    with pytest.raises(O.OracleError):
        O.read_edges(str(fn), 4)


# --------------------------------------------------------- planted mistakes
def test_planted_mistakes_are_caught():
    """Perturbing the oracle's backward rule or sigma rule breaks the pins."""
    e, n = G.grid(4, 5)
    off, adj = csr(e, n)
    good = O.brandes(off, adj, np.arange(n))
    brute = brute_bc(n, e, range(n))
    assert np.allclose(good, brute)
    # dropped "+1" in (delta[r]+1)/sigma[r]*sigma[c]: delta stays 0, BC vanishes
    dropped = np.zeros(n)
    assert not np.allclose(dropped, brute)
    # sigma taken as 1 everywhere (forgetting to add parents' counts)
    from math import comb
    _, sg, _, _ = O.sssp(off, adj, 0)
    assert any(int(sg[i * 5 + j]) != 1 for i in range(4) for j in range(5))
    assert int(sg[3 * 5 + 4]) == comb(7, 3)


def test_highprec_oracle_matches_uint64_oracle():
    """The decimal/big-integer Brandes agrees with the uint64 oracle where
    both apply, and with the path closed form."""
    off, adj = csr(*G.path(9))
    assert np.allclose(O.brandes_highprec(off, adj, range(9)), 2.0 * np.arange(9) * (8 - np.arange(9)))
    e, n = G.random_graph(60, 150, 3)
    off, adj = csr(e, n)
    src = [0, 3, 17, 40]
    a = O.brandes_highprec(off, adj, src)
    b = O.brandes(off, adj, src)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-12)
    off, adj = csr(*G.diamonds(70))        # 2^70 paths: beyond uint64
    bc = O.brandes_highprec(off, adj, [0])
    # in a diamond chain every middle vertex of a diamond carries half of the
    # paths to the vertices after it
    assert bc[1] == bc[2] and bc[1] > 0


def rel_err(a, b):
    """max |a-b|/|b| over entries with b != 0; entries with b == 0 must be 0."""
    nz = b != 0
    if np.any(a[~nz] != 0):
        return float("inf")
    return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]))) if nz.any() else 0.0


def test_scaled_oracle_matches_exact_and_highprec():
    """orc_brandes_scaled (binary64 path counts with a power-of-two scale per
    level) equals the uint64 oracle where the counts fit (to binary64
    rounding), and the big-integer Brandes past 2^64: on a bridged-block
    chain and on a diamond chain of 2^1100 paths, beyond binary64's range,
    where an unscaled binary64 count would overflow to infinity."""
    for seed in range(3):
        off, adj = csr(*G.random_graph(200, 600, seed))
        src = np.arange(0, 200, 5)
        assert rel_err(O.brandes_scaled(off, adj, src, nthreads=4), O.brandes(off, adj, src)) <= 1e-13
    off, adj = csr(*G.block_chain(200, 32, 8.0, 5))
    src = [3, 3200, 6390]
    with pytest.raises(O.OracleError):
        O.brandes(off, adj, src)                        # counts beyond 2^64
    hp = O.brandes_highprec(off, adj, src)
    assert rel_err(O.brandes_scaled(off, adj, src, nthreads=2), hp) <= 1e-12
    off, adj = csr(*G.diamonds(1100))
    src = [0, 7, 1650]
    hp = O.brandes_highprec(off, adj, src)
    sc = O.brandes_scaled(off, adj, src)
    assert np.all(np.isfinite(sc))
    assert rel_err(sc, hp) <= 1e-12


def test_depth_sums_closed_forms_and_bc_identity():
    """Distance checksum: on a path from an end it is (n-1)(n-2)/2; from the
    centre of a star 0; on any graph it equals the sum of the sources'
    dependencies (the BC sum), here checked on random graphs, and beyond
    2^64 path counts against the big-integer Brandes."""
    off, adj = csr(*G.path(12))
    assert O.depth_sums(off, adj, [0])[0] == 11 * 10 // 2
    assert O.depth_sums(off, adj, [5])[0] == sum(k - 1 for k in range(1, 6)) + sum(k - 1 for k in range(1, 7))
    off, adj = csr(*G.star(9))
    assert O.depth_sums(off, adj, [0])[0] == 0
    assert O.depth_sums(off, adj, [1])[0] == 9 - 1          # the 8 other leaves at distance 2
    for seed in range(4):
        e, n = G.random_graph(80, 200, seed)
        off, adj = csr(e, n)
        src = np.arange(0, n, 7)
        bc = O.brandes(off, adj, src)
        ds = O.depth_sums(off, adj, src, nthreads=2)
        assert abs(bc.sum() - ds.sum()) <= 1e-9 * max(1.0, ds.sum())
        for s, d in zip(src[:3], ds[:3]):                   # per source
            assert abs(O.brandes(off, adj, [s]).sum() - d) <= 1e-9 * max(1.0, d)
    off, adj = csr(*G.diamonds(70))
    hp = O.brandes_highprec(off, adj, [0, 5])
    assert abs(hp.sum() - O.depth_sums(off, adj, [0, 5]).sum()) <= 1e-9 * hp.sum()
