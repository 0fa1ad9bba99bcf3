"""GPU parity: the HIP engine against the CPU oracle on seeded graphs.

Depths and sigma must match bit for bit; delta and BC within 1e-9 relative
(the bound the engine promises for FP64 accumulation)."""
import numpy as np
import pytest

import bcmi
from oracle import oracle as O
import graphs as G

pytestmark = pytest.mark.gpu

BC_TOL = 1e-9


def csr(edges, n, directed=False):
    return bcmi.Graph.from_edges(edges, n, directed=directed)


CASES = {
    "path9": G.path(9),
    "star7": G.star(7),
    "cycle10": G.cycle(10),
    "k6": G.complete(6),
    "grid5x7": G.grid(5, 7),
    "cube6": G.hypercube(6),
    "rand60": G.random_graph(60, 150, 3),
    "rand200": G.random_graph(200, 500, 4),
    "forest": G.random_forest_plus(150, 5),
}


def _set_dir(engine, mode):
    # mode: "auto", "td" (never bottom-up), "bu" (always bottom-up)
    a = {"auto": 1000, "td": 1 << 60, "bu": 0}[mode]
    engine.set_param("alpha_plain", a)
    engine.set_param("alpha_batched", a)


@pytest.fixture(autouse=True)
def _restore(engine):
    yield
    _set_dir(engine, "auto")


@pytest.mark.parametrize("direction", ["td", "bu", "auto"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_plain_sssp_exact(engine, name, direction):
    e, n = CASES[name]
    g = csr(e, n)
    engine.upload(g)
    _set_dir(engine, direction)
    for s in range(0, n, max(1, n // 7)):
        d, sg, dl = engine.sssp(s)
        od, osg, odl, _ = O.sssp(g.off, g.adj, s)
        assert np.array_equal(d, od), (name, s)
        assert np.array_equal(sg, osg), (name, s)
        odl[s] = 0.0
        assert bcmi.rel_err(dl, odl) <= 1e-12, (name, s)


@pytest.mark.parametrize("direction", ["td", "bu", "auto"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_batched_msbfs_exact(engine, name, direction):
    e, n = CASES[name]
    g = csr(e, n)
    engine.upload(g)
    _set_dir(engine, direction)
    src = np.arange(0, n, max(1, n // 64))[:64].astype(np.int32)
    d, sg, dl = engine.msbfs(src)
    for q, s in enumerate(src):
        od, osg, odl, _ = O.sssp(g.off, g.adj, int(s))
        assert np.array_equal(d[q], od), (name, s)
        assert np.array_equal(sg[q], osg), (name, s)
        odl[s] = 0.0
        # lanes sum successors in adjacency order like the oracle
        assert bcmi.rel_err(dl[q], odl) <= 1e-12, (name, s)


@pytest.mark.parametrize("mode", [bcmi.MODE_PLAIN, bcmi.MODE_BATCHED])
@pytest.mark.parametrize("name", sorted(CASES))
def test_bc_all_sources(engine, name, mode):
    e, n = CASES[name]
    g = csr(e, n)
    engine.upload(g)
    engine.reset()
    engine.run(np.arange(n), mode=mode)
    bc = engine.bc()
    ref = O.brandes(g.off, g.adj, np.arange(n))
    assert bcmi.rel_err(bc, ref) <= BC_TOL, name


@pytest.mark.parametrize("mode", [bcmi.MODE_PLAIN, bcmi.MODE_BATCHED])
def test_rmat_scale10_config(engine, mode):
    """R-MAT scale 10, edge factor 16, all 1,024 sources (isolated skipped)."""
    g = bcmi.Graph.rmat(10, 16, seed=20240101)
    engine.upload(g)
    engine.reset()
    st = engine.run(np.arange(g.n), mode=mode)
    bc = engine.bc()
    ref, ost = O.brandes(g.off, g.adj, np.arange(g.n), nthreads=4, stats=True)
    assert bcmi.rel_err(bc, ref) <= BC_TOL
    deg = g.degrees()
    live = deg[np.arange(g.n)] > 0
    assert st["sources_run"] == int(live.sum())
    assert st["sum_m"] == int(ost["m_s"][live].sum())
    assert st["sum_a"] == int(ost["a_s"][live].sum())
    assert st["sum_n"] == int(ost["n_s"][live].sum())


def test_directed_in_arcs(engine):
    """-U: adjacency holds in-neighbours (CSC); BFS walks them like the reference."""
    e, n = G.random_graph(120, 400, 11)
    g = csr(e, n, directed=True)
    engine.upload(g)
    for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
        engine.reset()
        engine.run(np.arange(n), mode=mode)
        ref = O.brandes(g.off, g.adj, np.arange(n))
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL


def test_sigma_exact_near_2_64(engine):
    e, n = G.diamonds(63)
    g = csr(e, n)
    engine.upload(g)
    d, sg, _ = engine.sssp(0)
    assert int(sg[n - 1]) == 1 << 63
    dm, sgm, _ = engine.msbfs(np.array([0], np.int32))
    assert int(sgm[0, n - 1]) == 1 << 63


@pytest.mark.parametrize("mode", [bcmi.MODE_PLAIN, bcmi.MODE_BATCHED])
def test_sigma_overflow_reported(engine, mode):
    e, n = G.diamonds(64)
    g = csr(e, n)
    engine.upload(g)
    engine.reset()
    engine.set_param("sigma", 0)
    with pytest.raises(bcmi.BcmiError) as ex:
        engine.run(np.array([0], np.int32), mode=mode)
    engine.set_param("sigma", 2)
    assert ex.value.code == bcmi.ERR_OVERFLOW
    assert np.all(engine.bc() == 0)
    # the engine stays usable
    engine.reset()
    engine.run(np.array([1], np.int32), mode=mode)


def test_duplicate_sources_count_twice(engine):
    e, n = G.random_graph(80, 200, 9)
    g = csr(e, n)
    engine.upload(g)
    src = np.array([3, 3, 5, 7, 7, 7], np.int32)
    ref = O.brandes(g.off, g.adj, src)
    for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
        engine.reset()
        engine.run(src, mode=mode)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL


def test_batch_width_and_many_batches(engine):
    g = bcmi.Graph.rmat(11, 8, seed=5)
    engine.upload(g)
    src = g.sample_sources(300, seed=2)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    for width in (64, 17, 1):
        engine.set_param("width", width)
        engine.reset()
        engine.run(src, mode=bcmi.MODE_BATCHED)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL, width
    engine.set_param("width", 64)


@pytest.mark.parametrize("batching", [0, 1, 2, 3, 4, 5])
def test_batch_orders_and_tuned_shape(batching):
    """Every source order (as given, hub depths, highest and second
    highest-degree neighbours) with a large-graph launch shape (forward grid
    2048, backward grid 16384, records of 512 arcs, switch thresholds 1/8 and
    1/128, pushes below 1/4 of the arcs): BC within 1e-9 of the oracle; the
    order changes only how the sources are grouped."""
    g = bcmi.Graph.rmat(12, 16, seed=31)
    src = g.sample_sources(700, seed=5)
    ref = O.brandes(g.off, g.adj, src, nthreads=8)
    e = bcmi.Engine(batching=batching, grid=2048, grid_back=16384, alpha_batched=125, beta_batched=128,
                    gamma_push=4, steps=10, chunk=512)
    e.upload(g)
    st = e.run(src, mode=bcmi.MODE_BATCHED)
    assert st["batches"] == (len(src) + 63) // 64
    assert bcmi.rel_err(e.bc(), ref) <= BC_TOL
    e.close()


def test_small_chunks_split_hubs(engine):
    """Chunk 64 forces multi-record vertices (atomic partial sums)."""
    e2 = bcmi.Engine(chunk=64)
    g = bcmi.Graph.rmat(11, 16, seed=8)
    e2.upload(g)
    src = g.sample_sources(128, seed=4)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
        e2.reset()
        e2.run(src, mode=mode)
        assert bcmi.rel_err(e2.bc(), ref) <= BC_TOL
    d, sg, dl = e2.msbfs(src[:64])
    for q in range(0, 64, 9):
        od, osg, odl, _ = O.sssp(g.off, g.adj, int(src[q]))
        assert np.array_equal(d[q], od) and np.array_equal(sg[q], osg)
    e2.close()


def test_deep_chain_many_graph_replays(engine):
    """Diameter far above the levels captured per graph (device-driven loop)."""
    g = bcmi.Graph.block_chain(40, 64, 6.0, seed=3)
    engine.upload(g)
    src = np.array([0, 5, g.n // 2, g.n - 1], np.int32)
    ref = O.brandes(g.off, g.adj, src)
    for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
        engine.reset()
        st = engine.run(src, mode=mode)
        assert st["max_levels"] > 16
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL


def test_kernel_timer(engine):
    g = bcmi.Graph.rmat(12, 16, seed=3)
    engine.upload(g)
    src = g.sample_sources(128, seed=1)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    for kt in (1, 2, 3):
        engine.set_param("ktime", kt)
        engine.reset()
        st = engine.run(src, mode=bcmi.MODE_BATCHED)
        ms, launches = engine.kernel_time()
        assert launches >= st["batches"] * 2 and 0 < ms < st["wall_ms"], (kt, ms, launches)
        assert bcmi.rel_err(engine.bc(), ref) <= BC_TOL
    engine.set_param("ktime", 7)
    engine.reset()
    engine.run(src[:8], mode=bcmi.MODE_PLAIN)
    ms, launches = engine.kernel_time()
    assert launches > 0 and ms > 0
    engine.set_param("ktime", 0)


def test_level_loop_without_graphs():
    """graphs = 0 launches the captured graphs' kernels one by one: depths
    and path counts equal the captured run's bit for bit and dependencies
    within 1e-12 (split hubs add their partial sums atomically), one batch and
    one plain sweep; BC matches the oracle in both modes, and the device timer
    still sees the timed kernel."""
    g = bcmi.Graph.rmat(12, 16, seed=3)
    src = g.sample_sources(160, seed=5)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    out = {}
    for graphs in (1, 0):
        e = bcmi.Engine(graphs=graphs, ktime=2)
        assert e.get_param("graphs") == graphs
        e.upload(g)
        out[graphs] = (e.msbfs(src[:64]), e.sssp(int(src[0])))
        for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
            e.reset()
            e.run(src, mode=mode)
            assert bcmi.rel_err(e.bc(), ref) <= BC_TOL, (graphs, mode)
        ms, launches = e.kernel_time()
        assert launches > 0 and ms > 0, graphs
        e.close()
    for run1, run0 in zip(out[1], out[0]):
        assert np.array_equal(run1[0], run0[0]) and np.array_equal(run1[1], run0[1])
        assert np.allclose(run1[2], run0[2], rtol=1e-12, atol=0)


def test_relabel_off_and_on_agree():
    """Degree relabelling is internal: ids at the ABI are the caller's."""
    g = bcmi.Graph.rmat(11, 16, seed=21)
    src = g.sample_sources(100, seed=3)
    ref = O.brandes(g.off, g.adj, src, nthreads=4)
    for reorder in (0, 1):
        e2 = bcmi.Engine(reorder=reorder)
        e2.upload(g)
        for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
            e2.reset()
            e2.run(src, mode=mode)
            assert bcmi.rel_err(e2.bc(), ref) <= BC_TOL, (reorder, mode)
        d, sg, dl = e2.sssp(int(src[0]))
        od, osg, odl, _ = O.sssp(g.off, g.adj, int(src[0]))
        assert np.array_equal(d, od) and np.array_equal(sg, osg)
        dm, sgm, dlm = e2.msbfs(src[:8])
        for q in range(8):
            od, osg, odl, _ = O.sssp(g.off, g.adj, int(src[q]))
            odl[src[q]] = 0.0
            assert np.array_equal(dm[q], od) and np.array_equal(sgm[q], osg)
            assert bcmi.rel_err(dlm[q], odl) <= 1e-12
        e2.close()


@pytest.mark.parametrize("reorder", [0, 1])
def test_batched_delta_bit_exact_when_no_vertex_is_split(reorder):
    """With every degree <= the record size, each lane sums its successor
    terms in adjacency order with the oracle's operations (no FP contraction),
    so delta matches the oracle bit for bit."""
    e2 = bcmi.Engine(reorder=reorder)
    for g in (bcmi.Graph.block_chain(30, 128, 8.0, seed=4), bcmi.Graph.from_edges(*G.random_graph(3000, 12000, 5))):
        if g.degrees().max() > 256:
            continue
        e2.upload(g)
        src = g.sample_sources(64, seed=9)
        dm, sgm, dlm = e2.msbfs(src)
        for q in range(64):
            od, osg, odl, _ = O.sssp(g.off, g.adj, int(src[q]))
            odl[src[q]] = 0.0
            assert np.array_equal(dm[q], od) and np.array_equal(sgm[q], osg)
            assert np.array_equal(dlm[q], odl), q
    e2.close()


@pytest.mark.parametrize("strategy", [1, 2])
@pytest.mark.parametrize("mode", [bcmi.MODE_PLAIN, bcmi.MODE_BATCHED])
def test_degenerate_inputs(mode, strategy):
    """No edges, no sources, degree-0 sources only, a single edge, and a
    star with peeling: BC is what the oracle says (zero where no pair has an
    inner vertex), in every engine."""
    e = bcmi.Engine(strategy=strategy)
    g = bcmi.Graph.from_edges(np.zeros((0, 2), np.int32), 5)
    e.upload(g)
    e.reset()
    st = e.run(np.arange(5, dtype=np.int32), mode=mode)
    assert st["sources_run"] == 0 and np.all(e.bc() == 0)
    k2 = bcmi.Graph.from_edges([(0, 1)], 3)
    for h in (bcmi.H_NONE, bcmi.H_PEEL):
        e.upload(k2, heuristic=h)
        e.reset()
        e.run(np.array([], np.int32), mode=mode)
        assert np.all(e.bc() == 0)
        e.run(np.array([0, 1, 2, 1], np.int32), mode=mode)
        assert np.all(e.bc() == 0)
    star = bcmi.Graph.from_edges(*G.star(9))
    for h in (bcmi.H_NONE, bcmi.H_PEEL):
        e.upload(star, heuristic=h)
        e.reset()
        e.run(np.arange(star.n, dtype=np.int32), mode=mode)
        ref = O.brandes(star.off, star.adj, np.arange(star.n))
        assert bcmi.rel_err(e.bc(), ref) <= BC_TOL
        assert e.bc()[0] == 9 * 8                      # ordered pairs through the centre
    e.close()


@pytest.mark.parametrize("sigma", [2, 1])
@pytest.mark.parametrize("gamma_back,fuse_bu,chunk,pack_bu", [(0, 1, 256, 1), (0, 0, 64, 1), (2, 2, 256, 0),
                                                              (8, 2, 64, 1), (1 << 40, 1, 64, 0), (0, 2, 64, 0)])
def test_batched_level_variants(gamma_back, fuse_bu, chunk, pack_bu, sigma):
    """Bottom-up fusion (none, unsplit vertices, all: split vertices add
    partial sums into rows kept zero between sweeps) and pushed backward
    levels (gamma_back; 2^40 pushes every level with a successor level),
    bottom-up with and without the packed low-degree kernel, with records of
    64 and 256 arcs, exact and scaled path counts: BC within 1e-9 of the
    oracle, twice in a row on the same engine."""
    graphs = [bcmi.Graph.rmat(12, 16, seed=3), bcmi.Graph.block_chain(40, 128, 6.0, seed=5)]
    for g in graphs:
        src = g.sample_sources(200, seed=11)
        ref = O.brandes(g.off, g.adj, src, nthreads=8)
        for h in (bcmi.H_NONE, bcmi.H_PEEL):
            e = bcmi.Engine(gamma_back=gamma_back, fuse_bu=fuse_bu, chunk=chunk, sigma=sigma, pack_bu=pack_bu)
            e.upload(g, heuristic=h)
            for _ in range(2):
                e.reset()
                e.run(src, mode=bcmi.MODE_BATCHED)
                assert bcmi.rel_err(e.bc(), ref) <= BC_TOL
            e.close()


@pytest.mark.parametrize("d", [13, 9])
@pytest.mark.parametrize("narrow", [2, 1, 0])
def test_batched_counts_past_narrow_rows(narrow, d):
    """Hypercubes Q13 and Q9: sigma(s -> v) = popcount(s ^ v)!, with 13! >
    2^32 - 1 and 2^16 - 1 < 9! < 2^32.  With 16-bit (narrow_sigma 2) or
    32-bit rows (1) first, a batch whose counts pass the rows' range is
    redone with the next wider rows (Q13 from 16 bits: twice): depths and
    path counts still match the closed form exactly, BC matches the oracle,
    and a later run on the same graph agrees."""
    from math import factorial
    g = bcmi.Graph.from_edges(*G.hypercube(d))
    src = np.array([0, 5, 1234, 8191, 4096, 77], np.int32) % g.n
    e = bcmi.Engine(narrow_sigma=narrow)
    e.upload(g)
    dm, sgm, _ = e.msbfs(src)
    v = np.arange(g.n)
    fact = np.array([factorial(k) for k in range(d + 1)], np.uint64)
    for q, s in enumerate(src):
        pc = np.array([bin(x).count("1") for x in (v ^ s)])
        assert np.array_equal(dm[q], pc)
        assert np.array_equal(sgm[q], fact[pc])
    ref = O.brandes(g.off, g.adj, src, nthreads=8)
    for _ in range(2):
        e.reset()
        e.run(src, mode=bcmi.MODE_BATCHED)
        assert bcmi.rel_err(e.bc(), ref) <= BC_TOL
    e.close()


@pytest.mark.parametrize("fuse_bu,pack_bu", [(2, 1), (1, 0)])
def test_sixteen_bit_rows(fuse_bu, pack_bu):
    """16-bit path-count rows first (narrow_sigma 2) on graphs whose counts fit
    them: records of 64 arcs split the hubs (their partial sums are added
    atomically through 32-bit words), top-down and bottom-up levels: depths and
    path counts bit-exact, BC within 1e-9 of the oracle, twice in a row."""
    g = bcmi.Graph.rmat(12, 16, seed=3)
    src = g.sample_sources(200, seed=11)
    ref = O.brandes(g.off, g.adj, src, nthreads=8)
    e = bcmi.Engine(narrow_sigma=2, chunk=64, fuse_bu=fuse_bu, pack_bu=pack_bu)
    e.upload(g)
    d, sg, _ = e.msbfs(src[:64])
    for q in range(0, 64, 7):
        od, osg, _, _ = O.sssp(g.off, g.adj, int(src[q]))
        assert int(osg.max()) < 1 << 16
        assert np.array_equal(d[q], od) and np.array_equal(sg[q], osg), q
    for _ in range(2):
        e.reset()
        e.run(src, mode=bcmi.MODE_BATCHED)
        assert bcmi.rel_err(e.bc(), ref) <= BC_TOL
    e.close()


def test_batch_order_reproduces_the_batches(engine):
    """bcmi_batch_order lists the sources as a batched run groups them
    (degree-0 sources last): run in that order with "batching" 0 they form
    the same batches (same (vertex, level) pairs) and give the same BC, also
    when split over two calls of whole batches."""
    g = bcmi.Graph.rmat(12, 16, seed=5)
    iso = np.flatnonzero(g.degrees() == 0)[:5].astype(np.int32)
    src = np.concatenate([g.sample_sources(300, seed=6), iso]).astype(np.int32)
    engine.upload(g)
    order = engine.batch_order(src)
    assert sorted(order.tolist()) == sorted(src.tolist())
    assert set(order[-len(iso):].tolist()) == set(iso.tolist())
    engine.reset()
    st = engine.run(src, mode=bcmi.MODE_BATCHED)
    full = engine.bc()
    engine.set_param("batching", 0)
    try:
        engine.reset()
        st2 = engine.run(order[:300], mode=bcmi.MODE_BATCHED)
        assert st2["sum_pairs"] == st["sum_pairs"] and st2["batches"] == st["batches"]
        assert bcmi.rel_err(engine.bc(), full) <= 1e-9
        engine.reset()
        engine.run(order[:128], mode=bcmi.MODE_BATCHED)
        engine.run(order[128:300], mode=bcmi.MODE_BATCHED)
        assert bcmi.rel_err(engine.bc(), full) <= 1e-9
        assert bcmi.rel_err(full, O.brandes(g.off, g.adj, src, nthreads=4)) <= 1e-9
    finally:
        engine.set_param("batching", 4)
    h = bcmi.Graph.rmat(9, 8, seed=7)
    engine.upload(h, heuristic=bcmi.H_PEEL)
    with pytest.raises(bcmi.BcmiError):
        engine.batch_order(h.sample_sources(8, seed=1))
