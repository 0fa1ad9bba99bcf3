"""Serving: continuous batching over a pool of graphs.  Each query's result
must equal the same query run alone (1e-9 relative, the FP64 bound; sums of
partials split across work items may associate differently) and the oracle."""
import threading
import time

import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_pool():
    gs = [bcmi.Graph.rmat(9, 8, seed=1), bcmi.Graph.rmat(11, 16, seed=2),
          bcmi.Graph.block_chain(20, 64, 8.0, seed=3), bcmi.Graph.rmat(10, 32, seed=4)]
    return gs


class Server:
    def __init__(self, budget, slots):
        L = bcmi.lib()
        import ctypes
        self._c = ctypes
        self.h = ctypes.c_void_p()
        bcmi._check(L.bcmi_server_create(0, int(budget), int(slots), ctypes.byref(self.h)))

    def add(self, g):
        gid = self._c.c_int32()
        bcmi._check(bcmi.lib().bcmi_server_add_graph(self.h, g.n, bcmi._p(g.off), bcmi._p(g.adj), self._c.byref(gid)))
        return gid.value

    def submit(self, gid, src):
        src = np.ascontiguousarray(src, np.int32)
        q = self._c.c_int64()
        bcmi._check(bcmi.lib().bcmi_server_submit(self.h, gid, bcmi._p(src), len(src), self._c.byref(q)))
        return q.value

    def wait(self, q, n):
        out = np.zeros(n, np.float64)
        res = bcmi.QueryResult()
        rc = bcmi.lib().bcmi_server_wait(self.h, q, bcmi._p(out), self._c.byref(res))
        return rc, out, res

    def stats(self):
        a, b, c = self._c.c_int64(), self._c.c_int64(), self._c.c_int64()
        bcmi._check(bcmi.lib().bcmi_server_stats(self.h, self._c.byref(a), self._c.byref(b), self._c.byref(c)))
        return a.value, b.value, c.value

    def close(self):
        bcmi.lib().bcmi_server_destroy(self.h)


def test_alone_equals_mixed():
    gs = make_pool()
    rng = np.random.default_rng(7)
    queries = []
    for i in range(24):
        gi = int(rng.integers(0, len(gs)))
        k = int(rng.integers(16, 200))
        queries.append((gi, gs[gi].sample_sources(k, seed=100 + i)))
    # alone: one slot, one query at a time
    s1 = Server(8 << 30, 1)
    ids1 = [s1.add(g) for g in gs]
    alone = []
    for gi, src in queries:
        q = s1.submit(ids1[gi], src)
        rc, bc, _ = s1.wait(q, gs[gi].n)
        assert rc == 0
        alone.append(bc)
    s1.close()
    # mixed: all submitted at once into 4 slots
    s4 = Server(8 << 30, 4)
    ids4 = [s4.add(g) for g in gs]
    qs = [s4.submit(ids4[gi], src) for gi, src in queries]
    for (gi, src), q, a in zip(queries, qs, alone):
        rc, bc, res = s4.wait(q, gs[gi].n)
        assert rc == 0, bcmi.lib().bcmi_last_error()
        assert bcmi.rel_err(bc, a) <= 1e-9
        assert res.finish_ms >= res.start_ms >= res.arrive_ms
    done, peak_slots, peak_bytes = s4.stats()
    assert done == len(queries) and peak_slots >= 2
    s4.close()
    # and against the oracle
    for (gi, src), a in list(zip(queries, alone))[:6]:
        g = gs[gi]
        assert bcmi.rel_err(a, O.brandes(g.off, g.adj, src, nthreads=4)) <= 1e-9


def test_budget_admission_and_rejection():
    gs = make_pool()
    # budget fits the graphs and roughly one workspace at a time
    big = gs[1]
    budget = big.n * 1300 + (64 << 20)
    s = Server(budget, 4)
    ids = [s.add(g) for g in gs]
    qs = [s.submit(ids[i % len(gs)], gs[i % len(gs)].sample_sources(20, seed=i)) for i in range(12)]
    for i, q in enumerate(qs):
        rc, bc, res = s.wait(q, gs[i % len(gs)].n)
        assert rc == 0, bcmi.lib().bcmi_last_error()
    done, peak_slots, peak_bytes = s.stats()
    assert peak_bytes <= budget
    s.close()
    # a budget too small for any workspace rejects the query instead of hanging
    s = Server(300 << 10, 2)   # the 512-vertex graph fits, its ~0.7 MB workspace does not
    gid = s.add(gs[0])
    q = s.submit(gid, [0, 1, 2])
    rc, _, _ = s.wait(q, gs[0].n)
    assert rc == bcmi.ERR_NOMEM
    s.close()


def test_allocation_failure_with_nothing_running():
    """The budget admits the query but the device cannot hold its workspace
    (the test holds nearly all device memory itself).  With no other query
    running, nothing will free memory, so the query must complete with
    BCMI_ERR_NOMEM instead of waiting; once the memory is back, the same
    query runs."""
    import ctypes
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    g = bcmi.Graph.rmat(18, 16, seed=11)
    src = g.sample_sources(64, seed=2)
    s = Server(1 << 40, 2)
    gid = s.add(g)
    held = []
    try:
        # leave ~64 MiB: the 268 MB of path-count and coefficient rows cannot fit
        for size in (8 << 30, 1 << 30, 128 << 20, 16 << 20):
            while free_bytes() - size >= (64 << 20):
                p = ctypes.c_void_p()
                if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(size)) != 0:
                    hip.hipGetLastError()
                    break
                held.append(p)
        print(f"holding {sum(1 for _ in held)} blocks, {free_bytes() >> 20} MiB free", flush=True)
        rc, _, res = s.wait(s.submit(gid, src), g.n)
        assert rc == bcmi.ERR_NOMEM, rc
        assert "allocation" in bcmi.lib().bcmi_last_error().decode()
    finally:
        for p in held:
            hip.hipFree(p)
    rc, bc, _ = s.wait(s.submit(gid, src), g.n)
    assert rc == 0, bcmi.lib().bcmi_last_error()
    assert bcmi.rel_err(bc, O.brandes(g.off, g.adj, src, nthreads=4)) <= 1e-9
    s.close()
