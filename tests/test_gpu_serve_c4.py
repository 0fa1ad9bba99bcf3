"""Serving parity on the served-query workload itself: 32 graphs (24 R-MAT
graphs of scale 12-20 and edge factor 8-32, 8 bridged-block chains of
64-2,048 blocks of 256 vertices), 48 queries of 16-1,024 sources with
seeded Poisson arrivals, 8 slots, an HBM budget that admission must respect.

Every query runs twice on one server: alone (submitted after the previous
one finished) and in the mix (Poisson arrivals at 100 queries/s, so up to 8
run at once).  Each mixed result must equal its alone result (1e-9 relative,
the FP64 bound; deep chains add with atomics, whose order changes the last
bits).  The alone results are checked against the CPU oracle: the full
Brandes sum where the query is cheap enough for the CPU (sources x arcs <=
2e9), else a 32-source probe query run on the same server; chains, whose
path counts pass 2^64, through the distance checksum (the BC sum of a
source is the sum over its reached targets of d(s,t) - 1).
"""
import ctypes
import os
import sys
import threading
import time

import numpy as np
import pytest

import bcmi
import serve_pool
from oracle import oracle as O

pytestmark = pytest.mark.gpu
SEED = 20250101
NTH = min(16, os.cpu_count() or 1)
BUDGET = 48 << 30
SLOTS = 8
RATE = 100.0
FULL_ORACLE = 2e9  # sources x arcs below which the whole query is checked by the CPU oracle
PROBE = 32


def say(msg):
    print(f"[serve-c4] {msg}", file=sys.stderr, flush=True)


def _hip():
    return ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")


def free_device_bytes(hip):
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


@pytest.fixture(scope="module")
def c4():
    t = time.time()
    graphs, specs, rng = serve_pool.make_pool(SEED)
    gsel, qsrc, gaps = serve_pool.make_queries(rng, graphs, SEED, 48, RATE)
    say(f"pool of {len(graphs)} graphs, 48 queries built in {time.time() - t:.1f}s")
    hip = _hip()
    bcmi.lib()
    free0 = free_device_bytes(hip)
    srv = bcmi.Server(BUDGET, SLOTS)
    ids = [srv.add_graph(g) for g in graphs]
    say(f"graphs uploaded ({(free0 - free_device_bytes(hip)) / 2**30:.2f} GiB on the device)")
    # alone: one query at a time
    t = time.time()
    alone = []
    for i, (gi, src) in enumerate(zip(gsel, qsrc)):
        rc, bc, res = srv.wait(srv.submit(ids[gi], src), graphs[gi].n)
        assert rc == 0, (i, bcmi.lib().bcmi_last_error())
        alone.append(bc)
    say(f"alone pass: {time.time() - t:.1f}s")
    # probe queries (a prefix of each large R-MAT query), alone too
    probes = {}
    for i, (gi, src) in enumerate(zip(gsel, qsrc)):
        if specs[gi][0] == "rmat" and len(src) * graphs[gi].m > FULL_ORACLE:
            rc, bc, _ = srv.wait(srv.submit(ids[gi], src[:PROBE]), graphs[gi].n)
            assert rc == 0
            probes[i] = bc
    # mixed: Poisson arrivals; device memory sampled meanwhile
    peak_used = [0]
    stop = threading.Event()

    def sampler():
        while not stop.is_set():
            peak_used[0] = max(peak_used[0], free0 - free_device_bytes(hip))
            time.sleep(0.002)

    mixed = [None] * len(qsrc)
    qids = [None] * len(qsrc)

    def waiter():
        for i in range(len(qsrc)):
            while qids[i] is None:
                time.sleep(0.0005)
            mixed[i] = srv.wait(qids[i], graphs[gsel[i]].n)

    th_s = threading.Thread(target=sampler)
    th_w = threading.Thread(target=waiter)
    th_s.start()
    th_w.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(len(qsrc)):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        qids[i] = srv.submit(ids[gsel[i]], qsrc[i])
    th_w.join()
    stop.set()
    th_s.join()
    say(f"mixed pass: {time.perf_counter() - t0:.1f}s, peak device use {peak_used[0] / 2**30:.2f} GiB")
    stats = srv.stats()
    srv.close()
    return dict(graphs=graphs, specs=specs, gsel=gsel, qsrc=qsrc, alone=alone, mixed=mixed, probes=probes,
                stats=stats, peak_used=peak_used[0])


def test_mixed_equals_alone(c4):
    lat = []
    for i, (a, m) in enumerate(zip(c4["alone"], c4["mixed"])):
        rc, bc, res = m
        assert rc == 0, (i, rc)
        assert bcmi.rel_err(bc, a) <= 1e-9, i
        assert res.finish_ms >= res.start_ms >= res.arrive_ms
        lat.append(res.finish_ms - res.arrive_ms)
    st = c4["stats"]
    say(f"stats {st}, p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms")
    # alone + probes + mixed all completed; queries overlapped on the device
    assert st["completed"] == 2 * len(c4["alone"]) + len(c4["probes"])
    assert st["peak_slots"] >= 4


def test_budget_respected(c4):
    st = c4["stats"]
    assert st["peak_bytes"] <= BUDGET
    # what the device actually held (graphs + workspaces + HIP's own buffers)
    assert c4["peak_used"] <= BUDGET + (2 << 30), c4["peak_used"] / 2**30


def test_alone_against_oracle(c4):
    graphs, specs = c4["graphs"], c4["specs"]
    t = time.time()
    checked = {"full": 0, "probe": 0, "checksum": 0}
    for i, (gi, src, bc) in enumerate(zip(c4["gsel"], c4["qsrc"], c4["alone"])):
        g = graphs[gi]
        if specs[gi][0] == "chain":
            ds = O.depth_sums(g.off, g.adj, src, nthreads=NTH)
            assert abs(bc.sum() - ds.sum()) <= 1e-9 * max(1.0, ds.sum()), i
            assert np.all(bc >= 0)
            checked["checksum"] += 1
        elif i in c4["probes"]:
            ref = O.brandes(g.off, g.adj, src[:PROBE], nthreads=NTH)
            assert bcmi.rel_err(c4["probes"][i], ref) <= 1e-9, i
            checked["probe"] += 1
        else:
            ref = O.brandes(g.off, g.adj, src, nthreads=NTH)
            assert bcmi.rel_err(bc, ref) <= 1e-9, i
            checked["full"] += 1
    say(f"oracle checks {checked} in {time.time() - t:.1f}s")
    assert checked["full"] >= 24 and checked["checksum"] >= 8
