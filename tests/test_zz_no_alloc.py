"""The allocation-free entry point (bcmi_bc_run_static) against the entry
point it varies (bcmi_bc_run, plain mode) through the C ABI: two calls in a
row, the second on a smaller graph and source set (the static scratch keeps
the first call's contents), each equal to bcmi_bc_run within 1e-9 and to the
CPU oracle; device memory is the same before and after a call; graphs whose
sweep state exceeds the scratch, and peeled graphs, are refused."""
import ctypes

import numpy as np
import pytest

import bcmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _free_device_bytes():
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def _both(e, g, src):
    e.upload(g)
    e.reset()
    e.run(src, mode=bcmi.MODE_PLAIN)
    ref = e.bc()
    e.reset()
    st = e.run_static(src)
    got = e.bc()
    return ref, got, st


def test_static_matches_plain_run_twice():
    e = bcmi.Engine()
    # first call: R-MAT scale 14 (16,384 vertices), 96 sources
    g1 = bcmi.Graph.rmat(14, 16, seed=31)
    s1 = g1.sample_sources(96, seed=32)
    ref1, got1, st1 = _both(e, g1, s1)
    print(f"call 1: {st1['sources_run']} sweeps, max rel diff {bcmi.rel_err(got1, ref1):.2e}", flush=True)
    assert st1["sources_run"] == 96 and st1["te"] > 0
    assert bcmi.rel_err(got1, ref1) <= 1e-9
    assert bcmi.rel_err(got1, O.brandes(g1.off, g1.adj, s1, nthreads=4)) <= 1e-9
    # second call, smaller: R-MAT scale 11 with isolated vertices among the
    # sources (skipped, c:1026), on the scratch the first call left behind
    g2 = bcmi.Graph.rmat(11, 8, seed=33)
    iso = np.flatnonzero(g2.degrees() == 0)[:3].astype(np.int32)
    s2 = np.concatenate([g2.sample_sources(37, seed=34), iso]).astype(np.int32)
    ref2, got2, st2 = _both(e, g2, s2)
    print(f"call 2: {st2['sources_run']} sweeps, max rel diff {bcmi.rel_err(got2, ref2):.2e}", flush=True)
    assert st2["sources_run"] == 37
    assert bcmi.rel_err(got2, ref2) <= 1e-9
    assert bcmi.rel_err(got2, O.brandes(g2.off, g2.adj, s2, nthreads=4)) <= 1e-9
    e.close()


def test_static_allocates_nothing():
    e = bcmi.Engine()
    g = bcmi.Graph.rmat(13, 16, seed=35)
    src = g.sample_sources(32, seed=36)
    e.upload(g)
    e.run_static(src[:4])  # first launches of the kernels in this process
    before = _free_device_bytes()
    e.run_static(src)
    after = _free_device_bytes()
    print(f"device free before {before} after {after}", flush=True)
    assert after == before
    e.close()


def test_static_refuses_what_does_not_fit():
    e = bcmi.Engine()
    n = 2_000_000  # 36 bytes per vertex of sweep state + work records: over 64 MiB
    g = bcmi.Graph.from_edges([[0, 1], [1, 2]], n)
    e.upload(g)
    with pytest.raises(bcmi.BcmiError) as ex:
        e.run_static([0])
    assert ex.value.code == bcmi.ERR_ARG and "static scratch" in str(ex.value)
    h = bcmi.Graph.rmat(10, 16, seed=37)
    e.upload(h, heuristic=bcmi.H_PEEL)
    with pytest.raises(bcmi.BcmiError) as ex:
        e.run_static(h.sample_sources(4, seed=1))
    assert ex.value.code == bcmi.ERR_STATE
    e.close()
