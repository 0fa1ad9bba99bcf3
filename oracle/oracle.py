"""ctypes front end of the CPU oracle (liboracle.so, built from bc_oracle.c).

TEST INFRASTRUCTURE ONLY.  This module is the checker used by tests/, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``.
The engine never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        L.orc_read_edges.restype = ctypes.c_longlong
        L.orc_read_edges.argtypes = [ctypes.c_char_p, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_uint64)]
        L.orc_free.argtypes = [vp]
        L.orc_normalize.restype = ctypes.c_uint64
        L.orc_normalize.argtypes = [_u64p, ctypes.c_uint64]
        L.orc_build_csc.argtypes = [_u64p, ctypes.c_uint64, ctypes.c_uint64, _i64p, _i32p]
        L.orc_sssp.restype = ctypes.c_int64
        L.orc_sssp.argtypes = [ctypes.c_int64, _i64p, _i32p, ctypes.c_int32, vp, _i32p, _u64p, _f64p]
        L.orc_brandes.restype = ctypes.c_int
        L.orc_brandes.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, ctypes.c_int64, vp, vp,
                                  _f64p, ctypes.c_int, vp, vp, vp, vp, vp]
        L.orc_ref_degree_reduction.restype = ctypes.c_uint64
        L.orc_ref_degree_reduction.argtypes = [_u64p, ctypes.c_uint64, _u64p,
                                               ctypes.POINTER(ctypes.c_uint64),
                                               ctypes.POINTER(ctypes.c_uint64), _u32p]
        L.orc_peel.restype = ctypes.c_int32
        L.orc_peel.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _i32p, _u64p, _i32p]
        L.orc_components.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, _i64p]
        L.orc_bc_pipeline.restype = ctypes.c_int
        L.orc_bc_pipeline.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, ctypes.c_int64,
                                      ctypes.c_int, ctypes.c_int, _f64p, ctypes.c_int]
        L.orc_depth_sums.restype = ctypes.c_int
        L.orc_depth_sums.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, ctypes.c_int64, ctypes.c_int, _i64p]
        L.orc_brandes_scaled.restype = ctypes.c_int
        L.orc_brandes_scaled.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p, ctypes.c_int64, _f64p, ctypes.c_int]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class OracleError(RuntimeError):
    pass


def read_edges(path, n, directed=False):
    """Reference ingest (read_graph + mirror): returns an (k,2) uint64 array."""
    out = ctypes.c_void_p()
    bad = ctypes.c_uint64(0)
    k = lib().orc_read_edges(path.encode(), n, 1 if directed else 0, ctypes.byref(out), ctypes.byref(bad))
    if k < 0:
        raise OracleError(f"orc_read_edges failed ({-k}), line {bad.value}")
    buf = (ctypes.c_uint64 * (2 * k)).from_address(out.value) if k else []
    arr = np.array(buf, dtype=np.uint64).reshape(-1, 2) if k else np.zeros((0, 2), np.uint64)
    lib().orc_free(out)
    return arr


def normalize(arcs):
    a = np.ascontiguousarray(arcs, dtype=np.uint64).copy()
    l = lib().orc_normalize(a.reshape(-1), a.shape[0])
    return a[:l]


def build_csc(arcs, n):
    a = np.ascontiguousarray(arcs, dtype=np.uint64)
    off = np.zeros(n + 1, np.int64)
    adj = np.zeros(max(a.shape[0], 1), np.int32)
    lib().orc_build_csc(a.reshape(-1), a.shape[0], n, off, adj)
    return off, adj[:a.shape[0]]


def csr_from_edges(edges, n, directed=False):
    """Edge list (k,2) -> CSR via the reference pipeline (mirror, normalise, CSC)."""
    e = np.asarray(edges, dtype=np.uint64).reshape(-1, 2)
    if not directed:
        keep = e[:, 0] != e[:, 1]
        e = np.concatenate([e, e[keep][:, ::-1]], axis=0)
    e = normalize(e)
    return build_csc(e, n)


def sssp(off, adj, s, w=None):
    n = off.shape[0] - 1
    depth = np.empty(n, np.int32)
    sigma = np.empty(n, np.uint64)
    delta = np.empty(n, np.float64)
    wv = None if w is None else np.ascontiguousarray(w, dtype=np.uint64)
    L = lib().orc_sssp(n, off, adj, int(s), _ptr(wv), depth, sigma, delta)
    if L < 0:
        raise OracleError("sigma overflow")
    return depth, sigma, delta, int(L)


def brandes(off, adj, sources, w=None, mult=None, nthreads=1, stats=False, allow_overflow=False):
    """Ordered-pair BC (not halved) summed over the given sources."""
    n = off.shape[0] - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.zeros(n, np.float64)
    wv = None if w is None else np.ascontiguousarray(w, dtype=np.uint64)
    mv = None if mult is None else np.ascontiguousarray(mult, dtype=np.uint64)
    ns = np.zeros(len(src), np.int64)
    ms = np.zeros(len(src), np.int64)
    as_ = np.zeros(len(src), np.int64)
    nw = np.zeros(len(src), np.uint64)
    lv = np.zeros(len(src), np.int32)
    rc = lib().orc_brandes(n, off, adj, src, len(src), _ptr(wv), _ptr(mv), bc, nthreads,
                           _ptr(ns), _ptr(ms), _ptr(as_), _ptr(nw), _ptr(lv))
    if rc == 1 and not allow_overflow:
        raise OracleError("sigma overflow")
    if rc not in (0, 1):
        raise OracleError(f"orc_brandes rc={rc}")
    if stats:
        return bc, dict(n_s=ns, m_s=ms, a_s=as_, nw=nw, levels=lv, overflow=(rc == 1))
    return bc


def ref_degree_reduction(arcs, n):
    """Reference single-pass reduction on normalised arcs -> (kept, removed, pnod, ncouple, reach)."""
    a = np.ascontiguousarray(arcs, dtype=np.uint64).copy()
    rem = np.zeros(max(2 * a.shape[0], 2), np.uint64)
    nrem = ctypes.c_uint64(0)
    ncp = ctypes.c_uint64(0)
    reach = np.zeros(n, np.uint32)
    kept = lib().orc_ref_degree_reduction(a.reshape(-1), a.shape[0], rem, ctypes.byref(nrem), ctypes.byref(ncp), reach)
    return a[:kept], rem.reshape(-1, 2), int(nrem.value), int(ncp.value), reach


def peel(off, adj):
    n = off.shape[0] - 1
    rnd = np.empty(n, np.int32)
    par = np.empty(n, np.int32)
    w = np.empty(n, np.uint64)
    cdeg = np.empty(n, np.int32)
    nr = lib().orc_peel(n, off, adj, rnd, par, w, cdeg)
    return dict(round=rnd, parent=par, w=w, core_deg=cdeg, rounds=int(nr))


def components(off, adj):
    n = off.shape[0] - 1
    lab = np.empty(n, np.int32)
    sz = np.empty(n, np.int64)
    lib().orc_components(n, off, adj, lab, sz)
    return lab, sz


def bc_pipeline(off, adj, sources, heuristic=0, expand=False, nthreads=1):
    n = off.shape[0] - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.zeros(n, np.float64)
    rc = lib().orc_bc_pipeline(n, off, adj, src, len(src), int(heuristic), 1 if expand else 0, bc, nthreads)
    if rc != 0:
        raise OracleError(f"orc_bc_pipeline rc={rc}")
    return bc


def depth_sums(off, adj, sources, nthreads=1):
    """Per source: sum over reached t != s of (d(s,t) - 1), which equals the
    sum of that source's dependencies delta_s(v) over all v."""
    n = off.shape[0] - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    out = np.zeros(len(src), np.int64)
    rc = lib().orc_depth_sums(n, off, adj, src, len(src), nthreads, out)
    if rc != 0:
        raise OracleError(f"orc_depth_sums rc={rc}")
    return out


def brandes_scaled(off, adj, sources, nthreads=1):
    """Ordered-pair BC (not halved) over the sources with path counts as
    binary64 values and a power-of-two scale per BFS level, for graphs whose
    counts pass 2^64 (long bridged-block chains).  Not exact: binary64
    rounding over the levels (orc_brandes_scaled)."""
    n = off.shape[0] - 1
    src = np.ascontiguousarray(sources, dtype=np.int32)
    bc = np.zeros(n, np.float64)
    rc = lib().orc_brandes_scaled(n, off, adj, src, len(src), bc, nthreads)
    if rc != 0:
        raise OracleError(f"orc_brandes_scaled rc={rc}")
    return bc


def brandes_highprec(off, adj, sources, digits=50):
    """Brandes with exact big-integer path counts and a `digits`-digit
    decimal backward pass (pure Python; small graphs only).  Used where path
    counts exceed 2^64 and the uint64 oracle stops (long bridged-block
    chains).  Same rules as sssp/orc_brandes: delta[c] += (delta[r]+1)/
    sigma[r]*sigma[c] over arcs to the next level (cuda_kernels.cu:590-596).
    Returns ordered-pair BC as float64."""
    import decimal
    ctx = decimal.Context(prec=digits)
    n = len(off) - 1
    off = [int(x) for x in off]
    adj = [int(x) for x in adj]
    bc = [decimal.Decimal(0)] * n
    for s in sources:
        s = int(s)
        if off[s + 1] == off[s]:
            continue
        depth = {s: 0}
        sigma = {s: 1}
        levels = [[s]]
        while True:
            nxt = []
            d1 = len(levels)
            for c in levels[-1]:
                for k in range(off[c], off[c + 1]):
                    r = adj[k]
                    if r not in depth:
                        depth[r] = d1
                        sigma[r] = 0
                        nxt.append(r)
                    if depth[r] == d1:
                        sigma[r] += sigma[c]
            if not nxt:
                break
            levels.append(nxt)
        delta = {}
        for d in range(len(levels) - 1, 0, -1):
            for c in levels[d]:
                acc = decimal.Decimal(0)
                sc = decimal.Decimal(sigma[c])
                for k in range(off[c], off[c + 1]):
                    r = adj[k]
                    if depth.get(r, -1) == d + 1:
                        acc = ctx.add(acc, ctx.multiply(ctx.divide(ctx.add(delta[r], 1), decimal.Decimal(sigma[r])), sc))
                delta[c] = acc
                bc[c] = ctx.add(bc[c], acc)
    return np.array([float(x) for x in bc], np.float64)
