"""Python front end (ctypes) of libbcmi.so, the MI355X Brandes BC engine.

Mirrors the reference program's operations (program/bc-cuda/bc2d.c, paths
relative to the ck-bc root): graph ingest (read_graph/mirror/norm_graph/
build_csc, bc2d.c:152-254, 430-501), device upload and degree-1 reduction
(initcuda/degree_reduction, cuda_kernels.cu:1123, bc2d.c:314), the per-source
BC loop (bc2d.c:1010-1050) and the result getter (get_bc, cuda_kernels.cu:1039).
Errors raise BcmiError carrying the library's message; there is no fallback:
if libbcmi.so is missing or no GPU is present the calls fail.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# BCMI_LIB names another build of the library (A/B comparisons of build switches)
LIB_PATH = os.environ.get("BCMI_LIB") or os.path.join(HERE, "libbcmi.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_OVERFLOW, ERR_STATE, ERR_IO, ERR_RANGE, ERR_BUSY = range(1, 9)
MODE_PLAIN, MODE_BATCHED = 0, 1
H_NONE, H_PEEL = 0, 1
SRC_EXACT, SRC_EXPAND = 0, 1
SRC_DEG2 = 2  # OR-ed into the source flags: 2-degree heuristic (plain sweeps)

EXPORTS = [
    "bcmi_last_error", "bcmi_version",
    "bcmi_graph_read_edgelist", "bcmi_graph_save", "bcmi_graph_load",
    "bcmi_graph_from_edges", "bcmi_gen_rmat", "bcmi_gen_pendant_trees",
    "bcmi_gen_block_chain", "bcmi_sample_sources", "bcmi_graph_free",
    "bcmi_create", "bcmi_destroy", "bcmi_set_param", "bcmi_get_param", "bcmi_upload_graph",
    "bcmi_get_peel", "bcmi_get_labels", "bcmi_bc_reset", "bcmi_bc_run", "bcmi_get_bc", "bcmi_bc_run_static", "bcmi_batch_order",
    "bcmi_get_sweep_stats", "bcmi_get_sweep_times", "bcmi_get_kernel_time", "bcmi_sssp", "bcmi_msbfs",
    "bcmi_server_create", "bcmi_server_destroy", "bcmi_server_add_graph", "bcmi_server_submit",
    "bcmi_server_wait", "bcmi_server_stats", "bcmi_launch_plan",
]


class BcmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bcmi error {code}: {msg}")
        self.code = code


class _Graph(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("m", ctypes.c_int64),
                ("off", ctypes.POINTER(ctypes.c_int64)), ("adj", ctypes.POINTER(ctypes.c_int32)),
                ("directed", ctypes.c_int)]


class RunStats(ctypes.Structure):
    _fields_ = [("sources_run", ctypes.c_int64), ("batches", ctypes.c_int64), ("te", ctypes.c_double),
                ("bytes", ctypes.c_double), ("gpu_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("max_levels", ctypes.c_int64), ("sum_n", ctypes.c_int64), ("sum_m", ctypes.c_int64),
                ("sum_a", ctypes.c_int64), ("sum_pairs", ctypes.c_int64), ("sum_union", ctypes.c_int64),
                ("scaled_batches", ctypes.c_int64), ("wg_sources", ctypes.c_int64),
                ("deg2_sources", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Launch(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_char * 24), ("grid", ctypes.c_int32), ("block", ctypes.c_int32),
                ("aux", ctypes.c_int64), ("timed", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class QueryResult(ctypes.Structure):
    _fields_ = [("query_id", ctypes.c_int64), ("graph_id", ctypes.c_int32), ("status", ctypes.c_int32),
                ("arrive_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("finish_ms", ctypes.c_double),
                ("te", ctypes.c_double)]


_LIB = None


def lib():
    """Loads libbcmi.so (built in-tree by the package Makefile)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise BcmiError(ERR_STATE, f"{LIB_PATH} not built (run make -C {HERE})")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i64, i32, u64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64
        gpp = ctypes.POINTER(ctypes.POINTER(_Graph))
        L.bcmi_last_error.restype = ctypes.c_char_p
        L.bcmi_graph_read_edgelist.argtypes = [ctypes.c_char_p, u64, ctypes.c_int, gpp, ctypes.POINTER(u64),
                                               ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.bcmi_graph_from_edges.argtypes = [vp, vp, i64, i64, ctypes.c_int, gpp]
        L.bcmi_graph_save.argtypes = [ctypes.POINTER(_Graph), ctypes.c_char_p, u64, i64]
        L.bcmi_graph_load.argtypes = [ctypes.c_char_p, gpp, ctypes.POINTER(u64), ctypes.POINTER(i64)]
        L.bcmi_gen_rmat.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    u64, ctypes.c_int, gpp]
        L.bcmi_gen_pendant_trees.argtypes = [ctypes.POINTER(_Graph), ctypes.c_double, u64, gpp]
        L.bcmi_gen_block_chain.argtypes = [i64, ctypes.c_int, ctypes.c_double, u64, gpp]
        L.bcmi_sample_sources.argtypes = [ctypes.POINTER(_Graph), i64, u64, vp]
        L.bcmi_graph_free.argtypes = [ctypes.POINTER(_Graph)]
        L.bcmi_graph_free.restype = None
        L.bcmi_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.bcmi_destroy.argtypes = [vp]
        L.bcmi_set_param.argtypes = [vp, ctypes.c_char_p, i64]
        L.bcmi_get_param.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(i64)]
        L.bcmi_upload_graph.argtypes = [vp, i64, vp, vp, ctypes.c_int, ctypes.c_int]
        L.bcmi_get_peel.argtypes = [vp, vp, vp, vp, ctypes.POINTER(i32)]
        L.bcmi_get_labels.argtypes = [vp, vp]
        L.bcmi_bc_reset.argtypes = [vp]
        L.bcmi_bc_run.argtypes = [vp, vp, i64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(RunStats)]
        L.bcmi_get_bc.argtypes = [vp, vp]
        L.bcmi_bc_run_static.argtypes = [vp, vp, i64, ctypes.POINTER(RunStats)]
        L.bcmi_batch_order.argtypes = [vp, vp, i64, vp]
        L.bcmi_get_sweep_stats.argtypes = [vp, vp, vp, vp, vp]
        L.bcmi_get_sweep_times.argtypes = [vp, vp]
        L.bcmi_get_kernel_time.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64)]
        L.bcmi_sssp.argtypes = [vp, i32, vp, vp, vp]
        L.bcmi_msbfs.argtypes = [vp, vp, ctypes.c_int, vp, vp, vp]
        L.bcmi_server_create.argtypes = [ctypes.c_int, i64, ctypes.c_int, ctypes.POINTER(vp)]
        L.bcmi_server_destroy.argtypes = [vp]
        L.bcmi_server_add_graph.argtypes = [vp, i64, vp, vp, ctypes.POINTER(i32)]
        L.bcmi_server_submit.argtypes = [vp, i32, vp, i64, ctypes.POINTER(i64)]
        L.bcmi_server_wait.argtypes = [vp, i64, vp, ctypes.POINTER(QueryResult)]
        L.bcmi_server_stats.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]
        L.bcmi_launch_plan.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(i64), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(Launch), ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != OK:
        raise BcmiError(rc, lib().bcmi_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- graphs --
class Graph:
    """Host CSR graph (int64 offsets, int32 adjacency) as numpy arrays."""

    def __init__(self, off, adj, directed=False):
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.adj = np.ascontiguousarray(adj, dtype=np.int32)
        self.directed = bool(directed)
        self.header = None

    @property
    def n(self):
        return self.off.shape[0] - 1

    @property
    def m(self):
        return int(self.off[-1])

    def degrees(self):
        return np.diff(self.off)

    @staticmethod
    def _take(gp):
        g = gp.contents
        n, m = g.n, g.m
        off = np.ctypeslib.as_array(g.off, shape=(n + 1,)).copy()
        adj = np.ctypeslib.as_array(g.adj, shape=(max(m, 1),))[:m].copy() if m else np.zeros(0, np.int32)
        out = Graph(off, adj, g.directed)
        lib().bcmi_graph_free(gp)
        return out

    @staticmethod
    def read_edgelist(path, n, directed=False):
        gp = ctypes.POINTER(_Graph)()
        bad, hn, he = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        rc = lib().bcmi_graph_read_edgelist(os.fsencode(path), n, int(directed), ctypes.byref(gp),
                                            ctypes.byref(bad), ctypes.byref(hn), ctypes.byref(he))
        _check(rc)
        g = Graph._take(gp)
        g.header = (hn.value, he.value) if hn.value else None
        return g

    @staticmethod
    def load(path):
        """Reads a binary CSR cache; returns (graph, (source size, source mtime ns))."""
        gp = ctypes.POINTER(_Graph)()
        sz, mt = ctypes.c_uint64(0), ctypes.c_int64(0)
        _check(lib().bcmi_graph_load(os.fsencode(path), ctypes.byref(gp), ctypes.byref(sz), ctypes.byref(mt)))
        return Graph._take(gp), (sz.value, mt.value)

    def save(self, path, src_size=0, src_mtime_ns=0):
        v = self._view()
        _check(lib().bcmi_graph_save(ctypes.byref(v), os.fsencode(path), src_size, src_mtime_ns))

    @staticmethod
    def from_edges(edges, n, directed=False):
        e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
        t = np.ascontiguousarray(e[:, 0])
        h = np.ascontiguousarray(e[:, 1])
        gp = ctypes.POINTER(_Graph)()
        _check(lib().bcmi_graph_from_edges(_p(t), _p(h), e.shape[0], n, int(directed), ctypes.byref(gp)))
        return Graph._take(gp)

    @staticmethod
    def rmat(scale, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, permute=True):
        gp = ctypes.POINTER(_Graph)()
        _check(lib().bcmi_gen_rmat(scale, edgefactor, a, b, c, seed, int(permute), ctypes.byref(gp)))
        return Graph._take(gp)

    @staticmethod
    def block_chain(nblocks, block=256, avgdeg=8.0, seed=1):
        gp = ctypes.POINTER(_Graph)()
        _check(lib().bcmi_gen_block_chain(nblocks, block, avgdeg, seed, ctypes.byref(gp)))
        return Graph._take(gp)

    def _view(self):
        g = _Graph()
        g.n, g.m = self.n, self.m
        g.off = self.off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        g.adj = self.adj.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        g.directed = int(self.directed)
        return g

    def with_pendant_trees(self, frac=0.3, seed=1):
        v = self._view()
        gp = ctypes.POINTER(_Graph)()
        _check(lib().bcmi_gen_pendant_trees(ctypes.byref(v), frac, seed, ctypes.byref(gp)))
        return Graph._take(gp)

    def sample_sources(self, k, seed=1):
        v = self._view()
        out = np.zeros(k, np.int32)
        _check(lib().bcmi_sample_sources(ctypes.byref(v), k, seed, _p(out)))
        return out


# ---------------------------------------------------------------- engine --
class Engine:
    """One device context: a resident graph and a BC accumulator."""

    def __init__(self, device=0, **params):
        self._c = ctypes.c_void_p()
        _check(lib().bcmi_create(device, ctypes.byref(self._c)))
        self.n = 0
        for k, v in params.items():
            self.set_param(k, v)

    def close(self):
        if self._c:
            lib().bcmi_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_param(self, key, value):
        _check(lib().bcmi_set_param(self._c, key.encode(), int(value)))

    def get_param(self, key):
        v = ctypes.c_int64()
        _check(lib().bcmi_get_param(self._c, key.encode(), ctypes.byref(v)))
        return v.value

    def upload(self, g, heuristic=H_NONE):
        _check(lib().bcmi_upload_graph(self._c, g.n, _p(g.off), _p(g.adj), int(g.directed), int(heuristic)))
        self.n = g.n

    def peel_info(self):
        rnd = np.zeros(self.n, np.int32)
        par = np.zeros(self.n, np.int32)
        w = np.zeros(self.n, np.uint64)
        nr = ctypes.c_int32()
        _check(lib().bcmi_get_peel(self._c, _p(rnd), _p(par), _p(w), ctypes.byref(nr)))
        return dict(round=rnd, parent=par, w=w, rounds=nr.value)

    def labels(self):
        lab = np.zeros(self.n, np.int32)
        _check(lib().bcmi_get_labels(self._c, _p(lab)))
        return lab

    def reset(self):
        _check(lib().bcmi_bc_reset(self._c))

    def run(self, sources, mode=MODE_BATCHED, expand=False, deg2=False):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = RunStats()
        flags = (SRC_EXPAND if expand else SRC_EXACT) | (SRC_DEG2 if deg2 else 0)
        _check(lib().bcmi_bc_run(self._c, _p(src), src.shape[0], int(mode), flags, ctypes.byref(st)))
        return st.as_dict()

    def run_static(self, sources):
        """Plain sweeps over the sources with no device allocation (state in the
        library's static per-GPU scratch); accumulates like run(mode=MODE_PLAIN)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        st = RunStats()
        _check(lib().bcmi_bc_run_static(self._c, _p(src), src.shape[0], ctypes.byref(st)))
        return st.as_dict()

    def batch_order(self, sources):
        """The sources in the order a batched run groups them (degree-0 sources last)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        out = np.empty_like(src)
        _check(lib().bcmi_batch_order(self._c, _p(src), src.shape[0], _p(out)))
        return out

    def bc(self):
        out = np.zeros(self.n, np.float64)
        _check(lib().bcmi_get_bc(self._c, _p(out)))
        return out

    def sweep_stats(self, k):
        n_s = np.zeros(k, np.int64)
        m_s = np.zeros(k, np.int64)
        a_s = np.zeros(k, np.int64)
        lv = np.zeros(k, np.int32)
        _check(lib().bcmi_get_sweep_stats(self._c, _p(n_s), _p(m_s), _p(a_s), _p(lv)))
        return dict(n_s=n_s, m_s=m_s, a_s=a_s, levels=lv)

    def sweep_times(self, k):
        """Per-sweep time (ms) of the last run, in the order of sweep_stats."""
        t = np.zeros(k, np.float64)
        _check(lib().bcmi_get_sweep_times(self._c, _p(t)))
        return t

    def kernel_time(self):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(lib().bcmi_get_kernel_time(self._c, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def sssp(self, s):
        d = np.zeros(self.n, np.int32)
        sg = np.zeros(self.n, np.uint64)
        dl = np.zeros(self.n, np.float64)
        _check(lib().bcmi_sssp(self._c, int(s), _p(d), _p(sg), _p(dl)))
        return d, sg, dl

    def msbfs(self, sources):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        k = src.shape[0]
        d = np.zeros((k, self.n), np.int32)
        sg = np.zeros((k, self.n), np.uint64)
        dl = np.zeros((k, self.n), np.float64)
        _check(lib().bcmi_msbfs(self._c, _p(src), k, _p(d), _p(sg), _p(dl)))
        return d, sg, dl


def launch_plan(params, mode, graph):
    """Kernel launches of one captured level step for the given tunables
    (dict of bcmi_set_param keys); graph 0 forward, 1 backward, 2/3 the same
    on sparse levels.  Returns (list of dicts, steps per graph).  No device."""
    keys = list(params)
    karr = (ctypes.c_char_p * max(1, len(keys)))(*[k.encode() for k in keys])
    varr = (ctypes.c_int64 * max(1, len(keys)))(*[int(params[k]) for k in keys])
    out = (Launch * 64)()
    cnt, steps = ctypes.c_int(), ctypes.c_int()
    _check(lib().bcmi_launch_plan(karr, varr, len(keys), int(mode), int(graph), out, 64, ctypes.byref(cnt),
                                  ctypes.byref(steps)))
    assert cnt.value <= 64
    return [dict(kernel=out[i].kernel.decode(), grid=out[i].grid, block=out[i].block, aux=out[i].aux,
                 timed=bool(out[i].timed)) for i in range(cnt.value)], steps.value


def rel_err(a, b):
    """max |a-b|/|b| over entries with b != 0; entries with b == 0 must match exactly."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nz = b != 0
    zero_ok = np.all(a[~nz] == 0)
    if not zero_ok:
        return float("inf")
    if not nz.any():
        return 0.0
    return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz])))


# ---------------------------------------------------------------- server --
class Server:
    """Query server: device-resident graph pool, continuous batching."""

    def __init__(self, hbm_budget, max_slots=8, device=0):
        self._h = ctypes.c_void_p()
        _check(lib().bcmi_server_create(device, int(hbm_budget), int(max_slots), ctypes.byref(self._h)))
        self.sizes = []

    def add_graph(self, g):
        gid = ctypes.c_int32()
        _check(lib().bcmi_server_add_graph(self._h, g.n, _p(g.off), _p(g.adj), ctypes.byref(gid)))
        self.sizes.append(g.n)
        return gid.value

    def submit(self, gid, sources):
        src = np.ascontiguousarray(sources, dtype=np.int32)
        q = ctypes.c_int64()
        _check(lib().bcmi_server_submit(self._h, int(gid), _p(src), src.shape[0], ctypes.byref(q)))
        return q.value

    def wait(self, qid, n, want_bc=True):
        """Returns (status, bc or None, QueryResult); status != 0 is an error code."""
        out = np.zeros(n, np.float64) if want_bc else None
        res = QueryResult()
        rc = lib().bcmi_server_wait(self._h, int(qid), _p(out) if want_bc else None, ctypes.byref(res))
        return rc, out, res

    def stats(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().bcmi_server_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return dict(completed=a.value, peak_slots=b.value, peak_bytes=c.value)

    def close(self):
        if self._h:
            lib().bcmi_server_destroy(self._h)
            self._h = ctypes.c_void_p()
