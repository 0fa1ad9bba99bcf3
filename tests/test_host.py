"""CPU tests of the engine's host side: graph ingest and CSR build against the
oracle's restatement of the reference pipeline, the synthetic generators, and
the C ABI (library loads, exports every declared entry point, fails cleanly
without a GPU).  No device calls."""
import json
import os
import re
import sys

import numpy as np
import pytest

import bcmi
from oracle import oracle as O
import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    src = open(os.path.join(ROOT, "include", "bcmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bcmi_\w+)\s*\(", src)))


def test_abi_exports_every_declared_symbol():
    lib = bcmi.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert set(bcmi.EXPORTS) == set(names)


def test_engine_without_gpu_fails_cleanly():
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); import bcmi\n"
            "try:\n  bcmi.Engine()\nexcept bcmi.BcmiError as e:\n  print('ERR', e.code)\n"
            "else:\n  print('GPU')\n") % os.path.join(ROOT, "ck-bc_amd")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert out.strip() in ("GPU", "ERR %d" % bcmi.ERR_HIP)


def test_last_error_and_bad_args():
    with pytest.raises(bcmi.BcmiError) as ex:
        bcmi.Graph.rmat(0, 16)
    assert ex.value.code == bcmi.ERR_ARG
    with pytest.raises(bcmi.BcmiError) as ex:
        bcmi.Graph.from_edges([[0, 5]], 3)
    assert ex.value.code == bcmi.ERR_RANGE


@pytest.mark.parametrize("directed", [False, True])
def test_csr_matches_reference_pipeline(directed):
    for seed in range(5):
        e, n = G.random_graph(200, 700, seed)
        g = bcmi.Graph.from_edges(e, n, directed=directed)
        off, adj = O.csr_from_edges(e, n, directed=directed)
        assert np.array_equal(g.off, off) and np.array_equal(g.adj, adj)


def test_read_edgelist_facebook(tmp_path):
    f = np.load(os.path.join(GOLD, "facebook.npz"))
    n = int(f["n"])
    e = f["edges"].astype(np.int64)
    fn = tmp_path / "fb.txt"
    with open(fn, "w") as fp:
        fp.write("# Directed graph\n# Nodes: 4039 Edges: 88234\n")
        for a, b in e:
            fp.write(f"{a} {b}\n")
    g = bcmi.Graph.read_edgelist(str(fn), n)
    assert g.header == (4039, 88234)
    arcs = O.normalize(O.read_edges(str(fn), n))
    off, adj = O.build_csc(arcs, n)
    assert np.array_equal(g.off, off) and np.array_equal(g.adj, adj)
    assert g.m == 2 * 88234


def test_read_edgelist_quirks(tmp_path):
    fn = tmp_path / "g.txt"
    fn.write_text("# Nodes: 5 Edges: 6\n0 1\n1 2\n\n2 0\n# comment 3 4\n3 3\n1 0\n4 2\n")
    g = bcmi.Graph.read_edgelist(str(fn), 5)
    off, adj = O.build_csc(O.normalize(O.read_edges(str(fn), 5)), 5)
    assert np.array_equal(g.off, off) and np.array_equal(g.adj, adj)
    gd = bcmi.Graph.read_edgelist(str(fn), 5, directed=True)
    offd, adjd = O.build_csc(O.normalize(O.read_edges(str(fn), 5, directed=True)), 5)
    assert np.array_equal(gd.off, offd) and np.array_equal(gd.adj, adjd)
    with pytest.raises(bcmi.BcmiError) as ex:
        bcmi.Graph.read_edgelist(str(fn), 4)
    assert ex.value.code == bcmi.ERR_RANGE and "line 7" in str(ex.value)
    with pytest.raises(bcmi.BcmiError) as ex:
        bcmi.Graph.read_edgelist(str(tmp_path / "missing.txt"), 4)
    assert ex.value.code == bcmi.ERR_IO


def test_rmat_generator_is_stable_and_well_formed():
    f = np.load(os.path.join(GOLD, "rmat10.npz"))
    g = bcmi.Graph.rmat(10, 16, seed=20240101)
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    keep = src < g.adj
    assert np.array_equal(np.stack([src[keep], g.adj[keep]], 1), f["edges"].astype(np.int64))
    g2 = bcmi.Graph.rmat(12, 8, seed=3)
    s2 = np.repeat(np.arange(g2.n), np.diff(g2.off))
    assert np.all(s2 != g2.adj)                         # no self loops
    pairs = set(zip(s2.tolist(), g2.adj.tolist()))
    assert all((b, a) in pairs for a, b in pairs)       # symmetric
    assert all(np.all(np.diff(g2.adj[g2.off[v]:g2.off[v + 1]]) > 0) for v in range(0, g2.n, 97))
    deg = g2.degrees()
    assert deg.max() > 20 * deg[deg > 0].mean()         # skewed degrees (a = .57)
    assert g2.m <= 2 * 8 * g2.n


def test_pendant_trees_generator():
    core = bcmi.Graph.rmat(10, 16, seed=5)
    g = core.with_pendant_trees(0.3, seed=9)
    extra = g.n - core.n
    assert extra == round(0.3 * core.n)
    assert g.m == core.m + 2 * extra                    # each new vertex adds one edge
    p = O.peel(g.off, g.adj)
    removed = int((p["round"] > 0).sum())
    assert removed >= extra
    assert 1 <= p["rounds"] <= 8


def test_block_chain_generator():
    g = bcmi.Graph.block_chain(10, 64, 8.0, seed=2)
    assert g.n == 640
    src = np.repeat(np.arange(g.n), np.diff(g.off))
    cross = (src // 64) != (g.adj // 64)
    assert int(cross.sum()) == 2 * 9                    # one bridge between consecutive blocks
    assert np.all(np.abs(src[cross] // 64 - g.adj[cross] // 64) == 1)
    lab, _ = O.components(g.off, g.adj)                  # bridges join the blocks' large parts
    big = np.bincount(lab).max()
    assert big >= 0.99 * g.n
    for a, b in zip(src[cross], g.adj[cross]):
        assert lab[a] == lab[b] and np.sum(lab == lab[a]) == big


def test_sample_sources():
    g = bcmi.Graph.rmat(11, 4, seed=1)
    s = g.sample_sources(500, seed=7)
    assert len(set(s.tolist())) == 500
    assert np.all(g.degrees()[s] > 0)
    assert np.array_equal(s, g.sample_sources(500, seed=7))
    assert not np.array_equal(s, g.sample_sources(500, seed=8))


def test_csr_cache_roundtrip_and_checks(tmp_path):
    for g in (bcmi.Graph.rmat(10, 16, seed=3), bcmi.Graph.from_edges(*G.random_graph(300, 900, 2), directed=True)):
        p = tmp_path / "g.csr"
        g.save(str(p), 1234, 5678)
        h, prov = bcmi.Graph.load(str(p))
        assert prov == (1234, 5678)
        assert h.n == g.n and h.directed == g.directed
        assert np.array_equal(h.off, g.off) and np.array_equal(h.adj, g.adj)
    raw = bytearray(open(p, "rb").read())
    bad = tmp_path / "bad.csr"
    for mutate in (lambda b: b[:len(b) - 4],                                   # truncated
                   lambda b: b[:70] + bytes([b[70] ^ 1]) + b[71:],             # flipped bit
                   lambda b: b"XXXXXXXX" + b[8:]):                             # bad magic
        open(bad, "wb").write(bytes(mutate(raw)))
        with pytest.raises(bcmi.BcmiError) as ex:
            bcmi.Graph.load(str(bad))
        assert ex.value.code == bcmi.ERR_IO
    with pytest.raises(bcmi.BcmiError) as ex:
        bcmi.Graph.load(str(tmp_path / "missing.csr"))
    assert ex.value.code == bcmi.ERR_IO


def test_ck_dataset_entries_and_generator(tmp_path):
    """Every run command of the CK program finds a dataset entry carrying its
    tags; the generator writes the graph the benchmarks use."""
    import json
    import subprocess
    import sys
    pkg = os.path.join(ROOT, "ck-bc_amd")
    prog = json.load(open(os.path.join(pkg, "program", "bc-hip", ".cm", "meta.json")))
    entries = {}
    for d in os.listdir(os.path.join(pkg, "dataset")):
        entries[d] = json.load(open(os.path.join(pkg, "dataset", d, ".cm", "meta.json")))
    for name, rc in prog["run_cmds"].items():
        if name == "default":
            continue                       # the reference's facebook dataset entry
        tags = set(rc["dataset_tags"])
        assert any(tags <= set(m["tags"]) for m in entries.values()), name
    out = tmp_path / "r10.csr"
    subprocess.check_call([sys.executable, os.path.join(pkg, "gen_dataset.py"), "rmat", "--scale", "10",
                           "--edgefactor", "16", "--seed", "20250101", "--out", str(out)])
    h, _ = bcmi.Graph.load(str(out))
    g = bcmi.Graph.rmat(10, 16, seed=20250101)
    assert np.array_equal(h.off, g.off) and np.array_equal(h.adj, g.adj)
    txt = tmp_path / "r10.txt"
    subprocess.check_call([sys.executable, os.path.join(pkg, "gen_dataset.py"), "rmat", "--scale", "10",
                           "--edgefactor", "16", "--seed", "20250101", "--out", str(txt), "--text"])
    t = bcmi.Graph.read_edgelist(str(txt), g.n)
    assert np.array_equal(t.off, g.off) and np.array_equal(t.adj, g.adj)


def test_tuned_table_names_known_parameters():
    """Every entry of the autotuned table only holds parameters the tuner
    applies (plus its own notes), with values the engine accepts, and apply()
    sets exactly those on an engine."""
    import autotune
    table = autotune.load()
    assert {"default", "rmat-s18-ef16", "rmat-s22-ef16", "rmat-s25-ef16"} <= set(table)
    notes = set(autotune.NOTES)
    applied = {"chunk", "block", "grid", "grid_back", "grid_plain", "alpha_batched", "beta_batched", "gamma_push",
               "fuse_bu", "pack_bu", "gamma_back", "steps", "alpha_plain", "width", "narrow_sigma", "concurrency"}
    assert set(autotune.APPLIED) == applied

    class Rec:
        def __init__(self):
            self.set = {}

        def set_param(self, k, v):
            self.set[k] = v

    for key, entry in table.items():
        assert set(entry) <= applied | notes, (key, set(entry) - applied - notes)
        assert entry["block"] in (64, 128, 256) and entry["chunk"] >= 64 and 1 <= entry["width"] <= 64
        assert entry["steps"] >= 1 and entry["grid"] >= 1 and entry.get("narrow_sigma", 1) in (0, 1, 2)
        assert 0 <= entry.get("concurrency", 0) <= 16
        r = Rec()
        autotune.apply(r, table, key)
        assert r.set == {k: int(v) for k, v in entry.items() if k in applied}
    assert autotune.graph_key("rmat", scale=22, edgefactor=16) == "rmat-s22-ef16"


def test_cli_rejects_malformed_tunables():
    """-T needs key=value with an integer value; the check comes before any
    device call, so it runs without a GPU."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ck-bc_amd", "ck-bc-hip")
    for bad in ("grid", "=4", "grid=", "grid=12a"):
        r = subprocess.run([exe, "-p", "1x1", "-T", bad, "-f", "x", "-n", "1"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and "Invalid tunable" in r.stderr, (bad, r.stderr)


def test_read_tiny_files_many_threads(tmp_path):
    """Files with fewer bytes than parser threads (every early cut falls at
    byte 0): the parser must stay inside the mapped file and still read
    every edge, with or without a final newline."""
    import subprocess
    cases = {"a.txt": ("0 1\n", [[1], [0]]), "b.txt": ("0 1", [[1], [0]]), "c.txt": ("", [[], []]),
             "d.txt": ("1 0\n0 1\n", [[1], [0]]), "e.txt": ("#\n", [[], []])}
    for name, (text, _) in cases.items():
        (tmp_path / name).write_text(text)
    code = (
        "import sys, json; sys.path[:0] = [%r, %r]\n"
        "import bcmi\n"
        "out = {}\n"
        "for f in sys.argv[1:]:\n"
        "    g = bcmi.Graph.read_edgelist(f, 2)\n"
        "    out[f] = [g.adj[g.off[v]:g.off[v + 1]].tolist() for v in range(2)]\n"
        "print(json.dumps(out))\n") % (ROOT, os.path.join(ROOT, "ck-bc_amd"))
    env = dict(os.environ, OMP_NUM_THREADS="64")
    r = subprocess.run([sys.executable, "-c", code] + [str(tmp_path / n) for n in cases], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout.strip().splitlines()[-1])
    for name, (_, want) in cases.items():
        assert got[str(tmp_path / name)] == want, name


def test_ck_run_commands_use_the_tuned_shapes():
    """The CK run commands that pass launch shapes (-T key=value) pass the
    autotuned table's values for their graph's R-MAT scale (-n = 2^scale)."""
    import autotune
    meta = json.load(open(os.path.join(ROOT, "ck-bc_amd", "program", "bc-hip", ".cm", "meta.json")))
    table = autotune.load()
    seen = 0
    for name, cmd in meta["run_cmds"].items():
        line = cmd["run_time"]["run_cmd_main"]
        tun = re.findall(r"-T (\w+)=(\d+)", line)
        if not tun:
            continue
        n = int(re.search(r"-n (\d+)", line).group(1))
        scale = n.bit_length() - 1
        assert n == 1 << scale, name
        entry = table[autotune.graph_key("rmat", scale=scale, edgefactor=16)]
        for k, v in tun:
            assert k in autotune.APPLIED and int(v) == entry[k], (name, k, v, entry[k])
        seen += 1
    assert seen >= 2


def test_tuner_measurement_repeats_to_min_time():
    """A tuner measurement repeats its run (resetting the engine before each)
    until the timed runs add up to min_time, and reports total traversed
    edges over total time: a run of a few milliseconds on a small graph is not
    timed alone."""
    import time
    import autotune

    class FakeEngine:
        def __init__(self):
            self.runs = self.resets = 0

        def reset(self):
            self.resets += 1

        def run(self, src, mode):
            self.runs += 1
            time.sleep(0.01)
            return {"te": 1e7}

    t = autotune.Tuner(None, None, 1, 1, min_time=0.05)
    e = FakeEngine()
    rate = t._rate(e, [0], bcmi.MODE_BATCHED)
    assert e.runs >= 5 and e.resets == e.runs
    assert 0.0 < rate < 1.05          # 1e7 edges per >= 10 ms run: at most 1 GTEPS
    t = autotune.Tuner(None, None, 1, 1, min_time=0.0)
    e = FakeEngine()
    t._rate(e, [0], bcmi.MODE_BATCHED)
    assert e.runs == 1
