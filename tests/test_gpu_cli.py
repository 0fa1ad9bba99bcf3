"""The drop-in program end to end: the CK run of the reference
(-p 1x1 -f facebook_combined.txt -n 4039 -N 1000) writes <o>_1X1_0.log with
the reference's header and %.2f rows of BC/2, here checked against the
networkx golden vector for sources 0..999."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ck-bc_amd", "ck-bc-hip")


def write_facebook(tmp_path):
    f = np.load(os.path.join(ROOT, "tests", "golden", "facebook.npz"))
    fn = tmp_path / "facebook_combined.txt"
    with open(fn, "w") as fp:
        for a, b in f["edges"].astype(np.int64):
            fp.write(f"{a} {b}\n")
    return fn, f


def read_result(path):
    lines = open(path).read().splitlines()
    assert lines[0] == "BC RESULTS" and lines[1] == "ROWPP 4039" and lines[2] == "NodeId\tBC_VAL"
    rows = [l.split("\t") for l in lines[3:]]
    assert [int(r[0]) for r in rows] == list(range(4039))
    return np.array([float(r[1]) for r in rows])


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_ck_run_facebook(tmp_path, mode):
    fn, f = write_facebook(tmp_path)
    out = tmp_path / "res"
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "1000", "-o", str(out), "-B", mode],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "BC rounds: 1000" in r.stdout and "------------- RESULTS ---------------" in r.stdout
    got = read_result(str(out) + "_1X1_0.log")
    want = np.round(f["bc_src1000"], 2)
    assert np.max(np.abs(got - want)) <= 0.0100001


def test_ck_run_heuristic_full_precision(tmp_path):
    fn, f = write_facebook(tmp_path)
    out = tmp_path / "res"
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "4039", "-H", "1", "-P",
                        "-o", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = read_result(str(out) + "_1X1_0.log")
    want = f["bc_all"]
    nz = want != 0
    assert np.all(got[~nz] == 0)
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) < 1e-9


@pytest.mark.parametrize("heur", ["2", "3"])
def test_ck_run_two_degree_heuristic(tmp_path, heur):
    """-H 2 / -H 3 (the reference accepts them and prints REMOVED, bc2d.c:759-760,
    c:1015): plain sweeps take degree-2 sources from their neighbours' sweeps
    and the result file holds the same BC as networkx over all sources."""
    fn, f = write_facebook(tmp_path)
    out = tmp_path / "res"
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "4039", "-H", heur, "-B", "plain",
                        "-P", "-o", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "REMOVED" not in r.stdout
    line = [l for l in r.stdout.splitlines() if "taken from their neighbours" in l]
    assert line and int(line[0].split()[-1]) > 0
    got = read_result(str(out) + "_1X1_0.log")
    want = f["bc_all"]
    nz = want != 0
    assert np.all(got[~nz] == 0)
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) < 1e-9


def test_ck_run_tunables(tmp_path):
    """-T key=value reaches the engine: a launch shape of small records, few
    levels per captured graph and early bottom-up gives the same BC; an
    unknown key fails with exit status 1."""
    fn, f = write_facebook(tmp_path)
    out = tmp_path / "res"
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "1000", "-P", "-o", str(out),
                        "-T", "chunk=64", "-T", "steps=4", "-T", "alpha_batched=125", "-T", "grid=2048"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = read_result(str(out) + "_1X1_0.log")
    want = f["bc_src1000"]
    nz = want != 0
    assert np.all(got[~nz] == 0)
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) < 1e-9
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "10", "-T", "nosuch=1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "nosuch" in (r.stdout + r.stderr)


def test_ck_run_errors(tmp_path):
    fn, _ = write_facebook(tmp_path)
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "invalid edge" in r.stderr
    r = subprocess.run([EXE, "-p", "2x2", "-f", str(fn), "-n", "4039"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1


def test_ck_runner(tmp_path):
    fn, _ = write_facebook(tmp_path)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "ck-bc_amd", "ck_run.py"),
                        os.path.join(ROOT, "ck-bc_amd", "program", "bc-hip"), "--dataset", str(fn)],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "BC rounds: 1000" in r.stdout


def test_csr_cache(tmp_path):
    """-C: the first run parses the text and writes the binary CSR, the next
    reads it (same result file); a changed text file is parsed again."""
    fn, f = write_facebook(tmp_path)
    cache = tmp_path / "fb.csr"
    outs = []
    for k in range(3):
        if k == 2:
            os.utime(fn, ns=(0, 10**18))                # the text changed (mtime)
        out = tmp_path / f"res{k}"
        r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "4039", "-N", "200", "-o", str(out),
                            "-C", str(cache)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert ("Graph read from CSR cache" in r.stdout) == (k == 1)
        outs.append(open(str(out) + "_1X1_0.log").read())
    assert cache.exists()
    assert outs[0] == outs[1] == outs[2]


def test_ck_run_generated_dataset_entry(tmp_path):
    """CK-style run of a generated dataset entry (R-MAT scale 10, all 1,024
    sources): the runner creates the binary CSR from the entry's recipe and
    the program reads it through -f; BC matches the oracle."""
    import shutil
    sys.path.insert(0, os.path.join(ROOT, "ck-bc_amd"))
    import bcmi
    from oracle import oracle as O
    ent = tmp_path / "graph-rmat-s10"
    shutil.copytree(os.path.join(ROOT, "ck-bc_amd", "dataset", "graph-rmat-s10"), ent)
    out = tmp_path / "res"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "ck-bc_amd", "ck_run.py"),
                        os.path.join(ROOT, "ck-bc_amd", "program", "bc-hip"), "--cmd", "rmat10",
                        "--dataset-entry", str(ent), "--extra", f"-o {out} -P"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Graph read from CSR file" in r.stdout
    g = bcmi.Graph.rmat(10, 16, seed=20250101)
    lines = open(str(out) + "_1X1_0.log").read().splitlines()
    got = np.array([float(l.split("\t")[1]) for l in lines[3:]])
    ref = O.brandes(g.off, g.adj, np.arange(g.n)) / 2.0
    nz = ref != 0
    assert np.all(got[~nz] == 0) and np.max(np.abs(got[nz] - ref[nz]) / ref[nz]) < 1e-9


def parse_results(stdout):
    """The RESULTS block of the reference's stdout (bc2d.c:1074-1097)."""
    lines = stdout.splitlines()
    i = lines.index("------------- RESULTS ---------------")
    assert lines[i + 1] == "ClusterId\tSkip\tRounds\tExecTime\tRoundsTime\tMax\tMin\tMean\t1-dReduTime"
    row = lines[i + 2].split("\t")
    out = dict(skip=int(row[1]), rounds=int(row[2]), exec=float(row[3]), rounds_time=float(row[4]),
               max=float(row[5]), min=float(row[6]), mean=float(row[7]))
    for l in lines[i + 3:]:
        if l.startswith("task 0 BC M-TEPS: "):
            out["mteps"] = float(l.split(": ")[1].split()[0])
        if l.startswith("task 0 BC simulated time: "):
            rest = l.split(": ", 1)[1]
            out["simulated"] = float(rest.split()[0])
            out["vskipped"] = int(rest.split("virtual skipped ")[1].rstrip(")"))
    return out


@pytest.mark.parametrize("mode", ["plain", "batched"])
def test_results_block_statistics(tmp_path, mode):
    """Per-source statistics as the reference computes them: Max and Min
    over the sources' own times (c:1047-1048; a batched sweep's time shared
    by its sources), Mean = rounds time / rounds, M-TEPS the mean of the
    per-source rates (c:569-579, c:1091), and, with -N below the vertex
    count, the simulated whole-graph time avg * (N - skipped - vskipped) with
    vskipped = (N - nbfs) * skipped / nbfs (c:1092-1097)."""
    # a graph with isolated vertices among the first sources: 0..9 isolated,
    # a ring over 10..199
    fn = tmp_path / "g.txt"
    with open(fn, "w") as fp:
        for v in range(10, 200):
            fp.write(f"{v} {10 + (v - 9) % 190}\n")
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "200", "-N", "100", "-B", mode],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    s = parse_results(r.stdout)
    assert s["skip"] == 10 and s["rounds"] == 90
    # times print with 1 us resolution
    assert 0 < s["min"] <= s["max"] <= s["rounds_time"] + 1e-6
    assert abs(s["mean"] - s["rounds_time"] / 90) <= 2e-6
    # each source reaches the 190-vertex ring (degree 2): (380 / 2) / t / 2^20 per
    # source, so the mean rate lies between the rates at the max and min times
    lo, hi = 190 / (s["max"] + 1e-6) / 2**20, 190 / max(s["min"] - 1e-6, 1e-9) / 2**20
    assert lo * 0.999 <= s["mteps"] <= hi * 1.001
    vsk = (200 - 100) * 10 // 100
    assert s["vskipped"] == vsk
    # the engine multiplies the unrounded mean; the printed one is off by <= 0.5 us
    assert abs(s["simulated"] - s["mean"] * (200 - 10 - vsk)) <= (200 - 10 - vsk) * 0.5e-6 + 1e-6
    # all sources: no simulated line
    r = subprocess.run([EXE, "-p", "1x1", "-f", str(fn), "-n", "200", "-N", "200", "-B", mode],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "simulated" not in parse_results(r.stdout)
