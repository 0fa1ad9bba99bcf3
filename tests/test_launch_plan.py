"""The captured level loop's launch plan against the kernels' own resource
limits, on the CPU: every node of every captured graph (forward, backward,
sparse forward, sparse backward; plain and batched), for the default
tunables, every entry of the tuned table and the extremes bcmi_set_param
accepts, checked against the gfx950 code object's metadata (workgroup size
limit, static LDS, kernel-argument and scratch sizes) and the AQL dispatch
packet's 32-bit grid size.  The device timer's two nodes must follow the
timed kernel, fold exactly that launch's waves and fit the stamp buffer."""
import os
import subprocess

import pytest
import yaml

import autotune
import bcmi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "ck-bc_amd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
TIMER_MAX_WAVES = 65536 * 4   # per-wave stamp slots of the device timer
APPENDER_MAX_WAVES = 4        # LDS staging areas per block in the expansion kernels
STEP_KERNELS = {"k_plain_decide", "k_plain_td", "k_plain_bu", "k_bdecide", "k_plain_back", "k_plain_bfin",
                "k_ms_decide", "k_ms_td", "k_ms_bu", "k_ms_bu_pk", "k_ms_sigma", "k_ms_push", "k_ms_bdecide",
                "k_ms_back", "k_ms_bpush", "k_ms_bfin", "k_kt_accum", "k_kt_fold"}


@pytest.fixture(scope="module")
def kernel_meta(tmp_path_factory):
    """Kernel metadata of engine.hip's gfx950 code object, keyed by kernel name."""
    d = tmp_path_factory.mktemp("co")
    bundle, co = str(d / "engine.bundle"), str(d / "engine-gfx950.co")
    subprocess.check_call([f"{ROCM}/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
                           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                           f"-I{ROOT}/include", f"-I{PKG}/csrc", "--offload-device-only", "-c",
                           f"{PKG}/csrc/engine.hip", "-o", bundle], stderr=subprocess.DEVNULL)
    subprocess.check_call([f"{ROCM}/llvm/bin/clang-offload-bundler", "--unbundle", "--type=o", f"--input={bundle}",
                           "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"])
    notes = subprocess.check_output([f"{ROCM}/llvm/bin/llvm-readelf", "--notes", co], text=True)
    lines = notes.splitlines()
    a = next(i for i, l in enumerate(lines) if l.strip() == "---")
    b = next(i for i, l in enumerate(lines) if l.strip() == "..." and i > a)
    meta = yaml.safe_load("\n".join(lines[a + 1:b]))
    out = {}
    for k in meta["amdhsa.kernels"]:
        for name in STEP_KERNELS:
            # kernels live in namespace bcmi: _ZN4bcmi<len><name>E...
            if f"N4bcmi{len(name)}{name}E" in k[".name"]:
                out[name] = k
    return out


def configs():
    yield "defaults", {}
    table = autotune.load()
    for key in sorted(table):
        yield f"tuned {key}", {k: v for k, v in table[key].items()
                               if k in autotune.APPLIED}
    yield "largest grids", {"grid": 65536, "grid_back": 65536, "grid_plain": 65536, "sgrid": 65536, "block": 256}
    yield "smallest blocks", {"grid": 1, "grid_back": 1, "grid_plain": 1, "sgrid": 1, "block": 64}
    yield "pushed backward, one wide step", {"gamma_back": 8, "steps": 1, "block": 128}
    yield "unpacked bottom-up, longest graph", {"pack_bu": 0, "steps": 4096}


@pytest.mark.parametrize("ktime", range(8))
def test_launch_plan_fits_gfx950(kernel_meta, ktime):
    assert set(kernel_meta) == STEP_KERNELS
    for name, prm in configs():
        prm = dict(prm, ktime=ktime)
        timed_seen = 0
        for mode in (bcmi.MODE_PLAIN, bcmi.MODE_BATCHED):
            for graph in range(4):
                plan, steps = bcmi.launch_plan(prm, mode, graph)
                assert plan, (name, mode, graph)
                assert steps * len(plan) <= 4096 * 8
                for i, nd in enumerate(plan):
                    k = kernel_meta[nd["kernel"]]
                    where = (name, mode, graph, nd)
                    assert 1 <= nd["block"] <= k[".max_flat_workgroup_size"], where
                    assert nd["block"] == 1 or nd["block"] % 64 == 0, where
                    assert nd["grid"] >= 1 and nd["grid"] * nd["block"] < 2**32, where
                    assert k[".group_segment_fixed_size"] <= 64 * 1024, where
                    assert k[".kernarg_segment_size"] <= 4096, where
                    assert k[".private_segment_fixed_size"] <= 256, where
                    if nd["kernel"] not in ("k_kt_accum", "k_kt_fold"):
                        assert nd["block"] <= 64 * APPENDER_MAX_WAVES, where
                    if nd["timed"]:
                        timed_seen += 1
                        acc, fold = plan[i + 1], plan[i + 2]
                        assert acc["kernel"] == "k_kt_accum" and fold["kernel"] == "k_kt_fold", where
                        assert acc["aux"] == nd["grid"] * nd["block"] // 64 <= TIMER_MAX_WAVES, where
                        assert (acc["grid"], acc["block"], fold["grid"], fold["block"]) == (64, 1024, 1, 64)
                    if nd["kernel"] == "k_kt_accum":
                        assert i >= 1 and plan[i - 1]["timed"], where
                    if nd["kernel"] in ("k_plain_decide", "k_ms_decide"):
                        assert nd["grid"] == 1 and nd["block"] == 1 and nd["aux"] == (graph == 0), where
        # the selected kernel is timed in some graph, no other kernel is
        assert (timed_seen > 0) == (ktime != 0), (name, ktime)


def test_plan_matches_kernel_count_per_level():
    """Nodes per level step at the tuned scale-22 shape: the batched forward
    step is decide + top-down + two bottom-up kernels + sigma + push, the
    backward step decide + backward (+ the timer's two nodes) + finish."""
    p = {k: v for k, v in autotune.load()["rmat-s22-ef16"].items() if k in autotune.APPLIED}
    fwd, steps = bcmi.launch_plan(dict(p, ktime=2), bcmi.MODE_BATCHED, 0)
    bwd, _ = bcmi.launch_plan(dict(p, ktime=2), bcmi.MODE_BATCHED, 1)
    assert [n["kernel"] for n in fwd] == ["k_ms_decide", "k_ms_td", "k_ms_bu", "k_ms_bu_pk", "k_ms_sigma", "k_ms_push"]
    assert [n["kernel"] for n in bwd] == ["k_ms_bdecide", "k_ms_back", "k_kt_accum", "k_kt_fold", "k_ms_bfin"]
    assert steps == p["steps"]
    assert bwd[1]["grid"] == p["grid_back"] and fwd[1]["grid"] == p["grid"]


def test_plan_rejects_bad_parameters():
    with pytest.raises(bcmi.BcmiError):
        bcmi.launch_plan({"block": 512}, bcmi.MODE_BATCHED, 0)
    with pytest.raises(bcmi.BcmiError):
        bcmi.launch_plan({"grid": 0}, bcmi.MODE_PLAIN, 1)
    with pytest.raises(bcmi.BcmiError):
        bcmi.launch_plan({}, bcmi.MODE_PLAIN, 4)
