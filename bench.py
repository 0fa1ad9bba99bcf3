"""Benchmark: batched multi-source Brandes BC on R-MAT scale 22 (1 GPU).

A step is one batch of 64 sources (one forward + backward pass of the
batched sweep over the graph).  The workload is R-MAT scale 22, edge factor
16, a/b/c/d = .57/.19/.19/.05, ids permuted, 16,384 seeded sampled sources in
batches of 64.  The batches are the ones the engine forms over all 16,384
sources (it groups the sources of a run by similarity, so the whole set
batches better than a part of it; bcmi_batch_order gives that grouping).  By
default 2 warmup batches and the other 254 are timed; with --steps K below
that, K batches spread evenly over the schedule are timed.  The graph and
state are already resident in HBM.

Printed JSON line (one): GTEPS of the batched path (traversed edges = 2 *
sum over sources of the degree sum of the vertices each source reaches),
the plain one-source-per-sweep GTEPS on `--plain-sources` of the timed
sources (spread over the timed batches), the whole run's model bytes per
second against HBM peak, the roofline of the dominant kernel (algorithmic
bytes from the frozen byte model / kernel duration from a device-side timer
inside the captured level loop, on a pass of `--roof-batches` of the timed
batches run one at a time), the HBM traffic per launch of that kernel from the
committed rocprofv3 counter runs, the CPU oracle timed on a bounded sample
and a served-query run (p50/p99 latency).

Other workloads (not the default line):
  --scale 18 --pendant 0.3 --sources 8192 --steps 126
                                              R-MAT 18 + 30% pendant trees, peeled
  --scale 25 --sources 4096 --steps 62        R-MAT scale 25
  --graph chain --chain-blocks 32768 --sources 1024 --steps 12 --warmup 4 --kernel none --wg-mem 192
                                              bridged-block chain, 8.4M vertices
  --deg2 [--graph facebook | --scale 16 [--pendant 0.3]]
                                              2-degree heuristic: plain sweeps over every vertex as a
                                              source, with and without it (-H 2; -H 3 when peeled)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ck-bc_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def serve_bench(args, log):
    """Served BC queries: a pool of graphs of uneven size (R-MAT scales 12-20,
    edge factor 8-32, and bridged-block chains of 256-vertex Erdos-Renyi
    blocks), Poisson arrivals, each query = (graph, 16..1024 sampled sources).
    Returns p50/p99 latency (arrival to completion) and aggregate GTEPS."""
    import threading
    import bcmi
    import serve_pool
    graphs, _, rng = serve_pool.make_pool(args.seed, args.serve_graphs, args.serve_max_blocks)
    srv = bcmi.Server(args.serve_budget_gb << 30, args.serve_slots)
    ids = [srv.add_graph(g) for g in graphs]
    nq = args.serve_queries
    gsel, qsrc, gaps = serve_pool.make_queries(rng, graphs, args.seed, nq, args.serve_rate)
    # warm each graph's workspace once (first use captures the level loops)
    for gid, g in zip(ids, graphs):
        q = srv.submit(gid, g.sample_sources(1, seed=1))
        srv.wait(q, g.n, want_bc=False)
    qids = [None] * nq
    results = [None] * nq

    def waiter():
        for i in range(nq):
            while qids[i] is None:
                time.sleep(0.0005)
            results[i] = srv.wait(qids[i], graphs[gsel[i]].n, want_bc=False)

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    for i in range(nq):
        t_next += gaps[i]
        d = t_next - time.perf_counter()
        if d > 0:
            time.sleep(d)
        qids[i] = srv.submit(ids[gsel[i]], qsrc[i])
    th.join()
    wall = time.perf_counter() - t0
    lat = np.array([r[2].finish_ms - r[2].arrive_ms for r in results])
    ok = all(r[0] == 0 for r in results)
    te = sum(r[2].te for r in results)
    first = min(r[2].arrive_ms for r in results)
    last = max(r[2].finish_ms for r in results)
    st = srv.stats()
    srv.close()
    out = {"queries": nq, "graphs": len(graphs), "rate_qps": args.serve_rate, "slots": args.serve_slots,
           "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
           "gteps": round(te / ((last - first) / 1e3) / 1e9, 2), "all_ok": ok,
           "peak_slots": st["peak_slots"], "peak_gb": round(st["peak_bytes"] / 2**30, 2), "wall_s": round(wall, 1)}
    log(f"serving: {out}")
    return out


def deg2_bench(args):
    """Plain sweeps from every vertex of the graph (exact BC), once as they
    are and once with the 2-degree heuristic (a degree-2 source whose two
    neighbours are sources takes its forward state from their sweeps), each
    timed after a warm run.  Prints one JSON line: both GTEPS (the same
    traversed-edge count), the sources taken from their neighbours, the gain."""
    import bcmi
    if args.graph == "facebook":
        f = np.load(os.path.join(ROOT, "tests", "golden", "facebook.npz"))
        g = bcmi.Graph.from_edges(f["edges"], int(f["n"]))
        gdesc = "facebook_combined (the reference's dataset)"
    else:
        g = bcmi.Graph.rmat(args.scale, args.edgefactor, seed=args.seed)
        gdesc = f"R-MAT scale {args.scale} ef {args.edgefactor} (a=.57 b=.19 c=.19)"
    if args.pendant > 0:
        g = g.with_pendant_trees(args.pendant, seed=args.seed + 2)
        gdesc += f" + pendant trees ({args.pendant:.0%} more vertices, peeled)"
    peel = args.pendant > 0
    e = bcmi.Engine(0)
    for kv in args.set:
        k, v = kv.split("=", 1)
        e.set_param(k, int(v))
    e.upload(g, heuristic=bcmi.H_PEEL if peel else bcmi.H_NONE)
    src = np.arange(g.n, dtype=np.int32)
    res = {}
    for deg2 in (False, True):
        e.run(src[:64], mode=bcmi.MODE_PLAIN, expand=peel, deg2=deg2)  # warm (capture, slots)
        e.reset()
        t0 = time.perf_counter()
        st = e.run(src, mode=bcmi.MODE_PLAIN, expand=peel, deg2=deg2)
        dt = time.perf_counter() - t0
        res[deg2] = (dt, st)
        log(f"deg2={deg2}: {dt:.3f}s, {st['te'] / dt / 1e9:.3f} GTEPS, {st['deg2_sources']} of "
            f"{st['sources_run']} sweeps from their neighbours")
    (t0_, s0), (t1_, s1) = res[False], res[True]
    out = {"metric": "plain BC GTEPS with and without the 2-degree heuristic (all sources, 1 GPU)",
           "value": round(s1["te"] / t1_ / 1e9, 3), "unit": "GTEPS", "n_gpus": 1, "higher_is_better": True,
           "dtype": "u64 sigma / f64 delta", "data": "synthetic, seeded" if args.graph != "facebook" else "reference dataset",
           "config": {"workload": f"plain Brandes BC, every vertex a source ({'-H 3' if peel else '-H 2'}), {gdesc}",
                      "n": g.n, "arcs": g.m},
           "without_gteps": round(s0["te"] / t0_ / 1e9, 3), "sweeps": s1["sources_run"],
           "from_neighbours": s1["deg2_sources"], "gain": round(t0_ / t1_ - 1.0, 4)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=254,
                    help="timed batches of 64 sources (254 + 2 warmup = all 16,384 sources)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--sources", type=int, default=16384)
    ap.add_argument("--plain-sources", type=int, default=64)
    ap.add_argument("--cpu-sources", type=int, default=96)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roof-batches", type=int, default=32,
                    help="timed batches the kernel roofline pass runs one at a time")
    ap.add_argument("--kernel", default="back", choices=["back", "sigma", "td", "bu", "none"],
                    help="kernel timed on the device: back (the roofline kernel), sigma, td or bu "
                         "(time only), none (no device timer)")
    ap.add_argument("--graph", default="rmat", choices=["rmat", "chain", "facebook"],
                    help="chain: bridged-block chain of --chain-blocks Erdos-Renyi blocks of 256 vertices, degree 8; "
                         "facebook: the reference's dataset (the CK run's graph, from tests/golden; --deg2 only)")
    ap.add_argument("--deg2", action="store_true",
                    help="time plain sweeps over all sources with and without the 2-degree heuristic")
    ap.add_argument("--chain-blocks", type=int, default=32768)
    ap.add_argument("--wg-mem", type=int, default=0, metavar="GIB",
                    help="device memory for one-workgroup-per-source sweeps on deep graphs (0: engine default)")
    ap.add_argument("--pendant", type=float, default=0.0,
                    help="attach pendant trees adding this fraction of vertices and peel them (-H 1, exact sources)")
    ap.add_argument("--serve-queries", type=int, default=48, help="0 skips the serving run")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="Poisson arrival rate (queries/s)")
    ap.add_argument("--serve-graphs", type=int, default=32)
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-budget-gb", type=int, default=96)
    ap.add_argument("--serve-max-blocks", type=int, default=2048)
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="engine parameter applied after the tuned table (repeatable)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (PMC byte calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        a = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB five times")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "sources shard across GPUs with one FP64 all-reduce at the end, "
              "which this single-GPU build does not run.", file=sys.stderr)
        sys.exit(2)

    import bcmi

    if args.deg2:
        deg2_bench(args)
        return

    t0 = time.time()
    if args.graph == "chain":
        g = bcmi.Graph.block_chain(args.chain_blocks, 256, 8.0, seed=args.seed)
        gdesc = f"bridged-block chain of {args.chain_blocks} blocks x 256 vertices (degree 8)"
    else:
        g = bcmi.Graph.rmat(args.scale, args.edgefactor, seed=args.seed)
        gdesc = f"R-MAT scale {args.scale} ef {args.edgefactor} (a=.57 b=.19 c=.19)"
    if args.pendant > 0:
        g = g.with_pendant_trees(args.pendant, seed=args.seed + 2)
        gdesc += f" + pendant trees ({args.pendant:.0%} more vertices, peeled)"
    heur = bcmi.H_PEEL if args.pendant > 0 else bcmi.H_NONE
    log(f"graph {gdesc}: n={g.n} arcs={g.m} built in {time.time() - t0:.1f}s")
    width = 64
    src = g.sample_sources(args.sources, seed=args.seed + 1)

    e = bcmi.Engine(0)
    if args.wg_mem:
        e.set_param("wg_mem", args.wg_mem << 30)
    import autotune
    tuned = None
    if args.graph == "rmat":
        tuned = autotune.apply(e, autotune.load(), autotune.graph_key("rmat", scale=args.scale,
                                                                      edgefactor=args.edgefactor))
    if tuned:
        log(f"tuned parameters: {tuned}")
        width = int(tuned.get("width", 64))
    for kv in args.set:
        k, v = kv.split("=", 1)
        e.set_param(k, int(v))
        if k == "width":
            width = int(v)
        log(f"parameter {k} = {int(v)}")
    t0 = time.time()
    e.upload(g, heuristic=heur)
    log(f"uploaded in {time.time() - t0:.1f}s")

    need = (args.warmup + args.steps) * width
    if need > len(src):
        src = g.sample_sources(need, seed=args.seed + 1)
    kt = {"sigma": 1, "back": 2, "td": 3, "bu": 4, "none": 0}[args.kernel]
    nb = len(src) // width
    if heur == bcmi.H_NONE:
        # A step is one batch of the whole workload: the engine's grouping of all
        # the sampled sources into batches of `width` (bcmi_batch_order).  With
        # fewer steps than batches the timed ones are spread evenly over that
        # schedule and the warmup batches are others, so every timed batch is
        # one the full run would execute; "batching" 0 keeps them as formed.
        order = e.batch_order(src)
        batches = [order[i * width:(i + 1) * width] for i in range(nb)]
        if args.warmup + args.steps >= nb:
            warm_ix, timed_ix = range(args.warmup), range(args.warmup, args.warmup + args.steps)
        else:
            timed_ix = sorted({int((i + 0.5) * nb / args.steps) for i in range(args.steps)})
            rest = [i for i in range(nb) if i not in set(timed_ix)]
            warm_ix = rest[::max(1, len(rest) // max(1, args.warmup))][:args.warmup]
        e.set_param("batching", 0)
        warm = np.concatenate([batches[i] for i in warm_ix]) if args.warmup else np.zeros(0, np.int32)
        timed = np.concatenate([batches[i] for i in timed_ix])
        schedule = f"{len(timed_ix)} of the {nb} batches the engine forms over all {len(src)} sources"
    else:
        warm, timed = src[:args.warmup * width], src[args.warmup * width: need]
        schedule = f"{len(timed)} of {len(src)} sources, batched among themselves"
    # warmup (also captures the level-loop graphs); the timed run carries no
    # kernel timer (the roofline pass below does)
    e.set_param("ktime", 0)
    if args.warmup:
        e.run(warm, mode=bcmi.MODE_BATCHED)
    e.reset()
    # bcmi_bc_run returns after its last kernel and copy have completed
    t0 = time.perf_counter()
    st = e.run(timed, mode=bcmi.MODE_BATCHED)
    t_batched = time.perf_counter() - t0
    te = st["te"]
    gteps = te / t_batched / 1e9
    log(f"batched: {len(timed)} sources in {t_batched:.3f}s -> {gteps:.2f} GTEPS, "
        f"model bytes {st['bytes'] / 1e9:.1f} GB ({st['bytes'] / t_batched / 1e9:.0f} GB/s whole-run), "
        f"{st['sum_pairs'] / max(1, st['sum_union']):.3f} levels per reached vertex per batch")

    # Dominant-kernel roofline.  The batched byte model per batch is
    # B(S) = 24 m_S + 32 n_S + sum_s (24 a_s + 32 n_s); each of the two passes
    # (forward, backward) reads the CSR and probes masks once, so each gets half
    # of the union part; the backward pass (k_ms_back) gets sum_s (16 a_s + 16 n_s)
    # (coef of successors + sigma/coef of the vertex itself).  The forward
    # pass's share is spread over its kernels as the levels go (bottom-up levels
    # sum path counts in their own scan, top-down levels in k_ms_sigma or
    # k_ms_push), so a forward kernel gets its time but no byte figure.
    # The kernel is timed on a pass of its own over `--roof-batches` of the
    # timed batches (spread over them) run one at a time: batches that run at
    # once (the tuned concurrency) overlap their kernels, and each launch then
    # takes longer while it shares the device, so its duration would not be
    # the kernel's.  The timer: every wave of the timed kernel stamps entry and
    # exit; the launch lasts from the first entry to the last exit.
    k_ms, k_launch, kbytes, achieved, roof_n = 0.0, 0, None, None, 0
    if kt:
        nt = len(timed) // width
        rix = sorted({int((i + 0.5) * nt / min(args.roof_batches, nt)) for i in range(min(args.roof_batches, nt))})
        roof = np.concatenate([timed[i * width:(i + 1) * width] for i in rix])
        roof_n = len(rix)
        conc = e.get_param("concurrency")
        e.set_param("concurrency", 1)
        e.set_param("ktime", kt)
        e.run(roof[:width], mode=bcmi.MODE_BATCHED)  # captures the timed graphs
        e.reset()
        st_r = e.run(roof, mode=bcmi.MODE_BATCHED)
        k_ms, k_launch = e.kernel_time()
        ss = e.sweep_stats(len(roof))
        e.set_param("ktime", 0)
        e.set_param("concurrency", conc)
        n_s, a_s = ss["n_s"].astype(np.float64), ss["a_s"].astype(np.float64)
        per_src = float(np.sum(24.0 * a_s + 32.0 * n_s))
        union_part = st_r["bytes"] - per_src
        kbytes = union_part / 2.0 + float(np.sum(16.0 * a_s + 16.0 * n_s)) if args.kernel == "back" else None
        achieved = kbytes / (k_ms / 1e3) / 1e9 if kbytes and k_ms > 0 else None
    # with pushed backward levels the backward timer also covers k_ms_bpush and k_ms_bfin
    kname = f"k_ms_{args.kernel}"
    if args.kernel == "back" and e.get_param("gamma_back") > 0:
        kname = "k_ms_back+k_ms_bpush+k_ms_bfin"
    if args.kernel == "bu" and e.get_param("pack_bu"):
        kname = "k_ms_bu+k_ms_bu_pk"  # the bottom-up timer covers both kernels
    log(f"kernel {args.kernel}: {k_launch} launches, {k_ms:.1f} ms total, "
        f"avg {k_ms / max(k_launch, 1) * 1e3:.1f} us, achieved {achieved} GB/s")

    # plain path on the same (first) sources
    # plain sweeps and the CPU sample on sources spread over the timed batches
    plain_src = timed[np.linspace(0, len(timed) - 1, min(args.plain_sources, len(timed))).astype(np.int64)]
    gteps_plain = None
    try:
        e.run(plain_src[:4], mode=bcmi.MODE_PLAIN)  # warm (capture)
        e.reset()
        t0 = time.perf_counter()
        stp = e.run(plain_src, mode=bcmi.MODE_PLAIN)
        t_plain = time.perf_counter() - t0
        gteps_plain = stp["te"] / t_plain / 1e9
        log(f"plain: {len(plain_src)} of the timed sources in {t_plain:.3f}s "
            f"-> {gteps_plain:.2f} GTEPS")
    except bcmi.BcmiError as ex:
        if ex.code != bcmi.ERR_OVERFLOW:
            raise
        log("plain: path counts exceed 2^64 on this graph (the plain sweep is exact-only)")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        nth = args.cpu_threads or min(16, os.cpu_count() or 1)
        cs = timed[np.linspace(0, len(timed) - 1, min(args.cpu_sources, len(timed))).astype(np.int64)]
        # plain Brandes on the uploaded graph (no peeling); on chains the uint64
        # path counts wrap, which changes no work: the run is timed only
        O.brandes(g.off, g.adj, cs[:1], nthreads=1, allow_overflow=True)  # page in
        t0 = time.perf_counter()
        _, ost = O.brandes(g.off, g.adj, cs, nthreads=nth, stats=True, allow_overflow=True)
        t_cpu = time.perf_counter() - t0
        te_cpu = 2.0 * float(np.sum(ost["m_s"]))
        cpu = {"value": round(te_cpu / t_cpu / 1e9, 4), "unit": "GTEPS", "cores": nth, "kind": "port",
               "sample": f"{len(cs)} of the timed sources on the same graph ({gdesc}), "
                         f"{t_cpu:.1f}s, OpenMP over sources"}
        log(f"cpu oracle: {cpu}")

    # the serving run allocates its own graphs and slots: free this engine's
    # device memory first (a deep graph's workgroup sweeps may hold most of it)
    e.close()
    serving = serve_bench(args, log) if args.serve_queries > 0 else None

    # HBM traffic per launch of the roofline kernel, from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload (copy-calibrated)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_rmat{args.scale}.json")
    if os.path.exists(tpath) and args.edgefactor == 16 and args.graph == "rmat" and args.pendant == 0:
        t = json.load(open(tpath)).get("per_kernel", {}).get(f"k_ms_{args.kernel}")
        if t:
            traffic = round(t["read_bytes_per_dispatch"] + t["write_bytes_per_dispatch"])

    out = {
        "metric": "BC GTEPS on R-MAT scale-22 (batched vs plain, 1 GPU) + p50/p99 query latency",
        "value": round(gteps, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched / max(1, args.steps) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 sigma / f64 delta",
        "data": "synthetic, seeded",
        "config": {"workload": f"batched Brandes BC, {gdesc}, sampled sources in batches of {width}: {schedule}",
                   "n": g.n, "arcs": g.m, "batch": width,
                   "tuned": {k: v for k, v in (tuned or {}).items() if k in
                             ("chunk", "block", "grid", "grid_back", "alpha_batched", "beta_batched", "gamma_push", "fuse_bu", "pack_bu",
                              "gamma_back", "steps", "width", "narrow_sigma", "concurrency")}},
        "plain_gteps": round(gteps_plain, 3) if gteps_plain else None,
        "batched_over_plain": round(gteps / gteps_plain, 2) if gteps_plain else None,
        "model_bytes_gbs": round(st["bytes"] / t_batched / 1e9, 1),
        "model_bytes_frac": round(st["bytes"] / t_batched / 1e9 / HBM_PEAK_GBS, 4),
        "levels_per_vertex_batch": round(st["sum_pairs"] / max(1, st["sum_union"]), 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                     "traffic": traffic, "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE+WRITE_SIZE)",
                     "kernel": kname if kt else None,
                     "pass": f"{roof_n} of the timed batches, one at a time" if kt else None,
                     "avg_launch_us": round(k_ms / max(k_launch, 1) * 1e3, 2)},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
