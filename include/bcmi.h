/*
 * bcmi.h -- C ABI of the MI355X (gfx950) Brandes betweenness-centrality
 * engine.  Plain pointers and sizes only; every call returns a status code
 * (BCMI_OK == 0) and bcmi_last_error() gives the message for the calling
 * thread.  Nothing here exits the process (the reference's MY_CUDA_CHECK /
 * prexit did, program/bc-cuda/cudamacro.h:1-16, bc2d.c:64-72).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the ck-bc repository root; "cu" = program/bc-cuda/cuda_kernels.cu,
 * "c" = program/bc-cuda/bc2d.c, "h" = program/bc-cuda/bc2d.h).
 *
 * Conventions: vertex ids int32, CSR offsets int64, sigma uint64, delta/BC
 * double.  BC values are ordered-pair sums (the reference halves them only
 * when writing its result file, c:1065).
 */
#ifndef BCMI_H
#define BCMI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BCMI_OK 0
#define BCMI_ERR_ARG 1      /* invalid argument */
#define BCMI_ERR_HIP 2      /* HIP runtime failure */
#define BCMI_ERR_NOMEM 3    /* device or host allocation failed, host threads unavailable, or over budget */
#define BCMI_ERR_OVERFLOW 4 /* a shortest-path count exceeded 2^64-1 */
#define BCMI_ERR_STATE 5    /* call made in the wrong state (e.g. no graph) */
#define BCMI_ERR_IO 6       /* file could not be read or parsed */
#define BCMI_ERR_RANGE 7    /* vertex id out of range */
#define BCMI_ERR_BUSY 8     /* serving: queue full / query not finished */

/* BC execution modes */
#define BCMI_MODE_PLAIN 0   /* one source per sweep */
#define BCMI_MODE_BATCHED 1 /* up to 64 sources per sweep (64-bit source masks) */

/* degree-1 heuristic (-H) */
#define BCMI_H_NONE 0       /* exact Brandes on the whole graph */
#define BCMI_H_PEEL 1       /* iterative pendant peeling + closed-form terms */

/* source semantics when peeling (bcmi_bc_run flags) */
#define BCMI_SRC_EXACT 0    /* BC summed over exactly the listed sources */
#define BCMI_SRC_EXPAND 1   /* reference semantics: a listed core vertex stands for its whole
                               pruned tree, listed pruned vertices are skipped (c:1026, cu:274) */
/* 2-degree heuristic (-H 2, and -H 3 with peeling; declared but never defined
 * in the reference: setcuda_2degree h:133, set_mlp_cuda_2degree h:137,
 * scan_frt_csc_cuda_mono_2degree h:147, update_bc_cuda_2degree h:156, c:759-760,
 * c:1015).  OR-ed into src_flags: in BCMI_MODE_PLAIN a listed source of
 * (traversed) degree 2 whose two neighbours are listed too takes its depths
 * and path counts from the saved forward states of their sweeps instead of a
 * forward sweep of its own, and so does a listed source of degree 1 from its
 * neighbour's (without peeling; peeled cores have none); BC is unchanged.  Batched runs, directed graphs
 * and deep graphs (workgroup sweeps) ignore it. */
#define BCMI_SRC_DEG2 2

/* Host graph in CSR form: adjacency of v = off[v]..off[v+1] in adj, sorted,
 * no duplicates, no self loops.  For directed graphs adj lists in-neighbours
 * (the reference's CSC, c:472-501). */
typedef struct bcmi_graph {
    int64_t n;
    int64_t m;
    int64_t *off;
    int32_t *adj;
    int directed;
} bcmi_graph;

typedef struct bcmi_ctx bcmi_ctx;

/* Statistics of the last bcmi_bc_run (work counts use the traversed CSR). */
typedef struct bcmi_run_stats {
    int64_t sources_run;   /* BFS sweeps executed (distinct core roots when peeling) */
    int64_t batches;       /* batches (batched mode) or sources (plain mode) launched */
    double te;             /* traversed edges: sum over sweeps of 2 * m_s * multiplicity-free */
    double bytes;          /* algorithmic HBM bytes of the byte model for the run */
    double gpu_ms;         /* device time of the BC sweeps (HIP events) */
    double wall_ms;        /* host wall time of bcmi_bc_run */
    int64_t max_levels;    /* deepest BFS seen */
    int64_t sum_n;         /* sum of reached vertices */
    int64_t sum_m;         /* sum of degrees of reached vertices */
    int64_t sum_a;         /* sum of shortest-path DAG arcs */
    int64_t sum_pairs;     /* batched: (vertex, depth) pairs over batches, i.e. work-list vertices */
    int64_t sum_union;     /* batched: distinct vertices reached per batch, summed over batches */
    int64_t scaled_batches; /* batches (or workgroup sweeps) run with scaled FP64 path counts */
    int64_t wg_sources;    /* sources run one per workgroup (deep graphs, see "strategy") */
    int64_t deg2_sources;  /* plain sweeps whose forward state came from two neighbours' (BCMI_SRC_DEG2) */
} bcmi_run_stats;

/* ---- error reporting ---------------------------------------------------- */
const char *bcmi_last_error(void);
int bcmi_version(void);

/* ---- host graph construction (replaces read_graph c:175-254, mirror
 *      c:152-170, norm_graph c:430-453 and build_csc c:472-501) ---------- */
/* Text edge list "u v" per line; lines containing '#' are comments; ids >= n
 * fail with BCMI_ERR_RANGE and *bad_line set (c:232-237).  Unless directed,
 * the reversed copy of every non-loop edge is added (c:152-170).  The
 * "# Nodes: N Edges: E" header values are returned when present (c:220-226). */
int bcmi_graph_read_edgelist(const char *path, uint64_t n, int directed, bcmi_graph **out,
                             uint64_t *bad_line, uint64_t *hdr_nodes, uint64_t *hdr_edges);
/* Binary CSR cache (no counterpart in the reference, which re-parses the
 * text with read_graph/norm_graph/build_csc on every run, c:175-254,
 * c:430-453, c:472-501).  save writes header + int64 offsets + int32
 * adjacency with a checksum; src_size/src_mtime_ns record the text file the
 * CSR came from (0 if none) so a caller can tell a stale cache.  load checks
 * magic, size, checksum and CSR invariants (BCMI_ERR_IO otherwise). */
int bcmi_graph_save(const bcmi_graph *g, const char *path, uint64_t src_size, int64_t src_mtime_ns);
int bcmi_graph_load(const char *path, bcmi_graph **out, uint64_t *src_size, int64_t *src_mtime_ns);
/* Same pipeline from in-memory arcs tails[i] -> heads[i]. */
int bcmi_graph_from_edges(const int32_t *tails, const int32_t *heads, int64_t k, int64_t n,
                          int directed, bcmi_graph **out);
/* Seeded synthetic graphs (the reference's -S/-E generator is disabled,
 * c:664, c:802-810).  R-MAT with quadrant probabilities a,b,c (d = 1-a-b-c),
 * 2^scale vertices, edgefactor * 2^scale generated edges, optional random id
 * permutation; undirected. */
int bcmi_gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed,
                  int permute, bcmi_graph **out);
/* Adds pendant trees (depth 1-4) holding frac * n extra vertices. */
int bcmi_gen_pendant_trees(const bcmi_graph *core, double frac, uint64_t seed, bcmi_graph **out);
/* Chain of Erdos-Renyi blocks joined by single bridge edges. */
int bcmi_gen_block_chain(int64_t nblocks, int block, double avgdeg, uint64_t seed, bcmi_graph **out);
/* k distinct uniformly sampled vertices of degree > 0. */
int bcmi_sample_sources(const bcmi_graph *g, int64_t k, uint64_t seed, int32_t *out);
void bcmi_graph_free(bcmi_graph *g);

/* ---- device context (replaces initcuda cu:1123-1198, fincuda cu:1201-1232,
 *      setcuda cu:1100-1119; module globals cu:22-74 become context state) - */
int bcmi_create(int device, bcmi_ctx **out);
int bcmi_destroy(bcmi_ctx *ctx);
/* Tunables: "block" (threads per block), "grid" (blocks), "chunk" (edges per
 * work item), "alpha_plain"/"alpha_batched" (x1000: top-down -> bottom-up
 * switch when frontier arcs > alpha * unvisited arcs), "beta_plain"/
 * "beta_batched" (and only when frontier arcs >= all arcs / beta),
 * "fuse_bu" (bottom-up levels also sum path counts in their scan: 0 never,
 * 1 for vertices held in one work record, 2 for all - the default),
 * "grid_back" (blocks per batched backward kernel; 0, the default, uses
 * "grid"), "pack_bu" (1, the default: batched bottom-up levels scan vertices of
 * degree <= 16 four to a wave in a kernel of their own; 0: one per wave),
 * "narrow_sigma" (1, the default: exact batched sweeps first keep path
 * counts in 32-bit rows and redo a batch with 64-bit rows when a count
 * passes 2^32-1; 2: 16-bit rows first, then 32-bit, then 64-bit; 0: 64-bit
 * rows only),
 * "gamma_back" (batched backward levels whose successor level holds
 * fewer than 1/gamma_back of their arcs take their sums pushed from the
 * successors; 0, the default, always pulls),
 * "gamma_push" (batched top-down levels whose frontier holds fewer than
 * all arcs / gamma push path counts from the frontier; 0 always pulls),
 * "steps" (levels per
 * captured graph), "width" (sources per batch, 1..64), "ktime" (kernel timed
 * on the device, see bcmi_get_kernel_time), "sigma" (path counts: 0 exact
 * uint64, BCMI_ERR_OVERFLOW past 2^64-1; 1 FP64 scaled per level; 2 exact,
 * with batches that overflow recomputed scaled - the default; the plain
 * sweep is exact only), "batching" (source order: 0 as given, 1 hub-depth
 * signature, 2 highest-degree neighbour then signature, 3 that neighbour's
 * signature, the neighbour, the signature, 4 - the default - highest then
 * second highest-degree neighbour then signature, 5 highest-degree
 * neighbour, signature, second neighbour), "reorder"
 * (relabel by degree at upload), "kflags" (kernel variants), "sgrid"
 * (blocks per kernel on small levels), "sparse_verts" (levels with fewer
 * vertices use the small grid), "streams" (sweeps of one call run at once,
 * each on its own stream and workspace, on graphs whose hub BFS depth is
 * 64 or more), "strategy" (0 auto, 1 device-wide level loop, 2 one workgroup
 * per source: deep graphs, whose hub BFS depth reaches "wg_depth", run
 * hundreds of sources at once, one sweep per workgroup, within "wg_mem"
 * bytes of per-source state and at most "wg_slots" sources in flight),
 * "deg2_slots" (forward states of 2-degree sources' neighbours held at once
 * under BCMI_SRC_DEG2, 12 bytes per vertex each; 0, the default: as many as
 * "deg2_mem" bytes hold, 4 GiB by default; 1 turns the heuristic off), "trace",
 * "graphs" (1, the default: the level loop replays captured
 * graphs; 0: it launches the same kernels one by one, for profilers and
 * other tools that do not follow graph launches). */
int bcmi_set_param(bcmi_ctx *ctx, const char *key, int64_t value);
/* Reads a parameter above, or the read-only "label_passes": union and
 * labelling passes the uploaded graph's component labels took (1 when the
 * first pass already agrees with every arc; BCMI_ERR_STATE with no graph). */
int bcmi_get_param(bcmi_ctx *ctx, const char *key, int64_t *value);

/* Uploads a CSR graph (host pointers) and derives degrees and component
 * labels on the device (replaces initcuda's CSC upload and set_degree,
 * cu:1130-1136, cu:679-687).  heuristic = BCMI_H_PEEL runs the iterative
 * device peel (replaces degree_reduction c:314-419, init_bc_1degree
 * c:455-465, init_bc_1degree_device cu:1083-1087); undirected graphs only. */
int bcmi_upload_graph(bcmi_ctx *ctx, int64_t n, const int64_t *off, const int32_t *adj,
                      int directed, int heuristic);
/* Peeling results (valid after upload with BCMI_H_PEEL): removal round
 * (0 = core), parent (-1 = core), subtree weight w, and number of rounds. */
int bcmi_get_peel(bcmi_ctx *ctx, int32_t *round, int32_t *parent, uint64_t *w, int32_t *nrounds);
/* Component label (smallest vertex id of the component) per vertex. */
int bcmi_get_labels(bcmi_ctx *ctx, int32_t *label);

/* Zeroes the device BC accumulator. */
int bcmi_bc_reset(bcmi_ctx *ctx);
/* Accumulates BC over the listed sources (multiset) into the device
 * accumulator (replaces the source loop c:1010-1050 with setcuda,
 * set_mlp_cuda cu:689-703, scan_col_csc_cuda_mono cu:835-918,
 * scan_frt_csc_cuda_mono cu:766-827, pre_update_bc_cuda cu:330-336 and
 * update_bc_cuda cu:280-286).  src_flags: BCMI_SRC_EXACT / BCMI_SRC_EXPAND,
 * optionally OR-ed with BCMI_SRC_DEG2.  Sources of degree 0 contribute nothing (c:1026-1035).  st may be NULL. */
int bcmi_bc_run(bcmi_ctx *ctx, const int32_t *sources, int64_t nsrc, int mode, int src_flags,
                bcmi_run_stats *st);
/* Allocation-free variant of bcmi_bc_run in BCMI_MODE_PLAIN.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The sweep state lives in a static device array of
 * BCMI_STATIC_SCRATCH_BYTES per GPU instead of a workspace; the first call
 * on a GPU looks the array up, which loads it (held until the process ends),
 * and every call after that allocates nothing; graphs whose state does not fit
 * (about 1.5M vertices at 16 arcs each) fail with BCMI_ERR_ARG, peeled graphs
 * with BCMI_ERR_STATE.  It accumulates into the context's BC like
 * bcmi_bc_run (same sources, same results) and fills st's totals (per-sweep
 * statistics are not kept).  Every call on one GPU uses that GPU's array and
 * leaves its contents behind; calls that overlap on one GPU would race on it,
 * so the library runs them one at a time per GPU. */
#define BCMI_STATIC_SCRATCH_BYTES (64ull << 20)
int bcmi_bc_run_static(bcmi_ctx *ctx, const int32_t *sources, int64_t nsrc, bcmi_run_stats *st);
/* The order in which bcmi_bc_run (batched mode) takes these sources: out
 * receives the nsrc sources permuted by the "batching" similarity key,
 * computed on the host from the uploaded graph, with the sources of degree 0
 * (not swept) last; consecutive runs of "width" entries are the batches of
 * one batched run over the whole list.  Passing such runs to later calls with
 * "batching" = 0 reproduces those batches, so one workload can be split over
 * several calls or devices, whole batches each, without losing batch quality.
 * Unpeeled graphs only (BCMI_ERR_STATE).  No reference counterpart (its
 * source loop takes startv, startv+1, ... in turn, c:1017). */
int bcmi_batch_order(bcmi_ctx *ctx, const int32_t *sources, int64_t nsrc, int32_t *out);
/* Copies the BC accumulator to the host (replaces get_bc cu:1039-1041;
 * double instead of float, not halved). */
int bcmi_get_bc(bcmi_ctx *ctx, double *bc);
/* Per-sweep statistics of the last run: reached vertices, sum of their
 * degrees, DAG arcs, levels (arrays of length >= st->sources_run). */
int bcmi_get_sweep_stats(bcmi_ctx *ctx, int64_t *n_s, int64_t *m_s, int64_t *a_s, int32_t *levels);

/* Time per sweep of the last run, in ms, one entry per sweep in the order of
 * bcmi_get_sweep_stats (length st->sources_run): a plain sweep's own time, or
 * a batched sweep's (or a workgroup-per-source launch's) time divided by its
 * sources.  Replaces the per-source bc_time the reference accumulates into
 * its total, max and min (c:1037-1048, bc_func_mono c:504-580). */
int bcmi_get_sweep_times(bcmi_ctx *ctx, double *ms);

/* Duration (ms; per launch, first wave entry to last wave exit on the
 * device's constant-rate clock, summed on the device inside the captured
 * level loop)
 * and launch count of the kernel selected by parameter "ktime" during the
 * last run (1 batched sigma, 2 batched backward, 3/4 batched top-down/
 * bottom-up, 5/6/7 plain top-down/bottom-up/backward).  No reference
 * counterpart (the reference times kernels with cudaEvents and discards the
 * result, cu:887-899). */
int bcmi_get_kernel_time(bcmi_ctx *ctx, double *ms, int64_t *launches);

/* Debug: one plain sweep from s on the uploaded (core) graph, returning
 * depth (-1 unreached), sigma and delta per vertex (the reference's dead
 * get_lvl/get_sigma getters, cu:1004-1006, cu:1035-1037).  Does not touch
 * the BC accumulator. */
int bcmi_sssp(bcmi_ctx *ctx, int32_t s, int32_t *depth, uint64_t *sigma, double *delta);
/* Debug: one batched sweep over k <= 64 sources; outputs are k*n arrays laid
 * out [source][vertex]. */
int bcmi_msbfs(bcmi_ctx *ctx, const int32_t *sources, int k, int32_t *depth, uint64_t *sigma,
               double *delta);

/* Diagnostic: the kernel launches of one level step of the captured level
 * loop (a captured graph holds "steps" such steps), for the tunables given as
 * nparams key/value pairs (bcmi_set_param keys) applied to the defaults.  No
 * reference counterpart: the reference sizes and launches every kernel from
 * the host, level by level (cu:861-907).  mode: BCMI_MODE_PLAIN or
 * BCMI_MODE_BATCHED; graph: 0 forward, 1 backward, 2 forward on sparse
 * levels, 3 backward on sparse levels.  Writes at most cap entries, sets
 * *count to the step's node count and *steps to the steps per graph.  Needs
 * no device. */
typedef struct bcmi_launch {
    char kernel[24];   /* kernel name */
    int32_t grid;      /* workgroups */
    int32_t block;     /* work-items per workgroup */
    int64_t aux;       /* decide kernels: bottom-up allowed; k_kt_accum: waves of the timed launch */
    int32_t timed;     /* the launch stamps the device kernel timer */
    int32_t reserved;
} bcmi_launch;
int bcmi_launch_plan(const char *const *keys, const int64_t *values, int nparams, int mode, int graph,
                     bcmi_launch *out, int cap, int *count, int *steps);

/* ---- serving ------------------------------------------------------------ */
typedef struct bcmi_server bcmi_server;
typedef struct bcmi_query_result {
    int64_t query_id;
    int32_t graph_id;
    int32_t status;        /* BCMI_OK or error code */
    double arrive_ms;      /* submission time (server clock) */
    double start_ms;       /* admission time */
    double finish_ms;      /* completion time */
    double te;             /* traversed edges of the query */
} bcmi_query_result;

/* A server owns a device-resident pool of graphs and runs queries
 * (graph id, source multiset) with continuous batching: each admitted query
 * holds a slot (per-query sigma/delta/BC state sized to its graph) charged
 * against hbm_budget bytes; up to max_slots queries run concurrently on
 * their own streams.  No counterpart in the reference (single program run,
 * c:673-1139). */
int bcmi_server_create(int device, int64_t hbm_budget, int max_slots, bcmi_server **out);
int bcmi_server_destroy(bcmi_server *s);
int bcmi_server_add_graph(bcmi_server *s, int64_t n, const int64_t *off, const int32_t *adj,
                          int32_t *graph_id);
int bcmi_server_submit(bcmi_server *s, int32_t graph_id, const int32_t *sources, int64_t nsrc,
                       int64_t *query_id);
/* Waits for a query; bc (n of its graph) receives its BC (ordered pairs). */
int bcmi_server_wait(bcmi_server *s, int64_t query_id, double *bc, bcmi_query_result *res);
/* Server statistics: completed queries, peak concurrent slots, peak bytes. */
int bcmi_server_stats(bcmi_server *s, int64_t *completed, int64_t *peak_slots, int64_t *peak_bytes);

#ifdef __cplusplus
}
#endif
#endif /* BCMI_H */
