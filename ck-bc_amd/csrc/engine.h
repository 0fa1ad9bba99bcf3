// Host-side view of the device engine: device-resident graph, per-slot
// workspaces and the sweep drivers.  Shared by the context C ABI and the
// query server.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <vector>

#include "bcmi_internal.h"

namespace bcmi {

// Device control block of one workspace (one sweep in flight at a time).
struct Ctl {
  int32_t level;      // forward level being expanded
  int32_t started;
  int32_t fdone;      // forward finished
  int32_t last;       // deepest non-empty level
  int32_t blevel;     // backward level being processed
  int32_t bstarted;
  int32_t bdone;
  int32_t mode;       // 0 top-down, 1 bottom-up
  int32_t overflow;   // record capacity exceeded
  int32_t sigma_ovf;  // a path count wrapped past 2^64-1
  int32_t src_label;  // component of the plain-sweep source
  int32_t width;      // sources in the batch
  int32_t scaled;     // batched: path counts held as FP64 scaled per level
  int32_t push;       // batched: this level's path counts are pushed from the frontier (k_ms_push)
  int32_t bpush;      // batched: this backward level pulls nothing; the successor level pushes (k_ms_bpush)
  int32_t narrow;     // batched: path counts held in 32-bit rows
  int32_t cnt_back;   // plain: the backward pass counts the DAG arcs (forward state not swept)
  int32_t pad_;
  int64_t rec_count;  // records appended so far
  int64_t arcs_next;  // degree sum of vertices appended during this step
  int64_t verts_next;
  int64_t arcs_cur;   // degree sum of the current frontier
  int64_t verts_cur;
  int64_t arcs_seen;  // plain: degree sum of visited vertices; batched: of complete vertices
  int64_t arcs_reach; // degree sum of the source components
  int64_t sum_a;      // shortest-path DAG arcs (plain sweep)
  int64_t n_td, n_bu; // levels run top-down / bottom-up
  int64_t tot_verts;  // plain: reached vertices; batched: (vertex, level) pairs
  int64_t tot_arcs;   // degree sum over the same
  int64_t uni_verts;  // batched: distinct vertices reached by any source
  int64_t uni_arcs;   // batched: their degree sum
};

struct DevGraph {
  int64_t n = 0;
  int64_t m = 0;             // arcs of the traversed (core) graph
  int64_t fm = 0;            // arcs of the uploaded graph
  int directed = 0;
  int heuristic = 0;
  int chunk = 256;
  int64_t* off = nullptr;    // traversed CSR
  int32_t* adj = nullptr;
  int64_t* toff = nullptr;   // transposed traversed CSR (aliases off/adj when undirected)
  int32_t* tadj = nullptr;
  int64_t* foff = nullptr;   // uploaded CSR (aliases off/adj when nothing was peeled)
  int32_t* fadj = nullptr;
  int32_t* label = nullptr;     // component label (smallest id) per vertex
  int64_t* comp_size = nullptr; // vertices per label
  int64_t* comp_arcs = nullptr; // traversed arcs per label
  int32_t* all_v = nullptr;     // chunk records of every traversed vertex
  int32_t* all_j = nullptr;
  int64_t nall = 0;
  int32_t* split_v = nullptr;   // vertices held in more than one chunk record
  int64_t nsplit = 0;
  int32_t* sm_v = nullptr;      // vertices of degree 1..16 (packed bottom-up)
  int64_t nsm = 0;
  int32_t* lg_v = nullptr;      // chunk records of the other vertices
  int32_t* lg_j = nullptr;
  int64_t nlg = 0;
  // peeling
  int32_t nrounds = 0;
  int32_t* round = nullptr;
  int32_t* parent = nullptr;
  uint64_t* w = nullptr;
  int32_t* root = nullptr;
  int32_t* order = nullptr;      // removed vertices by round
  std::vector<int64_t> round_lo; // offsets of each round in order (size nrounds+1)
  std::vector<int64_t> h_off;    // traversed offsets on the host (degrees)
  std::vector<int32_t> h_round;  // removal round per vertex on the host
  std::vector<uint64_t> h_sig;   // batching key: BFS depths from the eight largest hubs
  std::vector<int32_t> h_topnb;  // batching key: highest-degree neighbour
  std::vector<int32_t> h_topnb2; // batching key: second highest-degree neighbour
  int depth_hint = 0;            // deepest level of the hubs' BFS
  int cc_passes = 0;             // union-find passes the component labels took
  // Vertices are relabelled by descending degree on upload (hubs' rows and
  // masks then share cache lines); the caller's ids are mapped at the ABI.
  bool relabeled = false;
  std::vector<int32_t> h_n2o, h_o2n;  // internal id -> caller id, and back
  int32_t* d_n2o = nullptr;
  int32_t to_int(int32_t v) const { return relabeled ? h_o2n[v] : v; }
  // out[v] = in[internal(v)] for every caller id v (in-place).
  template <class T>
  void to_caller(T* a) const {
    if (!relabeled) return;
    std::vector<T> t(a, a + n);
    for (int64_t v = 0; v < n; v++) a[v] = t[h_o2n[v]];
  }
  // set once a batched sweep's exact path counts wrapped (auto mode then
  // starts later batches in scaled FP64)
  mutable std::atomic<bool> sigma_overflowed{false};
  // narrowest path-count row width later batches start with (2 16-bit, 1
  // 32-bit, 0 64-bit), lowered when a batch's counts pass a row's range
  mutable std::atomic<int> narrow_cap{2};
  size_t bytes = 0;
  void release();
};

struct Workspace;

// Kernels of the captured level loop (one level step per graph node group).
enum StepKernel {
  SK_PLAIN_DECIDE, SK_PLAIN_TD, SK_PLAIN_BU, SK_BDECIDE, SK_PLAIN_BACK, SK_PLAIN_BFIN,
  SK_MS_DECIDE, SK_MS_TD, SK_MS_BU, SK_MS_BU_PK, SK_MS_SIGMA, SK_MS_PUSH,
  SK_MS_BDECIDE, SK_MS_BACK, SK_MS_BPUSH, SK_MS_BFIN, SK_KT_ACCUM, SK_KT_FOLD, SK_COUNT
};
struct StepNode {
  int kernel;    // StepKernel
  int grid;      // workgroups
  int block;     // work-items per workgroup
  int64_t aux;   // decide kernels: bottom-up allowed; k_kt_accum: waves of the timed launch
  bool timed;    // the kernel stamps the device timer (Params::ktime)
};
// The kernel nodes of one captured level step; graph 0 forward, 1 backward,
// 2/3 the same on sparse levels.  Host-only: needs no device.
std::vector<StepNode> step_plan(const Params& prm, bool plain, int graph);
const char* step_kernel_name(int kernel);
// Per-wave stamp slots of the device timer: a timed launch has at most this
// many waves (grid <= 65536 blocks of <= 256 work-items).
constexpr int64_t kTimerMaxWaves = 65536 * 4;

// Uploads and prepares a graph (peeling, components, work lists).
int devgraph_build(DevGraph& g, int64_t n, const int64_t* off, const int32_t* adj, int directed,
                   int heuristic, const Params& prm, hipStream_t st);

// Bytes a workspace for graph g needs in the given mode.
size_t workspace_bytes(const DevGraph& g, int mode, int width, const Params& prm);

struct RunResult {
  int64_t sweeps = 0, launches = 0, max_levels = 0;
  int64_t sum_n = 0, sum_m = 0, sum_a = 0;
  int64_t n_td = 0, n_bu = 0;
  int64_t pairs = 0;                  // batched: (vertex, depth) pairs
  int64_t scaled_batches = 0;         // batches (or workgroup sweeps) run with scaled FP64 path counts
  int64_t wg_sources = 0;             // sources run one per workgroup (deep graphs)
  int64_t deg2_sources = 0;           // plain: sweeps whose forward state came from two neighbours'
  std::vector<int64_t> uni_n, uni_m;  // batched: union reach per batch
  double te = 0, bytes = 0, gpu_ms = 0, wall_ms = 0;
  double ktime_ms = 0;      // summed duration of the timed kernel (Params::ktime)
  int64_t klaunches = 0;    // its launches (including ones that found no work)
  std::vector<int64_t> n_s, m_s, a_s;
  std::vector<int32_t> levels;
  // per sweep (one entry per source, in the order of n_s): ms of a plain
  // sweep, or of a batched / workgroup launch divided by its sources
  std::vector<double> t_s;
};

Workspace* workspace_create(const DevGraph& g, int mode, int width, const Params& prm, int* rc);
void workspace_destroy(Workspace* ws);
hipStream_t workspace_stream(Workspace* ws);
double* workspace_bc(Workspace* ws);
// Makes the workspace accumulate into an external BC array (not owned).
void workspace_share_bc(Workspace* ws, double* bc);
size_t workspace_size(Workspace* ws);
int workspace_bc_reset(Workspace* ws);

// Accumulates BC over the sources (multiset) into the workspace BC array.
int bc_run(const DevGraph& g, Workspace* ws, const Params& prm, const int32_t* sources, int64_t nsrc,
           int mode, int src_flags, RunResult* out);

// Allocation-free plain sweeps: per-sweep state in a static device array of
// kStaticScratchBytes per GPU (static_scratch.hip).  static_scratch_init
// (once per device, with that device current) looks up the array's address,
// static_scratch_ptr returns it; bc_run_static runs plain sweeps on stream st
// into bc without allocating (its first call on a device does the set-up).
constexpr size_t kStaticScratchBytes = BCMI_STATIC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
struct StaticRun {
  int64_t sweeps = 0, sum_n = 0, sum_m = 0, sum_a = 0, max_levels = 0;
  double te = 0, bytes = 0, wall_ms = 0;
};
int static_scratch_init(int device);
void* static_scratch_ptr(int device);
size_t static_scratch_need(const DevGraph& g);
int bc_run_static(const DevGraph& g, double* bc, hipStream_t st, int device, const Params& prm,
                  const int32_t* sources, int64_t nsrc, StaticRun* out);

// The sources in the order a batched run groups them (consecutive runs of
// Params::width are its batches), degree-0 sources last; unpeeled graphs.
int batch_order(const DevGraph& g, const Params& prm, const int32_t* src, int64_t nsrc, int32_t* out);

// Debug sweeps (no BC accumulation).
int sssp_debug(const DevGraph& g, Workspace* ws, const Params& prm, int32_t s, int32_t* depth,
               uint64_t* sigma, double* delta);
int msbfs_debug(const DevGraph& g, Workspace* ws, const Params& prm, const int32_t* src, int k,
                int32_t* depth, uint64_t* sigma, double* delta);

}  // namespace bcmi
