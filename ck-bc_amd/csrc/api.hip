// C ABI of the engine context (one device graph, plain and batched
// workspaces sharing one BC accumulator).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>
#include <mutex>
#include <string>

#include "bcmi.h"
#include "bcmi_internal.h"
#include "engine.h"

namespace bcmi {
const char* last_error();
}
using namespace bcmi;

struct bcmi_ctx {
  int device = 0;
  Params prm;
  DevGraph g;
  bool have_graph = false;
  Workspace* wsp = nullptr;  // plain
  Workspace* wsb = nullptr;  // batched
  double* bc = nullptr;
  hipStream_t st = nullptr;
  RunResult last;
  std::mutex mu;
};

#define CHECK_CTX(c) \
  if (!(c)) return set_error(BCMI_ERR_ARG, "null context")

static int use_device(bcmi_ctx* c) {
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return set_error(BCMI_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
  return BCMI_OK;
}

static void drop_workspaces(bcmi_ctx* c) {
  workspace_destroy(c->wsp);
  workspace_destroy(c->wsb);
  c->wsp = c->wsb = nullptr;
}

static int get_ws(bcmi_ctx* c, int mode, Workspace** out) {
  Workspace*& ws = mode == BCMI_MODE_PLAIN ? c->wsp : c->wsb;
  if (!ws) {
    int rc = BCMI_OK;
    ws = workspace_create(c->g, mode, c->prm.width, c->prm, &rc);
    if (!ws) return rc;
    workspace_share_bc(ws, c->bc);
  }
  *out = ws;
  return BCMI_OK;
}

// Validates and sets one tunable (bcmi_set_param's keys).  have_graph: a
// graph is uploaded (parameters fixed at upload may no longer change).
static int param_set(Params& p, const std::string& k, int64_t v, bool have_graph) {
  if (k == "block") {
    if (v < 64 || v > 256 || v % 64) return set_error(BCMI_ERR_ARG, "block must be 64, 128 or 256");
    p.block = (int)v;
  } else if (k == "grid") {
    if (v < 1 || v > 65536) return set_error(BCMI_ERR_ARG, "grid must be in [1,65536]");
    p.grid = (int)v;
  } else if (k == "grid_back") {
    if (v < 0 || v > 65536) return set_error(BCMI_ERR_ARG, "grid_back must be in [0,65536]");
    p.grid_back = (int)v;
  } else if (k == "grid_plain") {
    if (v < 1 || v > 65536) return set_error(BCMI_ERR_ARG, "grid_plain must be in [1,65536]");
    p.grid_plain = (int)v;
  } else if (k == "chunk") {
    if (v < 64 || v > (1 << 20) || v % 64) return set_error(BCMI_ERR_ARG, "chunk must be a multiple of 64");
    if (have_graph) return set_error(BCMI_ERR_STATE, "chunk must be set before uploading a graph");
    p.chunk = (int)v;
  } else if (k == "alpha_plain") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "alpha must be >= 0");
    p.alpha_plain = v;
  } else if (k == "alpha_batched") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "alpha must be >= 0");
    p.alpha_batched = v;
  } else if (k == "fuse_bu") {
    if (v < 0 || v > 2) return set_error(BCMI_ERR_ARG, "fuse_bu must be 0, 1 or 2");
    p.fuse_bu = (int)v;
  } else if (k == "pack_bu") {
    p.pack_bu = v ? 1 : 0;
  } else if (k == "narrow_sigma") {
    if (v < 0 || v > 2) return set_error(BCMI_ERR_ARG, "narrow_sigma must be 0 (64-bit rows), 1 (32-bit first) or 2 (16-bit first)");
    p.narrow_sigma = (int)v;
  } else if (k == "gamma_back") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "gamma_back must be >= 0");
    p.gamma_back = v;
  } else if (k == "gamma_push") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "gamma_push must be >= 0");
    p.gamma_push = v;
  } else if (k == "beta_plain" || k == "beta_batched") {
    if (v < 1) return set_error(BCMI_ERR_ARG, "beta must be >= 1");
    (k == "beta_plain" ? p.beta_plain : p.beta_batched) = v;
  } else if (k == "steps") {
    if (v < 1 || v > 4096) return set_error(BCMI_ERR_ARG, "steps must be in [1,4096]");
    p.steps = (int)v;
  } else if (k == "width") {
    if (v < 1 || v > 64) return set_error(BCMI_ERR_ARG, "width must be in [1,64]");
    p.width = (int)v;
  } else if (k == "sgrid") {
    if (v < 1 || v > 65536) return set_error(BCMI_ERR_ARG, "sgrid must be in [1,65536]");
    p.sgrid = (int)v;
  } else if (k == "sparse_verts") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "sparse_verts must be >= 0");
    p.sparse_verts = v;
  } else if (k == "strategy") {
    if (v < 0 || v > 2) return set_error(BCMI_ERR_ARG, "strategy must be 0 (auto), 1 (level loop) or 2 (workgroup per source)");
    p.strategy = (int)v;
  } else if (k == "wg_depth") {
    if (v < 1) return set_error(BCMI_ERR_ARG, "wg_depth must be >= 1");
    p.wg_depth = (int)std::min<int64_t>(v, INT32_MAX);
  } else if (k == "wg_mem") {
    if (v < (1 << 20)) return set_error(BCMI_ERR_ARG, "wg_mem must be >= 1 MiB");
    p.wg_mem = v;
  } else if (k == "wg_slots") {
    if (v < 1 || v > 65536) return set_error(BCMI_ERR_ARG, "wg_slots must be in [1,65536]");
    p.wg_slots = (int)v;
  } else if (k == "concurrency") {
    if (v < 0 || v > 16) return set_error(BCMI_ERR_ARG, "concurrency must be in [0,16]");
    p.concurrency = (int)v;
  } else if (k == "deg2_slots") {
    if (v < 0 || v > (1 << 20)) return set_error(BCMI_ERR_ARG, "deg2_slots must be in [0,2^20]");
    p.deg2_slots = (int)v;
  } else if (k == "deg2_mem") {
    if (v < 0) return set_error(BCMI_ERR_ARG, "deg2_mem must be >= 0");
    p.deg2_mem = v;
  } else if (k == "streams") {
    if (v < 1 || v > 16) return set_error(BCMI_ERR_ARG, "streams must be in [1,16]");
    p.streams = (int)v;
  } else if (k == "sigma") {
    if (v < 0 || v > 2) return set_error(BCMI_ERR_ARG, "sigma must be 0 (exact), 1 (scaled) or 2 (auto)");
    p.sigma = (int)v;
  } else if (k == "kflags") {
    if (v < 0 || v > 3) return set_error(BCMI_ERR_ARG, "kflags must be 0..3");
    p.kflags = (int)v;
  } else if (k == "batching") {
    if (v < 0 || v > 5) return set_error(BCMI_ERR_ARG, "batching must be 0..5");
    p.batching = (int)v;
  } else if (k == "reorder") {
    if (have_graph) return set_error(BCMI_ERR_STATE, "reorder must be set before uploading a graph");
    p.reorder = v ? 1 : 0;
  } else if (k == "trace") {
    p.trace = v ? 1 : 0;
  } else if (k == "graphs") {
    p.graphs = v ? 1 : 0;
  } else if (k == "ktime") {
    if (v < 0 || v > 7) return set_error(BCMI_ERR_ARG, "ktime must be in [0,7]");
    p.ktime = (int)v;
  } else {
    return set_error(BCMI_ERR_ARG, "unknown parameter " + k);
  }
  return BCMI_OK;
}

extern "C" {

const char* bcmi_last_error(void) { return last_error(); }
int bcmi_version(void) { return 1; }

int bcmi_create(int device, bcmi_ctx** out) {
  return guard([&]() -> int {
    if (!out) return set_error(BCMI_ERR_ARG, "null output");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return set_error(BCMI_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(BCMI_ERR_ARG, "device index out of range");
    bcmi_ctx* c = new bcmi_ctx();
    c->device = device;
    int rc = use_device(c);
    if (rc) { delete c; return rc; }
    e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return set_error(BCMI_ERR_HIP, "stream creation failed"); }
    *out = c;
    return BCMI_OK;
  });
}

int bcmi_destroy(bcmi_ctx* c) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    use_device(c);
    drop_workspaces(c);
    c->g.release();
    if (c->bc) (void)hipFree(c->bc);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return BCMI_OK;
  });
}

int bcmi_set_param(bcmi_ctx* c, const char* key, int64_t v) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!key) return set_error(BCMI_ERR_ARG, "null key");
    std::lock_guard<std::mutex> lk(c->mu);
    return param_set(c->prm, key, v, c->have_graph);
  });
}

int bcmi_launch_plan(const char* const* keys, const int64_t* values, int nparams, int mode, int graph,
                     bcmi_launch* out, int cap, int* count, int* steps) {
  return guard([&]() -> int {
    if (nparams < 0 || (nparams > 0 && (!keys || !values)) || !count || cap < 0 || (cap > 0 && !out))
      return set_error(BCMI_ERR_ARG, "bad argument");
    if (mode != BCMI_MODE_PLAIN && mode != BCMI_MODE_BATCHED) return set_error(BCMI_ERR_ARG, "bad mode");
    if (graph < 0 || graph > 3) return set_error(BCMI_ERR_ARG, "graph must be 0..3");
    Params p;
    for (int i = 0; i < nparams; i++) {
      if (!keys[i]) return set_error(BCMI_ERR_ARG, "null key");
      int rc = param_set(p, keys[i], values[i], false);
      if (rc) return rc;
    }
    const std::vector<StepNode> plan = step_plan(p, mode == BCMI_MODE_PLAIN, graph);
    *count = (int)plan.size();
    if (steps) *steps = p.steps;
    for (int i = 0; i < cap && i < (int)plan.size(); i++) {
      bcmi_launch& L = out[i];
      memset(&L, 0, sizeof L);
      strncpy(L.kernel, step_kernel_name(plan[i].kernel), sizeof L.kernel - 1);
      L.grid = plan[i].grid;
      L.block = plan[i].block;
      L.aux = plan[i].aux;
      L.timed = plan[i].timed ? 1 : 0;
    }
    return BCMI_OK;
  });
}

int bcmi_get_param(bcmi_ctx* c, const char* key, int64_t* v) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!key || !v) return set_error(BCMI_ERR_ARG, "null argument");
    std::string k(key);
    const Params& p = c->prm;
    if (k == "label_passes") {  // read-only: union-find passes the uploaded graph's labels took
      if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
      *v = c->g.cc_passes;
      return BCMI_OK;
    }
    if (k == "block") *v = p.block;
    else if (k == "grid") *v = p.grid;
    else if (k == "grid_plain") *v = p.grid_plain;
    else if (k == "grid_back") *v = p.grid_back;
    else if (k == "chunk") *v = p.chunk;
    else if (k == "alpha_plain") *v = p.alpha_plain;
    else if (k == "alpha_batched") *v = p.alpha_batched;
    else if (k == "beta_plain") *v = p.beta_plain;
    else if (k == "beta_batched") *v = p.beta_batched;
    else if (k == "gamma_push") *v = p.gamma_push;
    else if (k == "gamma_back") *v = p.gamma_back;
    else if (k == "narrow_sigma") *v = p.narrow_sigma;
    else if (k == "pack_bu") *v = p.pack_bu;
    else if (k == "fuse_bu") *v = p.fuse_bu;
    else if (k == "steps") *v = p.steps;
    else if (k == "width") *v = p.width;
    else if (k == "ktime") *v = p.ktime;
    else if (k == "trace") *v = p.trace;
    else if (k == "graphs") *v = p.graphs;
    else if (k == "reorder") *v = p.reorder;
    else if (k == "batching") *v = p.batching;
    else if (k == "kflags") *v = p.kflags;
    else if (k == "sigma") *v = p.sigma;
    else if (k == "sgrid") *v = p.sgrid;
    else if (k == "sparse_verts") *v = p.sparse_verts;
    else if (k == "streams") *v = p.streams;
    else if (k == "strategy") *v = p.strategy;
    else if (k == "wg_depth") *v = p.wg_depth;
    else if (k == "wg_mem") *v = p.wg_mem;
    else if (k == "wg_slots") *v = p.wg_slots;
    else if (k == "deg2_slots") *v = p.deg2_slots;
    else if (k == "concurrency") *v = p.concurrency;
    else if (k == "deg2_mem") *v = p.deg2_mem;
    else return set_error(BCMI_ERR_ARG, "unknown parameter " + k);
    return BCMI_OK;
  });
}

int bcmi_upload_graph(bcmi_ctx* c, int64_t n, const int64_t* off, const int32_t* adj, int directed, int heuristic) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!off || (!adj && n > 0 && off[n] > 0)) return set_error(BCMI_ERR_ARG, "null CSR arrays");
    if (heuristic != BCMI_H_NONE && heuristic != BCMI_H_PEEL) return set_error(BCMI_ERR_ARG, "heuristic must be 0 or 1");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = use_device(c);
    if (rc) return rc;
    drop_workspaces(c);
    if (c->bc) { (void)hipFree(c->bc); c->bc = nullptr; }
    c->have_graph = false;
    rc = devgraph_build(c->g, n, off, adj, directed, heuristic, c->prm, c->st);
    if (rc) { c->g.release(); return rc; }
    hipError_t e = hipMalloc((void**)&c->bc, sizeof(double) * n);
    if (e != hipSuccess) { c->g.release(); return set_error(BCMI_ERR_NOMEM, "BC accumulator allocation failed"); }
    e = hipMemsetAsync(c->bc, 0, sizeof(double) * n, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
    c->have_graph = true;
    return BCMI_OK;
  });
}

int bcmi_get_peel(bcmi_ctx* c, int32_t* round, int32_t* parent, uint64_t* w, int32_t* nrounds) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    use_device(c);
    int64_t n = c->g.n;
    hipError_t e = hipSuccess;
    if (round && e == hipSuccess) e = hipMemcpy(round, c->g.round, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    if (parent && e == hipSuccess) e = hipMemcpy(parent, c->g.parent, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    if (w && e == hipSuccess) e = hipMemcpy(w, c->g.w, sizeof(uint64_t) * n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
    if (round) c->g.to_caller(round);
    if (w) c->g.to_caller(w);
    if (parent) {
      c->g.to_caller(parent);
      if (c->g.relabeled)
        for (int64_t v = 0; v < n; v++)
          if (parent[v] >= 0) parent[v] = c->g.h_n2o[parent[v]];
    }
    if (nrounds) *nrounds = c->g.nrounds;
    return BCMI_OK;
  });
}

int bcmi_get_labels(bcmi_ctx* c, int32_t* label) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (!label) return set_error(BCMI_ERR_ARG, "null output");
    use_device(c);
    hipError_t e = hipMemcpy(label, c->g.label, sizeof(int32_t) * c->g.n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
    if (c->g.relabeled) {
      // internal labels are the smallest internal id; report the smallest caller id
      int64_t n = c->g.n;
      std::vector<int32_t> mn(n, INT32_MAX);
      for (int64_t i = 0; i < n; i++) mn[label[i]] = std::min(mn[label[i]], c->g.h_n2o[i]);
      std::vector<int32_t> out(n);
      for (int64_t v = 0; v < n; v++) out[v] = mn[label[c->g.h_o2n[v]]];
      memcpy(label, out.data(), sizeof(int32_t) * n);
    }
    return BCMI_OK;
  });
}

int bcmi_bc_reset(bcmi_ctx* c) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    use_device(c);
    hipError_t e = hipMemsetAsync(c->bc, 0, sizeof(double) * c->g.n, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
    return BCMI_OK;
  });
}

int bcmi_bc_run(bcmi_ctx* c, const int32_t* sources, int64_t nsrc, int mode, int src_flags, bcmi_run_stats* st) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (nsrc < 0 || (nsrc > 0 && !sources)) return set_error(BCMI_ERR_ARG, "bad source list");
    if (mode != BCMI_MODE_PLAIN && mode != BCMI_MODE_BATCHED) return set_error(BCMI_ERR_ARG, "bad mode");
    if (src_flags & ~(BCMI_SRC_EXPAND | BCMI_SRC_DEG2)) return set_error(BCMI_ERR_ARG, "bad source flags");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = use_device(c);
    if (rc) return rc;
    Workspace* ws = nullptr;
    rc = get_ws(c, mode, &ws);
    if (rc) return rc;
    RunResult R;
    rc = bc_run(c->g, ws, c->prm, sources, nsrc, mode, src_flags, &R);
    // a run stopped part way (device or host failure) may leave a sweep's
    // state uncleared: the next call starts from fresh workspaces
    if (rc == BCMI_ERR_HIP || rc == BCMI_ERR_NOMEM) drop_workspaces(c);
    c->last = R;
    if (st) {
      memset(st, 0, sizeof(*st));
      st->sources_run = R.sweeps;
      st->batches = R.launches;
      st->te = R.te;
      st->bytes = R.bytes;
      st->gpu_ms = R.gpu_ms;
      st->wall_ms = R.wall_ms;
      st->max_levels = R.max_levels;
      st->sum_n = R.sum_n;
      st->sum_m = R.sum_m;
      st->sum_a = R.sum_a;
      st->sum_pairs = R.pairs;
      st->scaled_batches = R.scaled_batches;
      st->wg_sources = R.wg_sources;
      st->deg2_sources = R.deg2_sources;
      for (int64_t u : R.uni_n) st->sum_union += u;
    }
    return rc;
  });
}

int bcmi_bc_run_static(bcmi_ctx* c, const int32_t* sources, int64_t nsrc, bcmi_run_stats* st) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (nsrc < 0 || (nsrc > 0 && !sources)) return set_error(BCMI_ERR_ARG, "bad source list");
    std::lock_guard<std::mutex> lk(c->mu);
    int rc = use_device(c);
    if (rc) return rc;
    StaticRun S;
    rc = bc_run_static(c->g, c->bc, c->st, c->device, c->prm, sources, nsrc, &S);
    // the per-sweep records describe bcmi_bc_run's last run: drop them (clear keeps
    // the vectors' storage, so nothing is freed or allocated)
    c->last.n_s.clear(); c->last.m_s.clear(); c->last.a_s.clear(); c->last.levels.clear(); c->last.t_s.clear();
    c->last.ktime_ms = 0; c->last.klaunches = 0;
    if (st) {
      memset(st, 0, sizeof(*st));
      st->sources_run = S.sweeps;
      st->batches = S.sweeps;
      st->te = S.te;
      st->bytes = S.bytes;
      st->gpu_ms = S.wall_ms;
      st->wall_ms = S.wall_ms;
      st->max_levels = S.max_levels;
      st->sum_n = S.sum_n;
      st->sum_m = S.sum_m;
      st->sum_a = S.sum_a;
    }
    return rc;
  });
}

int bcmi_batch_order(bcmi_ctx* c, const int32_t* sources, int64_t nsrc, int32_t* out) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (nsrc < 0 || (nsrc > 0 && (!sources || !out))) return set_error(BCMI_ERR_ARG, "bad source list");
    std::lock_guard<std::mutex> lk(c->mu);
    return batch_order(c->g, c->prm, sources, nsrc, out);
  });
}

int bcmi_get_bc(bcmi_ctx* c, double* bc) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (!bc) return set_error(BCMI_ERR_ARG, "null output");
    use_device(c);
    hipError_t e = hipMemcpy(bc, c->bc, sizeof(double) * c->g.n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
    c->g.to_caller(bc);
    return BCMI_OK;
  });
}

int bcmi_get_sweep_stats(bcmi_ctx* c, int64_t* n_s, int64_t* m_s, int64_t* a_s, int32_t* levels) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    const RunResult& R = c->last;
    size_t k = R.n_s.size();
    if (n_s) memcpy(n_s, R.n_s.data(), k * sizeof(int64_t));
    if (m_s) memcpy(m_s, R.m_s.data(), k * sizeof(int64_t));
    if (a_s) memcpy(a_s, R.a_s.data(), k * sizeof(int64_t));
    if (levels) memcpy(levels, R.levels.data(), R.levels.size() * sizeof(int32_t));
    return BCMI_OK;
  });
}

int bcmi_get_sweep_times(bcmi_ctx* c, double* ms) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!ms) return set_error(BCMI_ERR_ARG, "null output");
    const RunResult& R = c->last;
    if (!R.t_s.empty()) memcpy(ms, R.t_s.data(), R.t_s.size() * sizeof(double));
    return BCMI_OK;
  });
}

int bcmi_get_kernel_time(bcmi_ctx* c, double* ms, int64_t* launches) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (ms) *ms = c->last.ktime_ms;
    if (launches) *launches = c->last.klaunches;
    return BCMI_OK;
  });
}

int bcmi_sssp(bcmi_ctx* c, int32_t s, int32_t* depth, uint64_t* sigma, double* delta) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (!depth || !sigma || !delta) return set_error(BCMI_ERR_ARG, "null output");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    Workspace* ws = nullptr;
    int rc = get_ws(c, BCMI_MODE_PLAIN, &ws);
    if (rc) return rc;
    return sssp_debug(c->g, ws, c->prm, s, depth, sigma, delta);
  });
}

int bcmi_msbfs(bcmi_ctx* c, const int32_t* sources, int k, int32_t* depth, uint64_t* sigma, double* delta) {
  return guard([&]() -> int {
    CHECK_CTX(c);
    if (!c->have_graph) return set_error(BCMI_ERR_STATE, "no graph uploaded");
    if (!sources || !depth || !sigma || !delta) return set_error(BCMI_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    use_device(c);
    Workspace* ws = nullptr;
    int rc = get_ws(c, BCMI_MODE_BATCHED, &ws);
    if (rc) return rc;
    return msbfs_debug(c->g, ws, c->prm, sources, k, depth, sigma, delta);
  });
}

}  // extern "C"
