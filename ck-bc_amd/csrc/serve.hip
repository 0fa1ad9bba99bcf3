// BC query server: a device-resident pool of graphs and a continuous-batching
// scheduler.  A query is (graph id, source multiset); it needs one slot = a
// batched workspace (sigma/coef rows, masks, work lists, its own BC
// accumulator) sized to its graph.  Queries are admitted in arrival order as
// slots and HBM budget free up; each admitted query runs on its own worker
// thread and HIP stream, so the sweeps of different queries overlap on the
// device.  A query's result depends only on its own sources: it is computed
// in its own workspace, batch after batch (the level-loop engines sum in a
// fixed order; the workgroup engine of deep graphs adds with atomics, whose
// order changes only the last bits).
//
// No counterpart in the reference (one program run per graph, bc2d.c:673).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bcmi.h"
#include "bcmi_internal.h"
#include "engine.h"

using namespace bcmi;

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Query {
  int64_t id = 0;
  int32_t gid = 0;
  std::vector<int32_t> src;
  double arrive = 0, start = 0, finish = 0;
  double te = 0;
  int status = BCMI_OK;
  std::string err;
  bool done = false;
  std::vector<double> bc;
};

struct Slot {
  Workspace* ws = nullptr;
  int32_t gid = -1;
  size_t bytes = 0;
  bool busy = false;
  double last_used = 0;
};

}  // namespace

struct bcmi_server {
  int device = 0;
  int64_t budget = 0;
  int max_slots = 1;
  double t0 = 0;
  Params prm;
  std::vector<std::unique_ptr<DevGraph>> graphs;
  std::vector<size_t> graph_ws_bytes;
  int64_t graph_bytes = 0;
  std::mutex mu;
  std::condition_variable cv_sched;  // new query or freed slot
  std::condition_variable cv_done;   // a query finished
  std::deque<Query*> pending;
  std::map<int64_t, std::unique_ptr<Query>> queries;
  std::vector<std::unique_ptr<Slot>> slots;  // cached and busy workspaces
  int64_t slot_bytes = 0;
  int64_t next_id = 1;
  int64_t completed = 0, peak_active = 0, peak_bytes = 0;
  bool stop = false;
  std::thread sched;
  std::vector<std::thread> workers;
  std::vector<std::pair<Query*, Slot*>> assigned;  // consumed by the workers
  std::condition_variable cv_work;
};

static int64_t active_slots(bcmi_server* s) {
  int64_t a = 0;
  for (auto& sl : s->slots) a += sl->busy;
  return a;
}

// Frees the least recently used idle workspace; false when none is idle.
static bool evict_one(bcmi_server* s) {
  int victim = -1;
  for (size_t i = 0; i < s->slots.size(); i++)
    if (!s->slots[i]->busy && (victim < 0 || s->slots[i]->last_used < s->slots[victim]->last_used)) victim = (int)i;
  if (victim < 0) return false;
  workspace_destroy(s->slots[victim]->ws);
  s->slot_bytes -= (int64_t)s->slots[victim]->bytes;
  s->slots.erase(s->slots.begin() + victim);
  return true;
}

enum Take { TAKEN, WAIT, FAIL };

// Picks (or makes room for) a slot for graph gid.  WAIT: a running query must
// free memory first.  FAIL: the query cannot run (*rc and *why say why): it
// can never fit the budget, or its workspace cannot be allocated although no
// other query holds device memory.  A new slot is charged the larger of the
// graph's workspace estimate (which covers what the sweeps allocate later:
// helper workspaces, workgroup-sweep state) and what was allocated.
static Take take_slot(bcmi_server* s, int32_t gid, Slot** out, int* rc, std::string* why) {
  *out = nullptr;
  for (auto& sl : s->slots)
    if (!sl->busy && sl->gid == gid) { *out = sl.get(); return TAKEN; }
  const int64_t need = (int64_t)s->graph_ws_bytes[gid];
  if (need + s->graph_bytes > s->budget) {
    *rc = BCMI_ERR_NOMEM;
    *why = "query needs more device memory than the HBM budget";
    return FAIL;
  }
  // evict idle workspaces of other graphs, least recently used first
  while (s->graph_bytes + s->slot_bytes + need > s->budget)
    if (!evict_one(s)) return WAIT;  // a busy slot must finish first
  Workspace* ws = nullptr;
  for (;;) {
    int wrc = BCMI_OK;
    ws = workspace_create(*s->graphs[gid], BCMI_MODE_BATCHED, 64, s->prm, &wrc);
    if (ws) break;
    // The device could not hold it: idle workspaces give their memory back
    // first; then only a running query can free more.
    if (evict_one(s)) continue;
    if (active_slots(s) > 0) return WAIT;
    *rc = wrc;
    *why = std::string("workspace allocation failed: ") + bcmi_last_error();
    return FAIL;
  }
  std::unique_ptr<Slot> sl(new Slot());
  sl->ws = ws;
  sl->gid = gid;
  sl->bytes = std::max<size_t>(workspace_size(ws), (size_t)need);
  s->slot_bytes += (int64_t)sl->bytes;
  s->peak_bytes = std::max<int64_t>(s->peak_bytes, s->slot_bytes + s->graph_bytes);
  s->slots.push_back(std::move(sl));
  *out = s->slots.back().get();
  return TAKEN;
}

static void scheduler(bcmi_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    bool progressed = false;
    if (!s->pending.empty() && active_slots(s) < s->max_slots) {
      Query* q = s->pending.front();
      std::string why;
      int rc = BCMI_OK;
      Slot* sl = nullptr;
      const Take t = take_slot(s, q->gid, &sl, &rc, &why);
      if (t == FAIL) {
        s->pending.pop_front();
        q->status = rc;
        q->err = why;
        q->start = q->finish = now_ms() - s->t0;
        q->done = true;
        s->completed++;
        s->cv_done.notify_all();
        progressed = true;
      } else if (t == TAKEN) {
        s->pending.pop_front();
        sl->busy = true;
        q->start = now_ms() - s->t0;
        s->assigned.push_back({q, sl});
        s->peak_active = std::max(s->peak_active, active_slots(s));
        s->cv_work.notify_one();
        progressed = true;
      }
      // WAIT: a busy slot exists (take_slot checked), and its worker
      // notifies cv_sched when it finishes
    }
    if (!progressed) s->cv_sched.wait(lk);
  }
}

static void worker(bcmi_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] { return s->stop || !s->assigned.empty(); });
    if (s->stop && s->assigned.empty()) return;
    auto job = s->assigned.front();
    s->assigned.erase(s->assigned.begin());
    Query* q = job.first;
    Slot* sl = job.second;
    const DevGraph& g = *s->graphs[q->gid];
    lk.unlock();
    int rc = workspace_bc_reset(sl->ws);
    RunResult R;
    if (rc == BCMI_OK) rc = bc_run(g, sl->ws, s->prm, q->src.data(), (int64_t)q->src.size(), BCMI_MODE_BATCHED,
                                   BCMI_SRC_EXACT, &R);
    std::string err = rc == BCMI_OK ? "" : bcmi_last_error();
    std::vector<double> bc(g.n);
    if (rc == BCMI_OK) {
      hipError_t e = hipMemcpyAsync(bc.data(), workspace_bc(sl->ws), sizeof(double) * g.n, hipMemcpyDeviceToHost,
                                    workspace_stream(sl->ws));
      if (e == hipSuccess) e = hipStreamSynchronize(workspace_stream(sl->ws));
      if (e != hipSuccess) { rc = BCMI_ERR_HIP; err = hipGetErrorString(e); }
      g.to_caller(bc.data());
    }
    lk.lock();
    q->bc.swap(bc);
    q->status = rc;
    q->err = err;
    q->te = R.te;
    q->finish = now_ms() - s->t0;
    q->done = true;
    s->completed++;
    // the sweeps may have grown the workspace (work lists, level tables,
    // workgroup-sweep state): charge what it holds now
    const size_t held = workspace_size(sl->ws);
    if (held > sl->bytes) {
      s->slot_bytes += (int64_t)(held - sl->bytes);
      sl->bytes = held;
      s->peak_bytes = std::max<int64_t>(s->peak_bytes, s->slot_bytes + s->graph_bytes);
    }
    sl->busy = false;
    sl->last_used = now_ms();
    s->cv_done.notify_all();
    s->cv_sched.notify_one();
  }
}

extern "C" {

int bcmi_server_create(int device, int64_t hbm_budget, int max_slots, bcmi_server** out) {
  return guard([&]() -> int {
    if (!out || hbm_budget <= 0 || max_slots < 1 || max_slots > 64) return set_error(BCMI_ERR_ARG, "bad server arguments");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
      return set_error(BCMI_ERR_HIP, "no such HIP device");
    if (hipSetDevice(device) != hipSuccess) return set_error(BCMI_ERR_HIP, "hipSetDevice failed");
    bcmi_server* s = new bcmi_server();
    s->device = device;
    s->budget = hbm_budget;
    s->max_slots = max_slots;
    s->t0 = now_ms();
    // queries already overlap on the slots: each runs its batches one at a time
    // (no helper workspaces outside the budget's accounting of one per slot)
    s->prm.concurrency = 1;
    try {
      s->workers.reserve(max_slots);
      s->sched = std::thread(scheduler, s);
      for (int i = 0; i < max_slots; i++) s->workers.emplace_back(worker, s);
    } catch (const std::exception& e) {
      // a thread that cannot start: stop the ones that did, then report
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
      }
      s->cv_sched.notify_all();
      s->cv_work.notify_all();
      if (s->sched.joinable()) s->sched.join();
      for (auto& t : s->workers) t.join();
      delete s;
      return set_error(BCMI_ERR_NOMEM, std::string("server threads could not start: ") + e.what());
    }
    *out = s;
    return BCMI_OK;
  });
}

int bcmi_server_destroy(bcmi_server* s) {
  return guard([&]() -> int {
    if (!s) return set_error(BCMI_ERR_ARG, "null server");
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // finish what was admitted; drop what is still queued
      s->cv_done.wait(lk, [&] { return s->assigned.empty() && active_slots(s) == 0; });
      s->stop = true;
      // queries never admitted complete with an error (their waiters return)
      for (Query* q : s->pending) {
        q->status = BCMI_ERR_STATE;
        q->err = "server destroyed before the query was admitted";
        q->done = true;
      }
      s->pending.clear();
      s->cv_done.notify_all();
    }
    s->cv_sched.notify_all();
    s->cv_work.notify_all();
    s->sched.join();
    for (auto& t : s->workers) t.join();
    (void)hipSetDevice(s->device);
    for (auto& sl : s->slots) workspace_destroy(sl->ws);
    for (auto& g : s->graphs) g->release();
    delete s;
    return BCMI_OK;
  });
}

int bcmi_server_add_graph(bcmi_server* s, int64_t n, const int64_t* off, const int32_t* adj, int32_t* graph_id) {
  return guard([&]() -> int {
    if (!s || !off || !graph_id) return set_error(BCMI_ERR_ARG, "null argument");
    (void)hipSetDevice(s->device);
    std::unique_ptr<DevGraph> g(new DevGraph());
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return set_error(BCMI_ERR_HIP, "stream");
    int rc = devgraph_build(*g, n, off, adj, 0, BCMI_H_NONE, s->prm, st);
    (void)hipStreamDestroy(st);
    if (rc != BCMI_OK) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->graph_bytes + (int64_t)g->bytes > s->budget) {
      g->release();
      return set_error(BCMI_ERR_NOMEM, "graph does not fit in the HBM budget");
    }
    s->graph_bytes += (int64_t)g->bytes;
    s->graph_ws_bytes.push_back(workspace_bytes(*g, BCMI_MODE_BATCHED, 64, s->prm));
    s->graphs.push_back(std::move(g));
    *graph_id = (int32_t)s->graphs.size() - 1;
    return BCMI_OK;
  });
}

int bcmi_server_submit(bcmi_server* s, int32_t graph_id, const int32_t* sources, int64_t nsrc, int64_t* query_id) {
  return guard([&]() -> int {
    if (!s || !query_id || nsrc < 0 || (nsrc > 0 && !sources)) return set_error(BCMI_ERR_ARG, "bad argument");
    std::unique_ptr<Query> q(new Query());
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (graph_id < 0 || graph_id >= (int32_t)s->graphs.size()) return set_error(BCMI_ERR_ARG, "unknown graph id");
      int64_t n = s->graphs[graph_id]->n;
      for (int64_t i = 0; i < nsrc; i++)
        if (sources[i] < 0 || sources[i] >= n) return set_error(BCMI_ERR_RANGE, "source id out of range");
      q->id = s->next_id++;
      q->gid = graph_id;
      q->src.assign(sources, sources + nsrc);
      q->arrive = now_ms() - s->t0;
      *query_id = q->id;
      s->pending.push_back(q.get());
      s->queries[q->id] = std::move(q);
    }
    s->cv_sched.notify_one();
    return BCMI_OK;
  });
}

int bcmi_server_wait(bcmi_server* s, int64_t query_id, double* bc, bcmi_query_result* res) {
  return guard([&]() -> int {
    if (!s) return set_error(BCMI_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->queries.find(query_id);
    if (it == s->queries.end()) return set_error(BCMI_ERR_ARG, "unknown query id");
    Query* q = it->second.get();
    s->cv_done.wait(lk, [&] { return q->done; });
    if (res) {
      res->query_id = q->id;
      res->graph_id = q->gid;
      res->status = q->status;
      res->arrive_ms = q->arrive;
      res->start_ms = q->start;
      res->finish_ms = q->finish;
      res->te = q->te;
    }
    int rc = q->status;
    if (rc == BCMI_OK && bc) memcpy(bc, q->bc.data(), sizeof(double) * q->bc.size());
    if (rc != BCMI_OK) set_error(rc, q->err);
    s->queries.erase(it);
    return rc;
  });
}

int bcmi_server_stats(bcmi_server* s, int64_t* completed, int64_t* peak_slots, int64_t* peak_bytes) {
  return guard([&]() -> int {
    if (!s) return set_error(BCMI_ERR_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (peak_slots) *peak_slots = s->peak_active;
    if (peak_bytes) *peak_bytes = s->peak_bytes;
    return BCMI_OK;
  });
}

}  // extern "C"
