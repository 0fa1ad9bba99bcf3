// Internal declarations shared by the host graph code, the device engine and
// the query server.
#pragma once
#include <cstdint>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "bcmi.h"

namespace bcmi {

// Records the message for bcmi_last_error() (thread-local) and returns code.
int set_error(int code, const std::string& msg);

// Runs the body of a C entry point: a C++ exception from inside the library
// (host memory for a container, or a thread that cannot start) returns
// BCMI_ERR_NOMEM with its message instead of crossing the C ABI.
template <class F>
int guard(F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return set_error(BCMI_ERR_NOMEM, "out of host memory");
  } catch (const std::exception& e) {
    return set_error(BCMI_ERR_NOMEM, std::string("host resources exhausted: ") + e.what());
  } catch (...) {
    return set_error(BCMI_ERR_NOMEM, "host resources exhausted");
  }
}

bcmi_graph* graph_alloc(int64_t n, int64_t m, int directed);
int graph_save(const bcmi_graph* g, const char* path, uint64_t src_size, int64_t src_mtime_ns);
int graph_load(const char* path, bcmi_graph** out, uint64_t* src_size, int64_t* src_mtime_ns);
bcmi_graph* csr_from_arcs(int64_t n, const int32_t* tail, const int32_t* head, int64_t k, int mirror,
                          int directed);
void relabel_by_degree(int64_t n, const int64_t* off, const int32_t* adj, std::vector<int32_t>& n2o,
                       std::vector<int32_t>& o2n, std::vector<int64_t>& noff, std::vector<int32_t>& nadj);

// Tunable launch and algorithm parameters.
struct Params {
  int block = 256;          // threads per block (multiple of 64)
  int grid = 4096;          // blocks per batched expansion kernel (about twice what is resident)
  int grid_plain = 4096;    // blocks per plain-sweep kernel
  int grid_back = 0;        // blocks per batched backward kernel (0: grid)
  int chunk = 256;          // max edges per work item
  int64_t alpha_plain = 1000;   // x1000; bottom-up when frontier arcs > alpha * unvisited arcs
  int64_t alpha_batched = 1000; // x1000
  int64_t beta_plain = 16;      // bottom-up only when frontier arcs >= all arcs / beta
  int64_t beta_batched = 16;
  int fuse_bu = 2;              // bottom-up levels sum path counts in the same scan: 0 never, 1 of
                                // unsplit vertices, 2 of all (split vertices add per record atomically)
  int64_t gamma_push = 8;       // batched top-down levels with frontier arcs < all arcs / gamma push
                                // path counts from the frontier (0: always pull from parents)
  int pack_bu = 1;              // batched bottom-up packs vertices of degree <= 16 four to a wave
  int narrow_sigma = 1;         // exact batched sweeps first keep path counts in 32-bit (1) or
                                // 16-bit (2) rows, redone with the next wider rows when a count
                                // passes the rows' range (0: 64-bit rows only)
  int64_t gamma_back = 0;       // batched backward levels whose successor level holds fewer than
                                // 1/gamma_back of their arcs push coef from it (0: always pull)
  int steps = 16;           // levels per captured graph
  int sgrid = 512;          // blocks per kernel in the small-grid (sparse-level) graphs
  int64_t sparse_verts = 16384; // levels below this many vertices replay the small-grid graphs
  int streams = 4;          // sweeps of one call run at once on deep graphs (own workspace and stream each)
  int concurrency = 0;      // sweeps of one call run at once on any graph (0: `streams` on deep graphs, else 1)
  int strategy = 0;         // 0 auto, 1 device-wide level loop, 2 one workgroup per source
  int wg_depth = 128;       // auto: graphs whose hub BFS depth reaches this use a workgroup per source
  int64_t wg_mem = 16ll << 30;  // device bytes for the per-source state of the workgroup sweeps
  int wg_slots = 2048;      // at most this many sources in flight (workgroups; 8 per CU fit)
  int width = 64;           // sources per batch
  int sigma = 2;            // path counts: 0 exact uint64 (error on overflow), 1 scaled FP64,
                            // 2 exact, batches that overflow are recomputed scaled
  int kflags = 0;           // kernel variants: 1 gate mask probes with level bitmaps,
                            // 2 non-temporal adjacency loads
  int batching = 4;         // source order for batches: 0 as given, 1 hub-depth signature,
                            // 2 highest-degree neighbour then signature, 3 that neighbour's
                            // signature first, 4 highest then second highest-degree neighbour
                            // then signature, 5 highest-degree neighbour, signature, second
  int reorder = 1;          // relabel vertices by descending degree at upload (locality)
  int trace = 0;            // diagnostic: per-level statistics on stderr (use with steps = 1)
  int graphs = 1;           // 1: the level loop replays captured graphs; 0: it launches the same
                            // kernels one by one (for tools that do not follow graph launches)
  int deg2_slots = 0;       // 2-degree sources: forward states of their neighbours kept at once
                            // (0: as many as deg2_mem holds; 1 turns the heuristic off)
  int64_t deg2_mem = 4ll << 30;  // device bytes for those states when deg2_slots is 0
  int ktime = 0;            // kernel timed on the device (wave entry/exit stamps) inside the level loop:
                            // 0 none, 1 batched sigma, 2 batched backward, 3 batched top-down,
                            // 4 batched bottom-up, 5 plain top-down, 6 plain bottom-up, 7 plain backward
};

}  // namespace bcmi
