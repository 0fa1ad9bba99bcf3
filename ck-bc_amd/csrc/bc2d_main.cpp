// ck-bc-hip: drop-in replacement of the reference's bc2d program
// (program/bc-cuda/bc2d.c:673-1139) on one MI355X.
//
// Same command line (getopt "o:p:amhDR:Uf:n:r:S:E:N:H:", bc2d.c:752-823),
// same text edge-list input, same result file <o>_1X1_0.log with
// "BC RESULTS" / "ROWPP N" / "NodeId\tBC_VAL" and "%d\t%.2f" rows of BC/2
// (bc2d.c:907-911, 1057-1068) and the same RESULTS summary on stdout
// (bc2d.c:1070-1103).  Additive options: -B plain|batched (default batched),
// -s seed (sample -N random sources instead of startv..startv+N-1),
// -F file (sources, one id per line), -P (full precision %.17g in the result
// file), -X (with -H 1: exact BC over the listed sources instead of the
// reference's "a core source stands for its pruned tree" semantics),
// -S/-E generate an R-MAT graph (disabled in the reference, bc2d.c:664),
// -T key=value sets an engine tunable (repeatable; bcmi_set_param).
// Errors print a message and exit 1 like prexit (bc2d.c:64-72).
#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "bcmi.h"

static void die(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(EXIT_FAILURE);
}

static void usage(const char* p) {
  die("Usage (BC-1GPU):\n"
      "\t %s -p 1x1 [-o outfile] [-D] [-m] [-N <# of searches>] [-H 0|1|2|3]\n"
      "\t -> to visit a graph read from file:\n"
      "\t\t -f <graph file> -n <# vertices> [-r <start vert>]\n"
      "\t -> to visit an RMAT graph:\n"
      "\t\t -S <scale> [-E <edge factor>]\n"
      "\t Where:\n"
      "\t\t -D to ENABLE debug information\n"
      "\t\t -U DO NOT make graph Undirected\n"
      "\t\t -B plain|batched   one source per sweep, or 64 per sweep (default)\n"
      "\t\t -s <seed>          sample -N random sources\n"
      "\t\t -F <file>          read the sources from a file\n"
      "\t\t -P                 full-precision result file\n"
      "\t\t -X                 with -H 1|3: exact BC over exactly the listed sources\n"
      "\t\t -H 2|3             2-degree sources taken from their neighbours' sweeps\n"
      "\t\t                    (with -B plain; -H 3 also peels degree-1 vertices)\n"
      "\t\t -C <file>          binary CSR cache of the -f graph: used when it matches\n"
      "\t\t                    the text file's size and time, (re)written otherwise\n"
      "\t\t -T <key>=<value>   engine tunable (integer; repeatable), e.g. -T grid=2048\n"
      "\n",
      p);
}

#define CK(call)                                                              \
  do {                                                                        \
    int rc_ = (call);                                                         \
    if (rc_ != BCMI_OK) die("bcmi error %d: %s\n", rc_, bcmi_last_error());   \
  } while (0)

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
  if (argc == 1) usage(argv[0]);
  int R = 0, C = 0, heuristic = 0, debug = 0, undirected = 1, gread = -1, scale = 21, edgef = 16;
  uint64_t N = 0, startv = 0, nbfs = 1, seed = 0;
  int have_seed = 0, full_prec = 0, exact = 0, mode = BCMI_MODE_BATCHED;
  std::string gfile, ofile, sfile, cfile;
  std::vector<std::pair<std::string, long long>> tunables;  // -T key=value, applied before the upload
  printf("MPI Tasks 1\n");
  for (int i = 0; i < argc; i++) printf(" %s", argv[i]);
  printf("\n\n");
  int opt;
  while ((opt = getopt(argc, argv, "o:p:amhDR:Uf:n:r:S:E:N:H:B:s:F:PXC:T:")) != -1) {
    switch (opt) {
      case 'H':
        if (sscanf(optarg, "%d", &heuristic) != 1) die("Invalid Heuristic Option (-H): %s\n", optarg);
        break;
      case 'o': ofile = optarg; break;
      case 'p': {
        if (sscanf(optarg, "%dx%d", &R, &C) != 2) die("Invalid proc mesh field.\n");
        break;
      }
      case 'R': {
        int r = 0;
        if (sscanf(optarg, "%d", &r) != 1 || (r != 0 && r != 1)) die("Invalid random option (-R): %s\n", optarg);
        break;
      }
      case 'D': debug = 1; break;
      case 'm': break;  // the reference's multi-GPU switch; one GPU here
      case 'a': break;
      case 'f':
        if (gread == 0) die("Unexpected option -f!\n");
        gread = 1; gfile = optarg; break;
      case 'n':
        if (gread == 0) die("Unexpected option -n!\n");
        gread = 1;
        if (sscanf(optarg, "%" SCNu64, &N) != 1) die("Invalid number of vertices (-n): %s\n", optarg);
        break;
      case 'r':
        if (sscanf(optarg, "%" SCNu64, &startv) != 1) die("Invalid root vertex (-r): %s\n", optarg);
        break;
      case 'S':
        if (gread == 1) die("Unexpected option -S!\n");
        gread = 0;
        if (sscanf(optarg, "%d", &scale) != 1) die("Invalid scale (-S): %s\n", optarg);
        N = (uint64_t)1 << scale;
        break;
      case 'E':
        if (gread == 1) die("Unexpected option -E!\n");
        gread = 0;
        if (sscanf(optarg, "%d", &edgef) != 1) die("Invalid edge factor (-E): %s\n", optarg);
        break;
      case 'N':
        if (sscanf(optarg, "%" SCNu64, &nbfs) != 1) die("Invalid number of bfs (-N): %s\n", optarg);
        break;
      case 'U': undirected = 0; break;
      case 'B':
        if (!strcmp(optarg, "plain")) mode = BCMI_MODE_PLAIN;
        else if (!strcmp(optarg, "batched")) mode = BCMI_MODE_BATCHED;
        else die("Invalid mode (-B): %s\n", optarg);
        break;
      case 's':
        if (sscanf(optarg, "%" SCNu64, &seed) != 1) die("Invalid seed (-s): %s\n", optarg);
        have_seed = 1;
        break;
      case 'F': sfile = optarg; break;
      case 'P': full_prec = 1; break;
      case 'X': exact = 1; break;
      case 'C': cfile = optarg; break;
      case 'T': {
        const char* eq = strchr(optarg, '=');
        char* end = nullptr;
        long long v = eq ? strtoll(eq + 1, &end, 10) : 0;
        if (!eq || eq == optarg || !end || end == eq + 1 || *end)
          die("Invalid tunable (-T key=value with an integer value): %s\n", optarg);
        tunables.emplace_back(std::string(optarg, eq - optarg), v);
        break;
      }
      default: usage(argv[0]);
    }
  }
  if (gread == 1 && (gfile.empty() || !N))
    die("Graph file (-f) and number of vertices (-n) must be specified for file based bfs.\n");
  if (gread == -1) die("Graph file (-f) and number of vertices (-n) must be specified for file based bfs.\n");
  if (R <= 0 || R > 1024 || C <= 0 || C > 1024) die("R and C must be in range [1,1024] and [1,1024], respectively.\n");
  if (R * C != 1) die("This build runs on one GPU: use -p 1x1.\n");
  if (startv >= N) die("Invalid start vertex: %" PRIu64 ".\n", startv);
  if (heuristic < 0 || heuristic > 3) die("Invalid Heuristic Option (-H): %d\n", heuristic);
  if ((heuristic & 1) && !undirected) die("The 1-degree reduction (-H 1, -H 3) needs an undirected graph.\n");
  // -H 1 and -H 3 peel pendant vertices; -H 2 and -H 3 take degree-2 sources
  // from their neighbours' sweeps (plain sweeps; bc2d.c:759-760 accepts both and
  // prints "REMOVED" for them, c:1015)
  const bool peel = heuristic == 1 || heuristic == 3, deg2 = heuristic >= 2;

  printf("\n\n****** MI355X (gfx950) BC ******\n\n");
  printf("Total number of vertices (N): %" PRIu64 "\n", N);
  printf("Processor mesh rows (R): %d\n", R);
  printf("Processor mesh columns (C): %d\n", C);
  if (gread) printf("Reading graph from file: %s\n", gfile.c_str());
  else printf("RMAT graph scale: %d\nRMAT graph edge factor: %d\n", scale, edgef);
  printf("\n\n");
  if (heuristic == 0) printf("HEURISTICs: OFF: %d\n", heuristic);
  else if (heuristic == 1) printf("HEURISTICs: 1-degree reduction ON: %d\n", heuristic);
  else if (heuristic == 2) printf("HEURISTICs: 2-degree reduction ON: %d\n", heuristic);
  else printf("HEURISTICs: 1-degree and 2-degree reduction ON: %d\n", heuristic);
  printf("SWEEP: %s\n\n", mode == BCMI_MODE_PLAIN ? "plain (one source per sweep)" : "batched (64 sources per sweep)");
  std::string resname;
  if (!ofile.empty()) {
    printf("Result written to file: %s\n", ofile.c_str());
    resname = ofile + "_1X1_0.log";
  }

  double t0 = now_s();
  bcmi_graph* g = nullptr;
  // binary CSR cache: valid when it records the text file's size and mtime
  // and matches -n and -U
  uint64_t src_size = 0;
  int64_t src_mtime = 0;
  if (gread && !cfile.empty()) {
    struct stat stt;
    if (stat(gfile.c_str(), &stt) == 0) {
      src_size = (uint64_t)stt.st_size;
      src_mtime = (int64_t)stt.st_mtim.tv_sec * 1000000000ll + stt.st_mtim.tv_nsec;
      uint64_t cs = 0;
      int64_t cm = 0;
      bcmi_graph* cg = nullptr;
      if (bcmi_graph_load(cfile.c_str(), &cg, &cs, &cm) == BCMI_OK) {
        if (cs == src_size && cm == src_mtime && (uint64_t)cg->n == N && cg->directed == (undirected ? 0 : 1)) {
          g = cg;
          printf("Graph read from CSR cache %s\n", cfile.c_str());
        } else {
          bcmi_graph_free(cg);
        }
      }
    }
  }
  bool csr_input = false;
  if (gread && !g) {
    // -f may name a binary CSR file (as written by -C or gen_dataset.py)
    char mg[8] = {0};
    FILE* fp = fopen(gfile.c_str(), "rb");
    if (fp) {
      csr_input = fread(mg, 1, 8, fp) == 8 && memcmp(mg, "BCMICSR1", 8) == 0;
      fclose(fp);
    }
    if (csr_input) {
      CK(bcmi_graph_load(gfile.c_str(), &g, nullptr, nullptr));
      if ((uint64_t)g->n != N) die("CSR file %s holds %" PRId64 " vertices, -n says %" PRIu64 "\n", gfile.c_str(), g->n, N);
      if (g->directed != (undirected ? 0 : 1)) die("CSR file %s: directedness differs from -U\n", gfile.c_str());
      printf("Graph read from CSR file %s\n", gfile.c_str());
    }
  }
  if (g) {
    // loaded from the cache or a CSR file
  } else if (gread) {
    uint64_t bad = 0, hn = 0, he = 0;
    int rc = bcmi_graph_read_edgelist(gfile.c_str(), N, undirected ? 0 : 1, &g, &bad, &hn, &he);
    if (rc != BCMI_OK) {
      if (rc == BCMI_ERR_RANGE)
        die("[0] In file %s line %" PRIu64 " found invalid edge for N=%" PRIu64 ": %s\n", gfile.c_str(), bad, N,
            bcmi_last_error());
      die("%s\n", bcmi_last_error());
    }
    if (hn) printf("N=%" PRIu64 " E=%" PRIu64 "\n", hn, he);
    if (!cfile.empty() && src_size) CK(bcmi_graph_save(g, cfile.c_str(), src_size, src_mtime));
  } else {
    CK(bcmi_gen_rmat(scale, edgef, 0.57, 0.19, 0.19, seed ? seed : 1, 1, &g));
  }
  printf(" done in %f secs\n", now_s() - t0);

  bcmi_ctx* ctx = nullptr;
  CK(bcmi_create(0, &ctx));
  for (const auto& t : tunables) CK(bcmi_set_param(ctx, t.first.c_str(), (int64_t)t.second));
  t0 = now_s();
  CK(bcmi_upload_graph(ctx, g->n, g->off, g->adj, g->directed, peel ? BCMI_H_PEEL : BCMI_H_NONE));
  double redu_time = peel ? now_s() - t0 : 0.0;
  std::vector<int32_t> round;
  if (peel) {
    round.resize(g->n);
    int32_t nr = 0;
    CK(bcmi_get_peel(ctx, round.data(), nullptr, nullptr, &nr));
    int64_t removed = 0;
    for (int32_t r : round) removed += r > 0;
    printf("task 0 Total vertices removed %" PRId64 " in %d rounds\n", removed, nr);
  }

  // sources: startv .. startv+N-1 (bc2d.c:1017-1018), a seeded sample, or a file
  std::vector<int32_t> sources;
  uint64_t want = nbfs < N ? nbfs : N;
  if (!sfile.empty()) {
    FILE* fp = fopen(sfile.c_str(), "r");
    if (!fp) die("Cannot open file %s...\n", sfile.c_str());
    long long v;
    while (fscanf(fp, "%lld", &v) == 1) {
      if (v < 0 || (uint64_t)v >= N) die("Invalid source %lld\n", v);
      sources.push_back((int32_t)v);
    }
    fclose(fp);
  } else if (have_seed) {
    sources.resize(want);
    CK(bcmi_sample_sources(g, (int64_t)want, seed, sources.data()));
  } else {
    for (uint64_t ui = 0; ui < want && startv + ui < N; ui++) sources.push_back((int32_t)(startv + ui));
  }
  // skip rule of bc2d.c:1026-1035: degree 0 in the (reduced) graph
  int64_t skipped = 0;
  std::vector<int32_t> run;
  for (int32_t s : sources) {
    bool zero = g->off[s + 1] == g->off[s] || (peel && round[s] != 0);
    if (zero) skipped++;
    else run.push_back(s);
  }
  if (peel && !exact) run = sources;  // the engine applies the same rule when expanding
  printf("task 0: BC computation is running...\n");
  bcmi_run_stats st;
  memset(&st, 0, sizeof st);
  t0 = now_s();
  if (!run.empty())
    CK(bcmi_bc_run(ctx, run.data(), (int64_t)run.size(), mode,
                   (peel && !exact ? BCMI_SRC_EXPAND : BCMI_SRC_EXACT) | (deg2 ? BCMI_SRC_DEG2 : 0), &st));
  double all_time = now_s() - t0;
  std::vector<double> bc(g->n);
  t0 = now_s();
  CK(bcmi_get_bc(ctx, bc.data()));
  double bcred_time = now_s() - t0;

  if (!resname.empty()) {
    FILE* out = fopen(resname.c_str(), "w");
    if (!out) die("Cannot open file %s...\n", resname.c_str());
    fprintf(out, "BC RESULTS\n");
    fprintf(out, "ROWPP %" PRId64 "\n", g->n);
    fprintf(out, "NodeId\tBC_VAL\n");
    for (int64_t k = 0; k < g->n; k++) {
      if (full_prec) fprintf(out, "%" PRId64 "\t%.17g\n", k, bc[k] / 2.0);
      else fprintf(out, "%" PRId64 "\t%.2f\n", k, bc[k] / 2.0);
    }
    fclose(out);
  }
  if (debug) {
    FILE* dbg = fopen("debug_0.log", "w");
    if (dbg) {
      fprintf(dbg, "sweeps %" PRId64 " batches %" PRId64 " max_levels %" PRId64 " sum_n %" PRId64 " sum_m %" PRId64
                   " sum_a %" PRId64 " gpu_ms %.3f\n",
              st.sources_run, st.batches, st.max_levels, st.sum_n, st.sum_m, st.sum_a, st.gpu_ms);
      fclose(dbg);
    }
  }
  // Per-source statistics as the reference keeps them (c:1037-1048,
  // c:565-579): Max/Min over the per-source times (a batched sweep's time
  // divided by its sources, each source getting that share), Mean = total /
  // rounds, M-TEPS = mean over sources of (degree sum of the reached
  // vertices / 2) / time / 2^20.  Rounds are the listed sources that were not
  // skipped (c:1026-1037); under -H 1 a core source stands for its tree and
  // the engine sweeps each core root once, so the statistics average over
  // those sweeps.
  const int64_t nrounds = run.empty() ? 0 : (int64_t)sources.size() - skipped;
  const int64_t nsw = st.sources_run;
  std::vector<double> ts(nsw > 0 ? nsw : 1), tsum(nsw > 0 ? nsw : 1);
  std::vector<int64_t> ms(nsw > 0 ? nsw : 1);
  double tmax = 0.0, tmin = 0.0, mteps_sum = 0.0;
  if (nsw > 0) {
    CK(bcmi_get_sweep_times(ctx, ts.data()));
    CK(bcmi_get_sweep_stats(ctx, nullptr, ms.data(), nullptr, nullptr));
    tmin = 1e300;
    for (int64_t i = 0; i < nsw; i++) {
      const double t = ts[i] / 1e3;
      tmax = std::max(tmax, t);
      if (t > 0) tmin = std::min(tmin, t);
      if (t > 0) mteps_sum += (double)(ms[i] / 2) / t / 1048576.0;
    }
    if (tmin == 1e300) tmin = 0.0;
  }
  const double mean = nrounds ? all_time / nrounds : 0.0;
  const double mteps = nsw ? mteps_sum / nsw : 0.0;
  printf("\n------------- RESULTS ---------------\n");
  printf("ClusterId\tSkip\tRounds\tExecTime\tRoundsTime\tMax\tMin\tMean\t1-dReduTime\n");
  printf("%d\t%" PRId64 "\t%" PRId64 "\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", 0, skipped, nrounds, all_time + bcred_time,
         all_time, tmax, tmin, mean, redu_time);
  printf("task 0 BC skipped: %" PRId64 " \n", skipped);
  printf("task 0 BC rounds: %" PRId64 "  \n", nrounds);
  printf("task 0 BC execution total time: %lf sec\n", all_time + bcred_time);
  printf("task 0 BC rounds total time: %lf sec\n", all_time);
  printf("task 0 BC Max time: %lf sec\n", tmax);
  printf("task 0 BC Min time: %lf sec\n", tmin);
  printf("task 0 BC Mean time: %lf sec\n", mean);
  printf("task 0 BC M-TEPS: %lf sec\n", mteps);
  if (nbfs < N) {
    // the whole graph's run time estimated from the sampled sources (c:1092-1097)
    const uint64_t vskipped = (N - nbfs) * (uint64_t)skipped / nbfs;
    printf("task 0 BC simulated time: %lf sec (virtual skipped %u)\n", mean * (double)(N - skipped - vskipped),
           (unsigned)vskipped);
  }
  printf("task 0 BC GTEPS (forward+backward arcs): %lf\n", all_time > 0 ? st.te / all_time / 1e9 : 0.0);
  if (peel) printf("task 0 1-Degree reduction: %lf sec\n", redu_time);
  if (deg2)
    printf("task 0 2-Degree sources (and degree 1 when not peeled) taken from their neighbours: %" PRId64 "\n",
           st.deg2_sources);
  printf("\n");
  bcmi_destroy(ctx);
  bcmi_graph_free(g);
  return 0;
}
