// Host-side graph construction for the BC engine: text edge-list ingest,
// normalisation, CSR build and the seeded synthetic generators (R-MAT,
// pendant trees, bridged-block chains).  OpenMP-parallel; deterministic for
// a given seed regardless of thread count.
//
// Reference behaviour restated here (paths relative to the ck-bc root):
//   read_graph  program/bc-cuda/bc2d.c:175-254  ('#' comment lines, ids >= N
//               rejected with the line number, "# Nodes: N Edges: E" echoed)
//   mirror      bc2d.c:152-170  (reversed copy of every non-loop edge unless
//               the graph is directed, -U)
//   norm_graph  bc2d.c:430-453  (sort, drop duplicates and self loops)
//   build_csc   bc2d.c:472-501  (column c lists every u with an arc u->c,
//               ascending; for a mirrored graph this is the CSR adjacency)
// Offsets are int64 (the reference's uint32 offsets overflow past 2^32 arcs,
// bc2d.c:474) and vertex ids int32.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <new>
#include <numeric>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>
#include <omp.h>

#include "bcmi.h"
#include "bcmi_internal.h"

namespace bcmi {

// ---------------------------------------------------------------- RNG ----
static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// xoshiro256** seeded through splitmix64 (Blackman & Vigna).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) { x = splitmix64(x); s[i] = x; }
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t k) {  // unbiased enough for k << 2^64 (Lemire multiply-shift)
    return (uint64_t)(((unsigned __int128)next() * k) >> 64);
  }
};

// ------------------------------------------------------------ graph ------
bcmi_graph* graph_alloc(int64_t n, int64_t m, int directed) {
  bcmi_graph* g = (bcmi_graph*)calloc(1, sizeof(bcmi_graph));
  if (!g) return nullptr;
  g->n = n; g->m = m; g->directed = directed;
  g->off = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  g->adj = (int32_t*)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
  if (!g->off || !g->adj) { free(g->off); free(g->adj); free(g); return nullptr; }
  return g;
}

// Build a CSR from arcs (src[i] -> dst[i]); the adjacency of c lists the
// tails of the arcs into c (CSC, bc2d.c:472-501), sorted ascending with
// duplicates and self loops removed (bc2d.c:430-453).  If mirror is set the
// reversed copy of every non-loop arc is added first (bc2d.c:152-170).
bcmi_graph* csr_from_arcs(int64_t n, const int32_t* tail, const int32_t* head, int64_t k,
                          int mirror, int directed) {
  std::vector<int64_t> cnt(n + 1, 0);
  std::atomic<int64_t>* acnt = reinterpret_cast<std::atomic<int64_t>*>(cnt.data());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < k; i++) {
    int32_t u = tail[i], v = head[i];
    if (u == v) continue;
    acnt[v + 1].fetch_add(1, std::memory_order_relaxed);
    if (mirror) acnt[u + 1].fetch_add(1, std::memory_order_relaxed);
  }
  for (int64_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
  int64_t tot = cnt[n];
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (tot > 0 ? tot : 1));
  if (!tmp) return nullptr;
  std::vector<int64_t> cur(cnt.begin(), cnt.end() - 1);
  std::atomic<int64_t>* acur = reinterpret_cast<std::atomic<int64_t>*>(cur.data());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < k; i++) {
    int32_t u = tail[i], v = head[i];
    if (u == v) continue;
    tmp[acur[v].fetch_add(1, std::memory_order_relaxed)] = u;
    if (mirror) tmp[acur[u].fetch_add(1, std::memory_order_relaxed)] = v;
  }
  std::vector<int64_t> deg(n, 0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < n; v++) {
    int32_t* b = tmp + cnt[v];
    int32_t* e = tmp + cnt[v + 1];
    std::sort(b, e);
    deg[v] = std::unique(b, e) - b;
  }
  int64_t m = 0;
  std::vector<int64_t> noff(n + 1);
  for (int64_t v = 0; v < n; v++) { noff[v] = m; m += deg[v]; }
  noff[n] = m;
  bcmi_graph* g = graph_alloc(n, m, directed);
  if (!g) { free(tmp); return nullptr; }
  memcpy(g->off, noff.data(), sizeof(int64_t) * (n + 1));
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < n; v++)
    memcpy(g->adj + noff[v], tmp + cnt[v], sizeof(int32_t) * deg[v]);
  free(tmp);
  return g;
}

// ----------------------------------------------------------- parsing -----
// Parses one line [p, e).  Returns 0 = comment/blank, 1 = edge, -1 = malformed.
static int parse_line(const char* p, const char* e, uint64_t* u, uint64_t* v, int* nodes_hdr,
                      uint64_t* hn, uint64_t* he) {
  if (memchr(p, '#', e - p)) {
    const char* q = (const char*)memmem(p, e - p, "Nodes:", 6);
    if (q) {
      *nodes_hdr = 1;
      std::string s(p, e);
      unsigned long long a = 0, b = 0;
      if (sscanf(s.c_str(), "# Nodes: %llu Edges: %llu", &a, &b) >= 1) { *hn = a; *he = b; }
    }
    return 0;
  }
  const char* q = p;
  while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
  if (q == e) return 0;
  uint64_t x[2];
  for (int t = 0; t < 2; t++) {
    while (q < e && (*q == ' ' || *q == '\t')) q++;
    if (q == e || *q < '0' || *q > '9') return -1;
    uint64_t val = 0;
    while (q < e && *q >= '0' && *q <= '9') {
      uint64_t nv = val * 10 + (uint64_t)(*q - '0');
      if (nv < val) return -1;
      val = nv; q++;
    }
    x[t] = val;
  }
  *u = x[0]; *v = x[1];
  return 1;
}

int read_edgelist(const char* path, uint64_t n, int directed, bcmi_graph** out, uint64_t* bad_line,
                  uint64_t* hdr_n, uint64_t* hdr_e) {
  if (n == 0 || n > (uint64_t)INT32_MAX) return set_error(BCMI_ERR_ARG, "vertex count must be in [1, 2^31-1]");
  int fd = open(path, O_RDONLY);
  if (fd < 0) return set_error(BCMI_ERR_IO, std::string("cannot open file ") + path);
  struct stat st;
  if (fstat(fd, &st) != 0) { close(fd); return set_error(BCMI_ERR_IO, "cannot stat file"); }
  size_t sz = (size_t)st.st_size;
  const char* data = nullptr;
  if (sz > 0) {
    data = (const char*)mmap(nullptr, sz, PROT_READ, MAP_PRIVATE, fd, 0);
    if (data == MAP_FAILED) { close(fd); return set_error(BCMI_ERR_IO, "mmap failed"); }
  }
  int nt = omp_get_max_threads();
  std::vector<size_t> cut(nt + 1, sz);
  cut[0] = 0;
  for (int t = 1; t < nt; t++) {
    // a cut starts a line: scan from byte 1 at least (fewer bytes than
    // threads put early cuts at 0, and data[c - 1] must stay in the file)
    size_t c = std::max<size_t>(sz * t / nt, std::min<size_t>(sz, 1));
    while (c < sz && data[c - 1] != '\n') c++;
    cut[t] = std::max(c, cut[t - 1]);
  }
  std::vector<std::vector<int32_t>> tails(nt), heads(nt);
  std::vector<uint64_t> lines(nt, 0), badl(nt, 0), badkind(nt, 0);
  std::vector<uint64_t> hn(nt, 0), he(nt, 0);
  std::vector<int> hh(nt, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    const char* p = data + cut[t];
    const char* end = data + cut[t + 1];
    uint64_t lc = 0;
    while (p < end) {
      const char* nl = (const char*)memchr(p, '\n', end - p);
      const char* le = nl ? nl : end;
      uint64_t u, v;
      int r = parse_line(p, le, &u, &v, &hh[t], &hn[t], &he[t]);
      if (r != 0 || !memchr(p, '#', le - p)) lc++;
      if (r == 1) {
        if (u >= n || v >= n) { if (!badl[t]) { badl[t] = lc; badkind[t] = 1; } }
        else { tails[t].push_back((int32_t)u); heads[t].push_back((int32_t)v); }
      } else if (r < 0) {
        if (!badl[t]) { badl[t] = lc; badkind[t] = 2; }
      }
      p = le + 1;
    }
    lines[t] = lc;
  }
  if (data) munmap((void*)data, sz);
  close(fd);
  uint64_t base = 0;
  for (int t = 0; t < nt; t++) {
    if (badl[t]) {
      if (bad_line) *bad_line = base + badl[t];
      return badkind[t] == 1
                 ? set_error(BCMI_ERR_RANGE, "invalid edge (vertex id >= N) at data line " + std::to_string(base + badl[t]))
                 : set_error(BCMI_ERR_IO, "malformed edge line at data line " + std::to_string(base + badl[t]));
    }
    base += lines[t];
  }
  for (int t = 0; t < nt; t++)
    if (hh[t]) { if (hdr_n) *hdr_n = hn[t]; if (hdr_e) *hdr_e = he[t]; break; }
  size_t k = 0;
  for (int t = 0; t < nt; t++) k += tails[t].size();
  std::vector<int32_t> T(k), H(k);
  size_t o = 0;
  for (int t = 0; t < nt; t++) {
    std::copy(tails[t].begin(), tails[t].end(), T.begin() + o);
    std::copy(heads[t].begin(), heads[t].end(), H.begin() + o);
    o += tails[t].size();
  }
  bcmi_graph* g = csr_from_arcs((int64_t)n, T.data(), H.data(), (int64_t)k, directed ? 0 : 1, directed);
  if (!g) return set_error(BCMI_ERR_NOMEM, "out of host memory building CSR");
  *out = g;
  return BCMI_OK;
}

// ------------------------------------------------------- generators ------
static void permutation(int64_t n, uint64_t seed, std::vector<int32_t>& perm) {
  perm.resize(n);
  std::iota(perm.begin(), perm.end(), 0);
  Rng r(seed ^ 0x5851F42D4C957F2Dull);
  for (int64_t i = n - 1; i > 0; i--) {  // Fisher-Yates
    int64_t j = (int64_t)r.below((uint64_t)i + 1);
    std::swap(perm[i], perm[j]);
  }
}

// R-MAT edge i is drawn from its own xoshiro stream seeded with
// splitmix64(seed, i), so the edge set does not depend on thread count.
static void rmat_edges(int scale, int64_t k, double a, double b, double c, uint64_t seed,
                       std::vector<int32_t>& T, std::vector<int32_t>& H) {
  T.resize(k); H.resize(k);
  double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < k; i++) {
    Rng r(splitmix64(seed) ^ splitmix64((uint64_t)i * 0xD1B54A32D192ED03ull + 1));
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      double x = r.uniform();
      u <<= 1; v <<= 1;
      if (x < a) {
      } else if (x < ab) {
        v |= 1;
      } else if (x < abc) {
        u |= 1;
      } else {
        u |= 1; v |= 1;
      }
    }
    T[i] = (int32_t)u; H[i] = (int32_t)v;
  }
}

int gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed, int permute,
             bcmi_graph** out) {
  if (scale < 1 || scale > 30 || edgefactor < 1) return set_error(BCMI_ERR_ARG, "bad R-MAT scale/edge factor");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return set_error(BCMI_ERR_ARG, "bad R-MAT probabilities");
  int64_t n = (int64_t)1 << scale;
  int64_t k = n * (int64_t)edgefactor;
  std::vector<int32_t> T, H;
  rmat_edges(scale, k, a, b, c, seed, T, H);
  if (permute) {
    std::vector<int32_t> perm;
    permutation(n, seed, perm);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < k; i++) { T[i] = perm[T[i]]; H[i] = perm[H[i]]; }
  }
  bcmi_graph* g = csr_from_arcs(n, T.data(), H.data(), k, 1, 0);
  if (!g) return set_error(BCMI_ERR_NOMEM, "out of host memory");
  *out = g;
  return BCMI_OK;
}

// Pendant trees: trees of depth D ~ U{1..4} hang at uniformly chosen core
// vertices.  A tree of depth D has a chain of D-1 internal vertices and
// max(1, 2(D-1)) leaves attached to chain vertices (the deepest chain vertex
// always gets one), so about two thirds of the added vertices are leaves.
// Trees are added until frac * n extra vertices exist; all ids are then
// randomly permuted.
int gen_pendant_trees(const bcmi_graph* core, double frac, uint64_t seed, bcmi_graph** out) {
  if (!core || core->directed) return set_error(BCMI_ERR_ARG, "pendant trees need an undirected core graph");
  int64_t n0 = core->n;
  int64_t extra = (int64_t)(frac * (double)n0 + 0.5);
  Rng r(seed ^ 0xA0761D6478BD642Full);
  std::vector<int32_t> T, H;
  T.reserve(core->m / 2 + extra); H.reserve(core->m / 2 + extra);
  for (int64_t v = 0; v < n0; v++)
    for (int64_t e = core->off[v]; e < core->off[v + 1]; e++)
      if (core->adj[e] > v) { T.push_back((int32_t)v); H.push_back(core->adj[e]); }
  int64_t next = n0, end = n0 + extra;
  while (next < end) {
    int32_t at = (int32_t)r.below((uint64_t)n0);
    int depth = 1 + (int)r.below(4);
    std::vector<int32_t> chain;
    int32_t prev = at;
    for (int i = 0; i < depth - 1 && next < end; i++) {
      int32_t x = (int32_t)next++;
      T.push_back(prev); H.push_back(x);
      chain.push_back(x); prev = x;
    }
    int nleaves = chain.empty() ? 1 : 2 * (int)chain.size();
    for (int i = 0; i < nleaves && next < end; i++) {
      int32_t par = chain.empty() ? at : (i == 0 ? chain.back() : chain[r.below(chain.size())]);
      int32_t x = (int32_t)next++;
      T.push_back(par); H.push_back(x);
    }
  }
  int64_t n = end;
  std::vector<int32_t> perm;
  permutation(n, seed, perm);
  for (size_t i = 0; i < T.size(); i++) { T[i] = perm[T[i]]; H[i] = perm[H[i]]; }
  bcmi_graph* g = csr_from_arcs(n, T.data(), H.data(), (int64_t)T.size(), 1, 0);
  if (!g) return set_error(BCMI_ERR_NOMEM, "out of host memory");
  *out = g;
  return BCMI_OK;
}

// Bridged-block chain: nblocks Erdos-Renyi blocks of `block` vertices with
// round(block * avgdeg / 2) uniformly drawn internal edges each; block i and
// block i+1 are joined by one bridge between uniformly chosen members.
// Vertex ids are block-contiguous.
int gen_block_chain(int64_t nblocks, int block, double avgdeg, uint64_t seed, bcmi_graph** out) {
  if (nblocks < 1 || block < 2 || avgdeg < 0) return set_error(BCMI_ERR_ARG, "bad block chain parameters");
  int64_t n = nblocks * block;
  if (n > INT32_MAX) return set_error(BCMI_ERR_ARG, "block chain too large");
  int64_t per = (int64_t)(block * avgdeg / 2.0 + 0.5);
  int64_t k = nblocks * per + (nblocks - 1);
  std::vector<int32_t> T(k), H(k);
  // Bridge endpoints come from the largest connected part of their block
  // (an Erdos-Renyi block can leave a few vertices isolated), so the chain
  // is one component: entry[bi] joins the previous block, exit[bi] the next.
  std::vector<int32_t> entry(nblocks), exitv(nblocks);
#pragma omp parallel for schedule(static)
  for (int64_t bi = 0; bi < nblocks; bi++) {
    Rng r(splitmix64(seed) ^ splitmix64((uint64_t)bi + 0x1234567ull));
    int64_t base = bi * block;
    std::vector<int32_t> par(block), cnt(block, 0);
    for (int i = 0; i < block; i++) par[i] = i;
    auto find = [&](int x) {
      while (par[x] != x) x = par[x] = par[par[x]];
      return x;
    };
    for (int64_t e = 0; e < per; e++) {
      int a = (int)r.below(block), b = (int)r.below(block);
      T[bi * per + e] = (int32_t)(base + a);
      H[bi * per + e] = (int32_t)(base + b);
      int ra = find(a), rb = find(b);
      if (ra != rb) par[ra] = rb;
    }
    int best = 0;
    for (int i = 0; i < block; i++) {
      int q = find(i);
      if (++cnt[q] > cnt[best]) best = q;
    }
    std::vector<int32_t> big;
    for (int i = 0; i < block; i++)
      if (find(i) == best) big.push_back(i);
    entry[bi] = (int32_t)(base + big[r.below(big.size())]);
    exitv[bi] = (int32_t)(base + big[r.below(big.size())]);
  }
  for (int64_t bi = 0; bi + 1 < nblocks; bi++) {
    T[nblocks * per + bi] = exitv[bi];
    H[nblocks * per + bi] = entry[bi + 1];
  }
  bcmi_graph* g = csr_from_arcs(n, T.data(), H.data(), k, 1, 0);
  if (!g) return set_error(BCMI_ERR_NOMEM, "out of host memory");
  *out = g;
  return BCMI_OK;
}

// ------------------------------------------------------------ CSR cache
// Binary CSR file: a 64-byte header (magic, version, directed, n, m, size
// and mtime of the text file it was built from, checksum), then off[n+1]
// as int64 and adj[m] as int32, little-endian.  Loading checks the header,
// the file size, the checksum and the CSR invariants.
static const char kCsrMagic[8] = {'B', 'C', 'M', 'I', 'C', 'S', 'R', '1'};
struct CsrHeader {
  char magic[8];
  uint32_t version, directed;
  int64_t n, m;
  uint64_t src_size;
  int64_t src_mtime_ns;
  uint64_t checksum;
  uint64_t reserved;
};
static_assert(sizeof(CsrHeader) == 64, "header layout");

// Order-dependent hash of a byte range: FNV-1a per 1 MiB chunk, chunks mixed
// with their index (parallel over chunks).
static uint64_t range_hash(const unsigned char* p, size_t len) {
  const size_t C = 1 << 20;
  const int64_t nc = (int64_t)((len + C - 1) / C);
  uint64_t acc = 0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
  for (int64_t c = 0; c < nc; c++) {
    uint64_t h = 1469598103934665603ull;
    const size_t lo = (size_t)c * C, hi = std::min(len, lo + C);
    for (size_t i = lo; i < hi; i++) h = (h ^ p[i]) * 1099511628211ull;
    acc += splitmix64(h ^ (uint64_t)c);
  }
  return acc;
}

static uint64_t csr_checksum(const bcmi_graph* g) {
  return splitmix64(range_hash((const unsigned char*)g->off, sizeof(int64_t) * (g->n + 1)) ^
                    (range_hash((const unsigned char*)g->adj, sizeof(int32_t) * g->m) * 31));
}

int graph_save(const bcmi_graph* g, const char* path, uint64_t src_size, int64_t src_mtime_ns) {
  CsrHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kCsrMagic, 8);
  h.version = 1;
  h.directed = (uint32_t)g->directed;
  h.n = g->n;
  h.m = g->m;
  h.src_size = src_size;
  h.src_mtime_ns = src_mtime_ns;
  h.checksum = csr_checksum(g);
  std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return set_error(BCMI_ERR_IO, std::string("cannot write ") + tmp + ": " + strerror(errno));
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 &&
            fwrite(g->off, sizeof(int64_t), (size_t)g->n + 1, f) == (size_t)g->n + 1 &&
            (g->m == 0 || fwrite(g->adj, sizeof(int32_t), (size_t)g->m, f) == (size_t)g->m);
  ok = (fclose(f) == 0) && ok;
  if (!ok || rename(tmp.c_str(), path) != 0) {
    unlink(tmp.c_str());
    return set_error(BCMI_ERR_IO, std::string("writing ") + path + " failed");
  }
  return BCMI_OK;
}

int graph_load(const char* path, bcmi_graph** out, uint64_t* src_size, int64_t* src_mtime_ns) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return set_error(BCMI_ERR_IO, std::string("cannot open ") + path + ": " + strerror(errno));
  struct stat st;
  if (fstat(fd, &st) != 0 || (size_t)st.st_size < sizeof(CsrHeader)) {
    close(fd);
    return set_error(BCMI_ERR_IO, std::string(path) + ": not a CSR cache (too short)");
  }
  CsrHeader h;
  if (pread(fd, &h, sizeof h, 0) != (ssize_t)sizeof h || memcmp(h.magic, kCsrMagic, 8) != 0 || h.version != 1) {
    close(fd);
    return set_error(BCMI_ERR_IO, std::string(path) + ": not a CSR cache (bad magic or version)");
  }
  const uint64_t want = sizeof h + sizeof(int64_t) * (uint64_t)(h.n + 1) + sizeof(int32_t) * (uint64_t)h.m;
  if (h.n < 1 || h.n > INT32_MAX || h.m < 0 || (uint64_t)st.st_size != want) {
    close(fd);
    return set_error(BCMI_ERR_IO, std::string(path) + ": size does not match its header");
  }
  bcmi_graph* g = graph_alloc(h.n, h.m, (int)h.directed);
  if (!g) {
    close(fd);
    return set_error(BCMI_ERR_NOMEM, "out of host memory");
  }
  auto rd = [&](void* dst, size_t bytes, off_t at) {
    size_t got = 0;
    while (got < bytes) {
      ssize_t r = pread(fd, (char*)dst + got, bytes - got, at + (off_t)got);
      if (r <= 0) return false;
      got += (size_t)r;
    }
    return true;
  };
  bool ok = rd(g->off, sizeof(int64_t) * (size_t)(h.n + 1), sizeof h) &&
            rd(g->adj, sizeof(int32_t) * (size_t)h.m, (off_t)(sizeof h + sizeof(int64_t) * (size_t)(h.n + 1)));
  close(fd);
  if (!ok) {
    bcmi_graph_free(g);
    return set_error(BCMI_ERR_IO, std::string(path) + ": read failed");
  }
  if (csr_checksum(g) != h.checksum) {
    bcmi_graph_free(g);
    return set_error(BCMI_ERR_IO, std::string(path) + ": checksum mismatch");
  }
  bool valid = g->off[0] == 0 && g->off[h.n] == h.m;
  for (int64_t v = 0; valid && v < h.n; v++) valid = g->off[v] <= g->off[v + 1];
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (int64_t e = 0; e < h.m; e++) bad += (g->adj[e] < 0 || g->adj[e] >= h.n);
  if (!valid || bad) {
    bcmi_graph_free(g);
    return set_error(BCMI_ERR_IO, std::string(path) + ": invalid CSR");
  }
  if (src_size) *src_size = h.src_size;
  if (src_mtime_ns) *src_mtime_ns = h.src_mtime_ns;
  *out = g;
  return BCMI_OK;
}

// Relabels vertices by descending degree (ties by id): new id i holds old
// vertex n2o[i].  The adjacency of each vertex keeps its original order
// (entries renamed, not re-sorted), so per-vertex sums keep their order.
void relabel_by_degree(int64_t n, const int64_t* off, const int32_t* adj, std::vector<int32_t>& n2o,
                       std::vector<int32_t>& o2n, std::vector<int64_t>& noff, std::vector<int32_t>& nadj) {
  n2o.resize(n);
  o2n.resize(n);
  std::iota(n2o.begin(), n2o.end(), 0);
  std::stable_sort(n2o.begin(), n2o.end(), [&](int32_t a, int32_t b) {
    return off[a + 1] - off[a] > off[b + 1] - off[b];
  });
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) o2n[n2o[i]] = (int32_t)i;
  noff.assign(n + 1, 0);
  for (int64_t i = 0; i < n; i++) noff[i + 1] = noff[i] + (off[n2o[i] + 1] - off[n2o[i]]);
  nadj.resize(noff[n]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; i++) {
    int64_t s = off[n2o[i]], d = noff[i], k = noff[i + 1] - noff[i];
    for (int64_t e = 0; e < k; e++) nadj[d + e] = o2n[adj[s + e]];
  }
}

// Uniform sample of k distinct vertices with degree > 0 (isolated sources
// contribute nothing and the reference skips them, bc2d.c:1026-1035),
// returned in draw order.
int sample_sources(const bcmi_graph* g, int64_t k, uint64_t seed, int32_t* out) {
  std::vector<int32_t> cand;
  cand.reserve(g->n);
  for (int64_t v = 0; v < g->n; v++)
    if (g->off[v + 1] > g->off[v]) cand.push_back((int32_t)v);
  if (k > (int64_t)cand.size()) return set_error(BCMI_ERR_ARG, "more sources requested than non-isolated vertices");
  Rng r(seed ^ 0x2545F4914F6CDD1Dull);
  for (int64_t i = 0; i < k; i++) {  // partial Fisher-Yates
    int64_t j = i + (int64_t)r.below((uint64_t)(cand.size() - i));
    std::swap(cand[i], cand[j]);
    out[i] = cand[i];
  }
  return BCMI_OK;
}

}  // namespace bcmi

// ------------------------------------------------------------ C ABI -----
extern "C" {

int bcmi_graph_read_edgelist(const char* path, uint64_t n, int directed, bcmi_graph** out,
                             uint64_t* bad_line, uint64_t* hdr_nodes, uint64_t* hdr_edges) {
  return bcmi::guard([&]() -> int {
    if (!path || !out) return bcmi::set_error(BCMI_ERR_ARG, "null argument");
    return bcmi::read_edgelist(path, n, directed, out, bad_line, hdr_nodes, hdr_edges);
  });
}

int bcmi_graph_save(const bcmi_graph* g, const char* path, uint64_t src_size, int64_t src_mtime_ns) {
  return bcmi::guard([&]() -> int {
    if (!g || !path) return bcmi::set_error(BCMI_ERR_ARG, "null argument");
    return bcmi::graph_save(g, path, src_size, src_mtime_ns);
  });
}

int bcmi_graph_load(const char* path, bcmi_graph** out, uint64_t* src_size, int64_t* src_mtime_ns) {
  return bcmi::guard([&]() -> int {
    if (!path || !out) return bcmi::set_error(BCMI_ERR_ARG, "null argument");
    return bcmi::graph_load(path, out, src_size, src_mtime_ns);
  });
}

int bcmi_graph_from_edges(const int32_t* tails, const int32_t* heads, int64_t k, int64_t n, int directed,
                          bcmi_graph** out) {
  return bcmi::guard([&]() -> int {
    if ((!tails || !heads) && k > 0) return bcmi::set_error(BCMI_ERR_ARG, "null edge arrays");
    if (!out || n < 1 || n > INT32_MAX) return bcmi::set_error(BCMI_ERR_ARG, "bad vertex count");
    for (int64_t i = 0; i < k; i++)
      if (tails[i] < 0 || tails[i] >= n || heads[i] < 0 || heads[i] >= n)
        return bcmi::set_error(BCMI_ERR_RANGE, "edge endpoint out of range at index " + std::to_string(i));
    bcmi_graph* g = bcmi::csr_from_arcs(n, tails, heads, k, directed ? 0 : 1, directed);
    if (!g) return bcmi::set_error(BCMI_ERR_NOMEM, "out of host memory");
    *out = g;
    return BCMI_OK;
  });
}

int bcmi_gen_rmat(int scale, int edgefactor, double a, double b, double c, uint64_t seed, int permute,
                  bcmi_graph** out) {
  return bcmi::guard([&]() -> int {
    if (!out) return bcmi::set_error(BCMI_ERR_ARG, "null output");
    return bcmi::gen_rmat(scale, edgefactor, a, b, c, seed, permute, out);
  });
}

int bcmi_gen_pendant_trees(const bcmi_graph* core, double frac, uint64_t seed, bcmi_graph** out) {
  return bcmi::guard([&]() -> int {
    if (!out) return bcmi::set_error(BCMI_ERR_ARG, "null output");
    return bcmi::gen_pendant_trees(core, frac, seed, out);
  });
}

int bcmi_gen_block_chain(int64_t nblocks, int block, double avgdeg, uint64_t seed, bcmi_graph** out) {
  return bcmi::guard([&]() -> int {
    if (!out) return bcmi::set_error(BCMI_ERR_ARG, "null output");
    return bcmi::gen_block_chain(nblocks, block, avgdeg, seed, out);
  });
}

int bcmi_sample_sources(const bcmi_graph* g, int64_t k, uint64_t seed, int32_t* out) {
  return bcmi::guard([&]() -> int {
    if (!g || (!out && k > 0) || k < 0) return bcmi::set_error(BCMI_ERR_ARG, "bad argument");
    return bcmi::sample_sources(g, k, seed, out);
  });
}

void bcmi_graph_free(bcmi_graph* g) {
  if (!g) return;
  free(g->off); free(g->adj); free(g);
}

}  // extern "C"
