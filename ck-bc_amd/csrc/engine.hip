// Brandes betweenness centrality on MI355X (gfx950): device kernels and the
// sweep drivers.
//
// Layout in HBM (n vertices, m arcs of the traversed graph):
//   CSR        off int64[n+1], adj int32[m]              (one block, the 1x1 case of
//                                                          the reference's 2D grid, bc2d.h:28-66)
//   plain      depth int32[n], sigma uint64[n], delta f64[n], coef f64[n]
//   batched    seen/lm0/lm1/lm2 uint64[n] (one bit per source of the batch),
//              sigma uint64[n][64], coef f64[n][64]   (source-minor rows: a wave's
//              64 lanes are the 64 sources, one 512-byte row per vertex)
//   work lists chunk records (vertex, chunk index) of every BFS level, appended
//              level after level like the reference's concatenated frontier
//              queues (cuda_kernels.cu:910-913), each covering at most `chunk`
//              arcs of its vertex so that hubs are split across waves.
//
// Reference rules restated (paths relative to the ck-bc root, "cu" =
// program/bc-cuda/cuda_kernels.cu):
//   forward  cu:524-537  first visit fixes the depth; sigma[r] += sigma[c] for
//                        every arc c->r with depth[r] == depth[c]+1
//   backward cu:590-596  delta[c] += (delta[r]+1)/sigma[r]*sigma[c] for every
//                        arc c->r with depth[r] == depth[c]+1, depths D-2..1
//   reset    cu:1089-1119 delta starts at the pruned-subtree weight w(v)
//   update   cu:263-277  bc[v] += delta[v] * multiplicity of the source
//
// Every level is one step of a device-driven loop: a one-thread "decide"
// kernel closes the level's work list and picks top-down or bottom-up from
// device counters; the expansion kernels are persistent (fixed grid, wave
// strided over work items) and return at once when their direction was not
// chosen or the sweep has finished.  A hipGraph holds `steps` such levels
// and is replayed until the control block reports completion, so the host
// neither sizes launches nor reads counters per level (the reference does
// three device-to-host copies and two syncs per level, cu:861-907).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>
#include <system_error>
#include <thread>
#include <unordered_map>

#include "bcmi.h"
#include "bcmi_internal.h"
#include "engine.h"

namespace bcmi {

static thread_local std::string g_err;
// Graph capture must not overlap device allocation or freeing in another
// thread (HIP invalidates the capture), so both take this lock.
static std::recursive_mutex g_alloc_mu;
#define ALLOC_LOCK std::lock_guard<std::recursive_mutex> alloc_lock_(g_alloc_mu)
int set_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
const char* last_error() { return g_err.c_str(); }

#define HIP_TRY(call)                                                                              \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return set_error(e_ == hipErrorOutOfMemory ? BCMI_ERR_NOMEM : BCMI_ERR_HIP,                  \
                       std::string(#call) + ": " + hipGetErrorString(e_));                        \
  } while (0)

typedef unsigned long long u64;

// ------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ u64 rl64(u64 x, int l) {
  unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, l);
  unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), l);
  return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ int rl32(int x, int l) { return __builtin_amdgcn_readlane(x, l); }

// Takes the next U set bits (lanes) of the wave-uniform mask P, in lane
// order, and hands them to body(q, ok).  While P holds at least U bits the
// slots are all valid and need no per-slot test (FULL: a straight run of
// find-first-bit and bit-clear); the last, partial group tests each slot.
template <int U, bool FULL, class Body>
__device__ __forceinline__ void take_lanes(u64& P, Body&& body) {
  int q[U];
  bool ok[U];
#pragma unroll
  for (int t = 0; t < U; t++) {
    if (FULL) {
      ok[t] = true;
      q[t] = __builtin_ctzll(P);
      P &= ~(1ull << q[t]);
    } else {
      ok[t] = P != 0;
      q[t] = ok[t] ? __builtin_ctzll(P) : 0;
      if (ok[t]) P &= P - 1;
    }
  }
  body(q, ok);
}
// Runs body over every set bit of P, U at a time.
template <int U, class Body>
__device__ __forceinline__ void for_lanes(u64 P, Body&& body) {
  while (__popcll(P) >= U) take_lanes<U, true>(P, body);
  if (P) take_lanes<U, false>(P, body);
}

__device__ __forceinline__ int wave_excl_scan(int x, int* total) {
  int lane = lane_id();
  int inc = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  *total = __shfl(inc, 63, 64);
  return inc - x;
}
__device__ __forceinline__ double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ long long wave_sum_ll(long long x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
// Sum with wrap detection (sets *ovf when any partial sum wraps).
__device__ __forceinline__ u64 wave_sum_u64_ovf(u64 x, int* ovf) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    u64 y = __shfl_xor(x, o, 64);
    u64 s = x + y;
    if (s < x) *ovf = 1;
    x = s;
  }
  return x;
}
__device__ __forceinline__ u64 wave_or64(u64 x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ double ld_coherent(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_coherent_u64(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Adjacency lists are streamed once per level: keep them out of L2 so the
// masks and rows gathered at random stay resident.
#ifndef UNR
#define UNR 8      // row gathers in flight per wave in the sigma and backward loops
#endif
#ifndef UNR_BACK
#define UNR_BACK UNR  // row gathers in flight per wave in the backward loop
#endif
#ifndef BACK_PK
#define BACK_PK 1  // small backward records several to a wave
#endif
#define KF_GATE 1  // gate mask probes with the level bitmaps
#define KF_NT 2    // non-temporal adjacency loads
// One bit per vertex: "this vertex has a non-zero mask on this level".
// Three bitmaps mirror the rotating level masks; at 4M vertices each is
// 512 KB, small enough to stay in L2, and gates the 8-byte mask probes.
__device__ __forceinline__ bool lb_test(const unsigned* b, int32_t v) { return (b[v >> 5] >> (v & 31)) & 1u; }
__device__ __forceinline__ void lb_set(unsigned* b, int32_t v) { atomicOr(&b[v >> 5], 1u << (v & 31)); }
__device__ __forceinline__ void lb_clr(unsigned* b, int32_t v) { atomicAnd(&b[v >> 5], ~(1u << (v & 31))); }
__device__ __forceinline__ int nchunks(int64_t deg, int chunk) {
  return deg <= chunk ? 1 : (int)((deg + chunk - 1) / chunk);
}

// Device-side kernel timer on the constant-rate wall clock (s_memrealtime):
// every wave of the timed kernel stamps its entry and exit into its own slot
// (plain stores, no shared counters).  Right after it, k_kt_accum (KT_PARTS
// blocks) reduces slices of the stamps to partial (first entry, last exit)
// pairs and the one-wave k_kt_fold adds last exit - first entry to the total.
#define KT_PARTS 64
struct KT {
  u64 total, launches;
  u64* wstart;  // per wave of the launch
  u64* wend;
  u64 plo[KT_PARTS], phi[KT_PARTS];  // partial first entry / last exit per accumulating block
};
#define KT_MAX_WAVES kTimerMaxWaves
static_assert(KT_PARTS == 64, "k_kt_fold folds the partials in one wave");
__device__ __forceinline__ int64_t kt_wave() {
  return ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
}
__device__ __forceinline__ void kt_enter(KT* kt) {
  if (kt && __lane_id() == 0) kt->wstart[kt_wave()] = (u64)wall_clock64();
}
__device__ __forceinline__ void kt_exit(KT* kt) {
  if (kt && __lane_id() == 0) kt->wend[kt_wave()] = (u64)wall_clock64();
}
// grid: KT_PARTS blocks of 1024 threads
__global__ void __launch_bounds__(1024) k_kt_accum(KT* kt, int64_t nwaves) {
  __shared__ u64 s_lo[16], s_hi[16];
  u64 lo = ~0ull, hi = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwaves; i += (int64_t)gridDim.x * blockDim.x) {
    u64 a = kt->wstart[i], b = kt->wend[i];
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  for (int o = 32; o > 0; o >>= 1) {
    u64 a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if ((threadIdx.x & 63) == 0) { s_lo[threadIdx.x >> 6] = lo; s_hi[threadIdx.x >> 6] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      lo = s_lo[w] < lo ? s_lo[w] : lo;
      hi = s_hi[w] > hi ? s_hi[w] : hi;
    }
    kt->plo[blockIdx.x] = lo;
    kt->phi[blockIdx.x] = hi;
  }
}
// grid: one block of KT_PARTS (= 64) threads, after k_kt_accum
__global__ void __launch_bounds__(64) k_kt_fold(KT* kt) {
  u64 lo = kt->plo[threadIdx.x], hi = kt->phi[threadIdx.x];
  for (int o = 32; o > 0; o >>= 1) {
    u64 a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  if (threadIdx.x == 0) {
    if (hi > lo) kt->total += hi - lo;
    kt->launches += 1;
  }
}

// ------------------------------------------------------------ kernel args
struct GA {
  const int64_t* off;
  const int32_t* adj;
  const int64_t* toff;  // parents of v: v appears in adj of each (transpose; == off/adj if undirected)
  const int32_t* tadj;
  int directed;
  const int32_t* label;
  const int64_t* comp_arcs;
  const int32_t* all_v;
  const int32_t* all_j;
  int64_t nall;
  const int32_t* sm_v;  // vertices of degree 1..PACK_DEG (bottom-up packs PACK_W per wave)
  int64_t nsm;
  const int32_t* lg_v;  // chunk records of the other vertices
  const int32_t* lg_j;
  int64_t nlg;
  const int32_t* split_v;  // vertices held in more than one record
  int64_t nsplit;
  const uint64_t* w;
  int chunk;
  int64_t n;
  int64_t m;  // traversed arcs
  int flags;  // KF_*
};

__device__ __forceinline__ int32_t adj_at(const GA& g, const int32_t* a, int64_t e) {
  return (g.flags & KF_NT) ? __builtin_nontemporal_load(a + e) : a[e];
}

struct PA {
  Ctl* ctl;
  int32_t* depth;
  u64* sigma;
  double* delta;
  double* coef;
  int32_t* done;
  int32_t* rec_v;
  int32_t* rec_j;
  int64_t cap;
  int64_t* rec_lo;
  const int32_t* srcbuf;
  const double* multbuf;
  double* bc;
  int64_t alpha;
  int64_t beta;   // bottom-up only when the frontier holds >= 1/beta of all arcs
};

struct BA {
  Ctl* ctl;
  u64* seen;
  u64* lm;  // 3 arrays of n
  unsigned* lb;  // 3 bitmaps of nw words
  int64_t nw;
  u64* sigma;
  double* coef;
  int32_t* done;
  int32_t* rec_v;
  int32_t* rec_j;
  u64* rec_mask;
  int64_t cap;
  int64_t* rec_lo;
  u64* compmask;
  const int32_t* srcbuf;
  const double* multbuf;
  double* bc;
  u64* st_n;
  u64* st_m;
  u64* st_a;
  double* dbg_delta;
  int64_t alpha;
  int64_t beta;   // bottom-up only when the frontier holds >= 1/beta of all arcs
  int fuse;       // bottom-up levels also sum the path counts of unsplit vertices
  int pack;       // bottom-up packs low-degree vertices PACK_W to a wave
  int64_t gamma;  // top-down levels whose frontier holds < 1/gamma of all arcs push path counts
  int64_t gamma_b;     // backward levels whose successor level holds < 1/gamma_b of their arcs are pushed
  int64_t* lvl_arcs;   // [level] degree sum of the level's vertices
  int64_t n;
  double* lscale;      // [level][64] scale factor m of each level (scaled path counts)
  int64_t lscale_cap;  // levels it can hold
  u64* lvmax;          // [64] running max of the level being built (double bits)
};

// Path-count rows.  Exact sweeps first keep the counts of a batch in narrow
// rows: 16-bit ([v][64] of uint16, 128 bytes, one cache line per gather) or
// 32-bit (256 bytes), as narrow_sigma asks; a batch whose counts pass the
// row's range is redone with the next wider rows, up to 64-bit rows (exact
// integers, or FP64 bits when scaled).  Width codes: 0 64-bit, 1 32-bit, 2
// 16-bit; the kernels are built for each (`NW` is a template constant in
// their loops).
template <int NW> struct RowT { using T = u64; };
template <> struct RowT<1> { using T = uint32_t; };
template <> struct RowT<2> { using T = uint16_t; };
template <int NW>
__device__ __forceinline__ u64 sg_ld(const u64* sg, int64_t i) {
  return (u64)reinterpret_cast<const typename RowT<NW>::T*>(sg)[i];
}
template <int NW>
__device__ __forceinline__ void sg_st(u64* sg, int64_t i, u64 x) {
  reinterpret_cast<typename RowT<NW>::T*>(sg)[i] = (typename RowT<NW>::T)x;
}
// the same with the width known at run time
__device__ __forceinline__ void sg_st_w(u64* sg, int nw, int64_t i, u64 x) {
  if (nw == 2) sg_st<2>(sg, i, x);
  else if (nw == 1) sg_st<1>(sg, i, x);
  else sg_st<0>(sg, i, x);
}
// True on the lanes whose bit is set in the wave-uniform mask m: the mask
// becomes the exec mask as it is, with no per-lane shift and compare.
__device__ __forceinline__ bool in_mask(u64 m) { return __builtin_amdgcn_inverse_ballot_w64(m); }
// One gather of an unrolled group: the lanes of the wave-uniform mask m read
// their element of row `row` (exec = m), the others get 0.  The value keeps
// the row's width, so the load writes its destination register directly and
// the loads of a group stay in flight together until the first use.
template <int NW>
__device__ __forceinline__ typename RowT<NW>::T sg_gather(const u64* sg, u64 m, int32_t row) {
  using T = typename RowT<NW>::T;
  T x = 0;
  if (in_mask(m)) x = reinterpret_cast<const T*>(sg)[(int64_t)row * 64 + lane_id()];
  return x;
}
// Adds a narrow term to a 64-bit count kept as two 32-bit words (lo, hi).
// The term stays narrow up to the add, so the gather before it loads straight
// into its register (a value widened to 64 bits would have to wait for the
// load inside the gather's branch).
__device__ __forceinline__ void add_w(uint32_t& lo, uint32_t& hi, uint32_t x) {
  uint32_t t;
  hi += __builtin_add_overflow(lo, x, &t);
  lo = t;
}
template <int NW>
__device__ __forceinline__ constexpr u64 row_max() {
  return NW == 2 ? 0xffffull : NW == 1 ? 0xffffffffull : ~0ull;
}
// stores an exact count; true when it does not fit the row
template <int NW>
__device__ __forceinline__ bool sg_put(u64* sg, int64_t i, u64 x) {
  sg_st<NW>(sg, i, x);
  return NW > 0 && x > row_max<NW>();
}
// Adds an exact count atomically; true when the row's width overflows.  A
// 16-bit element is added through its 32-bit word (element 2k is the low
// half): a carry out of it lands in the neighbouring element, which does not
// matter, as the batch is then redone with wider rows.
template <int NW>
__device__ __forceinline__ bool sg_add(u64* sg, int64_t i, u64 x) {
  if constexpr (NW == 2) {
    const int sh = (int)(i & 1) * 16;
    const uint32_t o = atomicAdd(reinterpret_cast<uint32_t*>(sg) + (i >> 1), (uint32_t)(x & 0xffffu) << sh);
    return x > 0xffffull || ((o >> sh) & 0xffffu) + (uint32_t)x > 0xffffu;
  } else if constexpr (NW == 1) {
    const uint32_t o = atomicAdd(reinterpret_cast<uint32_t*>(sg) + i, (uint32_t)x);
    return x > 0xffffffffull || o + (uint32_t)x < o;
  } else {
    const u64 o = atomicAdd(sg + i, x);
    return o + x < o;
  }
}

// Work-list appends are staged per wave in LDS and flushed with one global
// atomic per WBUF records, so the shared record counter is not hammered by
// every wave iteration.  The frontier degree and vertex counts are kept in
// registers and added once per wave at kernel end.
#define WBUF 256
#define MAX_WAVES_PER_BLOCK 4

struct Appender {
  int32_t* sv;  // this wave's LDS staging area
  int32_t* sj;
  int fill;     // wave-uniform
  long long arcs, verts;  // per lane
};

__device__ __forceinline__ void app_init(Appender& a, int32_t (*lv)[WBUF], int32_t (*lj)[WBUF]) {
  int w = threadIdx.x >> 6;
  a.sv = lv[w];
  a.sj = lj[w];
  a.fill = 0;
  a.arcs = 0;
  a.verts = 0;
}

__device__ __forceinline__ void lds_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void app_flush(Appender& a, Ctl* c, int32_t* rec_v, int32_t* rec_j, int64_t cap) {
  if (a.fill == 0) return;
  lds_wave_sync();
  const int lane = lane_id();
  u64 base = 0;
  if (lane == 0) base = atomicAdd((u64*)&c->rec_count, (u64)a.fill);
  base = __shfl(base, 0, 64);
  for (int i = lane; i < a.fill; i += 64) {
    int64_t r = (int64_t)base + i;
    if (r < cap) {
      rec_v[r] = a.sv[i];
      rec_j[r] = a.sj[i];
    } else {
      c->overflow = 1;
    }
  }
  lds_wave_sync();
  a.fill = 0;
}

// Appends the vertices of the lanes with `want` set: each gets
// ceil(deg/chunk) records, contiguous.  Wave-wide call.
__device__ __forceinline__ void app_push(Appender& a, bool want, int32_t v, int64_t degv, int chunk, Ctl* c,
                                         int32_t* rec_v, int32_t* rec_j, int64_t cap) {
  int nch = want ? nchunks(degv, chunk) : 0;
  int tot;
  int pre = wave_excl_scan(nch, &tot);
  if (tot == 0) return;
  if (want) { a.arcs += degv; a.verts += 1; }
  if (a.fill + tot > WBUF) app_flush(a, c, rec_v, rec_j, cap);
  if (tot > WBUF) {  // a hub: straight to global memory
    const int lane = lane_id();
    u64 base = 0;
    if (lane == 0) base = atomicAdd((u64*)&c->rec_count, (u64)tot);
    base = __shfl(base, 0, 64);
    for (int k = 0; k < nch; k++) {
      int64_t r = (int64_t)base + pre + k;
      if (r < cap) {
        rec_v[r] = v;
        rec_j[r] = k;
      } else {
        c->overflow = 1;
      }
    }
    return;
  }
  for (int k = 0; k < nch; k++) {
    a.sv[a.fill + pre + k] = v;
    a.sj[a.fill + pre + k] = k;
  }
  a.fill += tot;
}

// End of an expansion kernel: one reservation per block (not per wave) for
// the records still staged in LDS, and one update per block of the frontier
// counters; same-address atomics from every wave serialise, which costs more
// than the work on small levels.  Every wave of the block must call it.
__device__ __forceinline__ void app_finish(Appender& a, Ctl* c, int32_t* rec_v, int32_t* rec_j, int64_t cap) {
  __shared__ int s_fill[MAX_WAVES_PER_BLOCK];
  __shared__ long long s_ar[MAX_WAVES_PER_BLOCK], s_vr[MAX_WAVES_PER_BLOCK];
  __shared__ u64 s_base;
  const int w = threadIdx.x >> 6, lane = lane_id(), nwv = blockDim.x >> 6;
  long long ar = wave_sum_ll(a.arcs), vr = wave_sum_ll(a.verts);
  if (lane == 0) { s_fill[w] = a.fill; s_ar[w] = ar; s_vr[w] = vr; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    long long A = 0, V = 0;
    for (int i = 0; i < nwv; i++) {
      int f = s_fill[i];
      s_fill[i] = tot;
      tot += f;
      A += s_ar[i];
      V += s_vr[i];
    }
    s_base = tot ? atomicAdd((u64*)&c->rec_count, (u64)tot) : 0;
    if (A) atomicAdd((u64*)&c->arcs_next, (u64)A);
    if (V) atomicAdd((u64*)&c->verts_next, (u64)V);
  }
  __syncthreads();
  const u64 base = s_base + (u64)s_fill[w];
  for (int i = lane; i < a.fill; i += 64) {
    int64_t r = (int64_t)base + i;
    if (r < cap) {
      rec_v[r] = a.sv[i];
      rec_j[r] = a.sj[i];
    } else {
      c->overflow = 1;
    }
  }
  a.fill = 0;
}

// Adds a per-lane value (lane i of every wave -> dst[i]) with one atomic per
// lane per block.  Every wave of the block must call it.
__device__ __forceinline__ void block_add_lanes(u64 v, u64* dst) {
  __shared__ u64 s_v[MAX_WAVES_PER_BLOCK][64];
  const int w = threadIdx.x >> 6, lane = lane_id(), nwv = blockDim.x >> 6;
  s_v[w][lane] = v;
  __syncthreads();
  if (w == 0) {
    u64 t = 0;
    for (int i = 0; i < nwv; i++) t += s_v[i][lane];
    if (t) atomicAdd(dst + lane, t);
  }
  __syncthreads();
}

// Per-lane maximum of non-negative doubles given as their bit patterns (which
// order like the values), one atomic per lane per block.
__device__ __forceinline__ void block_max_lanes(u64 v, u64* dst) {
  __shared__ u64 s_m[MAX_WAVES_PER_BLOCK][64];
  const int w = threadIdx.x >> 6, lane = lane_id(), nwv = blockDim.x >> 6;
  s_m[w][lane] = v;
  __syncthreads();
  if (w == 0) {
    u64 t = 0;
    for (int i = 0; i < nwv; i++) t = s_m[i][lane] > t ? s_m[i][lane] : t;
    if (t) atomicMax(dst + lane, t);
  }
  __syncthreads();
}

// Adds a wave-uniform value once per block.  Every wave must call it.
__device__ __forceinline__ void block_add_uniform(long long v, u64* dst) {
  __shared__ long long s_u[MAX_WAVES_PER_BLOCK];
  const int w = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  if (lane_id() == 0) s_u[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int i = 0; i < nwv; i++) t += s_u[i];
    if (t) atomicAdd(dst, (u64)t);
  }
  __syncthreads();
}

#define APPENDER(a)                                        \
  __shared__ int32_t app_lv_[MAX_WAVES_PER_BLOCK][WBUF];   \
  __shared__ int32_t app_lj_[MAX_WAVES_PER_BLOCK][WBUF];   \
  Appender a;                                              \
  app_init(a, app_lv_, app_lj_)

// ============================================================ plain sweep
// Records of small vertices (all arcs in one record of at most PLAIN_PK) run
// 64 / PLAIN_PK to a wave iteration, one sub-group of PLAIN_PK lanes each, so
// that a vertex of a few arcs does not hold a whole wave through its chain
// of dependent loads.
#ifndef PLAIN_PK
#define PLAIN_PK 16
#endif
static_assert(64 % PLAIN_PK == 0, "sub-groups tile the wave");
// The record (header lane) of this lane's sub-group: sub-group k takes the
// k-th of the next 64 / PLAIN_PK set bits of the wave-uniform mask P, which
// loses them; -1 when P runs out.
__device__ __forceinline__ int take_sub(u64& P) {
  const int k = lane_id() / PLAIN_PK;
  int q = -1;
#pragma unroll
  for (int i = 0; i < 64 / PLAIN_PK; i++) {
    const int b = P ? __builtin_ctzll(P) : -1;
    P &= P - 1;
    if (k == i) q = b;
  }
  return q;
}
// Sum over a sub-group: the last steps of wave_sum_d's butterfly, so a record
// whose terms sit on lanes 0 .. PLAIN_PK-1 gets the same additions in the
// same order either way.
__device__ __forceinline__ double sub_sum_d(double x) {
#pragma unroll
  for (int o = PLAIN_PK / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ long long sub_sum_ll(long long x) {
#pragma unroll
  for (int o = PLAIN_PK / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ u64 sub_sum_u64_ovf(u64 x, int* ovf) {
#pragma unroll
  for (int o = PLAIN_PK / 2; o > 0; o >>= 1) {
    const u64 y = __shfl_xor(x, o, 64);
    const u64 t = x + y;
    if (t < x) *ovf = 1;
    x = t;
  }
  return x;
}

__global__ void k_plain_init(GA g, PA p) {
  Ctl* c = p.ctl;
  int32_t s = p.srcbuf[0];
  int64_t deg = g.off[s + 1] - g.off[s];
  int nch = nchunks(deg, g.chunk);
  c->level = 0; c->started = 0; c->fdone = 0; c->last = 0;
  c->blevel = 0; c->bstarted = 0; c->bdone = 0; c->mode = 0;
  c->overflow = 0;
  c->src_label = g.label[s];
  c->rec_count = nch;
  c->arcs_next = deg; c->verts_next = 1; c->arcs_cur = 0; c->verts_cur = 0;
  c->arcs_seen = 0;
  c->arcs_reach = g.comp_arcs[g.label[s]];
  c->sum_a = 0; c->n_td = 0; c->n_bu = 0;
  c->tot_verts = 0; c->tot_arcs = 0; c->uni_verts = 1; c->uni_arcs = deg;
  c->sigma_ovf = 0;
  c->scaled = 0;
  c->cnt_back = 0;
  p.rec_lo[0] = 0;
  for (int k = 0; k < nch && k < p.cap; k++) { p.rec_v[k] = s; p.rec_j[k] = k; }
  if (nch > p.cap) c->overflow = 1;
  p.depth[s] = 0;
  p.sigma[s] = 1;
  p.delta[s] = g.w ? (double)g.w[s] : 0.0;
}

__device__ __forceinline__ bool decide_common(Ctl* c, int64_t* rec_lo, int64_t alpha) {
  if (c->fdone) return false;
  if (c->started) c->level++;
  c->started = 1;
  int d = c->level;
  rec_lo[d + 1] = c->rec_count;
  // stop early when the work list overflowed (retried with more room) or an
  // exact path count wrapped (the sweep is discarded or redone scaled)
  if (rec_lo[d + 1] == rec_lo[d] || c->overflow || (c->sigma_ovf && !c->scaled)) {
    c->fdone = 1;
    c->last = d - 1;
    return false;
  }
  c->arcs_cur = c->arcs_next;
  c->verts_cur = c->verts_next;
  c->tot_arcs += c->arcs_cur;
  c->tot_verts += c->verts_cur;
  c->arcs_next = 0;
  c->verts_next = 0;
  return true;
}

__global__ void k_plain_decide(GA g, PA p, int allow_bu) {
  Ctl* c = p.ctl;
  if (!decide_common(c, p.rec_lo, p.alpha)) return;
  // Top-down reads the frontier's arcs; bottom-up reads the arcs of every
  // unvisited vertex of the component (no early exit: BC needs all parents).
  c->arcs_seen += c->arcs_cur;
  int64_t unvisited = c->arcs_reach - c->arcs_seen;
  // bottom-up scans every vertex's records, so a small frontier stays top-down
  bool bu = allow_bu && !g.directed && (double)c->arcs_cur * 1000.0 > (double)p.alpha * (double)unvisited &&
            (double)c->arcs_cur * (double)p.beta >= (double)g.m;
  c->mode = bu ? 1 : 0;
  if (bu) c->n_bu++; else c->n_td++;
}

// Top-down: every arc u->v out of level d; first visitor of v claims it with
// a CAS on depth and appends it; sigma[v] += sigma[u] when depth[v] == d+1.
__device__ __forceinline__ void k_plain_td_body(GA g, PA p) {
  Ctl* c = p.ctl;
  if (c->fdone || c->mode != 0) return;
  const int d = c->level;
  const int64_t lo = p.rec_lo[d], hi = p.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  long long acnt = 0;
  int ovf = 0;
  APPENDER(ap);
  // This wave's records lo + gw, lo + gw + tw, ...: the headers of 64 of them
  // load at once, one lane each, then the records run in order.
  for (int64_t k0 = 0; lo + gw + k0 * tw < hi; k0 += 64) {
    const int64_t rq = lo + gw + (k0 + lane) * tw;
    int32_t uq = 0, jq = 0;
    int64_t o0q = 0, o1q = 0;
    u64 suq = 0;
    if (rq < hi) {
      uq = p.rec_v[rq];
      jq = p.rec_j[rq];
      o0q = g.off[uq];
      o1q = g.off[uq + 1];
      suq = p.sigma[uq];
    }
    // small frontier vertices first, 64 / PLAIN_PK at a time
    u64 P = __ballot(rq < hi && o1q - o0q <= PLAIN_PK);
    u64 R = __ballot(rq < hi) & ~P;
    while (P) {
      const int q = take_sub(P);
      const int t = lane & (PLAIN_PK - 1), qs = q < 0 ? 0 : q;
      const long long o0 = __shfl((long long)o0q, qs, 64);
      const int dg = __shfl((int)(o1q - o0q), qs, 64);
      const u64 su = __shfl((unsigned long long)suq, qs, 64);
      bool won = false;
      int64_t dgv = 0;
      int32_t v = 0;
      if (q >= 0 && t < dg) {
        v = adj_at(g, g.adj, o0 + t);
        int dv = p.depth[v];
        if (dv == -1) {
          const int old = atomicCAS(&p.depth[v], -1, d + 1);
          won = old == -1;
          dv = won ? d + 1 : old;
        }
        if (dv == d + 1) {
          const u64 o = atomicAdd(&p.sigma[v], su);
          if (o + su < o) ovf = 1;
          acnt++;
        }
        if (won) {
          p.delta[v] = g.w ? (double)g.w[v] : 0.0;
          dgv = g.off[v + 1] - g.off[v];
        }
      }
      app_push(ap, won, v, dgv, g.chunk, c, p.rec_v, p.rec_j, p.cap);
    }
    while (R) {
      const int q = __ffsll((long long)R) - 1;
      R &= R - 1;
      const int j = rl32(jq, q);
      const u64 su = rl64(suq, q);
      const int64_t e0 = (int64_t)rl64((u64)o0q, q) + (int64_t)j * g.chunk;
      const int64_t e1 = min((int64_t)rl64((u64)o1q, q), e0 + g.chunk);
      for (int64_t eb = e0; eb < e1; eb += 64) {
        int64_t e = eb + lane;
        bool valid = e < e1;
        int32_t v = valid ? adj_at(g, g.adj, e) : 0;
        bool won = false;
        int64_t dg = 0;
        if (valid) {
          int dv = p.depth[v];
          if (dv == -1) {
            int old = atomicCAS(&p.depth[v], -1, d + 1);
            won = (old == -1);
            dv = won ? d + 1 : old;
          }
          if (dv == d + 1) {
            u64 o = atomicAdd(&p.sigma[v], su);
            if (o + su < o) ovf = 1;
            acnt++;
          }
          if (won) {
            p.delta[v] = g.w ? (double)g.w[v] : 0.0;
            dg = g.off[v + 1] - g.off[v];
          }
        }
        app_push(ap, won, v, dg, g.chunk, c, p.rec_v, p.rec_j, p.cap);
      }
    }
  }
  app_finish(ap, c, p.rec_v, p.rec_j, p.cap);
  acnt = wave_sum_ll(acnt);
  if (lane == 0 && acnt) atomicAdd((u64*)&c->sum_a, (u64)acnt);
  if (ovf) c->sigma_ovf = 1;
}

__global__ void __launch_bounds__(256) k_plain_td(GA g, PA p, KT* kt) {
  kt_enter(kt);
  k_plain_td_body(g, p);
  kt_exit(kt);
}

// Bottom-up: every unvisited vertex v of the source's component sums sigma
// over its neighbours at depth d (all of them: no early exit).
__device__ __forceinline__ void k_plain_bu_body(GA g, PA p) {
  Ctl* c = p.ctl;
  if (c->fdone || c->mode != 1) return;
  const int d = c->level;
  const int lab = c->src_label;
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  long long acnt = 0;
  int ovf = 0;
  APPENDER(ap);
  // This wave's records are gw, gw + tw, ...: 64 at a time are tested (one
  // lane each) for a vertex of the source's component that is unvisited or
  // on level d+1 (depths only move from -1 to d+1 during this kernel), and
  // only those are scanned, in the same order.
  for (int64_t k0 = 0; gw + k0 * tw < g.nall; k0 += 64) {
    const int64_t rq = gw + (k0 + lane) * tw;
    int32_t vq = 0, jq = 0;
    bool open = false;
    if (rq < g.nall) {
      vq = g.all_v[rq];
      jq = g.all_j[rq];
      if (g.label[vq] == lab) {
        const int dv = p.depth[vq];
        open = dv == -1 || dv == d + 1;
      }
    }
    // small open vertices (one record of at most PLAIN_PK arcs, so unvisited)
    // first, 64 / PLAIN_PK at a time
    int64_t o0q = 0, o1q = 0;
    if (open) {
      o0q = g.off[vq];
      o1q = g.off[vq + 1];
    }
    u64 P = __ballot(open && o1q - o0q <= PLAIN_PK);
    u64 O = __ballot(open) & ~P;
    while (P) {
      const int q = take_sub(P);
      const int t = lane & (PLAIN_PK - 1), qs = q < 0 ? 0 : q;
      const int32_t v = __shfl(vq, qs, 64);
      const long long o0 = __shfl((long long)o0q, qs, 64);
      const int dg = __shfl((int)(o1q - o0q), qs, 64);
      u64 acc = 0;
      long long cn = 0;
      if (q >= 0 && t < dg) {
        const int32_t u = adj_at(g, g.adj, o0 + t);
        if (p.depth[u] == d) {
          acc = p.sigma[u];
          cn = 1;
        }
      }
      acc = sub_sum_u64_ovf(acc, &ovf);
      cn = sub_sum_ll(cn);
      bool won = false;
      if (q >= 0 && t == 0 && cn > 0) {
        // v's only record: the sole writer of its depth and path count in this step
        p.depth[v] = d + 1;
        p.sigma[v] = acc;
        p.delta[v] = g.w ? (double)g.w[v] : 0.0;
        won = true;
        acnt += cn;
      }
      app_push(ap, won, v, won ? dg : 0, g.chunk, c, p.rec_v, p.rec_j, p.cap);
    }
    while (O) {
      const int oq = __ffsll((long long)O) - 1;
      O &= O - 1;
      const int32_t v = rl32(vq, oq);
      const int j = rl32(jq, oq);
      int64_t e0 = g.off[v] + (int64_t)j * g.chunk;
      int64_t e1 = min(g.off[v + 1], e0 + g.chunk);
      u64 acc = 0;
      int cnt = 0;
      for (int64_t eb = e0; eb < e1; eb += 64) {
        int64_t e = eb + lane;
        if (e < e1) {
          int32_t u = adj_at(g, g.adj, e);
          if (p.depth[u] == d) {
            u64 su = p.sigma[u];
            u64 s = acc + su;
            if (s < acc) ovf = 1;
            acc = s;
            cnt++;
          }
        }
      }
      acc = wave_sum_u64_ovf(acc, &ovf);
      cnt = (int)wave_sum_ll(cnt);
      bool won = false;
      if (cnt > 0 && lane == 0) {
        if (g.off[v + 1] - g.off[v] <= g.chunk) {
          // v's only record: the sole writer of its depth and path count in
          // this step (unvisited: depth -1, count 0), so plain stores claim it
          p.depth[v] = d + 1;
          p.sigma[v] = acc;
          won = true;
        } else {
          int old = atomicCAS(&p.depth[v], -1, d + 1);
          won = (old == -1);
          u64 o = atomicAdd(&p.sigma[v], acc);
          if (o + acc < o) ovf = 1;
        }
        if (won) p.delta[v] = g.w ? (double)g.w[v] : 0.0;
        acnt += cnt;
      }
      app_push(ap, won, v, won ? g.off[v + 1] - g.off[v] : 0, g.chunk, c, p.rec_v, p.rec_j, p.cap);
    }
  }
  app_finish(ap, c, p.rec_v, p.rec_j, p.cap);
  acnt = wave_sum_ll(acnt);
  if (lane == 0 && acnt) atomicAdd((u64*)&c->sum_a, (u64)acnt);
  if (ovf) c->sigma_ovf = 1;
}

__global__ void __launch_bounds__(256) k_plain_bu(GA g, PA p, KT* kt) {
  kt_enter(kt);
  k_plain_bu_body(g, p);
  kt_exit(kt);
}

// Backward levels run from the deepest one (its vertices have no successors,
// so delta = w(v)) up to depth 1; the source's own delta is never needed.
__global__ void k_bdecide(Ctl* c) {
  if (!c->fdone || c->bdone) return;
  // A work-list overflow is retried; an exact path count that wrapped makes
  // the sweep invalid, so its backward pass (and BC update) is skipped.
  if (c->overflow || (c->sigma_ovf && !c->scaled)) {
    c->bdone = 1;
    return;
  }
  if (!c->bstarted) {
    c->blevel = c->last;
    c->bstarted = 1;
  } else {
    c->blevel--;
  }
  if (c->blevel < 1) c->bdone = 1;
}

// Batched backward decide: as k_bdecide, and on undirected graphs a level
// whose successor level holds fewer than 1/gamma_b of its arcs is pushed
// (k_ms_bpush scans the successors' arcs instead of the level's own).
__global__ void k_ms_bdecide(GA g, BA b) {
  Ctl* c = b.ctl;
  if (!c->fdone || c->bdone) return;
  if (c->overflow || (c->sigma_ovf && !c->scaled)) {
    c->bdone = 1;
    return;
  }
  if (!c->bstarted) {
    c->blevel = c->last;
    c->bstarted = 1;
  } else {
    c->blevel--;
  }
  const int d = c->blevel;
  if (d < 1) c->bdone = 1;
  c->bpush = d >= 1 && d < c->last && b.gamma_b > 0 && !g.directed &&
             (double)b.lvl_arcs[d + 1] * (double)b.gamma_b < (double)b.lvl_arcs[d];
}

// Backward level d: delta[v] = w(v) + sum over successors z of
// (delta[z]+1)/sigma[z]*sigma[v]; coef[z] = (delta[z]+1)/sigma[z] was stored
// when z was finished.  Vertices split over several records combine their
// partial sums with FP64 atomics; the last record finishes the vertex.
__device__ __forceinline__ void k_plain_back_body(GA g, PA p) {
  Ctl* c = p.ctl;
  if (!c->fdone || c->bdone) return;
  const int d = c->blevel;
  const int64_t lo = p.rec_lo[d], hi = p.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  const double mult = p.multbuf[0];
  // a sweep whose forward state was not swept (2-degree source) counts its
  // shortest-path DAG arcs here: every one of them is a successor arc of a
  // level >= 1, except the source's own
  const bool cnt = c->cnt_back;
  long long acnt = 0;
  // headers of 64 of this wave's records at once, as in k_plain_td
  for (int64_t k0 = 0; lo + gw + k0 * tw < hi; k0 += 64) {
    const int64_t rq = lo + gw + (k0 + lane) * tw;
    int32_t vq = 0, jq = 0;
    int64_t o0q = 0, o1q = 0;
    u64 sq = 0;
    if (rq < hi) {
      vq = p.rec_v[rq];
      jq = p.rec_j[rq];
      o0q = g.off[vq];
      o1q = g.off[vq + 1];
      sq = p.sigma[vq];
    }
    // small vertices first, 64 / PLAIN_PK at a time
    u64 P = __ballot(rq < hi && o1q - o0q <= PLAIN_PK);
    u64 R = __ballot(rq < hi) & ~P;
    while (P) {
      const int q = take_sub(P);
      const int t = lane & (PLAIN_PK - 1), qs = q < 0 ? 0 : q;
      const int32_t v = __shfl(vq, qs, 64);
      const long long o0 = __shfl((long long)o0q, qs, 64);
      const int dg = __shfl((int)(o1q - o0q), qs, 64);
      const double sv = (double)__shfl((unsigned long long)sq, qs, 64);
      double acc = 0.0;
      if (q >= 0 && t < dg) {
        const int32_t z = adj_at(g, g.adj, o0 + t);
        if (p.depth[z] == d + 1) {
          acc = acc + p.coef[z] * sv;
          acnt++;
        }
      }
      acc = sub_sum_d(acc);
      if (q >= 0 && t == 0) {
        const double dl = p.delta[v] + acc;
        p.delta[v] = dl;
        p.coef[v] = (dl + 1.0) / sv;
        p.bc[v] += mult * dl;
      }
    }
    while (R) {
      const int q = __ffsll((long long)R) - 1;
      R &= R - 1;
      const int32_t v = rl32(vq, q);
      const int j = rl32(jq, q);
      const int64_t o0 = (int64_t)rl64((u64)o0q, q), o1 = (int64_t)rl64((u64)o1q, q);
      const int nch = nchunks(o1 - o0, g.chunk);
      const int64_t e0 = o0 + (int64_t)j * g.chunk;
      const int64_t e1 = min(o1, e0 + g.chunk);
      const double sv = (double)rl64(sq, q);
      double acc = 0.0;
      for (int64_t eb = e0; eb < e1; eb += 64) {
        int64_t e = eb + lane;
        if (e < e1) {
          int32_t z = adj_at(g, g.adj, e);
          if (p.depth[z] == d + 1) {
            acc = acc + p.coef[z] * sv;
            acnt++;
          }
        }
      }
      acc = wave_sum_d(acc);
      if (lane == 0) {
        if (nch == 1) {
          double dl = p.delta[v] + acc;
          p.delta[v] = dl;
          p.coef[v] = (dl + 1.0) / sv;
          p.bc[v] += mult * dl;
        } else {
          atomicAdd(&p.delta[v], acc);  // finished by k_plain_bfin
        }
      }
    }
  }
  if (cnt) {
    acnt = wave_sum_ll(acnt);
    if (lane_id() == 0 && acnt) atomicAdd((u64*)&c->sum_a, (u64)acnt);
  }
}

__global__ void __launch_bounds__(256) k_plain_back(GA g, PA p, KT* kt) {
  kt_enter(kt);
  k_plain_back_body(g, p);
  kt_exit(kt);
}

// Finishes the vertices of backward level d whose arcs were split over
// several records (their partial sums were added atomically).
__global__ void k_plain_bfin(GA g, PA p) {
  Ctl* c = p.ctl;
  if (!c->fdone || c->bdone) return;
  const int d = c->blevel;
  const int64_t lo = p.rec_lo[d], hi = p.rec_lo[d + 1];
  const double mult = p.multbuf[0];
  for (int64_t r = lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < hi; r += (int64_t)gridDim.x * blockDim.x) {
    if (p.rec_j[r] != 0) continue;
    int32_t v = p.rec_v[r];
    int64_t deg = g.off[v + 1] - g.off[v];
    if (deg <= g.chunk) continue;
    double dl = p.delta[v];
    p.coef[v] = (dl + 1.0) / (double)p.sigma[v];
    p.bc[v] += mult * dl;
  }
}

__global__ void k_plain_cleanup(PA p) {
  Ctl* c = p.ctl;
  if (!c->fdone || !c->bdone) return;
  int64_t n = c->rec_count < p.cap ? c->rec_count : p.cap;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (p.rec_j[r] == 0) {
      int32_t v = p.rec_v[r];
      p.depth[v] = -1;
      p.sigma[v] = 0;
    }
  }
}

// ============================================ 2-degree sources (plain sweeps)
// A source v0 of degree 2 whose neighbours a and b are swept too: every
// shortest path out of v0 leaves through a or b, so for t != v0
//   d0(t)     = 1 + min(da(t), db(t))
//   sigma0(t) = [da(t) = d0(t)-1] sigma_a(t) + [db(t) = d0(t)-1] sigma_b(t)
// (a shortest path from a to t through v0 is 2 longer than b's distance, so
// it is never counted on a minimal side).  v0's forward state is one pass
// over the saved depths and path counts of a's and b's sweeps, its level
// lists come from a counting sort on depth, and its backward pass is the
// plain one, which also counts the DAG arcs that no forward pass counted.
// The reference declares this heuristic (setcuda_2degree h:133,
// set_mlp_cuda_2degree h:137, scan_frt_csc_cuda_mono_2degree h:147,
// update_bc_cuda_2degree h:156) and prints "REMOVED" for it (c:1015).
#define D2_LDS_LEVELS 64

struct D2A {
  const int32_t* da;  // neighbour a's saved depths and path counts
  const u64* sa;
  const int32_t* db;  // neighbour b's
  const u64* sb;
  int32_t v0;
  int64_t* lcnt;      // [level] records of the level, then its write cursor
  int64_t nlev;       // levels lcnt holds: no depth reaches it (min of the two sweeps' levels + 2)
  int single;         // v0 has degree 1: a and b are its one neighbour (da == db), its side counted once
};

__global__ void k_d2_init(GA g, PA p, D2A d) {
  Ctl* c = p.ctl;
  c->level = 0; c->started = 1; c->fdone = 0; c->last = 0;
  c->blevel = 0; c->bstarted = 0; c->bdone = 0; c->mode = 0;
  c->overflow = 0; c->sigma_ovf = 0; c->scaled = 0;
  c->cnt_back = 1;
  c->src_label = g.label[d.v0];
  c->rec_count = 0;
  c->arcs_next = 0; c->verts_next = 0; c->arcs_cur = 0; c->verts_cur = 0; c->arcs_seen = 0;
  c->arcs_reach = g.comp_arcs[g.label[d.v0]];
  c->sum_a = d.single ? 1 : 2;  // v0's own arcs, to its neighbours on level 1
  c->n_td = 0; c->n_bu = 0;
  c->tot_verts = 0; c->tot_arcs = 0; c->uni_verts = 0; c->uni_arcs = 0;
  for (int64_t i = 0; i < d.nlev; i++) d.lcnt[i] = 0;
}

// Depths, path counts and delta = w(t) of every vertex v0 reaches, records
// per level (block histograms in LDS), reached vertices and their degree sum.
__global__ void __launch_bounds__(256) k_d2_combine(GA g, PA p, D2A d) {
  __shared__ unsigned long long h[D2_LDS_LEVELS];
  __shared__ int s_max, s_flags;
  for (int i = threadIdx.x; i < D2_LDS_LEVELS; i += blockDim.x) h[i] = 0;
  if (threadIdx.x == 0) { s_max = 0; s_flags = 0; }
  __syncthreads();
  Ctl* c = p.ctl;
  long long verts = 0, arcs = 0;
  int ovf = 0, bad = 0, dmax = 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.n; t += (int64_t)gridDim.x * blockDim.x) {
    const int x = d.da[t], y = d.db[t];
    if (x < 0 && y < 0) continue;
    int dep;
    u64 sg;
    if (t == d.v0) {
      dep = 0;
      sg = 1;
    } else {
      const int mn = x < 0 ? y : y < 0 ? x : min(x, y);
      dep = mn + 1;
      sg = x == mn ? d.sa[t] : 0;
      if (y == mn && !d.single) {
        const u64 s2 = sg + d.sb[t];
        if (s2 < sg) ovf = 1;
        sg = s2;
      }
    }
    if (dep + 1 >= d.nlev) {
      bad = 1;
      continue;
    }
    p.depth[t] = dep;
    p.sigma[t] = sg;
    p.delta[t] = g.w ? (double)g.w[t] : 0.0;
    const int64_t deg = g.off[t + 1] - g.off[t];
    const int nch = nchunks(deg, g.chunk);
    verts++;
    arcs += deg;
    dmax = max(dmax, dep);
    if (dep < D2_LDS_LEVELS) atomicAdd(&h[dep], (unsigned long long)nch);
    else atomicAdd((u64*)&d.lcnt[dep], (u64)nch);
  }
  // one global update per block for each total (same-address atomics serialise)
  if (dmax) atomicMax(&s_max, dmax);
  if (ovf | bad) atomicOr(&s_flags, ovf | (bad << 1));
  __syncthreads();
  for (int i = threadIdx.x; i < D2_LDS_LEVELS && i < d.nlev; i += blockDim.x)
    if (h[i]) atomicAdd((u64*)&d.lcnt[i], (u64)h[i]);
  block_add_uniform(wave_sum_ll(verts), (u64*)&c->tot_verts);
  block_add_uniform(wave_sum_ll(arcs), (u64*)&c->tot_arcs);
  if (threadIdx.x == 0) {
    if (s_max) atomicMax(&c->last, s_max);
    if (s_flags & 1) c->sigma_ovf = 1;
    if (s_flags & 2) c->overflow = 1;
  }
}

// Level offsets (rec_lo) and write cursors from the per-level record counts.
__global__ void k_d2_scan(PA p, D2A d) {
  Ctl* c = p.ctl;
  const int last = c->last;
  int64_t acc = 0;
  for (int l = 0; l <= last; l++) {
    const int64_t k = d.lcnt[l];
    p.rec_lo[l] = acc;
    d.lcnt[l] = acc;
    acc += k;
  }
  p.rec_lo[last + 1] = acc;
  c->rec_count = acc;
  if (acc > p.cap) c->overflow = 1;
  c->fdone = 1;
}

// Writes every reached vertex's records into its level.  Each block counts
// its vertices' records per level, reserves one range per level, then hands
// out positions in the range with LDS atomics (the order inside a level does
// not matter to the backward pass).
__global__ void __launch_bounds__(256) k_d2_scatter(GA g, PA p, D2A d) {
  if (p.ctl->overflow) return;
  __shared__ unsigned long long h[D2_LDS_LEVELS];
  for (int i = threadIdx.x; i < D2_LDS_LEVELS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x, ts = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = t0; t < g.n; t += ts) {
    const int dep = p.depth[t];
    if (dep >= 0 && dep < D2_LDS_LEVELS) atomicAdd(&h[dep], (unsigned long long)nchunks(g.off[t + 1] - g.off[t], g.chunk));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < D2_LDS_LEVELS && i < d.nlev; i += blockDim.x)
    if (h[i]) h[i] = atomicAdd((u64*)&d.lcnt[i], (u64)h[i]);
  __syncthreads();
  for (int64_t t = t0; t < g.n; t += ts) {
    const int dep = p.depth[t];
    if (dep < 0) continue;
    const int nch = nchunks(g.off[t + 1] - g.off[t], g.chunk);
    const int64_t pos = dep < D2_LDS_LEVELS ? (int64_t)atomicAdd(&h[dep], (unsigned long long)nch)
                                            : (int64_t)atomicAdd((u64*)&d.lcnt[dep], (u64)nch);
    for (int k = 0; k < nch; k++) {
      p.rec_v[pos + k] = (int32_t)t;
      p.rec_j[pos + k] = k;
    }
  }
}

// Undoes k_d2_combine's stores (a combined state that cannot be used).
__global__ void k_d2_reset(GA g, PA p, D2A d) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.n; t += (int64_t)gridDim.x * blockDim.x) {
    if (d.da[t] < 0 && d.db[t] < 0) continue;
    p.depth[t] = -1;
    p.sigma[t] = 0;
  }
}

// The neighbours of each listed vertex of degree 1 or 2 (degree 1: the one
// neighbour twice).
__global__ void k_d2_nb(const int64_t* off, const int32_t* adj, const int32_t* v, int64_t k, int32_t* out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < k; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = off[v[i]], dg = off[v[i] + 1] - o;
    out[2 * i] = adj[o];
    out[2 * i + 1] = dg > 1 ? adj[o + 1] : adj[o];
  }
}

// ========================================================== batched sweep
// Sources of a batch are the lanes of a wave: bit i of a vertex mask and
// column i of its sigma/coef row belong to source i.  Level lists hold
// (vertex, mask of the sources that reach it at that depth).  Dense masks:
// seen (sources that reached v so far) and three rotating level arrays lm[k]
// with lm[d%3] = mask of sources for which v has depth d, kept exact for the
// levels a kernel reads.
__global__ void k_ms_init(GA g, BA b, int k, int scaled, int narrow) {
  Ctl* c = b.ctl;
  const int i = threadIdx.x;  // one block of 64 threads
  u64* lm0 = b.lm;
  if (i == 0) {
    c->level = 0; c->started = 0; c->fdone = 0; c->last = 0;
    c->blevel = 0; c->bstarted = 0; c->bdone = 0; c->mode = 0;
    c->overflow = 0;
    c->rec_count = 0;
    c->arcs_next = 0; c->verts_next = 0; c->arcs_cur = 0; c->verts_cur = 0;
    c->arcs_seen = 0; c->sum_a = 0; c->n_td = 0; c->n_bu = 0;
    c->tot_verts = 0; c->tot_arcs = 0; c->uni_verts = 0; c->uni_arcs = 0;
    c->sigma_ovf = 0; c->width = k; c->scaled = scaled; c->push = 0; c->bpush = 0;
    c->narrow = scaled ? 0 : narrow;
    b.rec_lo[0] = 0;
  }
  b.st_n[i] = 0; b.st_m[i] = 0; b.st_a[i] = 0;
  if (scaled) {
    b.lvmax[i] = 0;
    b.lscale[i] = 1.0;
  }
  int32_t s = i < k ? b.srcbuf[i] : 0;
  if (i < k) {
    atomicOr(&lm0[s], 1ull << i);
    if (g.flags & KF_GATE) lb_set(b.lb, s);
    atomicOr(&b.seen[s], 1ull << i);
    atomicOr(&b.compmask[g.label[s]], 1ull << i);
    sg_st_w(b.sigma, scaled ? 0 : narrow, (int64_t)s * 64 + i, scaled ? (u64)__double_as_longlong(1.0) : 1ull);
  }
  __syncthreads();
  if (i < k) {
    u64 msk = lm0[s];
    int64_t deg = g.off[s + 1] - g.off[s];
    b.st_n[i] = 1;
    b.st_m[i] = deg;
    if (__ffsll((long long)msk) - 1 == i) {  // first lane holding this vertex appends it
      int nch = nchunks(deg, g.chunk);
      u64 base = atomicAdd((u64*)&c->rec_count, (u64)nch);
      atomicAdd((u64*)&c->arcs_next, (u64)deg);
      atomicAdd((u64*)&c->verts_next, 1ull);
      for (int q = 0; q < nch; q++) {
        if ((int64_t)base + q < b.cap) {
          b.rec_v[base + q] = s;
          b.rec_j[base + q] = q;
        } else {
          c->overflow = 1;
        }
      }
      if ((int64_t)base < b.cap) b.rec_mask[base] = msk;
      atomicAdd((u64*)&c->uni_verts, 1ull);
      atomicAdd((u64*)&c->uni_arcs, (u64)deg);
    }
  }
  __syncthreads();
  // Degree sum of the components holding sources, and of the sources already
  // complete (reached by every source of their component): lane i counts its
  // source's component and vertex if no lower lane holds the same one.
  {
    const int32_t lq = i < k ? g.label[s] : -1;
    bool first = i < k, firstv = i < k;
    for (int t = 0; t < 64; t++) {
      const int32_t lt = __shfl(lq, t, 64), st = __shfl(s, t, 64);
      if (t < i) {
        if (lt == lq) first = false;
        if (st == s) firstv = false;
      }
    }
    long long reach = first ? (long long)g.comp_arcs[lq] : 0;
    long long done = (firstv && b.seen[s] == b.compmask[lq]) ? (long long)(g.off[s + 1] - g.off[s]) : 0;
    reach = wave_sum_ll(reach);
    done = wave_sum_ll(done);
    if (i == 0) {
      c->arcs_reach = reach;
      c->arcs_seen = done;
    }
  }
}

__global__ void k_ms_decide(GA g, BA b, int allow_bu) {
  Ctl* c = b.ctl;
  if (!decide_common(c, b.rec_lo, b.alpha)) return;
  b.lvl_arcs[c->level] = c->arcs_cur;
  if (c->scaled) {
    // Scaled path counts: level d holds raw(v) = sigma(v)/S(d-1); its factor
    // m(d) (the largest raw value, 1 if none) gives S(d) = S(d-1) m(d).
    const int d = c->level;
    if (d >= b.lscale_cap) {
      c->overflow = 2;
      c->fdone = 1;
      c->last = d - 1;
      return;
    }
    for (int q = 0; q < 64; q++) {
      double mx = __longlong_as_double((long long)b.lvmax[q]);
      b.lscale[(int64_t)d * 64 + q] = mx > 0.0 ? mx : 1.0;
      b.lvmax[q] = 0;
    }
  }
  int64_t incomplete = c->arcs_reach - c->arcs_seen;
  // bottom-up scans the records of every incomplete vertex (and tests all
  // others), so a small frontier stays top-down
  bool bu = allow_bu && !g.directed && (double)c->arcs_cur * 1000.0 > (double)b.alpha * (double)incomplete &&
            (double)c->arcs_cur * (double)b.beta >= (double)g.m;
  c->mode = bu ? 1 : 0;
  if (bu) c->n_bu++; else c->n_td++;
  // Small top-down frontiers push path counts along the arcs they expand
  // (the next level's own arcs, many more, are then not scanned for parents).
  c->push = !bu && b.gamma > 0 && (double)c->arcs_cur * (double)b.gamma < (double)g.m;
}

// Top-down discovery: for every arc u->v out of level d, the sources of u
// that have not reached v are OR-ed into the next-level mask of v; the lane
// that turns that mask non-zero appends v.
__device__ __forceinline__ void k_ms_td_body(GA g, BA b) {
  Ctl* c = b.ctl;
  if (c->fdone || c->mode != 0) return;
  const int d = c->level;
  u64* nxt = b.lm + (int64_t)((d + 1) % 3) * b.n;
  unsigned* nlb = b.lb + (int64_t)((d + 1) % 3) * b.nw;
  const int64_t lo = b.rec_lo[d], hi = b.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  APPENDER(ap);
  for (int64_t r = lo + gw; r < hi; r += tw) {
    int32_t u = b.rec_v[r];
    int j = b.rec_j[r];
    u64 F = b.rec_mask[r - j];
    int64_t e0 = g.off[u] + (int64_t)j * g.chunk;
    int64_t e1 = min(g.off[u + 1], e0 + g.chunk);
    for (int64_t eb = e0; eb < e1; eb += 64) {
      int64_t e = eb + lane;
      bool valid = e < e1;
      int32_t v = valid ? adj_at(g, g.adj, e) : 0;
      bool won = false;
      int64_t dg = 0;
      if (valid) {
        u64 bits = F & ~b.seen[v];
        if (bits && (nxt[v] & bits) != bits) {
          u64 old = atomicOr(&nxt[v], bits);
          won = (old == 0);
          if (won) {
            dg = g.off[v + 1] - g.off[v];
            if (g.flags & KF_GATE) lb_set(nlb, v);
          }
        }
      }
      app_push(ap, won, v, dg, g.chunk, c, b.rec_v, b.rec_j, b.cap);
    }
  }
  app_finish(ap, c, b.rec_v, b.rec_j, b.cap);
}

__global__ void __launch_bounds__(256) k_ms_td(GA g, BA b, KT* kt) {
  kt_enter(kt);
  k_ms_td_body(g, b);
  kt_exit(kt);
}

// Bottom-up discovery: every vertex not yet reached by all sources of its
// component ORs the level-d masks of its neighbours.  With fusion on, each
// record also sums its parents' path counts in the same arc scan (lanes =
// sources, parents in adjacency order, as k_ms_sigma would) and the sigma
// kernel skips its parent scan: a vertex held in one record stores its sum,
// the records of a split vertex (fuse_bu 2) add theirs atomically into its
// row, whose columns are zero until the source reaches it.
#ifndef BU_PIPE
#define BU_PIPE 1  // bottom-up slabs' adjacency and mask loads run ahead of the gathers
#endif
#ifndef BU_WAVES
#define BU_WAVES 6
#endif
#ifndef UNR_BU
#define UNR_BU 12  // parent gathers in flight per wave in the bottom-up kernel
#endif
#ifndef UNR_PK
#define UNR_PK 4  // parent gathers in flight per wave in the packed bottom-up kernel
#endif
#define PACK_DEG 16  // packed bottom-up: vertices of degree <= PACK_DEG ...
#define PACK_W 4     // ... PACK_W to a wave (PACK_DEG * PACK_W = 64)
static_assert(PACK_DEG * PACK_W == 64 && PACK_W == 4, "the packed bottom-up scan picks one vertex per sub-group of 16 lanes");
// Headers of 64 records for each of a block's four waves, held in LDS while
// the wave runs the records (bottom-up and backward kernels): vertex, record index,
// a mask, first and end arc.
struct RecHdr {
  int32_t v[MAX_WAVES_PER_BLOCK][64], j[MAX_WAVES_PER_BLOCK][64];
  u64 L[MAX_WAVES_PER_BLOCK][64];
  int64_t o0[MAX_WAVES_PER_BLOCK][64], o1[MAX_WAVES_PER_BLOCK][64];
};
// LDS of the bottom-up kernels, declared once per kernel (not once per
// template instance): the appender's staging and the record headers.
struct BuLds {
  int32_t app_lv[MAX_WAVES_PER_BLOCK][WBUF];
  int32_t app_lj[MAX_WAVES_PER_BLOCK][WBUF];
  RecHdr hdr;
};
template <int NW, bool PACKED>
__device__ __forceinline__ void k_ms_bu_body(GA g, BA b, BuLds* sh) {
  Ctl* c = b.ctl;
  if (c->fdone || c->mode != 1) return;
  const int d = c->level;
  const u64* cur = b.lm + (int64_t)(d % 3) * b.n;
  u64* nxt = b.lm + (int64_t)((d + 1) % 3) * b.n;
  const unsigned* clb = b.lb + (int64_t)(d % 3) * b.nw;
  unsigned* nlb = b.lb + (int64_t)((d + 1) % 3) * b.nw;
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  const bool fuse = b.fuse;
  const bool scaled = c->scaled;
  const double inv_m = (fuse && scaled) ? 1.0 / b.lscale[(int64_t)d * 64 + lane] : 1.0;
  u64 ca = 0;
  int ovf = 0;
  double mx = 0.0;
  Appender ap;
  app_init(ap, sh->app_lv, sh->app_lj);
  const bool pk = b.pack;
  const int32_t* rv = pk ? g.lg_v : g.all_v;
  const int32_t* rjv = pk ? g.lg_j : g.all_j;
  const int64_t nr = pk ? g.nlg : g.nall;
  // This wave's records are gw, gw + tw, gw + 2 tw, ...: 64 of them at a time
  // are tested for unfinished sources, one lane each, and only the unfinished
  // ones are scanned (in the same order), so a late level where most vertices
  // are finished costs one load chain per 64 records instead of one per record.
  if (!PACKED)
  for (int64_t k0 = 0; gw + k0 * tw < nr; k0 += 64) {
    const int64_t rq = gw + (k0 + lane) * tw;
    int32_t vq = 0, jq = 0;
    u64 svq = 0;
    int64_t o0q = 0, o1q = 0;
    bool open = false;
    if (rq < nr) {
      vq = rv[rq];
      jq = rjv[rq];
      svq = b.seen[vq];
      o0q = g.off[vq];
      o1q = g.off[vq + 1];
      open = svq != b.compmask[g.label[vq]];
    }
    u64 O = __ballot(open);
    // the headers wait in LDS (a wave reads back only its own slots, and a
    // wave's LDS accesses complete in issue order)
    const int wb = threadIdx.x >> 6;
    sh->hdr.v[wb][lane] = vq; sh->hdr.j[wb][lane] = jq; sh->hdr.L[wb][lane] = svq;
    sh->hdr.o0[wb][lane] = o0q; sh->hdr.o1[wb][lane] = o1q;
    __builtin_amdgcn_wave_barrier();
    while (O) {
      const int oq = __ffsll((long long)O) - 1;
      O &= O - 1;
      const int32_t v = __builtin_amdgcn_readfirstlane(sh->hdr.v[wb][oq]);
      const u64 sv = rl64(sh->hdr.L[wb][oq], 0);
      const int j = __builtin_amdgcn_readfirstlane(sh->hdr.j[wb][oq]);
      const int64_t o0 = (int64_t)rl64((u64)sh->hdr.o0[wb][oq], 0), o1 = (int64_t)rl64((u64)sh->hdr.o1[wb][oq], 0);
      const int64_t deg = o1 - o0;
      int64_t e0 = o0 + (int64_t)j * g.chunk;
      int64_t e1 = min(o1, e0 + g.chunk);
      const bool split = deg > g.chunk;
      const bool fr = fuse && (!split || b.fuse > 1);
      u64 acc = 0, sacc = 0;
      double facc = 0.0;
      uint32_t slo = 0, shi = 0;  // 32-bit rows: the exact sum as two words
      uint32_t cnt = 0;  // this record's parents per source (at most its arcs)
#if BU_PIPE
      // The slabs' dependent loads run ahead of the sums: slab k+1's level
      // masks and slab k+2's adjacency are in flight while slab k's parents
      // are gathered.
      int32_t u_c = 0, u_n = 0;
      u64 F_c = 0;
      if (e0 + lane < e1) u_c = g.adj[e0 + lane];
      if (e0 + 64 + lane < e1) u_n = g.adj[e0 + 64 + lane];
      if (e0 + lane < e1) F_c = cur[u_c];
#endif
      for (int64_t eb = e0; eb < e1; eb += 64) {
#if BU_PIPE
        const int32_t u = u_c;
        const u64 F = F_c;
        F_c = 0;
        if (eb + 64 + lane < e1) F_c = cur[u_n];
        u_c = u_n;
        u_n = 0;
        if (eb + 128 + lane < e1) u_n = g.adj[eb + 128 + lane];
        acc |= F;
#else
        int64_t e = eb + lane;
        int32_t u = 0;
        u64 F = 0;
        if (e < e1) {
          u = g.adj[e];
          F = cur[u];
          acc |= F;
        }
#endif
        if (!fr) continue;
        const u64 need = F & ~sv;  // u is a parent of v for these sources
        for_lanes<UNR_BU>(__ballot(need != 0), [&](const int (&q)[UNR_BU], const bool (&okq)[UNR_BU]) {
          typename RowT<NW>::T x[UNR_BU];
#pragma unroll
          for (int t = 0; t < UNR_BU; t++) {
            const int32_t uq = rl32(u, q[t]);
            const u64 nq = okq[t] ? rl64(need, q[t]) : 0ull;
            x[t] = sg_gather<NW>(b.sigma, nq, uq);
            cnt += in_mask(nq);
          }
          // the lanes outside a gather's mask add 0 (exact in both forms;
          // scaled counts always use 64-bit rows)
          if (scaled && NW == 0) {
#pragma unroll
            for (int t = 0; t < UNR_BU; t++) facc = facc + __longlong_as_double((long long)x[t]);
          } else if constexpr (NW > 0) {
            // 32-bit terms: a 64-bit sum of fewer than 2^32 of them cannot wrap
#pragma unroll
            for (int t = 0; t < UNR_BU; t++) add_w(slo, shi, x[t]);
          } else {
#pragma unroll
            for (int t = 0; t < UNR_BU; t++) {
              u64 s2 = sacc + x[t];
              if (s2 < sacc) ovf = 1;
              sacc = s2;
            }
          }
        });
      }
      ca += cnt;
      if (NW > 0) sacc = ((u64)shi << 32) | slo;
      acc = wave_or64(acc) & ~sv;
      bool won = false;
      int64_t dg = 0;
      if (acc && lane == 0) {
        if (!split) {
          // v's only record: no other writer of its next-level mask in this
          // step, and the mask is zero at the step's start (k_ms_sigma cleared
          // the array), so a plain store claims it without a returning atomic
          nxt[v] = acc;
          won = true;
        } else if ((nxt[v] & acc) != acc) {
          u64 old = atomicOr(&nxt[v], acc);
          won = (old == 0);
        }
        if (won) {
          dg = deg;
          if (g.flags & KF_GATE) lb_set(nlb, v);
        }
      }
      if (fr && ((acc >> lane) & 1)) {
        u64* row = b.sigma + (int64_t)v * 64 + lane;
        if (scaled) {
          const double raw = facc * inv_m;
          if (!split) {
            *row = (u64)__double_as_longlong(raw);
            mx = fmax(mx, raw);
          } else {
            // every returned sum is a lower bound of a final one; the last is final
            const double o = atomicAdd((double*)row, raw);
            mx = fmax(mx, o + raw);
          }
        } else if (!split) {
          if (sg_put<NW>(b.sigma, (int64_t)v * 64 + lane, sacc)) ovf = 1;
        } else {
          if (sg_add<NW>(b.sigma, (int64_t)v * 64 + lane, sacc)) ovf = 1;
        }
      }
      app_push(ap, won, v, dg, g.chunk, c, b.rec_v, b.rec_j, b.cap);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (PACKED) {
    // Vertices of degree <= PACK_DEG, PACK_W per wave: sub-group q (lanes
    // 16q..16q+15) probes the arcs of its vertex, one per lane; the path
    // count sums then run with lanes = sources as above, one accumulator per
    // sub-group, the parents of each vertex in adjacency order.
    const int sub = lane / PACK_DEG, sl = lane % PACK_DEG;
    const int64_t npk = (g.nsm + PACK_W - 1) / PACK_W;
    // This wave's groups are gw, gw + tw, ... (PACK_W vertices each): the
    // vertices of 64 / PACK_W groups are tested one lane each, and only the
    // unfinished ones are scanned, PACK_W at a time in index order.
    for (int64_t k0 = 0; gw + k0 * tw < npk; k0 += 64 / PACK_W) {
      const int64_t xq = (gw + (k0 + lane / PACK_W) * tw) * PACK_W + lane % PACK_W;
      int32_t vq = 0;
      u64 svq = 0;
      int64_t o0q = 0, o1q = 0;
      bool open = false;
      if (xq < g.nsm) {
        vq = g.sm_v[xq];
        svq = b.seen[vq];
        o0q = g.off[vq];
        o1q = g.off[vq + 1];
        open = svq != b.compmask[g.label[vq]];
      }
      u64 O = __ballot(open);
      const int wb = threadIdx.x >> 6;
      sh->hdr.v[wb][lane] = vq; sh->hdr.L[wb][lane] = svq;
      sh->hdr.o0[wb][lane] = o0q; sh->hdr.o1[wb][lane] = o1q;
      __builtin_amdgcn_wave_barrier();
      while (O) {
        // the next PACK_W unfinished vertices, one per sub-group (-1: none)
        int pick[PACK_W];
#pragma unroll
        for (int t = 0; t < PACK_W; t++) {
          pick[t] = O ? __ffsll((long long)O) - 1 : -1;
          O &= O - 1;
        }
        const int my = sub == 0 ? pick[0] : sub == 1 ? pick[1] : sub == 2 ? pick[2] : pick[3];
        const int myk = my < 0 ? 0 : my;
        const int32_t v = sh->hdr.v[wb][myk];
        const u64 svs = sh->hdr.L[wb][myk];
        const u64 sv = my < 0 ? 0ull : svs;
        const int64_t o0s = sh->hdr.o0[wb][myk];
        const int64_t o1s = sh->hdr.o1[wb][myk];
        const int64_t e0 = my < 0 ? 0 : o0s, dgv = my < 0 ? 0 : o1s - o0s;
        int32_t u = 0;
        u64 need = 0;
        if (sl < dgv) {
          u = g.adj[e0 + sl];
          need = cur[u] & ~sv;  // u is a parent of v for these sources
        }
        u64 acc = need;
#pragma unroll
        for (int o = PACK_DEG / 2; o > 0; o >>= 1) acc |= __shfl_xor(acc, o, 64);
        u64 sq[PACK_W];  // per sub-group: exact sum, or FP64 bits when scaled
        uint32_t wl[PACK_W], wh[PACK_W];  // 32-bit rows: the exact sums as two words
#pragma unroll
        for (int q = 0; q < PACK_W; q++) sq[q] = 0, wl[q] = 0, wh[q] = 0;
        if (fuse) {
          u64 P = __ballot(need != 0);
          while (P) {
            int qa[UNR_PK];
            bool okq[UNR_PK];
#pragma unroll
            for (int t = 0; t < UNR_PK; t++) {
              okq[t] = P != 0;
              qa[t] = okq[t] ? __ffsll((long long)P) - 1 : 0;
              if (okq[t]) P &= P - 1;
            }
            typename RowT<NW>::T x[UNR_PK];
#pragma unroll
            for (int t = 0; t < UNR_PK; t++) {
              const int32_t uq = rl32(u, qa[t]);
              const u64 nq = okq[t] ? rl64(need, qa[t]) : 0ull;
              x[t] = sg_gather<NW>(b.sigma, nq, uq);
              ca += in_mask(nq);
            }
#pragma unroll
            for (int t = 0; t < UNR_PK; t++) {
              const int gq = qa[t] / PACK_DEG;  // wave-uniform
#pragma unroll
              for (int q = 0; q < PACK_W; q++) {
                if (gq != q || !okq[t]) continue;
                if constexpr (NW > 0) {
                  add_w(wl[q], wh[q], x[t]);  // 32-bit terms cannot wrap a 64-bit sum
                } else if (scaled) {
                  sq[q] = (u64)__double_as_longlong(__longlong_as_double((long long)sq[q]) +
                                                    __longlong_as_double((long long)x[t]));
                } else {
                  const u64 s2 = sq[q] + x[t];
                  if (s2 < sq[q]) ovf = 1;
                  sq[q] = s2;
                }
              }
            }
          }
        }
        if (NW > 0) {
#pragma unroll
          for (int q = 0; q < PACK_W; q++) sq[q] = ((u64)wh[q] << 32) | wl[q];
        }
        bool won = false;
        int64_t dg = 0;
        if (sl == 0 && acc) {
          // degree <= PACK_DEG: v's only record, a plain store claims it (as above)
          nxt[v] = acc;
          won = true;
          dg = dgv;
          if (g.flags & KF_GATE) lb_set(nlb, v);
        }
        if (fuse) {
#pragma unroll
          for (int q = 0; q < PACK_W; q++) {
            const u64 aq = rl64(acc, q * PACK_DEG);
            if (!((aq >> lane) & 1)) continue;
            const int64_t row = (int64_t)rl32(v, q * PACK_DEG) * 64 + lane;
            if (scaled) {
              const double raw = __longlong_as_double((long long)sq[q]) * inv_m;
              b.sigma[row] = (u64)__double_as_longlong(raw);
              mx = fmax(mx, raw);
            } else if (sg_put<NW>(b.sigma, row, sq[q])) {
              ovf = 1;
            }
          }
        }
        app_push(ap, won, v, dg, g.chunk, c, b.rec_v, b.rec_j, b.cap);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  app_finish(ap, c, b.rec_v, b.rec_j, b.cap);
  if (fuse) {
    block_add_lanes(ca, b.st_a);
    if (scaled) block_max_lanes((u64)__double_as_longlong(mx), b.lvmax);
    if (ovf) c->sigma_ovf = 1;
  }
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BU_WAVES))) k_ms_bu(GA g, BA b, KT* kt) {
  kt_enter(kt);
  __shared__ BuLds sh;
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_bu_body<2, false>(g, b, &sh);
  else if (nw == 1) k_ms_bu_body<1, false>(g, b, &sh);
  else k_ms_bu_body<0, false>(g, b, &sh);
  kt_exit(kt);
}

// The packed low-degree vertices, in a kernel of their own (own register budget:
// 6 waves per SIMD, 80 VGPRs and a few spilled outside the gather loop; its
// load chains need the extra resident waves).
#ifndef BU_PK_WAVES
#define BU_PK_WAVES 6
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BU_PK_WAVES))) k_ms_bu_pk(GA g, BA b, KT* kt) {
  kt_enter(kt);
  __shared__ BuLds sh;
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_bu_body<2, true>(g, b, &sh);
  else if (nw == 1) k_ms_bu_body<1, true>(g, b, &sh);
  else k_ms_bu_body<0, true>(g, b, &sh);
  kt_exit(kt);
}

// Path counts of level d+1: lane i of the wave sums sigma[u][i] over the
// neighbours u that are at depth d for source i (the parents).  The first
// record of v publishes its mask, marks it seen and counts statistics.  Also
// clears level d-1 from its rotating array (not read at this step).
template <int NW>
__device__ __forceinline__ void k_ms_sigma_body(GA g, BA b) {
  Ctl* c = b.ctl;
  if (c->fdone) return;
  const int d = c->level;
  const u64* cur = b.lm + (int64_t)(d % 3) * b.n;
  const u64* nxt = b.lm + (int64_t)((d + 1) % 3) * b.n;
  const unsigned* clb = b.lb + (int64_t)(d % 3) * b.nw;
  const int64_t lo = b.rec_lo[d + 1];
  const int64_t hi = min(c->rec_count, b.cap);
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  u64 cn = 0, cm = 0, ca = 0;
  int ovf = 0;
  long long done_arcs = 0, uv = 0, ua = 0;
  const bool scaled = c->scaled;
  const bool push = c->push;  // k_ms_push sums the parents
  const bool fused = c->mode == 1 && b.fuse;  // k_ms_bu summed the parents (of unsplit vertices if fuse 1)
  const double inv_m = scaled ? 1.0 / b.lscale[(int64_t)d * 64 + lane] : 1.0;
  double mx = 0.0;
  // After a bottom-up level that summed every vertex's parents, or before a
  // push (which only needs zeroed rows), no parents are scanned: the
  // bookkeeping runs one lane per record, 64 records per wave.
  const bool book = (push || (fused && b.fuse > 1)) && b.pack;
  if (book) {
    for (int64_t base = lo + gw * 64; base < hi; base += tw * 64) {
      const int64_t r = base + lane;
      int32_t v = 0;
      bool first = false;
      if (r < hi) {
        v = b.rec_v[r];
        first = b.rec_j[r] == 0;
      }
      u64 N = 0;
      int64_t deg = 0;
      if (first) {
        N = nxt[v];
        deg = g.off[v + 1] - g.off[v];
        const u64 so = b.seen[v];
        const u64 sn = so | N;
        b.seen[v] = sn;
        b.rec_mask[r] = N;
        if (so == 0) { uv++; ua += deg; }
        if (sn == b.compmask[g.label[v]]) done_arcs += deg;
      }
      // per-source counts (lanes = sources again) and the backward
      // accumulators of split vertices
      u64 F = __ballot(first);
      while (F) {
        const int q = __ffsll((long long)F) - 1;
        F &= F - 1;
        const u64 Nq = rl64(N, q);
        const int64_t dq = (int64_t)rl64((u64)deg, q);
        if ((Nq >> lane) & 1) {
          cn++;
          cm += dq;
          const int64_t row = (int64_t)rl32(v, q) * 64 + lane;
          if (dq > g.chunk) b.coef[row] = 0.0;
          if (push) sg_st<NW>(b.sigma, row, 0);  // k_ms_push adds the parents' counts next
        }
      }
    }
    done_arcs = wave_sum_ll(done_arcs);
    uv = wave_sum_ll(uv);
    ua = wave_sum_ll(ua);
  } else
  for (int64_t r = lo + gw; r < hi; r += tw) {
    int32_t v = b.rec_v[r];
    int j = b.rec_j[r];
    u64 N = nxt[v];
    int64_t deg = g.off[v + 1] - g.off[v];
    int nch = nchunks(deg, g.chunk);
    int64_t e0, e1;
    const bool done_by_bu = fused && (nch == 1 || b.fuse > 1);
    if (push || done_by_bu) {
      e0 = e1 = 0;  // no parent scan
    } else if (!g.directed) {
      e0 = g.off[v] + (int64_t)j * g.chunk;
      e1 = min(g.off[v + 1], e0 + g.chunk);
    } else {  // parents come from the transposed lists, split evenly over v's records
      int64_t t0 = g.toff[v], tl = g.toff[v + 1] - t0;
      e0 = t0 + tl * j / nch;
      e1 = t0 + tl * (j + 1) / nch;
    }
    const bool mine = (N >> lane) & 1;
    u64 acc = 0;
    uint32_t alo = 0, ahi = 0;  // 32-bit rows: the exact sum as two words
    double accf = 0.0;
    for (int64_t eb = e0; eb < e1; eb += 64) {
      int64_t e = eb + lane;
      int32_t u = 0;
      u64 F = 0;
      if (e < e1) {
        u = adj_at(g, g.tadj, e);
        if (!(g.flags & KF_GATE) || lb_test(clb, u)) F = cur[u] & N;
      }
      // Parents UNR at a time: the row loads are issued before any is
      // consumed, and the sums still follow adjacency order.
      for_lanes<UNR>(__ballot(F != 0), [&](const int (&q)[UNR], const bool (&ok)[UNR]) {
        typename RowT<NW>::T x[UNR];
#pragma unroll
        for (int t = 0; t < UNR; t++) {
          const int32_t uq = rl32(u, q[t]);
          const u64 fq = ok[t] ? rl64(F, q[t]) : 0ull;
          x[t] = sg_gather<NW>(b.sigma, fq, uq);
          ca += in_mask(fq);
        }
        if (scaled && NW == 0) {
#pragma unroll
          for (int t = 0; t < UNR; t++) accf = accf + __longlong_as_double((long long)x[t]);
        } else if constexpr (NW > 0) {
#pragma unroll
          for (int t = 0; t < UNR; t++) add_w(alo, ahi, x[t]);
        } else {
#pragma unroll
          for (int t = 0; t < UNR; t++) {
            u64 s = acc + x[t];
            if (s < acc) ovf = 1;
            acc = s;
          }
        }
      });
    }
    if (NW > 0) acc = ((u64)ahi << 32) | alo;
    u64* row = b.sigma + (int64_t)v * 64 + lane;
    if (done_by_bu) {
      // row written by k_ms_bu
    } else if (push) {
      if (mine && j == 0) sg_st<NW>(b.sigma, (int64_t)v * 64 + lane, 0);  // k_ms_push adds the parents' counts next
    } else if (scaled) {
      double raw = accf * inv_m;
      if (mine) {
        if (nch == 1) *row = (u64)__double_as_longlong(raw);
        else atomicAdd((double*)row, raw);
        mx = fmax(mx, raw);
      }
    } else if (nch == 1) {
      if (mine && sg_put<NW>(b.sigma, (int64_t)v * 64 + lane, acc)) ovf = 1;
    } else if (mine) {
      if (sg_add<NW>(b.sigma, (int64_t)v * 64 + lane, acc)) ovf = 1;
    }
    if (j == 0) {
      u64 so = b.seen[v];
      u64 sn = so | N;
      b.seen[v] = sn;
      if (lane == 0 && so == 0) { uv++; ua += deg; }
      if (lane == 0) b.rec_mask[r] = N;
      if (mine) {
        cn++;
        cm += deg;
        if (nch > 1) b.coef[(int64_t)v * 64 + lane] = 0.0;
      }
      if (lane == 0 && sn == b.compmask[g.label[v]]) done_arcs += deg;
    }
  }
  // clear level d-1 (array (d+2)%3)
  if (d >= 1) {
    u64* old = b.lm + (int64_t)((d + 2) % 3) * b.n;
    unsigned* olb = b.lb + (int64_t)((d + 2) % 3) * b.nw;
    const int64_t plo = b.rec_lo[d - 1], phi = b.rec_lo[d];
    for (int64_t r = plo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < phi;
         r += (int64_t)gridDim.x * blockDim.x)
      if (b.rec_j[r] == 0) {
        old[b.rec_v[r]] = 0;
        if (g.flags & KF_GATE) lb_clr(olb, b.rec_v[r]);
      }
  }
  block_add_lanes(cn, b.st_n);
  block_add_lanes(cm, b.st_m);
  block_add_lanes(ca, b.st_a);
  block_add_uniform(done_arcs, (u64*)&c->arcs_seen);
  block_add_uniform(uv, (u64*)&c->uni_verts);
  block_add_uniform(ua, (u64*)&c->uni_arcs);
  if (ovf) c->sigma_ovf = 1;
  if (scaled) block_max_lanes((u64)__double_as_longlong(mx), b.lvmax);
}

__global__ void __launch_bounds__(256) k_ms_sigma(GA g, BA b, KT* kt) {
  kt_enter(kt);
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_sigma_body<2>(g, b);
  else if (nw == 1) k_ms_sigma_body<1>(g, b);
  else k_ms_sigma_body<0>(g, b);
  kt_exit(kt);
}

// Pushed path counts (levels the decide kernel marked): for every arc u->v
// out of a level-d frontier record, sigma[v][i] += sigma[u][i] for the
// sources i with u on level d and v on level d+1 (rows of level d+1 were
// zeroed by k_ms_sigma).  The frontier's arcs are far fewer than the next
// level's, which the pull in k_ms_sigma would scan.  Exact counts use
// returning atomics to detect a wrap; scaled counts add raw values and keep
// the level maximum (every returned sum is a lower bound of a final one).
template <int NW>
__device__ __forceinline__ void k_ms_push_body(GA g, BA b) {
  Ctl* c = b.ctl;
  if (c->fdone || !c->push) return;
  const int d = c->level;
  const u64* nxt = b.lm + (int64_t)((d + 1) % 3) * b.n;
  const int64_t lo = b.rec_lo[d], hi = b.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  const bool scaled = c->scaled;
  const double inv_m = scaled ? 1.0 / b.lscale[(int64_t)d * 64 + lane] : 1.0;
  u64 ca = 0;
  int ovf = 0;
  double mx = 0.0;
  for (int64_t r = lo + gw; r < hi; r += tw) {
    const int32_t u = b.rec_v[r];
    const int j = b.rec_j[r];
    const u64 F = b.rec_mask[r - j];
    const int64_t e0 = g.off[u] + (int64_t)j * g.chunk;
    const int64_t e1 = min(g.off[u + 1], e0 + g.chunk);
    const u64 sb = ((F >> lane) & 1) ? sg_ld<NW>(b.sigma, (int64_t)u * 64 + lane) : 0ull;
    const double sf = scaled ? __longlong_as_double((long long)sb) * inv_m : 0.0;
    for (int64_t eb = e0; eb < e1; eb += 64) {
      const int64_t e = eb + lane;
      int32_t v = 0;
      u64 bits = 0;
      if (e < e1) {
        v = adj_at(g, g.adj, e);
        bits = F & nxt[v];
      }
      u64 P = __ballot(bits != 0);
      while (P) {
        const int q = __ffsll((long long)P) - 1;
        P &= P - 1;
        const int32_t vq = rl32(v, q);
        const u64 bq = rl64(bits, q);
        if ((bq >> lane) & 1) {
          u64* row = b.sigma + (int64_t)vq * 64 + lane;
          ca++;
          if (scaled) {
            double o = atomicAdd((double*)row, sf);
            mx = fmax(mx, o + sf);
          } else if (sg_add<NW>(b.sigma, (int64_t)vq * 64 + lane, sb)) {
            ovf = 1;
          }
        }
      }
    }
  }
  block_add_lanes(ca, b.st_a);
  if (scaled) block_max_lanes((u64)__double_as_longlong(mx), b.lvmax);
  if (ovf) c->sigma_ovf = 1;
}

__global__ void __launch_bounds__(256) k_ms_push(GA g, BA b) {
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_push_body<2>(g, b);
  else if (nw == 1) k_ms_push_body<1>(g, b);
  else k_ms_push_body<0>(g, b);
}

// Backward level d for a batch: lane i accumulates, in adjacency order,
// coef[z][i] * sigma[v][i] over the successors z of v for source i
// (neighbours with bit i in the level d+1 mask), starting from w(v).  Then
// coef[v][i] = (delta+1)/sigma and bc[v] += sum_i mult_i * delta_i.  Also
// publishes level d into lm[d%3] for the next (shallower) step and clears
// level d+2 from lm[(d+2)%3].
// Backward level d, one record (the pull): see k_ms_back_body.
template <int NW>
__device__ __forceinline__ void back_one(const GA& g, const BA& b, int32_t v, int j, u64 L, int64_t o0, int64_t o1,
                                         int d, const u64* lmn, u64* lmd, const unsigned* nlb, unsigned* dlb,
                                         double mult, bool scaled, double inv_md) {
  const int lane = lane_id();
  int64_t deg = o1 - o0;
  int nch = nchunks(deg, g.chunk);
  int64_t e0 = o0 + (int64_t)j * g.chunk;
  int64_t e1 = min(o1, e0 + g.chunk);
  const bool mine = (L >> lane) & 1;
  const double wv = g.w ? (double)g.w[v] : 0.0;
  const u64 sbits = mine ? sg_ld<NW>(b.sigma, (int64_t)v * 64 + lane) : 1ull;
  const double svr = !mine ? 1.0 : scaled ? __longlong_as_double((long long)sbits) : (double)sbits;
  const double sv = scaled ? svr * inv_md : svr;
  double acc = (nch == 1) ? wv : 0.0;
  for (int64_t eb = e0; eb < e1; eb += 64) {
    int64_t e = eb + lane;
    int32_t z = 0;
    u64 M = 0;
    if (e < e1) {
      z = adj_at(g, g.adj, e);
      if (!(g.flags & KF_GATE) || lb_test(nlb, z)) M = lmn[z] & L;
    }
    // Successors UNR at a time (loads first, then the sums in adjacency
    // order; FP contraction is off, so each step rounds like the oracle).
    for_lanes<UNR_BACK>(__ballot(M != 0), [&](const int (&q)[UNR_BACK], const bool (&ok)[UNR_BACK]) {
      double x[UNR_BACK];
#pragma unroll
      for (int t = 0; t < UNR_BACK; t++) {
        const int32_t zq = rl32(z, q[t]);
        const u64 mq = ok[t] ? rl64(M, q[t]) : 0ull;
        x[t] = 0.0;
        if (in_mask(mq)) x[t] = b.coef[(int64_t)zq * 64 + lane];
      }
      // a lane outside a successor's mask adds 0 * sv = +0 (sv is positive
      // and finite, acc never -0), which leaves acc unchanged
#pragma unroll
      for (int t = 0; t < UNR_BACK; t++) acc = acc + x[t] * sv;
    });
  }
  double* crow = b.coef + (int64_t)v * 64 + lane;
  if (nch > 1) {
    if (mine) atomicAdd(crow, acc);  // finished by k_ms_bfin
  } else {
    double contrib = 0.0;
    if (mine) {
      *crow = (acc + 1.0) / svr;
      contrib = mult * acc;
      if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = acc;
    }
    contrib = wave_sum_d(contrib);
    // a non-returning atomic add leaves no load of bc[v] to wait for; a
    // vertex held in one record is the only adder of bc[v] at this level,
    // so the sum is the one a load and store would give
    if (lane == 0) atomicAdd(&b.bc[v], contrib);
  }
  if (j == 0 && lane == 0) {
    lmd[v] = L;
    if (g.flags & KF_GATE) lb_set(dlb, v);
  }
}

// Backward level d for up to BPK_W records of at most BPK_DEG arcs (a vertex
// of one record each), given by their header slots pick[] (-1: none):
// sub-group k (lanes k*BPK_DEG ..) loads the arcs of record k and probes their
// ends' level d+1 masks, one arc per lane; the row gathers then run with
// lanes = sources over the successors of all of them, each record with its
// own accumulator, its terms added in adjacency order as back_one adds them.
#ifndef BPK_W
#define BPK_W 2
#endif
#define BPK_DEG (64 / BPK_W)
#ifndef BPK_MAX
#define BPK_MAX BPK_DEG  // records of at most this many arcs take the packed path
#endif
#ifndef UNR_BPK
#define UNR_BPK 3  // row gathers in flight per wave in back_pk
#endif
template <int NW>
__device__ __forceinline__ void back_pk(const GA& g, const BA& b, const RecHdr* hdr, int wb, const int (&pick)[BPK_W],
                                        const u64* lmn, u64* lmd, const unsigned* nlb, unsigned* dlb,
                                        double mult, bool scaled, double inv_md) {
  const int lane = lane_id();
  const int sub = lane / BPK_DEG, sl = lane % BPK_DEG;
  int32_t vr[BPK_W];
  u64 Lr[BPK_W];
  int my = -1;
#pragma unroll
  for (int r = 0; r < BPK_W; r++) {
    const int k = pick[r] < 0 ? 0 : pick[r];
    vr[r] = __builtin_amdgcn_readfirstlane(hdr->v[wb][k]);
    Lr[r] = pick[r] < 0 ? 0ull : rl64(hdr->L[wb][k], 0);
    if (sub == r) my = pick[r];
  }
  int32_t z = 0;
  u64 M = 0;
  if (my >= 0) {
    const int64_t o0 = hdr->o0[wb][my], o1 = hdr->o1[wb][my];
    const u64 Ls = hdr->L[wb][my];
    if (sl < o1 - o0) {
      z = adj_at(g, g.adj, o0 + sl);
      if (!(g.flags & KF_GATE) || lb_test(nlb, z)) M = lmn[z] & Ls;
    }
  }
  double svr[BPK_W], acc[BPK_W];
#pragma unroll
  for (int r = 0; r < BPK_W; r++) {
    const bool mine = (Lr[r] >> lane) & 1;
    const u64 sbits = mine ? sg_ld<NW>(b.sigma, (int64_t)vr[r] * 64 + lane) : 1ull;
    svr[r] = !mine ? 1.0 : scaled ? __longlong_as_double((long long)sbits) : (double)sbits;
    acc[r] = (pick[r] >= 0 && g.w) ? (double)g.w[vr[r]] : 0.0;
  }
  for_lanes<UNR_BPK>(__ballot(M != 0), [&](const int (&q)[UNR_BPK], const bool (&ok)[UNR_BPK]) {
    double x[UNR_BPK];
#pragma unroll
    for (int t = 0; t < UNR_BPK; t++) {
      const int32_t zq = rl32(z, q[t]);
      const u64 mq = ok[t] ? rl64(M, q[t]) : 0ull;
      x[t] = 0.0;
      if (in_mask(mq)) x[t] = b.coef[(int64_t)zq * 64 + lane];
    }
#pragma unroll
    for (int t = 0; t < UNR_BPK; t++) {
      const int rt = q[t] / BPK_DEG;  // wave-uniform
#pragma unroll
      for (int r = 0; r < BPK_W; r++)
        if (ok[t] && r == rt) acc[r] = acc[r] + x[t] * (scaled ? svr[r] * inv_md : svr[r]);
    }
  });
#pragma unroll
  for (int r = 0; r < BPK_W; r++) {
    if (pick[r] < 0) continue;
    const int32_t v = vr[r];
    double contrib = 0.0;
    if ((Lr[r] >> lane) & 1) {
      b.coef[(int64_t)v * 64 + lane] = (acc[r] + 1.0) / svr[r];
      contrib = mult * acc[r];
      if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = acc[r];
    }
    contrib = wave_sum_d(contrib);
    if (lane == 0) {
      atomicAdd(&b.bc[v], contrib);  // as in back_one
      lmd[v] = Lr[r];
      if (g.flags & KF_GATE) lb_set(dlb, v);
    }
  }
}

template <int NW>
__device__ __forceinline__ void k_ms_back_body(GA g, BA b, RecHdr* hdr) {
  Ctl* c = b.ctl;
  if (!c->fdone || c->bdone) return;
  const int d = c->blevel;
  const u64* lmn = b.lm + (int64_t)((d + 1) % 3) * b.n;
  u64* lmd = b.lm + (int64_t)(d % 3) * b.n;
  const unsigned* nlb = b.lb + (int64_t)((d + 1) % 3) * b.nw;
  unsigned* dlb = b.lb + (int64_t)(d % 3) * b.nw;
  const int64_t lo = b.rec_lo[d], hi = b.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  const double mult = b.multbuf[lane];
  // Scaled path counts: sigma_v/sigma_w = raw(v) / (raw(w) m(d)) for v on
  // level d and w on level d+1, so the terms use raw(v)/m(d) and coef keeps raw.
  const bool scaled = c->scaled;
  const double inv_md = scaled ? 1.0 / b.lscale[(int64_t)d * 64 + lane] : 1.0;
  if (c->bpush) {
    // pushed level: only zero the accumulators and publish the level
    for (int64_t r = lo + gw; r < hi; r += tw) {
      if (b.rec_j[r] != 0) continue;
      const int32_t v = b.rec_v[r];
      const u64 L = b.rec_mask[r];
      if ((L >> lane) & 1) b.coef[(int64_t)v * 64 + lane] = 0.0;
      if (lane == 0) {
        lmd[v] = L;
        if (g.flags & KF_GATE) lb_set(dlb, v);
      }
    }
  } else {
    // This wave's records lo + gw, lo + gw + tw, ...: the headers of 64 of
    // them load at once, one lane each, then the records run in order.
    // the headers wait in LDS (a wave reads back only its own slots, and a
    // wave's LDS accesses complete in issue order), which leaves their
    // registers to the row gathers
    int32_t (&h_v)[MAX_WAVES_PER_BLOCK][64] = hdr->v;
    int32_t (&h_j)[MAX_WAVES_PER_BLOCK][64] = hdr->j;
    u64 (&h_L)[MAX_WAVES_PER_BLOCK][64] = hdr->L;
    int64_t (&h_o0)[MAX_WAVES_PER_BLOCK][64] = hdr->o0;
    int64_t (&h_o1)[MAX_WAVES_PER_BLOCK][64] = hdr->o1;
    const int wb = threadIdx.x >> 6;
    for (int64_t k0 = 0; lo + gw + k0 * tw < hi; k0 += 64) {
      const int64_t rq = lo + gw + (k0 + lane) * tw;
      int32_t vq = 0, jq = 0;
      u64 Lq = 0;
      int64_t o0q = 0, o1q = 0;
      if (rq < hi) {
        vq = b.rec_v[rq];
        jq = b.rec_j[rq];
        Lq = b.rec_mask[rq - jq];
        o0q = g.off[vq];
        o1q = g.off[vq + 1];
      }
      u64 R = __ballot(rq < hi);
      h_v[wb][lane] = vq; h_j[wb][lane] = jq; h_L[wb][lane] = Lq; h_o0[wb][lane] = o0q; h_o1[wb][lane] = o1q;
      __builtin_amdgcn_wave_barrier();
#if BACK_PK
      {
        // records of at most BPK_MAX arcs, BPK_W at a time
        const bool small = rq < hi && jq == 0 && o1q - o0q <= BPK_MAX && o1q - o0q <= g.chunk;
        u64 P = __ballot(small);
        R &= ~P;
        while (P) {
          int pick[BPK_W];
#pragma unroll
          for (int t = 0; t < BPK_W; t++) {
            pick[t] = P ? __ffsll((long long)P) - 1 : -1;
            P &= P - 1;
          }
          back_pk<NW>(g, b, hdr, wb, pick, lmn, lmd, nlb, dlb, mult, scaled, inv_md);
        }
      }
#endif
      while (R) {
        const int q = __ffsll((long long)R) - 1;
        R &= R - 1;
        const int32_t v = __builtin_amdgcn_readfirstlane(h_v[wb][q]);
        const int j = __builtin_amdgcn_readfirstlane(h_j[wb][q]);
        const u64 L = rl64(h_L[wb][q], 0);
        const int64_t o0 = (int64_t)rl64((u64)h_o0[wb][q], 0), o1 = (int64_t)rl64((u64)h_o1[wb][q], 0);
        back_one<NW>(g, b, v, j, L, o0, o1, d, lmn, lmd, nlb, dlb, mult, scaled, inv_md);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  {
    u64* old = b.lm + (int64_t)((d + 2) % 3) * b.n;
    unsigned* olb = b.lb + (int64_t)((d + 2) % 3) * b.nw;
    const int64_t plo = b.rec_lo[d + 2], phi = b.rec_lo[d + 3];
    if (d + 2 <= c->last)
      for (int64_t r = plo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < phi;
           r += (int64_t)gridDim.x * blockDim.x)
        if (b.rec_j[r] == 0) {
          old[b.rec_v[r]] = 0;
          if (g.flags & KF_GATE) lb_clr(olb, b.rec_v[r]);
        }
  }
}

// 8 waves per SIMD (64 VGPRs; a few bytes spill outside the gather loop): more
// row gathers in flight per CU than 7 waves with no spill.
#ifndef BACK_WAVES
#define BACK_WAVES 8
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BACK_WAVES))) k_ms_back(GA g, BA b, KT* kt) {
  kt_enter(kt);
  __shared__ RecHdr hdr;
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_back_body<2>(g, b, &hdr);
  else if (nw == 1) k_ms_back_body<1>(g, b, &hdr);
  else k_ms_back_body<0>(g, b, &hdr);
  kt_exit(kt);
}

// Pushed backward level d (undirected): for every arc z-u out of a level-d+1
// record, coef[u][i] += coef[z][i] for the sources i with z on level d+1 and
// u on level d.  k_ms_bfin then multiplies the sums by sigma[u][i].  The
// successor level's arcs are far fewer than level d's, which the pull in
// k_ms_back would scan; the sums are added in arrival order.
__device__ __forceinline__ void k_ms_bpush_body(GA g, BA b) {
  Ctl* c = b.ctl;
  if (!c->fdone || c->bdone || !c->bpush) return;
  const int d = c->blevel;
  const u64* lmd = b.lm + (int64_t)(d % 3) * b.n;
  const unsigned* dlb = b.lb + (int64_t)(d % 3) * b.nw;
  const int64_t lo = b.rec_lo[d + 1], hi = b.rec_lo[d + 2];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  for (int64_t r = lo + gw; r < hi; r += tw) {
    const int32_t z = b.rec_v[r];
    const int j = b.rec_j[r];
    const u64 F = b.rec_mask[r - j];
    const int64_t e0 = g.off[z] + (int64_t)j * g.chunk;
    const int64_t e1 = min(g.off[z + 1], e0 + g.chunk);
    const double cz = ((F >> lane) & 1) ? b.coef[(int64_t)z * 64 + lane] : 0.0;
    for (int64_t eb = e0; eb < e1; eb += 64) {
      const int64_t e = eb + lane;
      int32_t u = 0;
      u64 M = 0;
      if (e < e1) {
        u = adj_at(g, g.adj, e);
        if (!(g.flags & KF_GATE) || lb_test(dlb, u)) M = lmd[u] & F;
      }
      u64 P = __ballot(M != 0);
      while (P) {
        const int q = __ffsll((long long)P) - 1;
        P &= P - 1;
        const int32_t uq = rl32(u, q);
        const u64 mq = rl64(M, q);
        if ((mq >> lane) & 1) atomicAdd(b.coef + (int64_t)uq * 64 + lane, cz);
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_ms_bpush(GA g, BA b, KT* kt) {
  kt_enter(kt);
  k_ms_bpush_body(g, b);
  kt_exit(kt);
}

// Finishes the level-d vertices whose arcs were split over several records
// (delta = w(v) + the atomically added partial sums held in coef) or, on a
// pushed level, every level-d vertex (delta = w(v) + sigma(v) x the pushed
// sum of its successors' coef).
// One such vertex (wave-wide; lanes = sources).
template <int NW>
__device__ __forceinline__ void bfin_one(const GA& g, const BA& b, int64_t r, int32_t v, bool bpush, bool scaled,
                                         double inv_md, double mult) {
  const int lane = lane_id();
  const u64 L = b.rec_mask[r];
  const bool mine = (L >> lane) & 1;
  double contrib = 0.0;
  if (mine) {
    double* crow = b.coef + (int64_t)v * 64 + lane;
    u64 sbits = sg_ld<NW>(b.sigma, (int64_t)v * 64 + lane);
    double svr = scaled ? __longlong_as_double((long long)sbits) : (double)sbits;
    double part = *crow;
    if (bpush) part = part * (scaled ? svr * inv_md : svr);
    double dl = (g.w ? (double)g.w[v] : 0.0) + part;
    *crow = (dl + 1.0) / svr;
    contrib = mult * dl;
    if (b.dbg_delta) b.dbg_delta[(int64_t)v * 64 + lane] = dl;
  }
  contrib = wave_sum_d(contrib);
  if (lane == 0) b.bc[v] += contrib;
}

template <int NW>
__device__ __forceinline__ void k_ms_bfin_body(GA g, BA b) {
  Ctl* c = b.ctl;
  if (!c->fdone || c->bdone) return;
  const int d = c->blevel;
  const int64_t lo = b.rec_lo[d], hi = b.rec_lo[d + 1];
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  const double mult = b.multbuf[lane];
  const bool bpush = c->bpush;
  const bool scaled = c->scaled;
  const double inv_md = scaled ? 1.0 / b.lscale[(int64_t)d * 64 + lane] : 1.0;
  if (bpush) {
    for (int64_t r = lo + gw; r < hi; r += tw)
      if (b.rec_j[r] == 0) bfin_one<NW>(g, b, r, b.rec_v[r], true, scaled, inv_md, mult);
    return;
  }
  // only split vertices are left: find their first records 64 at a time
  for (int64_t base = lo + gw * 64; base < hi; base += tw * 64) {
    const int64_t r = base + lane;
    int32_t v = 0;
    bool split = false;
    if (r < hi && b.rec_j[r] == 0) {
      v = b.rec_v[r];
      split = g.off[v + 1] - g.off[v] > g.chunk;
    }
    u64 S = __ballot(split);
    while (S) {
      const int q = __ffsll((long long)S) - 1;
      S &= S - 1;
      bfin_one<NW>(g, b, base + q, rl32(v, q), false, scaled, inv_md, mult);
    }
  }
}

__global__ void __launch_bounds__(256) k_ms_bfin(GA g, BA b, KT* kt) {
  kt_enter(kt);
  const int nw = b.ctl->narrow;
  if (nw == 2) k_ms_bfin_body<2>(g, b);
  else if (nw == 1) k_ms_bfin_body<1>(g, b);
  else k_ms_bfin_body<0>(g, b);
  kt_exit(kt);
}

// Resets the batch state for the next sweep.  seen is cleared for every
// reached vertex and compmask for the sources' components.  After a complete
// backward pass only levels 1 and 2 are left in the level masks (each step
// clears the level two below it), so only their records clear lm; a sweep
// whose backward pass was skipped clears lm for every record.
__global__ void k_ms_cleanup(GA g, BA b) {
  Ctl* c = b.ctl;
  if (!c->fdone || !c->bdone) return;
  int64_t n = c->rec_count < b.cap ? c->rec_count : b.cap;
  const bool full = c->overflow || (c->sigma_ovf && !c->scaled) || (g.flags & KF_GATE);
  const int64_t src_hi = b.rec_lo[1];
  const int64_t lm_hi = full ? n : b.rec_lo[min(3, c->last + 1)];
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    if (b.rec_j[r] == 0) {
      int32_t v = b.rec_v[r];
      b.seen[v] = 0;
      if (r < lm_hi) {
        b.lm[v] = 0;
        b.lm[b.n + v] = 0;
        b.lm[2 * b.n + v] = 0;
      }
      if (g.flags & KF_GATE) lb_clr(b.lb, v);
      if (g.flags & KF_GATE) lb_clr(b.lb + b.nw, v);
      if (g.flags & KF_GATE) lb_clr(b.lb + 2 * b.nw, v);
      if (r < src_hi) b.compmask[g.label[v]] = 0;
    }
  }
  // the path-count rows of split vertices are zero outside a sweep
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.nsplit * 64;
       t += (int64_t)gridDim.x * blockDim.x)
    sg_st_w(b.sigma, c->narrow, (int64_t)g.split_v[t >> 6] * 64 + (t & 63), 0);
}

// ====================================== workgroup-per-source sweeps
// Deep, narrow graphs (long chains) give every level a few hundred vertices
// per source: a device-wide level loop then pays kernel boundaries on tens
// of thousands of levels.  Here one workgroup runs a whole Brandes sweep for
// one source (levels separated by workgroup barriers, not kernels) and
// hundreds of workgroups run different sources at once.  Per-slot arrays are
// private to their workgroup; within it they are accessed with
// workgroup-scope atomics (loads served by the L2, which the RMWs update).
#define WGT 256
#ifndef WG_ILP_B
#define WG_ILP_B 8  // successors' loads in flight in the backward pass (8: +3 % on chains against 4)
#endif
#ifndef WG_ILP
#define WG_ILP 4  // arcs of a vertex whose loads (then claims, then adds) a workgroup-sweep thread keeps
                  // in flight (4 at 8 waves per SIMD: +6 % on chains against 8 claims one at a time)
#endif
struct WA {
  const int32_t* src;
  const double* mult;
  int64_t nsrc;
  unsigned long long* next;  // next source index
  int32_t* depth;            // slot-major, n per slot; -1 between sweeps
  u64* sigma;                // uint64 counts, or FP64 bits when scaled; 0 between sweeps.
                             // A finished vertex's slot holds coef = (delta+1)/sigma instead
  int32_t* queue;            // BFS order; level k = queue[lvl[k] .. lvl[k+1])
  int32_t* lvl;              // lcap per slot
  double* scale;             // scaled: per-level factor m(k), lcap per slot
  int64_t lcap;
  double* bc;
  int64_t* st_n;             // per source: reached vertices, their degree sum, DAG arcs
  int64_t* st_m;
  int64_t* st_a;
  int32_t* st_levels;
  int32_t* st_flag;          // 1: exact path count wrapped, 2: more levels than lcap
  int scaled;
};

template <class T>
__device__ __forceinline__ T wg_ld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class T>
__device__ __forceinline__ void wg_st(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ double ub2d(u64 x) { return __longlong_as_double((long long)x); }
__device__ __forceinline__ u64 d2ub(double x) { return (u64)__double_as_longlong(x); }

__device__ __forceinline__ double block_max_d(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  if (lane_id() == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  double m = red[0];
  for (int i = 1; i < WGT / 64; i++) m = fmax(m, red[i]);
  __syncthreads();
  return m;
}
__device__ __forceinline__ long long block_sum_ll(long long x, long long* red) {
  x = wave_sum_ll(x);
  if (lane_id() == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  long long t = 0;
  for (int i = 0; i < WGT / 64; i++) t += red[i];
  __syncthreads();
  return t;
}

// waves per SIMD asked of the compiler: 2048 workgroups of 4 waves fill 256
// CUs at 8 (build switch WG_WAVES)
#ifndef WG_WAVES
#define WG_WAVES 8
#endif
#define WG_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(WG_WAVES)))
__global__ void __launch_bounds__(WGT) WG_WAVES_ATTR k_wg_sweep(GA g, WA a) {
  const int64_t n = g.n;
  const int tid = threadIdx.x;
  const int64_t slot = blockIdx.x;
  int32_t* depth = a.depth + slot * n;
  u64* sig = a.sigma + slot * n;
  int32_t* q = a.queue + slot * n;
  int32_t* lvl = a.lvl + slot * a.lcap;
  double* scl = a.scale + slot * a.lcap;
  __shared__ int64_t s_i;
  __shared__ int s_tail, s_flag;
  __shared__ double s_redd[WGT / 64];
  __shared__ long long s_redl[WGT / 64];
  const bool sc = a.scaled != 0;
  for (;;) {
    if (tid == 0) s_i = (int64_t)atomicAdd(a.next, 1ull);
    __syncthreads();
    const int64_t i = s_i;
    if (i >= a.nsrc) break;
    const int32_t s = a.src[i];
    if (tid == 0) {
      wg_st(depth + s, 0);
      wg_st(sig + s, sc ? d2ub(1.0) : 1ull);
      wg_st(q, s);
      wg_st(lvl, 0);
      wg_st(lvl + 1, 1);
      s_tail = 1;
      s_flag = 0;
    }
    __syncthreads();
    long long m_cnt = 0, a_cnt = 0;
    // Forward: expand level L into level L+1 (first visit claims the vertex
    // with a CAS on depth; sigma of level L+1 adds up its parents').
    int L = 0;
    for (;;) {
      const int lo = wg_ld(lvl + L), hi = wg_ld(lvl + L + 1);
      if (lo == hi) break;
      if (L + 2 >= a.lcap) {
        if (tid == 0) s_flag = 2;
        __syncthreads();
        break;
      }
      const int32_t d1 = L + 1;
      for (int qq = lo + tid; qq < hi; qq += WGT) {
        const int32_t v = wg_ld(q + qq);
        const u64 sv = wg_ld(sig + v);
        const int64_t e0 = g.off[v], e1 = g.off[v + 1];
        m_cnt += e1 - e0;
        // arcs WG_ILP at a time: their neighbour ids and depths load together
        for (int64_t eb = e0; eb < e1; eb += WG_ILP) {
          int32_t ws[WG_ILP], ds[WG_ILP];
#pragma unroll
          for (int t = 0; t < WG_ILP; t++) ws[t] = eb + t < e1 ? g.adj[eb + t] : -1;
#pragma unroll
          for (int t = 0; t < WG_ILP; t++) ds[t] = ws[t] >= 0 ? wg_ld(depth + ws[t]) : INT32_MAX;
          // The claims of unvisited neighbours are all issued before any
          // result is used, then the path-count adds likewise: two atomic
          // round trips per WG_ILP arcs instead of up to two per arc.
          int32_t dw[WG_ILP];
#pragma unroll
          for (int t = 0; t < WG_ILP; t++) {
            dw[t] = ds[t];
            if (ws[t] >= 0 && ds[t] < 0) {
              int32_t e = -1;  // stays -1 when this claim wins, else the claimer's depth
              __hip_atomic_compare_exchange_strong(depth + ws[t], &e, d1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
              dw[t] = e;
            }
          }
#pragma unroll
          for (int t = 0; t < WG_ILP; t++) {
            if (ws[t] >= 0 && ds[t] < 0 && dw[t] == -1) {
              wg_st(q + atomicAdd(&s_tail, 1), ws[t]);
              dw[t] = d1;
            }
          }
          u64 old[WG_ILP];
#pragma unroll
          for (int t = 0; t < WG_ILP; t++) {
            old[t] = 0;
            if (ws[t] >= 0 && dw[t] == d1) {
              a_cnt++;
              if (sc)
                __hip_atomic_fetch_add((double*)(sig + ws[t]), ub2d(sv), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
              else
                old[t] = __hip_atomic_fetch_add(sig + ws[t], sv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
          }
          if (!sc) {
#pragma unroll
            for (int t = 0; t < WG_ILP; t++)
              if (ws[t] >= 0 && dw[t] == d1 && old[t] + sv < old[t]) s_flag = 1;
          }
        }
      }
      __syncthreads();
      const int nt = s_tail;
      if (sc && nt > hi) {
        // factor of level L+1: its largest raw count; stored counts <= 1
        double mx = 0.0;
        for (int qq = hi + tid; qq < nt; qq += WGT) mx = fmax(mx, ub2d(wg_ld(sig + wg_ld(q + qq))));
        const double m = block_max_d(mx, s_redd);
        const double f = m > 0.0 ? m : 1.0;
        for (int qq = hi + tid; qq < nt; qq += WGT) {
          const int32_t w = wg_ld(q + qq);
          wg_st(sig + w, d2ub(ub2d(wg_ld(sig + w)) / f));
        }
        if (tid == 0) wg_st(scl + d1, f);
      }
      if (tid == 0) wg_st(lvl + L + 2, nt);
      __syncthreads();
      if (s_flag) break;
      L++;
    }
    const int reached = s_tail;
    const int flag = s_flag;
    // Backward, deepest level first: delta(v) = w(v) + sum over successors z
    // of (delta(z)+1)/sigma(z)*sigma(v), in adjacency order (the oracle's).
    // A finished vertex's slot then holds coef(v) = (delta(v)+1)/sigma(v),
    // the quotient its predecessors use (computed once, same operations).
    if (!flag) {
      const double mult = a.mult[i];
      for (int d = L - 1; d >= 1; d--) {
        const int lo = wg_ld(lvl + d), hi = wg_ld(lvl + d + 1);
        const double inv_m = (sc && d + 1 < L) ? 1.0 / wg_ld(scl + d + 1) : 1.0;
        for (int qq = lo + tid; qq < hi; qq += WGT) {
          const int32_t v = wg_ld(q + qq);
          const u64 sb = wg_ld(sig + v);
          const double svr = sc ? ub2d(sb) : (double)sb;
          const double sv = sc ? svr * inv_m : svr;
          double acc = g.w ? (double)g.w[v] : 0.0;
          const int64_t f0 = g.off[v], f1 = g.off[v + 1];
          for (int64_t eb = f0; eb < f1; eb += WG_ILP_B) {
            // WG_ILP_B successors' loads in flight, the sum still in adjacency order
            int32_t zs[WG_ILP_B];
            bool ok[WG_ILP_B];
            double zc[WG_ILP_B];
#pragma unroll
            for (int t = 0; t < WG_ILP_B; t++) zs[t] = eb + t < f1 ? g.adj[eb + t] : -1;
#pragma unroll
            for (int t = 0; t < WG_ILP_B; t++) ok[t] = zs[t] >= 0 && wg_ld(depth + zs[t]) == d + 1;
#pragma unroll
            for (int t = 0; t < WG_ILP_B; t++) zc[t] = ok[t] ? ub2d(wg_ld(sig + zs[t])) : 0.0;
#pragma unroll
            for (int t = 0; t < WG_ILP_B; t++)
              if (ok[t]) acc = acc + zc[t] * sv;
          }
          if (acc != 0.0) atomicAdd(a.bc + v, mult * acc);
          // level d is read by level d-1 only after the barrier below
          wg_st(sig + v, d2ub((acc + 1.0) / svr));
        }
        __syncthreads();
      }
    }
    const long long msum = block_sum_ll(m_cnt, s_redl);
    const long long asum = block_sum_ll(a_cnt, s_redl);
    if (tid == 0) {
      a.st_n[i] = reached;
      a.st_m[i] = msum;
      a.st_a[i] = asum;
      a.st_levels[i] = L;
      a.st_flag[i] = flag;
    }
    // reset the visited entries for the next source of this slot
    for (int qq = tid; qq < reached; qq += WGT) {
      const int32_t v = wg_ld(q + qq);
      wg_st(depth + v, -1);
      wg_st(sig + v, 0ull);
    }
    __syncthreads();
  }
}

// ================================================= graph preparation kernels
__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(u64* p, u64 v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_iota(int32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

// Union-find with lock-free hooking of the larger root under the smaller, so
// the final label of a component is its smallest vertex id.
__device__ int32_t uf_find(int32_t* par, int32_t x) {
  for (;;) {
    int32_t p = __hip_atomic_load(&par[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == x) return x;
    int32_t gp = __hip_atomic_load(&par[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (gp != p) __hip_atomic_store(&par[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
  }
}
// Weak components: undirected lists hold each edge twice, so only the arc
// to the larger id is used; directed lists use every arc.
__global__ void k_cc_union(const int64_t* off, const int32_t* adj, const int32_t* rv, const int32_t* rj,
                           int64_t nrec, int chunk, int32_t* par, int directed) {
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  for (int64_t r = gw; r < nrec; r += tw) {
    int32_t v = rv[r];
    int j = rj[r];
    int64_t e0 = off[v] + (int64_t)j * chunk;
    int64_t e1 = min(off[v + 1], e0 + chunk);
    for (int64_t e = e0 + lane; e < e1; e += 64) {
      int32_t x = adj[e];
      if (x == v || (x < v && !directed)) continue;
      int32_t a = v, bb = x;
      for (;;) {
        a = uf_find(par, a);
        bb = uf_find(par, bb);
        if (a == bb) break;
        int32_t hi = a > bb ? a : bb, lo = a > bb ? bb : a;
        if (atomicCAS(&par[hi], hi, lo) == hi) break;
      }
    }
  }
}
// Labels from the union-find forest (par keeps its links; only atomic
// operations touch it).
__global__ void k_cc_label(int32_t* par, int64_t n, int32_t* label) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x)
    label[v] = uf_find(par, (int32_t)v);
}
// Flags an arc whose endpoints carry different labels; the host then runs
// another union and labelling pass, until every arc agrees.
__global__ void k_cc_check(const int64_t* off, const int32_t* adj, const int32_t* rv, const int32_t* rj,
                           int64_t nrec, int chunk, const int32_t* label, int* bad) {
  const int lane = lane_id();
  const int64_t gw = (int64_t)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t tw = (int64_t)(gridDim.x * blockDim.x) >> 6;
  int any = 0;
  for (int64_t r = gw; r < nrec; r += tw) {
    int32_t v = rv[r];
    int j = rj[r];
    int64_t e0 = off[v] + (int64_t)j * chunk;
    int64_t e1 = min(off[v + 1], e0 + chunk);
    const int32_t lv = label[v];
    for (int64_t e = e0 + lane; e < e1; e += 64) any |= label[adj[e]] != lv;
  }
  if (__ballot(any) && lane_id() == 0) atomicOr(bad, 1);
}
// Vertices and traversed arcs per component.  Most vertices share one label
// (the giant component): a wave whose lanes all carry lane 0's label adds
// its sums with one atomic, other waves per lane.
__global__ void k_cc_sizes(const int32_t* label, int64_t n, const int64_t* coff, int64_t* csize, int64_t* carcs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    const bool ok = v < n;
    const int32_t l = ok ? label[v] : -1;
    const int64_t dg = ok ? coff[v + 1] - coff[v] : 0;
    const int32_t l0 = rl32(l, 0);
    if (__all(!ok || l == l0)) {
      long long c = wave_sum_ll(ok ? 1 : 0), a = wave_sum_ll(dg);
      if (lane_id() == 0 && l0 >= 0) {
        atomicAdd((u64*)&csize[l0], (u64)c);
        if (a) atomicAdd((u64*)&carcs[l0], (u64)a);
      }
    } else if (ok) {
      atomicAdd((u64*)&csize[l], 1ull);
      if (dg) atomicAdd((u64*)&carcs[l], (u64)dg);
    }
  }
}

// Peeling, round r.  Candidates: live vertices of live degree 1.
__global__ void k_peel_first(const int64_t* off, int64_t n, int32_t* deg, int32_t* list, u64* cnt) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    int32_t dg = (int32_t)(off[v + 1] - off[v]);
    deg[v] = dg;
    if (dg == 1) list[atomicAdd(cnt, 1ull)] = (int32_t)v;
  }
}
__global__ void k_peel_mark(const int32_t* list, int64_t nl, const int32_t* deg, const int32_t* round,
                            int32_t* cand, int r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = list[i];
    if (round[u] == 0 && deg[u] == 1) cand[u] = r;
  }
}
// Decides removals against start-of-round state: u leaves unless its live
// neighbour p is also a candidate with p > u (an isolated edge keeps its
// smaller end).  Removed vertices go to order[base..].
__global__ void k_peel_decide(const int64_t* off, const int32_t* adj, const int32_t* list, int64_t nl,
                              const int32_t* round, const int32_t* cand, int r, int32_t* parent,
                              int32_t* order, u64* ocnt, const int32_t* n2o) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nl; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = list[i];
    if (cand[u] != r) continue;
    int32_t p = -1;
    for (int64_t e = off[u]; e < off[u + 1]; e++)
      if (round[adj[e]] == 0) { p = adj[e]; break; }
    if (p < 0) continue;
    if (cand[p] == r && (n2o ? n2o[p] > n2o[u] : p > u)) continue;  // ties in the caller's ids
    parent[u] = p;
    order[atomicAdd(ocnt, 1ull)] = u;
  }
}
__global__ void k_peel_apply(const int32_t* rem, int64_t nr, int r, int32_t* round, const int32_t* parent,
                             u64* w, int32_t* deg, int32_t* next, u64* ncnt) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = rem[i];
    int32_t p = parent[u];
    round[u] = r;
    atomicAdd(&w[p], w[u] + 1);
    int old = atomicSub(&deg[p], 1);
    if (old == 2) next[atomicAdd(ncnt, 1ull)] = p;
    deg[u] = 0;
  }
}
__global__ void k_peel_root(const int32_t* rem, int64_t nr, const int32_t* round, const int32_t* parent,
                            int32_t* root) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = rem[i];
    int32_t p = parent[u];
    root[u] = round[p] == 0 ? p : root[p];
  }
}
__global__ void k_core_deg(const int64_t* off, const int32_t* adj, int64_t n, const int32_t* round, int64_t* cdeg) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    int64_t c = 0;
    if (round[v] == 0)
      for (int64_t e = off[v]; e < off[v + 1]; e++) c += round[adj[e]] == 0;
    cdeg[v] = c;
  }
}
__global__ void k_core_fill(const int64_t* off, const int32_t* adj, int64_t n, const int32_t* round,
                            const int64_t* coff, int32_t* cadj) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    if (round[v] != 0) continue;
    int64_t o = coff[v];
    for (int64_t e = off[v]; e < off[v + 1]; e++)
      if (round[adj[e]] == 0) cadj[o++] = adj[e];
  }
}

// ===================================================== per-run source terms
// Multiset counts of the sources, optional expansion to whole pruned trees,
// and per-component source counts.
__global__ void k_src_count(const int32_t* src, int64_t ns, u64* cnt0) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ns; i += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt0[src[i]], 1ull);
}
__global__ void k_src_expand(const int32_t* root, const int32_t* round, int64_t n, const u64* listed, u64* cnt0) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    int32_t rt = root[v];
    cnt0[v] = (round[rt] == 0 && listed[rt] > 0) ? 1ull : 0ull;
  }
}
__global__ void k_src_init(const u64* cnt0, const int32_t* label, int64_t n, u64* cntS, u64* sa, u64* srcK) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    cntS[v] = cnt0[v];
    sa[v] = 0;
    if (cnt0[v]) atomicAdd(&srcK[label[v]], cnt0[v]);
  }
}
// One peel round, leaves first: cntS(p) += cntS(u), sa(p) += cntS(u)*(1+w(u)).
__global__ void k_tree_acc(const int32_t* rem, int64_t nr, const int32_t* parent, const u64* w, u64* cntS, u64* sa) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nr; i += (int64_t)gridDim.x * blockDim.x) {
    int32_t u = rem[i];
    int32_t p = parent[u];
    u64 cs = cntS[u];
    if (cs) {
      atomicAdd(&cntS[p], cs);
      atomicAdd(&sa[p], cs * (1 + w[u]));
    }
  }
}
// Closed-form dependencies carried by pruned trees (exact-subset semantics):
// for any vertex x with nK vertices in its component,
//   bc[x] += desc(x)*(nK-1) - sa(x)                 (sources below x)
//   bc[x] += w(x)*(|S cap K| - cntS(x))  if x pruned (sources outside x's subtree)
__global__ void k_finalize(const int32_t* label, const int32_t* round, const u64* w, const u64* cnt0,
                           const u64* cntS, const u64* sa, const u64* srcK, const int64_t* csize, int64_t n,
                           double* bc) {
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    int32_t l = label[v];
    u64 nK = (u64)csize[l];
    u64 desc = cntS[v] - cnt0[v];
    long long t = (long long)(desc * (nK - 1)) - (long long)sa[v];
    if (round[v] > 0) t += (long long)(w[v] * (srcK[l] - cntS[v]));
    if (t) bc[v] += (double)t;
  }
}

// ============================================================ host side ===
static int grid_for(int64_t n, int block = 256) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

template <class T>
static int dalloc(T** p, size_t count, size_t* acc = nullptr) {
  size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return set_error(BCMI_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  if (acc) *acc += bytes;
  return BCMI_OK;
}
#define TRY(x)               \
  do {                       \
    int rc_ = (x);           \
    if (rc_ != BCMI_OK) return rc_; \
  } while (0)

void DevGraph::release() {
  ALLOC_LOCK;
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  if (toff != off) { fr(toff); fr(tadj); }
  if (off != foff) { fr(off); fr(adj); }
  fr(foff); fr(fadj); fr(label); fr(comp_size); fr(comp_arcs); fr(all_v); fr(all_j); fr(split_v); fr(sm_v); fr(lg_v); fr(lg_j);
  fr(round); fr(parent); fr(w); fr(root); fr(order); fr(d_n2o);
  d_n2o = nullptr; relabeled = false; h_n2o.clear(); h_o2n.clear();
  sigma_overflowed.store(false);
  narrow_cap.store(2);
  off = foff = toff = nullptr; adj = fadj = tadj = nullptr; label = nullptr; comp_size = comp_arcs = nullptr;
  all_v = all_j = split_v = sm_v = lg_v = lg_j = nullptr; nsplit = nsm = nlg = 0; round = parent = root = order = nullptr; w = nullptr;
  n = m = fm = nall = 0; nrounds = 0; round_lo.clear(); h_off.clear(); h_round.clear(); h_sig.clear(); bytes = 0;
}

// Batching key per vertex: BFS depths (8 bits each, 255 = unreached) from the
// H highest-degree vertices, the largest hub in the most significant byte.
// Sources with equal keys have similar depth profiles, so a batch of them
// shares levels: its per-vertex masks are dense and each vertex appears on
// few levels.  max_depth: deepest level seen from the first HD hubs (sizing
// hint for the work list and the engine choice).
static void hub_signature(int64_t n, const int64_t* off, const int32_t* adj, std::vector<uint64_t>& sig,
                          int* max_depth) {
  const int H = 8, HD = 4;
  std::vector<int32_t> hubs;
  {
    std::vector<std::pair<int64_t, int32_t>> best;
    for (int64_t v = 0; v < n; v++) {
      int64_t d = off[v + 1] - off[v];
      if ((int)best.size() < H || d > best.back().first) {
        best.push_back({d, (int32_t)v});
        std::sort(best.begin(), best.end(), [](auto& a, auto& b) { return a.first > b.first; });
        if ((int)best.size() > H) best.pop_back();
      }
    }
    for (auto& x : best) hubs.push_back(x.second);
  }
  std::vector<std::vector<uint8_t>> dep(H, std::vector<uint8_t>(n, 255));
  std::vector<int> hmax(H, 0);
  std::vector<std::thread> th;
  th.reserve(hubs.size());
  std::atomic<bool> hub_fail{false};
  auto hub_bfs = [&](int h) {
    try {
      // level-synchronous BFS: depths are stored saturated at 254 (255 =
      // unreached), the level count stays exact
      std::vector<uint8_t>& D = dep[h];
      std::vector<int32_t> q;
      q.reserve(1024);
      q.push_back(hubs[h]);
      D[hubs[h]] = 0;
      size_t lo = 0;
      int level = 0;
      while (lo < q.size()) {
        size_t hi = q.size();
        uint8_t dn = (uint8_t)std::min(level + 1, 254);
        for (size_t i = lo; i < hi; i++) {
          int32_t u = q[i];
          for (int64_t e = off[u]; e < off[u + 1]; e++)
            if (D[adj[e]] == 255) { D[adj[e]] = dn; q.push_back(adj[e]); }
        }
        lo = hi;
        if (q.size() > hi) level++;
      }
      hmax[h] = level;
    } catch (...) {
      hub_fail = true;  // host memory: reported after the join
    }
  };
  // one thread per hub; a hub whose thread cannot start is swept here
  for (int h = 0; h < (int)hubs.size(); h++) {
    try {
      th.emplace_back(hub_bfs, h);
    } catch (const std::system_error&) {
      hub_bfs(h);
    }
  }
  for (auto& t : th) t.join();
  if (hub_fail) throw std::bad_alloc();
  *max_depth = 0;
  for (int h = 0; h < std::min(HD, (int)hubs.size()); h++) *max_depth = std::max(*max_depth, hmax[h]);
  sig.assign(n, 0);
  for (int64_t v = 0; v < n; v++) {
    uint64_t k = 0;
    for (int h = 0; h < H; h++) k = (k << 8) | (h < (int)hubs.size() ? dep[h][v] : 0);
    sig[v] = k;
  }
}

static void chunk_records(const std::vector<int64_t>& off, int chunk, std::vector<int32_t>& rv,
                          std::vector<int32_t>& rj) {
  int64_t n = (int64_t)off.size() - 1;
  rv.clear(); rj.clear();
  for (int64_t v = 0; v < n; v++) {
    int64_t dg = off[v + 1] - off[v];
    if (!dg) continue;
    int nch = (int)((dg + chunk - 1) / chunk);
    for (int k = 0; k < nch; k++) { rv.push_back((int32_t)v); rj.push_back(k); }
  }
}

int devgraph_build(DevGraph& g, int64_t n, const int64_t* off, const int32_t* adj, int directed,
                   int heuristic, const Params& prm, hipStream_t st) {
  ALLOC_LOCK;
  g.release();
  if (n < 1 || n > INT32_MAX) return set_error(BCMI_ERR_ARG, "vertex count out of range");
  if (off[0] != 0) return set_error(BCMI_ERR_ARG, "offsets must start at 0");
  for (int64_t v = 0; v < n; v++)
    if (off[v + 1] < off[v]) return set_error(BCMI_ERR_ARG, "offsets must be non-decreasing");
  int64_t m = off[n];
  for (int64_t e = 0; e < m; e++)
    if (adj[e] < 0 || adj[e] >= n) return set_error(BCMI_ERR_RANGE, "adjacency entry out of range");
  if (heuristic && directed) return set_error(BCMI_ERR_ARG, "degree-1 peeling needs an undirected graph");
  std::vector<int64_t> roff;
  std::vector<int32_t> radj;
  if (prm.reorder) {
    relabel_by_degree(n, off, adj, g.h_n2o, g.h_o2n, roff, radj);
    off = roff.data();
    adj = radj.data();
    g.relabeled = true;
  }
  g.n = n; g.fm = m; g.directed = directed; g.heuristic = heuristic; g.chunk = prm.chunk;
  size_t acc = 0;
  TRY(dalloc(&g.foff, n + 1, &acc));
  TRY(dalloc(&g.fadj, m, &acc));
  HIP_TRY(hipMemcpyAsync(g.foff, off, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
  if (m) HIP_TRY(hipMemcpyAsync(g.fadj, adj, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
  TRY(dalloc(&g.label, n, &acc));
  TRY(dalloc(&g.comp_size, n, &acc));
  TRY(dalloc(&g.comp_arcs, n, &acc));
  TRY(dalloc(&g.round, n, &acc));
  TRY(dalloc(&g.parent, n, &acc));
  TRY(dalloc(&g.w, n, &acc));
  TRY(dalloc(&g.root, n, &acc));
  TRY(dalloc(&g.order, n, &acc));
  if (g.relabeled) {
    TRY(dalloc(&g.d_n2o, n, &acc));
    HIP_TRY(hipMemcpyAsync(g.d_n2o, g.h_n2o.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st));
  }
  const int B = 256;
  int G = grid_for(n);
  hipLaunchKernelGGL(k_fill_i32, G, B, 0, st, g.round, 0, n);
  hipLaunchKernelGGL(k_fill_i32, G, B, 0, st, g.parent, -1, n);
  hipLaunchKernelGGL(k_fill_u64, G, B, 0, st, (u64*)g.w, 0ull, n);
  hipLaunchKernelGGL(k_iota, G, B, 0, st, g.root, n);
  g.round_lo.assign(1, 0);
  std::vector<int64_t> hoff(off, off + n + 1);
  if (heuristic) {
    int32_t *deg = nullptr, *cand = nullptr, *la = nullptr, *lb = nullptr;
    u64* cnts = nullptr;  // [0] list a count, [1] next count, [2] order count
    TRY(dalloc(&deg, n)); TRY(dalloc(&cand, n)); TRY(dalloc(&la, n)); TRY(dalloc(&lb, n));
    TRY(dalloc(&cnts, 4));
    u64* hc = nullptr;
    HIP_TRY(hipHostMalloc((void**)&hc, 4 * sizeof(u64), hipHostMallocDefault));
    hipLaunchKernelGGL(k_fill_i32, G, B, 0, st, cand, 0, n);
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, st, cnts, 0ull, (int64_t)4);
    hipLaunchKernelGGL(k_peel_first, G, B, 0, st, g.foff, n, deg, la, cnts);
    HIP_TRY(hipMemcpyAsync(hc, cnts, 4 * sizeof(u64), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    int64_t nl = (int64_t)hc[0];
    int64_t ocount = 0;
    int r = 0;
    while (nl > 0) {
      r++;
      int Gl = grid_for(nl);
      hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, st, cnts + 1, 0ull, (int64_t)1);
      hipLaunchKernelGGL(k_peel_mark, Gl, B, 0, st, la, nl, deg, g.round, cand, r);
      hipLaunchKernelGGL(k_peel_decide, Gl, B, 0, st, g.foff, g.fadj, la, nl, g.round, cand, r, g.parent,
                         g.order, cnts + 2, g.d_n2o);
      HIP_TRY(hipMemcpyAsync(hc + 2, cnts + 2, sizeof(u64), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      int64_t nrem = (int64_t)hc[2] - ocount;
      if (nrem > 0)
        hipLaunchKernelGGL(k_peel_apply, grid_for(nrem), B, 0, st, g.order + ocount, nrem, r, g.round, g.parent,
                           (u64*)g.w, deg, lb, cnts + 1);
      ocount = (int64_t)hc[2];
      g.round_lo.push_back(ocount);
      HIP_TRY(hipMemcpyAsync(hc + 1, cnts + 1, sizeof(u64), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      nl = (int64_t)hc[1];
      std::swap(la, lb);
      if (nrem == 0 && nl == 0) break;
    }
    while (g.round_lo.size() > 1 && g.round_lo[g.round_lo.size() - 1] == g.round_lo[g.round_lo.size() - 2])
      g.round_lo.pop_back();  // drop trailing rounds that removed nothing
    g.nrounds = (int32_t)(g.round_lo.size() - 1);
    // roots, processing rounds from the last one
    for (int q = g.nrounds; q >= 1; q--) {
      int64_t a = g.round_lo[q - 1], b = g.round_lo[q];
      if (b > a) hipLaunchKernelGGL(k_peel_root, grid_for(b - a), B, 0, st, g.order + a, b - a, g.round, g.parent, g.root);
    }
    (void)hipFree(deg); (void)hipFree(cand); (void)hipFree(la); (void)hipFree(lb); (void)hipFree(cnts);
    (void)hipHostFree(hc);
    if (ocount > 0) {
      int64_t* cdeg = nullptr;
      TRY(dalloc(&cdeg, n + 1));
      hipLaunchKernelGGL(k_core_deg, G, B, 0, st, g.foff, g.fadj, n, g.round, cdeg);
      std::vector<int64_t> hd(n);
      HIP_TRY(hipMemcpyAsync(hd.data(), cdeg, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      hoff[0] = 0;
      for (int64_t v = 0; v < n; v++) hoff[v + 1] = hoff[v] + hd[v];
      g.m = hoff[n];
      TRY(dalloc(&g.off, n + 1, &acc));
      TRY(dalloc(&g.adj, g.m, &acc));
      HIP_TRY(hipMemcpyAsync(g.off, hoff.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_core_fill, G, B, 0, st, g.foff, g.fadj, n, g.round, g.off, g.adj);
      HIP_TRY(hipStreamSynchronize(st));
      (void)hipFree(cdeg);
    }
  }
  if (!g.off) { g.off = g.foff; g.adj = g.fadj; g.m = m; }
  g.toff = g.off; g.tadj = g.adj;
  if (directed && m > 0) {
    // transpose on the host: parents of v are the u with v in adj(u)
    std::vector<int64_t> to(n + 1, 0);
    for (int64_t e = 0; e < m; e++) to[adj[e] + 1]++;
    for (int64_t v = 0; v < n; v++) to[v + 1] += to[v];
    std::vector<int32_t> ta(m);
    std::vector<int64_t> cur(to.begin(), to.end() - 1);
    for (int64_t u = 0; u < n; u++)
      for (int64_t e = off[u]; e < off[u + 1]; e++) ta[cur[adj[e]]++] = (int32_t)u;
    g.toff = nullptr; g.tadj = nullptr;
    TRY(dalloc(&g.toff, n + 1, &acc));
    TRY(dalloc(&g.tadj, m, &acc));
    HIP_TRY(hipMemcpyAsync(g.toff, to.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(g.tadj, ta.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  g.h_off = hoff;
  hub_signature(n, off, adj, g.h_sig, &g.depth_hint);
  // first neighbour in descending-degree order (the smallest internal id once relabelled)
  // and the second one (INT32_MAX for vertices of degree < 2)
  g.h_topnb.assign(n, INT32_MAX);
  g.h_topnb2.assign(n, INT32_MAX);
  for (int64_t v = 0; v < n; v++)
    for (int64_t e = off[v]; e < off[v + 1]; e++) {
      const int32_t u = adj[e];
      if (u < g.h_topnb[v]) {
        g.h_topnb2[v] = g.h_topnb[v];
        g.h_topnb[v] = u;
      } else if (u < g.h_topnb2[v]) {
        g.h_topnb2[v] = u;
      }
    }
  g.h_round.assign(n, 0);
  if (heuristic) {
    HIP_TRY(hipMemcpyAsync(g.h_round.data(), g.round, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  // work records of the traversed graph (bottom-up) and of the uploaded graph (components)
  std::vector<int32_t> rv, rj;
  chunk_records(hoff, prm.chunk, rv, rj);
  g.nall = (int64_t)rv.size();
  TRY(dalloc(&g.all_v, g.nall, &acc));
  TRY(dalloc(&g.all_j, g.nall, &acc));
  if (g.nall) {
    HIP_TRY(hipMemcpyAsync(g.all_v, rv.data(), sizeof(int32_t) * g.nall, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(g.all_j, rj.data(), sizeof(int32_t) * g.nall, hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  {
    // packed bottom-up lists: low-degree vertices, and the records of the rest
    std::vector<int32_t> smv, lgv, lgj;
    for (size_t r = 0; r < rv.size(); r++) {
      const int64_t dg = hoff[rv[r] + 1] - hoff[rv[r]];
      if (dg <= PACK_DEG) smv.push_back(rv[r]);
      else { lgv.push_back(rv[r]); lgj.push_back(rj[r]); }
    }
    g.nsm = (int64_t)smv.size();
    g.nlg = (int64_t)lgv.size();
    TRY(dalloc(&g.sm_v, std::max<int64_t>(g.nsm, 1), &acc));
    TRY(dalloc(&g.lg_v, std::max<int64_t>(g.nlg, 1), &acc));
    TRY(dalloc(&g.lg_j, std::max<int64_t>(g.nlg, 1), &acc));
    if (g.nsm) HIP_TRY(hipMemcpyAsync(g.sm_v, smv.data(), sizeof(int32_t) * g.nsm, hipMemcpyHostToDevice, st));
    if (g.nlg) {
      HIP_TRY(hipMemcpyAsync(g.lg_v, lgv.data(), sizeof(int32_t) * g.nlg, hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(g.lg_j, lgj.data(), sizeof(int32_t) * g.nlg, hipMemcpyHostToDevice, st));
    }
    HIP_TRY(hipStreamSynchronize(st));
  }
  {
    std::vector<int32_t> sv;
    for (size_t r = 0; r < rv.size(); r++)
      if (rj[r] == 1) sv.push_back(rv[r]);
    g.nsplit = (int64_t)sv.size();
    TRY(dalloc(&g.split_v, std::max<int64_t>(g.nsplit, 1), &acc));
    if (g.nsplit) {
      HIP_TRY(hipMemcpyAsync(g.split_v, sv.data(), sizeof(int32_t) * g.nsplit, hipMemcpyHostToDevice, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  {
    std::vector<int64_t> fo(off, off + n + 1);
    std::vector<int32_t> fv, fj;
    chunk_records(fo, prm.chunk, fv, fj);
    int32_t *dv = nullptr, *dj = nullptr;
    int64_t nf = (int64_t)fv.size();
    TRY(dalloc(&dv, nf)); TRY(dalloc(&dj, nf));
    if (nf) {
      HIP_TRY(hipMemcpyAsync(dv, fv.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
      HIP_TRY(hipMemcpyAsync(dj, fj.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
    }
    // Connected components: lock-free union-find, labels = smallest id of
    // each component, verified against every arc (and redone until they agree).
    int32_t* par = nullptr;
    int* bad = nullptr;
    int* hbad = nullptr;
    TRY(dalloc(&par, n));
    TRY(dalloc(&bad, 1));
    HIP_TRY(hipHostMalloc((void**)&hbad, sizeof(int), hipHostMallocDefault));
    hipLaunchKernelGGL(k_iota, G, B, 0, st, par, n);
    hipLaunchKernelGGL(k_fill_u64, G, B, 0, st, (u64*)g.comp_size, 0ull, n);
    hipLaunchKernelGGL(k_fill_u64, G, B, 0, st, (u64*)g.comp_arcs, 0ull, n);
    int pass = 0;
    for (;; pass++) {
      if (nf) hipLaunchKernelGGL(k_cc_union, 2048, 256, 0, st, g.foff, g.fadj, dv, dj, nf, prm.chunk, par, directed);
      hipLaunchKernelGGL(k_cc_label, G, B, 0, st, par, n, g.label);
      hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, st, bad, 0, (int64_t)1);
      if (nf) hipLaunchKernelGGL(k_cc_check, 2048, 256, 0, st, g.foff, g.fadj, dv, dj, nf, prm.chunk, g.label, bad);
      HIP_TRY(hipMemcpyAsync(hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (!*hbad) break;
      if (pass >= 64) {
        (void)hipFree(par); (void)hipFree(bad); (void)hipHostFree(hbad);
        return set_error(BCMI_ERR_HIP, "component labelling did not converge");
      }
    }
    hipLaunchKernelGGL(k_cc_sizes, G, B, 0, st, g.label, n, g.off, g.comp_size, g.comp_arcs);
    HIP_TRY(hipStreamSynchronize(st));
    g.cc_passes = pass + 1;
    (void)hipFree(par); (void)hipFree(bad); (void)hipHostFree(hbad);
    (void)hipFree(dv); (void)hipFree(dj);
  }
  HIP_TRY(hipGetLastError());
  g.bytes = acc;
  return BCMI_OK;
}

// ------------------------------------------------------------ workspaces
struct Workspace {
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int mode = 0;
  int width = 64;
  int64_t n = 0;
  double* bc = nullptr;
  bool own_bc = true;
  Ctl* ctl = nullptr;
  Ctl* hctl = nullptr;  // pinned mirror
  int32_t* rec_v = nullptr;
  int32_t* rec_j = nullptr;
  u64* rec_mask = nullptr;
  int64_t cap = 0;
  int64_t* rec_lo = nullptr;
  int32_t* done = nullptr;
  int32_t* srcbuf = nullptr;
  double* multbuf = nullptr;
  int32_t* hsrc = nullptr;   // pinned
  double* hmult = nullptr;   // pinned
  // plain
  int32_t* depth = nullptr;
  u64* sigma1 = nullptr;
  double* delta1 = nullptr;
  double* coef1 = nullptr;
  // batched
  u64* seen = nullptr;
  u64* lm = nullptr;
  unsigned* lb = nullptr;
  int64_t nw = 0;
  u64* sigma = nullptr;
  double* coef = nullptr;
  u64* compmask = nullptr;
  u64* stats = nullptr;   // 3 x 64
  u64* hstats = nullptr;  // pinned
  double* dbg_delta = nullptr;
  int64_t* lvl_arcs = nullptr;  // [level] degree sum of the level's vertices
  double* lscale = nullptr;  // scaled path counts: per-level factors
  int64_t lscale_cap = 0;
  u64* lvmax = nullptr;
  // per-run source terms (peeling)
  u64* cnt0 = nullptr;
  u64* cntS = nullptr;
  u64* sa = nullptr;
  u64* srcK = nullptr;
  int32_t* srcdev = nullptr;
  int64_t srcdev_cap = 0;
  // captured level loops
  hipGraphExec_t gfwd = nullptr, gbwd = nullptr;      // large grids, both directions
  hipGraphExec_t gfwd_s = nullptr, gbwd_s = nullptr;  // small grids, top-down only (sparse levels)
  int cap_steps = 0;
  Params cap_prm;
  KT* kt = nullptr;             // device timer of the kernel selected by Params::ktime
  KT* hkt = nullptr;            // pinned mirror
  u64* kt_ws = nullptr;         // per-wave entry / exit stamps of the timed kernel
  u64* kt_we = nullptr;
  // Deep graphs (long chains): further workspaces, each with its own stream
  // and BC array, that run sweeps of the same call concurrently.
  std::vector<Workspace*> helpers;
  const DevGraph* helpers_graph = nullptr;
  // workgroup-per-source sweeps (deep graphs): per-slot state and per-run buffers
  int sig_layout = -1;  // row width code of the last batched sweep's path counts (2, 1, 0), -1 all zero
  int64_t wg_slots = 0, wg_lcap = 0;
  int64_t wg_req = 0;                 // slots asked for (wg_slots may be fewer: device memory)
  int32_t* wg_depth = nullptr;
  u64* wg_sigma = nullptr;
  int32_t* wg_queue = nullptr;
  int32_t* wg_lvl = nullptr;
  double* wg_scale = nullptr;
  double* wg_bc = nullptr;            // run result, folded into bc when the run succeeds
  unsigned long long* wg_next = nullptr;
  int64_t wg_cap = 0;                 // per-run buffers below hold wg_cap sources
  int32_t* wg_src = nullptr;
  double* wg_mult = nullptr;
  int64_t* wg_st = nullptr;           // 3 x wg_cap: n, m, a
  int32_t* wg_st2 = nullptr;          // 2 x wg_cap: levels, flag
  size_t wg_bytes = 0;
  // 2-degree sources (plain): saved forward states of their neighbours and
  // the per-level record counts of a combined state
  int32_t* d2_depth = nullptr;        // d2_slots x n
  u64* d2_sigma = nullptr;            // d2_slots x n
  int d2_slots = 0;
  int64_t* d2_lcnt = nullptr;
  int64_t d2_lcap = 0;
  const DevGraph* cap_graph = nullptr;
  size_t bytes = 0;
};

static GA make_ga(const DevGraph& g, const Params& prm) {
  GA a;
  a.off = g.off; a.adj = g.adj; a.toff = g.toff; a.tadj = g.tadj; a.directed = g.directed;
  a.label = g.label; a.comp_arcs = g.comp_arcs;
  a.all_v = g.all_v; a.all_j = g.all_j; a.nall = g.nall; a.split_v = g.split_v; a.nsplit = g.nsplit;
  a.sm_v = g.sm_v; a.nsm = g.nsm; a.lg_v = g.lg_v; a.lg_j = g.lg_j; a.nlg = g.nlg;
  a.w = g.heuristic ? (const uint64_t*)g.w : nullptr;
  a.chunk = g.chunk; a.n = g.n; a.m = g.m;
  a.flags = prm.kflags;
  return a;
}
static PA make_pa(Workspace* ws, const Params& prm) {
  PA p;
  p.ctl = ws->ctl; p.depth = ws->depth; p.sigma = ws->sigma1; p.delta = ws->delta1; p.coef = ws->coef1;
  p.done = ws->done; p.rec_v = ws->rec_v; p.rec_j = ws->rec_j; p.cap = ws->cap; p.rec_lo = ws->rec_lo;
  p.srcbuf = ws->srcbuf; p.multbuf = ws->multbuf; p.bc = ws->bc; p.alpha = prm.alpha_plain; p.beta = prm.beta_plain;
  return p;
}
static BA make_ba(Workspace* ws, const Params& prm) {
  BA b;
  b.ctl = ws->ctl; b.seen = ws->seen; b.lm = ws->lm; b.lb = ws->lb; b.nw = ws->nw; b.sigma = ws->sigma; b.coef = ws->coef; b.done = ws->done;
  b.rec_v = ws->rec_v; b.rec_j = ws->rec_j; b.rec_mask = ws->rec_mask; b.cap = ws->cap; b.rec_lo = ws->rec_lo;
  b.compmask = ws->compmask; b.srcbuf = ws->srcbuf; b.multbuf = ws->multbuf; b.bc = ws->bc;
  b.st_n = ws->stats; b.st_m = ws->stats + 64; b.st_a = ws->stats + 128; b.dbg_delta = ws->dbg_delta;
  b.alpha = prm.alpha_batched; b.beta = prm.beta_batched; b.gamma = prm.gamma_push; b.fuse = prm.fuse_bu;
  b.gamma_b = prm.gamma_back; b.lvl_arcs = ws->lvl_arcs; b.pack = prm.pack_bu; b.n = ws->n;
  b.lscale = ws->lscale; b.lscale_cap = ws->lscale_cap; b.lvmax = ws->lvmax;
  return b;
}

static int64_t record_capacity(const DevGraph& g, int mode) {
  // Every vertex appears at most once per distinct depth; plain sweeps at
  // most once.  A batch's sources sit at up to 64 distinct distances from a
  // vertex: few on shallow graphs (their depths coincide), up to all 64 on
  // deep ones (long chains).  Start with room for depth/2 depths per vertex
  // (the hubs' BFS depth, at least 4); a sweep that overflows grows it.
  int64_t base = g.n + g.m / g.chunk + 1024;
  int64_t mult = std::min<int64_t>(64, std::max<int64_t>(4, g.depth_hint / 2));
  return mode == BCMI_MODE_PLAIN ? base : mult * base;
}

// Sweeps of one call that run at once, each on its own workspace and
// stream.  Deep graphs (hub BFS depth >= 64) have small, latency-bound levels,
// so several sweeps overlap well (`streams`).  On shallow graphs a second
// sweep fills what one leaves idle (its small first and last levels, the
// kernel boundaries, the tail of each level): R-MAT 22 batches 1,245 -> 1,367
// GTEPS with two, R-MAT 18 + trees 733 -> 956.
static int concurrent_sweeps(const DevGraph& g, const Params& prm) {
  if (prm.concurrency > 0) return std::min(prm.concurrency, 16);
  return g.depth_hint >= 64 ? std::max(1, std::min(prm.streams, 16)) : 2;
}

static bool use_wg(const DevGraph& g, const Params& prm);
static size_t wg_bytes_estimate(const DevGraph& g, const Params& prm);

size_t workspace_bytes(const DevGraph& g, int mode, int width, const Params& prm) {
  (void)width;
  const bool wg = use_wg(g, prm);
  const size_t copies = wg ? 1 : (size_t)concurrent_sweeps(g, prm);
  int64_t n = g.n;
  int64_t cap = record_capacity(g, mode);
  size_t b = sizeof(double) * n + sizeof(int64_t) * (n + 3) + sizeof(int32_t) * n + 2 * sizeof(int32_t) * cap;
  b += 4 * sizeof(u64) * n;  // per-run source terms
  if (mode == BCMI_MODE_PLAIN) b += (sizeof(int32_t) + sizeof(u64) + 2 * sizeof(double)) * n;
  else b += sizeof(u64) * cap + (5 * sizeof(u64) + 64 * (sizeof(u64) + sizeof(double))) * n;
  return b * copies + (wg ? wg_bytes_estimate(g, prm) : 0);
}

static void free_graphs(Workspace* ws) {
  for (hipGraphExec_t* x : {&ws->gfwd, &ws->gbwd, &ws->gfwd_s, &ws->gbwd_s}) {
    if (*x) (void)hipGraphExecDestroy(*x);
    *x = nullptr;
  }
  ws->cap_graph = nullptr;
}

__global__ void k_kt_reset(KT* kt, u64* ws, u64* we) {
  kt->total = 0;
  kt->launches = 0;
  kt->wstart = ws;
  kt->wend = we;
}

void workspace_destroy(Workspace* ws) {
  ALLOC_LOCK;
  if (!ws) return;
  for (Workspace* h : ws->helpers) workspace_destroy(h);
  ws->helpers.clear();
  free_graphs(ws);
  {
    auto fw = [](void* p) { if (p) (void)hipFree(p); };
    fw(ws->wg_depth); fw(ws->wg_sigma); fw(ws->wg_queue); fw(ws->wg_lvl); fw(ws->wg_scale);
    fw(ws->wg_bc); fw(ws->wg_next); fw(ws->wg_src); fw(ws->wg_mult); fw(ws->wg_st); fw(ws->wg_st2);
  }
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  if (ws->own_bc) fr(ws->bc);
  fr(ws->ctl); fr(ws->rec_v); fr(ws->rec_j); fr(ws->rec_mask); fr(ws->rec_lo); fr(ws->done);
  fr(ws->srcbuf); fr(ws->multbuf); fr(ws->depth); fr(ws->sigma1); fr(ws->delta1); fr(ws->coef1);
  fr(ws->seen); fr(ws->lm); fr(ws->lb); fr(ws->lscale); fr(ws->lvmax); fr(ws->sigma); fr(ws->coef); fr(ws->compmask); fr(ws->stats); fr(ws->dbg_delta); fr(ws->lvl_arcs);
  fr(ws->cnt0); fr(ws->cntS); fr(ws->sa); fr(ws->srcK); fr(ws->srcdev);
  fr(ws->d2_depth); fr(ws->d2_sigma); fr(ws->d2_lcnt);
  if (ws->hctl) (void)hipHostFree(ws->hctl);
  if (ws->hkt) (void)hipHostFree(ws->hkt);
  if (ws->kt) (void)hipFree(ws->kt);
  if (ws->kt_ws) (void)hipFree(ws->kt_ws);
  if (ws->kt_we) (void)hipFree(ws->kt_we);
  if (ws->hsrc) (void)hipHostFree(ws->hsrc);
  if (ws->hmult) (void)hipHostFree(ws->hmult);
  if (ws->hstats) (void)hipHostFree(ws->hstats);
  if (ws->ev0) (void)hipEventDestroy(ws->ev0);
  if (ws->ev1) (void)hipEventDestroy(ws->ev1);
  if (ws->st) (void)hipStreamDestroy(ws->st);
  delete ws;
}

static int alloc_records(Workspace* ws, int64_t cap) {
  ALLOC_LOCK;
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  fr(ws->rec_v); fr(ws->rec_j); fr(ws->rec_mask);
  ws->rec_v = nullptr; ws->rec_j = nullptr; ws->rec_mask = nullptr;
  ws->cap = cap;
  TRY(dalloc(&ws->rec_v, cap, &ws->bytes));
  TRY(dalloc(&ws->rec_j, cap, &ws->bytes));
  if (ws->mode == BCMI_MODE_BATCHED) TRY(dalloc(&ws->rec_mask, cap, &ws->bytes));
  free_graphs(ws);
  return BCMI_OK;
}

static int ws_init(Workspace* ws, const DevGraph& g, int mode, int width, const Params& prm) {
  int64_t n = g.n;
  ws->mode = mode; ws->width = width; ws->n = n;
  HIP_TRY(hipStreamCreateWithFlags(&ws->st, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&ws->ev0));
  HIP_TRY(hipEventCreate(&ws->ev1));
  size_t* acc = &ws->bytes;
  TRY(dalloc(&ws->bc, n, acc));
  TRY(dalloc(&ws->ctl, 1, acc));
  HIP_TRY(hipHostMalloc((void**)&ws->hctl, sizeof(Ctl), hipHostMallocDefault));
  TRY(dalloc(&ws->kt, 1, acc));
  TRY(dalloc(&ws->kt_ws, KT_MAX_WAVES, acc));
  TRY(dalloc(&ws->kt_we, KT_MAX_WAVES, acc));
  HIP_TRY(hipHostMalloc((void**)&ws->hkt, sizeof(KT), hipHostMallocDefault));
  TRY(dalloc(&ws->rec_lo, n + 3, acc));
  TRY(dalloc(&ws->done, n, acc));
  TRY(dalloc(&ws->srcbuf, 64, acc));
  TRY(dalloc(&ws->multbuf, 64, acc));
  HIP_TRY(hipHostMalloc((void**)&ws->hsrc, 64 * sizeof(int32_t), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&ws->hmult, 64 * sizeof(double), hipHostMallocDefault));
  TRY(dalloc(&ws->cnt0, n, acc));
  TRY(dalloc(&ws->cntS, n, acc));
  TRY(dalloc(&ws->sa, n, acc));
  TRY(dalloc(&ws->srcK, n, acc));
  int G = grid_for(n);
  hipStream_t st = ws->st;
  hipLaunchKernelGGL(k_fill_f64, G, 256, 0, st, ws->bc, 0.0, n);
  hipLaunchKernelGGL(k_fill_i32, G, 256, 0, st, ws->done, 0, n);
  hipLaunchKernelGGL(k_kt_reset, 1, 1, 0, st, ws->kt, ws->kt_ws, ws->kt_we);
  hipLaunchKernelGGL(k_fill_u64, grid_for(n + 3), 256, 0, st, (u64*)ws->rec_lo, 0ull, n + 3);
  if (mode == BCMI_MODE_PLAIN) {
    TRY(dalloc(&ws->depth, n, acc));
    TRY(dalloc(&ws->sigma1, n, acc));
    TRY(dalloc(&ws->delta1, n, acc));
    TRY(dalloc(&ws->coef1, n, acc));
    hipLaunchKernelGGL(k_fill_i32, G, 256, 0, st, ws->depth, -1, n);
    hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->sigma1, 0ull, n);
    hipLaunchKernelGGL(k_fill_f64, G, 256, 0, st, ws->delta1, 0.0, n);
    hipLaunchKernelGGL(k_fill_f64, G, 256, 0, st, ws->coef1, 0.0, n);
  } else {
    TRY(dalloc(&ws->seen, n, acc));
    TRY(dalloc(&ws->lm, 3 * n, acc));
    TRY(dalloc(&ws->lvl_arcs, n + 3, acc));
    // a sweep has at most 2 x (hub BFS depth) + 1 levels within the hubs'
    // component; deeper components elsewhere grow the table on demand
    ws->lscale_cap = std::min<int64_t>(n + 2, std::max<int64_t>(4096, 2 * (int64_t)g.depth_hint + 8));
    TRY(dalloc(&ws->lscale, (size_t)ws->lscale_cap * 64, acc));
    TRY(dalloc(&ws->lvmax, 64, acc));
    ws->nw = (n + 31) / 32;
    TRY(dalloc(&ws->lb, 3 * ws->nw, acc));
    hipLaunchKernelGGL(k_fill_i32, grid_for(3 * ws->nw), 256, 0, st, (int32_t*)ws->lb, 0, 3 * ws->nw);
    TRY(dalloc(&ws->sigma, 64 * n, acc));
    TRY(dalloc(&ws->coef, 64 * n, acc));
    TRY(dalloc(&ws->compmask, n, acc));
    TRY(dalloc(&ws->stats, 3 * 64, acc));
    HIP_TRY(hipHostMalloc((void**)&ws->hstats, 3 * 64 * sizeof(u64), hipHostMallocDefault));
    hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->seen, 0ull, n);
    hipLaunchKernelGGL(k_fill_u64, grid_for(3 * n), 256, 0, st, ws->lm, 0ull, 3 * n);
    hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->compmask, 0ull, n);
    hipLaunchKernelGGL(k_fill_u64, grid_for(64 * n), 256, 0, st, ws->sigma, 0ull, 64 * n);
  }
  TRY(alloc_records(ws, record_capacity(g, mode)));
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipGetLastError());
  return BCMI_OK;
}

Workspace* workspace_create(const DevGraph& g, int mode, int width, const Params& prm, int* rc) {
  ALLOC_LOCK;
  Workspace* ws = new Workspace();
  *rc = ws_init(ws, g, mode, width, prm);
  if (*rc != BCMI_OK) {
    workspace_destroy(ws);
    return nullptr;
  }
  return ws;
}
hipStream_t workspace_stream(Workspace* ws) { return ws->st; }
size_t workspace_size(Workspace* ws) {
  size_t b = ws->bytes + ws->wg_bytes;
  for (Workspace* h : ws->helpers) b += h->bytes;
  return b;
}
void workspace_share_bc(Workspace* ws, double* bc) {
  ALLOC_LOCK;
  if (ws->own_bc && ws->bc) (void)hipFree(ws->bc);
  ws->bc = bc;
  ws->own_bc = false;
  free_graphs(ws);
}
double* workspace_bc(Workspace* ws) { return ws->bc; }
int workspace_bc_reset(Workspace* ws) {
  hipLaunchKernelGGL(k_fill_f64, grid_for(ws->n), 256, 0, ws->st, ws->bc, 0.0, ws->n);
  HIP_TRY(hipStreamSynchronize(ws->st));
  return BCMI_OK;
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

const char* step_kernel_name(int k) {
  static const char* names[] = {"k_plain_decide", "k_plain_td", "k_plain_bu", "k_bdecide", "k_plain_back",
                                "k_plain_bfin", "k_ms_decide", "k_ms_td", "k_ms_bu", "k_ms_bu_pk", "k_ms_sigma",
                                "k_ms_push", "k_ms_bdecide", "k_ms_back", "k_ms_bpush", "k_ms_bfin", "k_kt_accum",
                                "k_kt_fold"};
  static_assert(sizeof(names) / sizeof(names[0]) == SK_COUNT, "one name per step kernel");
  return k >= 0 && k < SK_COUNT ? names[k] : "?";
}

// The kernel nodes of one level step of a captured graph (the graph holds
// Params::steps copies).  graph: 0 forward, 1 backward, 2 and 3 the same on
// sparse levels (small grid, top-down only).  The kernel selected by
// Params::ktime is followed by the two timer nodes that fold its per-wave
// stamps (k_kt_accum over grid x block / 64 waves, then k_kt_fold).
std::vector<StepNode> step_plan(const Params& prm, bool plain, int graph) {
  const bool fwd = (graph & 1) == 0;
  const bool sparse = graph >= 2;
  const int B = prm.block;
  const int Gk = sparse ? prm.sgrid : plain ? prm.grid_plain : fwd || prm.grid_back <= 0 ? prm.grid : prm.grid_back;
  std::vector<StepNode> v;
  auto add = [&](int k, int grid, int block, int64_t aux, int timer_id) {
    const bool timed = timer_id != 0 && prm.ktime == timer_id;
    v.push_back({k, grid, block, aux, timed});
    if (timed) {
      v.push_back({SK_KT_ACCUM, KT_PARTS, 1024, (int64_t)grid * block / 64, false});
      v.push_back({SK_KT_FOLD, 1, KT_PARTS, 0, false});
    }
  };
  if (plain && fwd) {
    add(SK_PLAIN_DECIDE, 1, 1, sparse ? 0 : 1, 0);
    add(SK_PLAIN_TD, Gk, B, 0, 5);
    if (!sparse) add(SK_PLAIN_BU, Gk, B, 0, 6);
  } else if (plain) {
    add(SK_BDECIDE, 1, 1, 0, 0);
    add(SK_PLAIN_BACK, Gk, B, 0, 7);
    add(SK_PLAIN_BFIN, Gk, B, 0, 0);
  } else if (fwd) {
    add(SK_MS_DECIDE, 1, 1, sparse ? 0 : 1, 0);
    add(SK_MS_TD, Gk, B, 0, 3);
    if (!sparse) add(SK_MS_BU, Gk, B, 0, 4);
    if (!sparse && prm.pack_bu) add(SK_MS_BU_PK, Gk, B, 0, 4);
    add(SK_MS_SIGMA, Gk, B, 0, 1);
    add(SK_MS_PUSH, Gk, B, 0, 0);
  } else {
    add(SK_MS_BDECIDE, 1, 1, 0, 0);
    add(SK_MS_BACK, Gk, B, 0, 2);
    // with pushed levels the backward work is shared by three kernels, so
    // the backward timer covers all of them
    if (prm.gamma_back > 0) {
      add(SK_MS_BPUSH, Gk, B, 0, 2);
      add(SK_MS_BFIN, Gk, B, 0, 2);
    } else {
      add(SK_MS_BFIN, Gk, B, 0, 0);
    }
  }
  return v;
}

// Launches `steps` levels of graph `which` (0 forward, 1 backward, 2 and 3
// their small-grid forms) on the workspace's stream, each step as step_plan
// lists it.  Captured into graphs by capture(); launched as it is when the
// level loop runs without graphs.
static void launch_steps(const DevGraph& g, Workspace* ws, const Params& prm, int which) {
  const GA ga = make_ga(g, prm);
  const PA pa = make_pa(ws, prm);
  const BA ba = make_ba(ws, prm);
  hipStream_t st = ws->st;
  const bool plain = ws->mode == BCMI_MODE_PLAIN;
  KT* const T = ws->kt;
  const std::vector<StepNode> plan = step_plan(prm, plain, which);
  for (int s = 0; s < prm.steps; s++) {
    for (const StepNode& nd : plan) {
      KT* const kt = nd.timed ? T : nullptr;
      const dim3 G(nd.grid), B(nd.block);
      switch (nd.kernel) {
        case SK_PLAIN_DECIDE: hipLaunchKernelGGL(k_plain_decide, G, B, 0, st, ga, pa, (int)nd.aux); break;
        case SK_PLAIN_TD: hipLaunchKernelGGL(k_plain_td, G, B, 0, st, ga, pa, kt); break;
        case SK_PLAIN_BU: hipLaunchKernelGGL(k_plain_bu, G, B, 0, st, ga, pa, kt); break;
        case SK_BDECIDE: hipLaunchKernelGGL(k_bdecide, G, B, 0, st, pa.ctl); break;
        case SK_PLAIN_BACK: hipLaunchKernelGGL(k_plain_back, G, B, 0, st, ga, pa, kt); break;
        case SK_PLAIN_BFIN: hipLaunchKernelGGL(k_plain_bfin, G, B, 0, st, ga, pa); break;
        case SK_MS_DECIDE: hipLaunchKernelGGL(k_ms_decide, G, B, 0, st, ga, ba, (int)nd.aux); break;
        case SK_MS_TD: hipLaunchKernelGGL(k_ms_td, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_BU: hipLaunchKernelGGL(k_ms_bu, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_BU_PK: hipLaunchKernelGGL(k_ms_bu_pk, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_SIGMA: hipLaunchKernelGGL(k_ms_sigma, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_PUSH: hipLaunchKernelGGL(k_ms_push, G, B, 0, st, ga, ba); break;
        case SK_MS_BDECIDE: hipLaunchKernelGGL(k_ms_bdecide, G, B, 0, st, ga, ba); break;
        case SK_MS_BACK: hipLaunchKernelGGL(k_ms_back, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_BPUSH: hipLaunchKernelGGL(k_ms_bpush, G, B, 0, st, ga, ba, kt); break;
        case SK_MS_BFIN: hipLaunchKernelGGL(k_ms_bfin, G, B, 0, st, ga, ba, kt); break;
        case SK_KT_ACCUM: hipLaunchKernelGGL(k_kt_accum, G, B, 0, st, T, nd.aux); break;
        case SK_KT_FOLD: hipLaunchKernelGGL(k_kt_fold, G, B, 0, st, T); break;
        default: break;
      }
    }
  }
}

// Captures `steps` forward levels and `steps` backward levels into graphs
// (kernel nodes only).
static int capture(const DevGraph& g, Workspace* ws, const Params& prm) {
  ALLOC_LOCK;
  free_graphs(ws);
  hipStream_t st = ws->st;
  // Four graphs: {large grid, small grid} x {forward, backward}.  The small
  // one serves levels with few vertices (long chains): its launches are cheap
  // and it never considers bottom-up (which scans every vertex).
  for (int which = 0; which < 4; which++) {
    hipGraph_t graph;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_steps(g, ws, prm, which);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t ex;
    e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    (which == 0 ? ws->gfwd : which == 1 ? ws->gbwd : which == 2 ? ws->gfwd_s : ws->gbwd_s) = ex;
  }
  ws->cap_steps = prm.steps;
  ws->cap_prm = prm;
  ws->cap_graph = &g;
  return BCMI_OK;
}

static bool same_launch(const Params& a, const Params& b) {
  return a.block == b.block && a.grid == b.grid && a.grid_plain == b.grid_plain && a.grid_back == b.grid_back &&
         a.chunk == b.chunk &&
         a.alpha_plain == b.alpha_plain &&
         a.alpha_batched == b.alpha_batched && a.beta_plain == b.beta_plain && a.beta_batched == b.beta_batched &&
         a.gamma_push == b.gamma_push && a.gamma_back == b.gamma_back && a.fuse_bu == b.fuse_bu && a.pack_bu == b.pack_bu &&
         a.steps == b.steps && a.ktime == b.ktime &&
         a.kflags == b.kflags && a.sgrid == b.sgrid;
}

// One replay of graph `which` (0 forward, 1 backward, 2 and 3 their
// small-grid forms): the captured graph, or with prm.graphs == 0 the same
// kernels launched one by one.
static int replay(const DevGraph& g, Workspace* ws, const Params& prm, int which) {
  if (!prm.graphs) {
    launch_steps(g, ws, prm, which);
    HIP_TRY(hipGetLastError());
    return BCMI_OK;
  }
  hipGraphExec_t ex = which == 0 ? ws->gfwd : which == 1 ? ws->gbwd : which == 2 ? ws->gfwd_s : ws->gbwd_s;
  HIP_TRY(hipGraphLaunch(ex, ws->st));
  return BCMI_OK;
}

// Backward levels of a sweep whose forward pass has finished (the host
// mirror of the control block is current), replayed until the control block
// reports the pass done.
static int backward_loop(const DevGraph& g, Workspace* ws, const Params& prm) {
  if (ws->hctl->bdone) return BCMI_OK;
  hipStream_t st = ws->st;
  // pick the grid of each backward replay from the sizes of its levels
  int last = ws->hctl->last;
  std::vector<int64_t> lo(last + 2);
  HIP_TRY(hipMemcpyAsync(lo.data(), ws->rec_lo, sizeof(int64_t) * (last + 2), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  while (!ws->hctl->bdone) {
    int top = ws->hctl->bstarted ? ws->hctl->blevel - 1 : last;
    int bot = std::max(0, top - prm.steps + 1);
    int64_t recs = lo[std::min(top + 1, last + 1)] - lo[bot];
    bool sparse = recs < prm.sparse_verts * (int64_t)(top - bot + 1);
    TRY(replay(g, ws, prm, sparse ? 3 : 1));
    HIP_TRY(hipMemcpyAsync(ws->hctl, ws->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return BCMI_OK;
}

// Runs one sweep (plain: srcbuf[0]; batched: srcbuf[0..k)) whose inputs are
// already in the workspace buffers.  keep != 0 skips the cleanup (debug
// readback), the caller then calls sweep_cleanup.
static int sweep(const DevGraph& g, Workspace* ws, const Params& prm, int k, bool keep, int* levels, int scaled = 0,
                 int narrow = 0) {
  if (prm.graphs && (!ws->gfwd || ws->cap_graph != &g || !same_launch(ws->cap_prm, prm))) TRY(capture(g, ws, prm));
  GA ga = make_ga(g, prm);
  hipStream_t st = ws->st;
  for (int attempt = 0;; attempt++) {
    if (ws->mode == BCMI_MODE_PLAIN) {
      PA pa = make_pa(ws, prm);
      hipLaunchKernelGGL(k_plain_init, 1, 1, 0, st, ga, pa);
    } else {
      // Split vertices' rows are zero between sweeps in the layout in use;
      // a change of layout starts from an all-zero array.
      const int lay = scaled ? 0 : narrow;
      if (ws->sig_layout >= 0 && ws->sig_layout != lay)
        hipLaunchKernelGGL(k_fill_u64, grid_for(64 * ws->n), 256, 0, st, ws->sigma, 0ull, 64 * ws->n);
      ws->sig_layout = lay;
      BA ba = make_ba(ws, prm);
      hipLaunchKernelGGL(k_ms_init, 1, 64, 0, st, ga, ba, k, scaled, narrow);
    }
    if (prm.trace) {
      // Diagnostic: one graph launch at a time, level statistics on stderr.
      HIP_TRY(hipStreamSynchronize(st));
      for (int it = 0; it < 100000; it++) {
        double t0 = now_ms();
        TRY(replay(g, ws, prm, 0));
        HIP_TRY(hipMemcpyAsync(ws->hctl, ws->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const Ctl& c = *ws->hctl;
        fprintf(stderr, "[trace] fwd level %d mode %s verts %lld arcs %lld pending-arcs %lld %.3f ms%s\n", c.level,
                c.mode ? "BU" : "TD", (long long)c.verts_cur, (long long)c.arcs_cur,
                (long long)(c.arcs_reach - c.arcs_seen), now_ms() - t0, c.fdone ? " (done)" : "");
        if (c.fdone) break;
      }
      for (int it = 0; it < 100000; it++) {
        double t0 = now_ms();
        TRY(replay(g, ws, prm, 1));
        HIP_TRY(hipMemcpyAsync(ws->hctl, ws->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        fprintf(stderr, "[trace] bwd level %d %.3f ms\n", ws->hctl->blevel, now_ms() - t0);
        if (ws->hctl->bdone) break;
      }
    }
    TRY(replay(g, ws, prm, 0));
    TRY(replay(g, ws, prm, 1));
    HIP_TRY(hipMemcpyAsync(ws->hctl, ws->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    while (!ws->hctl->fdone && !ws->hctl->overflow) {
      // small levels (long chains) replay the small-grid graph
      int64_t front = std::max<int64_t>(ws->hctl->verts_cur, ws->hctl->verts_next);
      TRY(replay(g, ws, prm, front < prm.sparse_verts ? 2 : 0));
      HIP_TRY(hipMemcpyAsync(ws->hctl, ws->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    if (ws->hctl->overflow) {
      // Work list too small: reset every state array and retry with four times the room.
      int64_t n = ws->n;
      int G = grid_for(n);
      if (ws->mode == BCMI_MODE_PLAIN) {
        hipLaunchKernelGGL(k_fill_i32, G, 256, 0, st, ws->depth, -1, n);
        hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->sigma1, 0ull, n);
      } else {
        hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->seen, 0ull, n);
        hipLaunchKernelGGL(k_fill_u64, grid_for(3 * n), 256, 0, st, ws->lm, 0ull, 3 * n);
        hipLaunchKernelGGL(k_fill_i32, grid_for(3 * ws->nw), 256, 0, st, (int32_t*)ws->lb, 0, 3 * ws->nw);
        hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->compmask, 0ull, n);
        hipLaunchKernelGGL(k_fill_u64, grid_for(64 * n), 256, 0, st, ws->sigma, 0ull, 64 * n);
        ws->sig_layout = -1;
      }
      hipLaunchKernelGGL(k_fill_i32, G, 256, 0, st, ws->done, 0, n);
      HIP_TRY(hipStreamSynchronize(st));
      if (attempt > 40) return set_error(BCMI_ERR_NOMEM, "work list keeps overflowing");
      if (ws->hctl->overflow == 2) {  // more levels than the scale table holds
        ALLOC_LOCK;
        int64_t cap = std::min<int64_t>(ws->n + 2, ws->lscale_cap * 4);
        (void)hipFree(ws->lscale);
        ws->lscale = nullptr;
        TRY(dalloc(&ws->lscale, (size_t)cap * 64, &ws->bytes));
        ws->lscale_cap = cap;
      } else {
        TRY(alloc_records(ws, ws->cap * 4));
      }
      if (prm.graphs) TRY(capture(g, ws, prm));
      continue;
    }
    TRY(backward_loop(g, ws, prm));
    break;
  }
  if (levels) *levels = ws->hctl->last + 1;
  if (!keep) {
    if (ws->mode == BCMI_MODE_PLAIN) {
      PA pa = make_pa(ws, prm);
      hipLaunchKernelGGL(k_plain_cleanup, grid_for(std::min<int64_t>(ws->hctl->rec_count, ws->cap)), 256, 0, st, pa);
    } else {
      BA ba = make_ba(ws, prm);
      hipLaunchKernelGGL(k_ms_cleanup, grid_for(std::max(std::min<int64_t>(ws->hctl->rec_count, ws->cap), 64 * g.nsplit)),
                         256, 0, st, ga, ba);
    }
  }
  return BCMI_OK;
}

static int sweep_cleanup(const DevGraph& g, Workspace* ws, const Params& prm) {
  GA ga = make_ga(g, prm);
  if (ws->mode == BCMI_MODE_PLAIN) {
    PA pa = make_pa(ws, prm);
    hipLaunchKernelGGL(k_plain_cleanup, grid_for(std::min<int64_t>(ws->hctl->rec_count, ws->cap)), 256, 0, ws->st, pa);
  } else {
    BA ba = make_ba(ws, prm);
    hipLaunchKernelGGL(k_ms_cleanup, grid_for(std::max(std::min<int64_t>(ws->hctl->rec_count, ws->cap), 64 * g.nsplit)),
                       256, 0, ws->st, ga, ba);
  }
  HIP_TRY(hipStreamSynchronize(ws->st));
  return BCMI_OK;
}

// Source preparation.  Without peeling every listed source of degree > 0 is
// swept with multiplicity 1.  With peeling (or expansion) sources are first
// pushed down to the core roots of their trees.
static int prepare_sources(const DevGraph& g, Workspace* ws, const int32_t* src, int64_t nsrc, int src_flags,
                           std::vector<int32_t>& cs, std::vector<double>& cm, bool& finalize) {
  cs.clear(); cm.clear();
  finalize = false;
  const std::vector<int64_t>& ho = g.h_off;
  for (int64_t i = 0; i < nsrc; i++)
    if (src[i] < 0 || src[i] >= g.n) return set_error(BCMI_ERR_RANGE, "source id out of range");
  std::vector<int32_t> mapped;
  if (g.relabeled) {
    mapped.resize(nsrc);
    for (int64_t i = 0; i < nsrc; i++) mapped[i] = g.h_o2n[src[i]];
    src = mapped.data();
  }
  if (!g.heuristic) {
    for (int64_t i = 0; i < nsrc; i++)
      if (ho[src[i] + 1] > ho[src[i]]) { cs.push_back(src[i]); cm.push_back(1.0); }
    return BCMI_OK;
  }
  finalize = true;
  hipStream_t st = ws->st;
  int64_t n = g.n;
  if (nsrc > ws->srcdev_cap) {
    ALLOC_LOCK;
    if (ws->srcdev) (void)hipFree(ws->srcdev);
    ws->srcdev = nullptr;
    TRY(dalloc(&ws->srcdev, nsrc));
    ws->srcdev_cap = nsrc;
  }
  if (nsrc) HIP_TRY(hipMemcpyAsync(ws->srcdev, src, sizeof(int32_t) * nsrc, hipMemcpyHostToDevice, st));
  int G = grid_for(n);
  hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->cnt0, 0ull, n);
  hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, ws->srcK, 0ull, n);
  if (nsrc) hipLaunchKernelGGL(k_src_count, grid_for(nsrc), 256, 0, st, ws->srcdev, nsrc, ws->cnt0);
  if (src_flags & BCMI_SRC_EXPAND) {
    hipLaunchKernelGGL(k_src_expand, G, 256, 0, st, g.root, g.round, n, ws->cnt0, ws->cntS);
    HIP_TRY(hipMemcpyAsync(ws->cnt0, ws->cntS, sizeof(u64) * n, hipMemcpyDeviceToDevice, st));
  }
  hipLaunchKernelGGL(k_src_init, G, 256, 0, st, ws->cnt0, g.label, n, ws->cntS, ws->sa, ws->srcK);
  for (int q = 1; q <= g.nrounds; q++) {
    int64_t a = g.round_lo[q - 1], b = g.round_lo[q];
    if (b > a) hipLaunchKernelGGL(k_tree_acc, grid_for(b - a), 256, 0, st, g.order + a, b - a, g.parent, (const u64*)g.w, ws->cntS, ws->sa);
  }
  std::vector<u64> hc(n);
  HIP_TRY(hipMemcpyAsync(hc.data(), ws->cntS, sizeof(u64) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // core roots in increasing id order; pruned vertices have no traversed arcs
  const std::vector<int32_t>& rnd = g.h_round;
  for (int64_t v = 0; v < n; v++)
    if (rnd[v] == 0 && hc[v] > 0 && ho[v + 1] > ho[v]) { cs.push_back((int32_t)v); cm.push_back((double)hc[v]); }
  return BCMI_OK;
}

// Statistics and outcome of the sweeps one thread drove.
struct RunPart {
  RunResult R;
  int ovf = 0;
  int rc = BCMI_OK;
  std::string msg;
};

// One plain sweep from source s (internal id) with BC multiplier mult.
// Records the statistics of the plain sweep just finished (host mirror of
// its control block) that took t_ms and reached `lv` levels.
static void plain_stats(const Ctl& c, int lv, double t_ms, RunPart& P) {
  RunResult& R = P.R;
  R.n_s.push_back(c.tot_verts);
  R.m_s.push_back(c.tot_arcs);
  R.a_s.push_back(c.sum_a);
  R.n_td += c.n_td; R.n_bu += c.n_bu;
  R.sweeps++; R.launches++;
  R.max_levels = std::max<int64_t>(R.max_levels, lv);
  R.levels.push_back(lv);
  R.t_s.push_back(t_ms);
  if (c.sigma_ovf) P.ovf = 1;
}

// One plain sweep from source s (internal id) with BC multiplier mult.  With
// keep the sweep's depths and path counts stay in the workspace (the caller
// saves them, then calls sweep_cleanup); *levels gets its level count.
static int run_plain_one(const DevGraph& g, Workspace* w, const Params& prm, int32_t s, double mult, RunPart& P,
                         bool keep = false, int* levels = nullptr) {
  const double t0 = now_ms();
  hipStream_t st = w->st;
  w->hsrc[0] = s;
  w->hmult[0] = mult;
  HIP_TRY(hipMemcpyAsync(w->srcbuf, w->hsrc, sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(w->multbuf, w->hmult, sizeof(double), hipMemcpyHostToDevice, st));
  int lv = 0;
  TRY(sweep(g, w, prm, 1, keep, &lv));
  if (levels) *levels = lv;
  plain_stats(*w->hctl, lv, now_ms() - t0, P);
  return BCMI_OK;
}

// Saved forward states of 2-degree sources' neighbours: slot i holds depths
// at d2_depth + i*n and path counts at d2_sigma + i*n.
static int d2_reserve(Workspace* ws, int slots, int64_t nlev) {
  ALLOC_LOCK;
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  if (slots > ws->d2_slots) {
    fr(ws->d2_depth); fr(ws->d2_sigma);
    ws->d2_depth = nullptr; ws->d2_sigma = nullptr; ws->d2_slots = 0;
    TRY(dalloc(&ws->d2_depth, (size_t)slots * ws->n, &ws->bytes));
    TRY(dalloc(&ws->d2_sigma, (size_t)slots * ws->n, &ws->bytes));
    ws->d2_slots = slots;
  }
  if (nlev > ws->d2_lcap) {
    fr(ws->d2_lcnt);
    ws->d2_lcnt = nullptr; ws->d2_lcap = 0;
    const int64_t cap = std::max<int64_t>(nlev, 256);
    TRY(dalloc(&ws->d2_lcnt, (size_t)cap, &ws->bytes));
    ws->d2_lcap = cap;
  }
  return BCMI_OK;
}

// Plain sweep of a 2-degree source v0 whose neighbours' forward states sit
// in slots sa and sb (lva, lvb levels): the forward state is combined from
// them (k_d2_*), then the plain backward pass runs.  *done is false when the
// combined state could not be used (it is undone; v0 is then swept normally).
static int run_plain_derived(const DevGraph& g, Workspace* w, const Params& prm, int32_t v0, double mult, int sa,
                             int sb, int lva, int lvb, RunPart& P, bool* done) {
  const double t0 = now_ms();
  hipStream_t st = w->st;
  const int64_t n = g.n;
  const int64_t nlev = (int64_t)std::min(lva, lvb) + 2;
  *done = false;
  TRY(d2_reserve(w, w->d2_slots, nlev));
  if (prm.graphs && (!w->gfwd || w->cap_graph != &g || !same_launch(w->cap_prm, prm))) TRY(capture(g, w, prm));
  w->hmult[0] = mult;
  HIP_TRY(hipMemcpyAsync(w->multbuf, w->hmult, sizeof(double), hipMemcpyHostToDevice, st));
  GA ga = make_ga(g, prm);
  PA pa = make_pa(w, prm);
  D2A d;
  d.da = w->d2_depth + (size_t)sa * n; d.sa = w->d2_sigma + (size_t)sa * n;
  d.db = w->d2_depth + (size_t)sb * n; d.sb = w->d2_sigma + (size_t)sb * n;
  d.v0 = v0; d.lcnt = w->d2_lcnt; d.nlev = nlev; d.single = sa == sb;
  const int G = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + 255) / 256));
  hipLaunchKernelGGL(k_d2_init, 1, 1, 0, st, ga, pa, d);
  hipLaunchKernelGGL(k_d2_combine, G, 256, 0, st, ga, pa, d);
  hipLaunchKernelGGL(k_d2_scan, 1, 1, 0, st, pa, d);
  hipLaunchKernelGGL(k_d2_scatter, G, 256, 0, st, ga, pa, d);
  HIP_TRY(hipMemcpyAsync(w->hctl, w->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (w->hctl->overflow) {
    hipLaunchKernelGGL(k_d2_reset, G, 256, 0, st, ga, pa, d);
    HIP_TRY(hipStreamSynchronize(st));
    return BCMI_OK;
  }
  TRY(backward_loop(g, w, prm));
  const int lv = w->hctl->last + 1;
  hipLaunchKernelGGL(k_plain_cleanup, grid_for(std::min<int64_t>(w->hctl->rec_count, w->cap)), 256, 0, st, pa);
  HIP_TRY(hipStreamSynchronize(st));
  plain_stats(*w->hctl, lv, now_ms() - t0, P);
  P.R.deg2_sources++;
  *done = true;
  return BCMI_OK;
}

// Plain sweeps with the 2-degree heuristic (BCMI_SRC_DEG2).  A listed vertex
// v0 of traversed degree 2 whose neighbours a and b are listed is taken from
// them when neither neighbour is itself taken that way and v0 is nobody's
// neighbour-to-be-saved (first come, in increasing id order); a and b become
// anchors.  Anchors run first, ordered so that a pair's anchors run close
// together (breadth-first over the pairs, from the anchors with most pairs);
// after an anchor's sweep its forward state goes into a slot, and every pair
// whose other anchor is in a slot runs its derived sources.  A full set of
// slots gives up the anchor with fewest pairs left; a derived
// source whose anchor left the slots before its partner ran is swept
// normally.  Other listed sources are not the plan's: they run as usual, on
// the other workspaces of the call at the same time.
struct Deg2Pair {
  int32_t a, b;
  std::vector<size_t> items;  // indices into the source list of the derived sources
  bool done = false;
};
struct Deg2Plan {
  std::vector<Deg2Pair> pairs;
  std::unordered_map<int32_t, std::vector<int>> pairs_at;  // anchor -> its pairs
  std::unordered_map<int32_t, size_t> anchor_item;         // anchor -> its first listed occurrence
  std::vector<int32_t> order;                              // anchors in run order
  std::vector<uint8_t> mine;                               // source-list entries the plan runs
};

static int deg2_plan(const DevGraph& g, Workspace* w, const std::vector<int32_t>& cs, Deg2Plan& D) {
  const int64_t n = g.n;
  const std::vector<int64_t>& ho = g.h_off;
  hipStream_t st = w->st;
  std::vector<int32_t> uniq(cs);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  auto listed = [&](int32_t v) { return std::binary_search(uniq.begin(), uniq.end(), v); };
  std::vector<int32_t> cand;
  for (int32_t v : uniq)
    if (ho[v + 1] - ho[v] == 2 || ho[v + 1] - ho[v] == 1) cand.push_back(v);
  std::vector<int32_t> nb(2 * cand.size());
  if (!cand.empty()) {
    int32_t *dc = nullptr, *dn = nullptr;
    {
      ALLOC_LOCK;
      TRY(dalloc(&dc, cand.size()));
      if (dalloc(&dn, 2 * cand.size()) != BCMI_OK) { (void)hipFree(dc); return BCMI_ERR_NOMEM; }
    }
    hipError_t e = hipMemcpyAsync(dc, cand.data(), sizeof(int32_t) * cand.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_d2_nb, grid_for((int64_t)cand.size()), 256, 0, st, g.off, g.adj, dc, (int64_t)cand.size(), dn);
      e = hipMemcpyAsync(nb.data(), dn, sizeof(int32_t) * nb.size(), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    {
      ALLOC_LOCK;
      (void)hipFree(dc);
      (void)hipFree(dn);
    }
    if (e != hipSuccess) return set_error(BCMI_ERR_HIP, hipGetErrorString(e));
  }
  // roles: 1 anchor, 2 derived; pairs of anchors with the derived vertices they serve
  std::vector<uint8_t> role(n, 0);
  std::vector<Deg2Pair>& pairs = D.pairs;
  std::map<std::pair<int32_t, int32_t>, int> pid;
  std::unordered_map<int32_t, int> pair_of;  // derived vertex -> pair
  for (size_t i = 0; i < cand.size(); i++) {
    const int32_t v0 = cand[i], a = nb[2 * i], b = nb[2 * i + 1];
    if (role[v0] || role[a] == 2 || role[b] == 2 || !listed(a) || !listed(b)) continue;  // (a == b: degree 1)
    role[v0] = 2;
    role[a] = role[b] = 1;
    auto key = std::make_pair(std::min(a, b), std::max(a, b));
    auto it = pid.find(key);
    if (it == pid.end()) {
      it = pid.emplace(key, (int)pairs.size()).first;
      pairs.push_back(Deg2Pair{key.first, key.second, {}, false});
    }
    pair_of[v0] = it->second;
  }
  std::unordered_map<int32_t, std::vector<int>>& pairs_at = D.pairs_at;
  std::unordered_map<int32_t, size_t>& anchor_item = D.anchor_item;
  D.mine.assign(cs.size(), 0);
  for (size_t i = 0; i < cs.size(); i++) {
    const int32_t v = cs[i];
    if (role[v] == 2) {
      pairs[pair_of[v]].items.push_back(i);
      D.mine[i] = 1;
    } else if (role[v] == 1 && !anchor_item.count(v)) {
      anchor_item[v] = i;
      D.mine[i] = 1;
    }
  }
  for (size_t q = 0; q < pairs.size(); q++) {
    pairs_at[pairs[q].a].push_back((int)q);
    if (pairs[q].b != pairs[q].a) pairs_at[pairs[q].b].push_back((int)q);
  }
  // anchor order: breadth-first over the pairs, starting from the anchors with most pairs
  std::vector<int32_t> anchors;
  for (auto& kv : pairs_at) anchors.push_back(kv.first);
  std::sort(anchors.begin(), anchors.end(), [&](int32_t x, int32_t y) {
    size_t px = pairs_at[x].size(), py = pairs_at[y].size();
    return px != py ? px > py : x < y;
  });
  std::vector<int32_t>& order = D.order;
  {
    std::unordered_map<int32_t, char> seen;
    for (int32_t r : anchors) {
      if (seen.count(r)) continue;
      std::deque<int32_t> q{r};
      seen[r] = 1;
      while (!q.empty()) {
        const int32_t x = q.front();
        q.pop_front();
        order.push_back(x);
        std::vector<int32_t> nx;
        for (int pq : pairs_at[x]) {
          const int32_t y = pairs[pq].a == x ? pairs[pq].b : pairs[pq].a;
          if (!seen.count(y)) { seen[y] = 1; nx.push_back(y); }
        }
        std::sort(nx.begin(), nx.end(), [&](int32_t u, int32_t v) {
          size_t pu = pairs_at[u].size(), pv = pairs_at[v].size();
          return pu != pv ? pu > pv : u < v;
        });
        for (int32_t y : nx) q.push_back(y);
      }
    }
  }
  return BCMI_OK;
}

// Runs a plan's anchors and derived sources on workspace w (the one whose
// slots hold the anchors' states); derived sources whose pair did not
// complete are swept normally here.
static int deg2_run(const DevGraph& g, Workspace* w, const Params& prm, const std::vector<int32_t>& cs,
                    const std::vector<double>& cm, Deg2Plan& D, RunPart& P) {
  const int64_t n = g.n;
  hipStream_t st = w->st;
  std::vector<Deg2Pair>& pairs = D.pairs;
  std::unordered_map<int32_t, std::vector<int>>& pairs_at = D.pairs_at;
  std::unordered_map<int32_t, size_t>& anchor_item = D.anchor_item;
  const std::vector<int32_t>& order = D.order;
  std::vector<uint8_t> ran(cs.size(), 0);
  int S = (int)std::min<int64_t>(prm.deg2_slots > 0 ? prm.deg2_slots
                                                    : std::max<int64_t>(2, prm.deg2_mem / (12 * std::max<int64_t>(1, n))),
                                 (int64_t)order.size());
  // fewer slots while the device cannot hold them (one slot still serves
  // degree-1 sources, whose group has a single anchor)
  while (S >= 1 && d2_reserve(w, S, 256) != BCMI_OK) S /= 2;
  if (S >= 1) {
    std::vector<int32_t> slot_v(S, -1);
    std::vector<int> slot_lv(S, 0);
    std::vector<int64_t> slot_t(S, 0);
    std::unordered_map<int32_t, int> slot_of;
    std::unordered_map<int32_t, int> pending;
    for (auto& kv : pairs_at) pending[kv.first] = (int)kv.second.size();
    int64_t clock = 0;
    for (int32_t a : order) {
      const size_t ia = anchor_item[a];
      int lv = 0;
      TRY(run_plain_one(g, w, prm, a, cm[ia], P, true, &lv));
      ran[ia] = 1;
      const bool usable = !w->hctl->sigma_ovf;
      if (usable) {
        // a free slot, else one whose anchor has no pairs left, else, among
        // those not holding a partner a is about to pair with, the one whose
        // anchor has fewest pairs left (least recently used first): hubs with
        // many pairs stay (on the peeled R-MAT 18 core with 1,050 slots this
        // completes 17.2k of 19.5k pairs, least recently used alone 9.2k)
        int s = -1;
        for (int i = 0; i < S && s < 0; i++)
          if (slot_v[i] < 0) s = i;
        for (int i = 0; i < S && s < 0; i++)
          if (pending[slot_v[i]] == 0) s = i;
        if (s < 0) {
          std::unordered_map<int32_t, char> partners;
          for (int pq : pairs_at[a])
            if (!pairs[pq].done) partners[pairs[pq].a == a ? pairs[pq].b : pairs[pq].a] = 1;
          std::vector<int> pend(S);
          for (int i = 0; i < S; i++) pend[i] = pending[slot_v[i]];
          auto better = [&](int i, int k) {  // slot i a better victim than slot k
            return pend[i] != pend[k] ? pend[i] < pend[k] : slot_t[i] < slot_t[k];
          };
          for (int pass = 0; pass < 2 && s < 0; pass++)
            for (int i = 0; i < S; i++)
              if ((pass == 1 || !partners.count(slot_v[i])) && (s < 0 || better(i, s))) s = i;
        }
        if (slot_v[s] >= 0) slot_of.erase(slot_v[s]);
        HIP_TRY(hipMemcpyAsync(w->d2_depth + (size_t)s * n, w->depth, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(w->d2_sigma + (size_t)s * n, w->sigma1, sizeof(u64) * n, hipMemcpyDeviceToDevice, st));
        slot_v[s] = a;
        slot_lv[s] = lv;
        slot_t[s] = ++clock;
        slot_of[a] = s;
      }
      TRY(sweep_cleanup(g, w, prm));
      if (!usable) continue;
      for (int pq : pairs_at[a]) {
        Deg2Pair& pr = pairs[pq];
        const int32_t b = pr.a == a ? pr.b : pr.a;
        if (pr.done || !slot_of.count(b)) continue;
        const int sa = slot_of[a], sb = slot_of[b];
        for (size_t j : pr.items) {
          bool ok = false;
          TRY(run_plain_derived(g, w, prm, cs[j], cm[j], sa, sb, slot_lv[sa], slot_lv[sb], P, &ok));
          if (ok) ran[j] = 1;
        }
        pr.done = true;
        pending[a]--;
        if (b != a) pending[b]--;
        slot_t[sb] = ++clock;
        slot_t[sa] = ++clock;
      }
    }
  }
  for (size_t i = 0; i < cs.size(); i++)
    if (D.mine[i] && !ran[i]) TRY(run_plain_one(g, w, prm, cs[i], cm[i], P));
  return BCMI_OK;
}

// One batched sweep over k <= 64 sources (internal ids) with multipliers.
// Later batches of g start with rows no narrower than width code nw.
static void lower_narrow_cap(const DevGraph& g, int nw) {
  int cap = g.narrow_cap.load();
  while (cap > nw && !g.narrow_cap.compare_exchange_weak(cap, nw)) {
  }
}

static int run_batch(const DevGraph& g, Workspace* w, const Params& prm, const int32_t* src, const double* mult,
                     int k, RunPart& P) {
  const double t0 = now_ms();
  hipStream_t st = w->st;
  for (int q = 0; q < 64; q++) {
    w->hsrc[q] = q < k ? src[q] : 0;
    w->hmult[q] = q < k ? mult[q] : 0.0;
  }
  HIP_TRY(hipMemcpyAsync(w->srcbuf, w->hsrc, 64 * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(w->multbuf, w->hmult, 64 * sizeof(double), hipMemcpyHostToDevice, st));
  int lv = 0;
  // auto mode: once a batch of this graph has overflowed, later ones start scaled
  int scaled = prm.sigma == 1 || (prm.sigma == 2 && g.sigma_overflowed.load());
  int narrow = scaled ? 0 : std::min(prm.narrow_sigma, g.narrow_cap.load());
  TRY(sweep(g, w, prm, k, false, &lv, scaled, narrow));
  while (narrow > 0 && w->hctl->sigma_ovf) {
    // a count passed the row's range (the backward pass was skipped): redo the
    // batch with the next wider rows, which later batches of this graph start with
    narrow--;
    lower_narrow_cap(g, narrow);
    TRY(sweep(g, w, prm, k, false, &lv, 0, narrow));
  }
  if (w->hctl->sigma_ovf && prm.sigma == 2) {
    // exact counts wrapped: the backward pass was skipped, redo the batch scaled
    g.sigma_overflowed.store(true);
    TRY(sweep(g, w, prm, k, false, &lv, 1));
  }
  RunResult& R = P.R;
  if (scaled || w->hctl->scaled) R.scaled_batches++;
  HIP_TRY(hipMemcpyAsync(w->hstats, w->stats, 3 * 64 * sizeof(u64), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const double per_src = (now_ms() - t0) / k;
  for (int q = 0; q < k; q++) {
    R.n_s.push_back((int64_t)w->hstats[q]);
    R.m_s.push_back((int64_t)w->hstats[64 + q]);
    R.a_s.push_back((int64_t)w->hstats[128 + q]);
    R.levels.push_back(lv);
    R.t_s.push_back(per_src);
  }
  if (w->hctl->sigma_ovf) P.ovf = 1;
  R.uni_n.push_back(w->hctl->uni_verts);
  R.pairs += w->hctl->tot_verts;
  R.uni_m.push_back(w->hctl->uni_arcs);
  R.n_td += w->hctl->n_td; R.n_bu += w->hctl->n_bu;
  R.sweeps += k;
  R.launches++;
  R.max_levels = std::max<int64_t>(R.max_levels, lv);
  return BCMI_OK;
}

__global__ void k_add_take_f64(double* dst, double* src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    dst[i] += src[i];
    src[i] = 0.0;
  }
}

// Makes sure ws has `count` helper workspaces for graph g.
static int ensure_helpers(const DevGraph& g, Workspace* ws, const Params& prm, int count) {
  if (ws->helpers_graph != &g) {
    for (Workspace* h : ws->helpers) workspace_destroy(h);
    ws->helpers.clear();
    ws->helpers_graph = &g;
  }
  while ((int)ws->helpers.size() < count) {
    int rc = BCMI_OK;
    Workspace* h = workspace_create(g, ws->mode, ws->width, prm, &rc);
    if (!h) return rc;
    ws->helpers.push_back(h);
  }
  return BCMI_OK;
}

// Workgroup-per-source sweeps serve deep graphs (hub BFS depth of at least
// Params::wg_depth) unless Params::strategy forces one engine.
static bool use_wg(const DevGraph& g, const Params& prm) {
  if (prm.strategy == 1) return false;
  if (prm.strategy == 2) return true;
  return g.depth_hint >= prm.wg_depth;
}

static size_t wg_slot_bytes(int64_t n, int64_t lcap) {
  return (size_t)n * (2 * sizeof(int32_t) + sizeof(u64)) +
         (size_t)lcap * (sizeof(int32_t) + sizeof(double));
}

static int64_t wg_slot_count(int64_t n, const Params& prm, int64_t lcap) {
  int64_t k = (int64_t)(prm.wg_mem / std::max<size_t>(wg_slot_bytes(n, lcap), 1));
  return std::max<int64_t>(1, std::min<int64_t>(k, prm.wg_slots));
}

static int64_t wg_initial_lcap(const DevGraph& g) {
  // a sweep inside the hubs' component has at most 2 x depth + 1 levels
  return std::min<int64_t>(g.n + 2, std::max<int64_t>(1024, 2 * (int64_t)g.depth_hint + 64));
}

static size_t wg_bytes_estimate(const DevGraph& g, const Params& prm) {
  const int64_t lcap = wg_initial_lcap(g);
  return (size_t)wg_slot_count(g.n, prm, lcap) * wg_slot_bytes(g.n, lcap) + sizeof(double) * g.n + (1 << 20);
}

static int wg_alloc_slots(const DevGraph& g, Workspace* ws, const Params& prm, int64_t lcap) {
  ALLOC_LOCK;
  auto fw = [](void* p) { if (p) (void)hipFree(p); };
  fw(ws->wg_depth); fw(ws->wg_sigma); fw(ws->wg_queue); fw(ws->wg_lvl); fw(ws->wg_scale);
  ws->wg_depth = nullptr; ws->wg_sigma = nullptr; ws->wg_queue = nullptr;
  ws->wg_lvl = nullptr; ws->wg_scale = nullptr;
  const int64_t n = g.n;
  int64_t S = wg_slot_count(n, prm, lcap);
  size_t b = 0;
  // fewer slots when the device cannot hold the requested per-source state
  for (;;) {
    b = 0;
    int rc = dalloc(&ws->wg_depth, (size_t)S * n, &b);
    if (rc == BCMI_OK) rc = dalloc(&ws->wg_sigma, (size_t)S * n, &b);
    if (rc == BCMI_OK) rc = dalloc(&ws->wg_queue, (size_t)S * n, &b);
    if (rc == BCMI_OK) rc = dalloc(&ws->wg_lvl, (size_t)S * lcap, &b);
    if (rc == BCMI_OK) rc = dalloc(&ws->wg_scale, (size_t)S * lcap, &b);
    if (rc == BCMI_OK) break;
    fw(ws->wg_depth); fw(ws->wg_sigma); fw(ws->wg_queue); fw(ws->wg_lvl); fw(ws->wg_scale);
    ws->wg_depth = nullptr; ws->wg_sigma = nullptr; ws->wg_queue = nullptr;
    ws->wg_lvl = nullptr; ws->wg_scale = nullptr;
    if (S == 1) return rc;
    S = (S + 1) / 2;
  }
  if (!ws->wg_bc) {
    TRY(dalloc(&ws->wg_bc, n, &b));
    TRY(dalloc(&ws->wg_next, 1, &b));
    hipLaunchKernelGGL(k_fill_f64, grid_for(n), 256, 0, ws->st, ws->wg_bc, 0.0, n);
  } else {
    b += sizeof(double) * n + sizeof(unsigned long long);
  }
  hipLaunchKernelGGL(k_fill_i32, grid_for(S * n), 256, 0, ws->st, ws->wg_depth, -1, S * n);
  hipLaunchKernelGGL(k_fill_u64, grid_for(S * n), 256, 0, ws->st, ws->wg_sigma, 0ull, S * n);
  HIP_TRY(hipStreamSynchronize(ws->st));
  ws->wg_slots = S;
  ws->wg_req = wg_slot_count(n, prm, lcap);
  ws->wg_lcap = lcap;
  ws->wg_bytes = b + ws->wg_cap * (sizeof(int32_t) * 3 + sizeof(double) + sizeof(int64_t) * 3);
  return BCMI_OK;
}

static int wg_alloc_run(Workspace* ws, int64_t k) {
  if (k <= ws->wg_cap) return BCMI_OK;
  ALLOC_LOCK;
  auto fw = [](void* p) { if (p) (void)hipFree(p); };
  fw(ws->wg_src); fw(ws->wg_mult); fw(ws->wg_st); fw(ws->wg_st2);
  ws->wg_src = nullptr; ws->wg_mult = nullptr; ws->wg_st = nullptr; ws->wg_st2 = nullptr;
  int64_t cap = std::max<int64_t>(k, 1024);
  TRY(dalloc(&ws->wg_src, cap));
  TRY(dalloc(&ws->wg_mult, cap));
  TRY(dalloc(&ws->wg_st, 3 * cap));
  TRY(dalloc(&ws->wg_st2, 2 * cap));
  ws->wg_bytes += (cap - ws->wg_cap) * (sizeof(int32_t) * 3 + sizeof(double) + sizeof(int64_t) * 3);
  ws->wg_cap = cap;
  return BCMI_OK;
}

// Runs every source as its own sweep on the workgroup-per-source kernel.
// BC goes to a scratch array first and is folded in only when the whole run
// succeeded (an exact path count that wraps leaves the result untouched).
static int run_wg(const DevGraph& g, Workspace* ws, const Params& prm, const std::vector<int32_t>& cs,
                  const std::vector<double>& cm, RunResult& R, int& ovf) {
  const double t0 = now_ms();
  const size_t first = R.n_s.size();
  hipStream_t st = ws->st;
  const int64_t lcap0 = std::max(wg_initial_lcap(g), ws->wg_lcap);
  if (!ws->wg_depth || ws->wg_lcap < lcap0 || ws->wg_req != wg_slot_count(g.n, prm, lcap0))
    TRY(wg_alloc_slots(g, ws, prm, lcap0));
  std::vector<int32_t> src(cs);
  std::vector<double> mult(cm);
  int scaled = prm.sigma == 1 || (prm.sigma == 2 && g.sigma_overflowed.load());
  const GA ga = make_ga(g, prm);
  for (int attempt = 0; !src.empty(); attempt++) {
    if (attempt > 12) return set_error(BCMI_ERR_STATE, "workgroup sweeps did not complete");
    const int64_t k = (int64_t)src.size();
    TRY(wg_alloc_run(ws, k));
    HIP_TRY(hipMemcpyAsync(ws->wg_src, src.data(), sizeof(int32_t) * k, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ws->wg_mult, mult.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_u64, 1, 64, 0, st, (u64*)ws->wg_next, 0ull, (int64_t)1);
    WA a;
    a.src = ws->wg_src; a.mult = ws->wg_mult; a.nsrc = k; a.next = ws->wg_next;
    a.depth = ws->wg_depth; a.sigma = ws->wg_sigma; a.queue = ws->wg_queue;
    a.lvl = ws->wg_lvl; a.scale = ws->wg_scale; a.lcap = ws->wg_lcap; a.bc = ws->wg_bc;
    a.st_n = ws->wg_st; a.st_m = ws->wg_st + ws->wg_cap; a.st_a = ws->wg_st + 2 * ws->wg_cap;
    a.st_levels = ws->wg_st2; a.st_flag = ws->wg_st2 + ws->wg_cap;
    a.scaled = scaled;
    const int grid = (int)std::min<int64_t>(ws->wg_slots, k);
    hipLaunchKernelGGL(k_wg_sweep, grid, WGT, 0, st, ga, a);
    std::vector<int64_t> hn(k), hm(k), ha(k);
    std::vector<int32_t> hl(k), hf(k);
    HIP_TRY(hipMemcpyAsync(hn.data(), a.st_n, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hm.data(), a.st_m, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(ha.data(), a.st_a, sizeof(int64_t) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hl.data(), a.st_levels, sizeof(int32_t) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hf.data(), a.st_flag, sizeof(int32_t) * k, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    R.launches++;
    std::vector<int32_t> rs;
    std::vector<double> rm;
    bool wrapped = false, deep = false;
    for (int64_t j = 0; j < k; j++) {
      if (hf[j] == 0) {
        R.n_s.push_back(hn[j]);
        R.m_s.push_back(hm[j]);
        R.a_s.push_back(ha[j]);
        R.levels.push_back(hl[j]);
        R.max_levels = std::max<int64_t>(R.max_levels, hl[j]);
        R.sweeps++;
        R.wg_sources++;
        if (scaled) R.scaled_batches++;
      } else {
        rs.push_back(src[j]);
        rm.push_back(mult[j]);
        wrapped |= hf[j] == 1;
        deep |= hf[j] == 2;
      }
    }
    if (wrapped) {
      if (prm.sigma == 0) {
        ovf = 1;
        hipLaunchKernelGGL(k_fill_f64, grid_for(g.n), 256, 0, st, ws->wg_bc, 0.0, g.n);
        HIP_TRY(hipStreamSynchronize(st));
        return BCMI_OK;
      }
      g.sigma_overflowed.store(true);
      scaled = 1;
    }
    if (deep) TRY(wg_alloc_slots(g, ws, prm, std::min<int64_t>(g.n + 2, ws->wg_lcap * 4)));
    src.swap(rs);
    mult.swap(rm);
  }
  hipLaunchKernelGGL(k_add_take_f64, grid_for(g.n), 256, 0, st, ws->bc, ws->wg_bc, g.n);
  // the sources ran together: each gets an equal share of the run
  const size_t done = R.n_s.size() - first;
  if (done) R.t_s.insert(R.t_s.end(), done, (now_ms() - t0) / (double)done);
  return BCMI_OK;
}

// Order in which a batched run takes its sources (internal ids, degree > 0):
// a permutation idx of cs, grouped by the Params::batching similarity key so
// that consecutive runs of `width` sources share levels.
static void batch_sort(const DevGraph& g, const Params& prm, const std::vector<int32_t>& cs,
                       std::vector<size_t>& idx) {
  idx.resize(cs.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  const std::vector<uint64_t>& sg = g.h_sig;
  const std::vector<int32_t>& tn = g.h_topnb;
  const std::vector<int32_t>& tn2 = g.h_topnb2;
  if (prm.batching == 3)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      int32_t ta = tn[cs[a]], tb = tn[cs[b]];
      uint32_t ka = ta < INT32_MAX ? sg[ta] : ~0u, kb = tb < INT32_MAX ? sg[tb] : ~0u;
      if (ka != kb) return ka < kb;
      if (ta != tb) return ta < tb;
      return sg[cs[a]] < sg[cs[b]];
    });
  else if (prm.batching == 4)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      int32_t ta = tn[cs[a]], tb = tn[cs[b]];
      if (ta != tb) return ta < tb;
      int32_t ua = tn2[cs[a]], ub = tn2[cs[b]];
      if (ua != ub) return ua < ub;
      return sg[cs[a]] < sg[cs[b]];
    });
  else if (prm.batching == 5)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      int32_t ta = tn[cs[a]], tb = tn[cs[b]];
      if (ta != tb) return ta < tb;
      if (sg[cs[a]] != sg[cs[b]]) return sg[cs[a]] < sg[cs[b]];
      return tn2[cs[a]] < tn2[cs[b]];
    });
  else if (prm.batching == 2)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
      int32_t ta = tn[cs[a]], tb = tn[cs[b]];
      if (ta != tb) return ta < tb;
      return sg[cs[a]] < sg[cs[b]];
    });
  else if (prm.batching == 1)
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return sg[cs[a]] < sg[cs[b]]; });
}

int batch_order(const DevGraph& g, const Params& prm, const int32_t* src, int64_t nsrc, int32_t* out) {
  if (g.heuristic) return set_error(BCMI_ERR_STATE, "batch order needs an unpeeled graph");
  for (int64_t i = 0; i < nsrc; i++)
    if (src[i] < 0 || src[i] >= g.n) return set_error(BCMI_ERR_RANGE, "source id out of range");
  std::vector<int32_t> cs;
  std::vector<int64_t> pos;
  for (int64_t i = 0; i < nsrc; i++) {
    const int32_t s = g.to_int(src[i]);
    if (g.h_off[s + 1] > g.h_off[s]) { cs.push_back(s); pos.push_back(i); }
  }
  std::vector<size_t> idx;
  batch_sort(g, prm, cs, idx);
  int64_t k = 0;
  for (size_t i : idx) out[k++] = src[pos[i]];
  // sources of degree 0 are not swept: last, in the given order
  for (int64_t i = 0; i < nsrc; i++) {
    const int32_t s = g.to_int(src[i]);
    if (g.h_off[s + 1] == g.h_off[s]) out[k++] = src[i];
  }
  return BCMI_OK;
}

int bc_run(const DevGraph& g, Workspace* ws, const Params& prm, const int32_t* sources, int64_t nsrc, int mode,
           int src_flags, RunResult* out) {
  if (mode != ws->mode) return set_error(BCMI_ERR_STATE, "workspace was created for another mode");
  if ((src_flags & BCMI_SRC_EXPAND) && !g.heuristic)
    return set_error(BCMI_ERR_ARG, "source expansion needs a peeled graph");
  double t0 = now_ms();
  std::vector<int32_t> cs;
  std::vector<double> cm;
  bool fin = false;
  TRY(prepare_sources(g, ws, sources, nsrc, src_flags, cs, cm, fin));
  hipStream_t st = ws->st;
  hipLaunchKernelGGL(k_kt_reset, 1, 1, 0, st, ws->kt, ws->kt_ws, ws->kt_we);
  HIP_TRY(hipEventRecord(ws->ev0, st));
  const bool plain = mode == BCMI_MODE_PLAIN;
  const int W = std::max(1, std::min(64, prm.width));
  const bool wg = use_wg(g, prm);
  RunResult R;
  int ovf = 0;
  if (wg) TRY(run_wg(g, ws, prm, cs, cm, R, ovf));
  if (!plain && !wg) {
    // Batch sources with similar depth profiles together (BC is a sum over
    // sources, so the order only changes the rounding of the final sums).
    std::vector<size_t> idx;
    batch_sort(g, prm, cs, idx);
    std::vector<int32_t> cs2(cs.size());
    std::vector<double> cm2(cm.size());
    for (size_t i = 0; i < idx.size(); i++) { cs2[i] = cs[idx[i]]; cm2[i] = cm[idx[i]]; }
    cs.swap(cs2);
    cm.swap(cm2);
  }
  // 2-degree heuristic: the plan's anchors and derived sources run on this
  // workspace (whose slots hold the anchors' states), the other sources on
  // every workspace of the call
  const bool deg2 = (src_flags & BCMI_SRC_DEG2) && plain && !wg && !g.directed && prm.deg2_slots != 1;
  Deg2Plan plan;
  std::vector<size_t> items;  // work items of the shared queue: sources (plain) or batches
  if (deg2) {
    TRY(deg2_plan(g, ws, cs, plan));
    for (size_t i = 0; i < cs.size(); i++)
      if (!plan.mine[i]) items.push_back(i);
  } else if (!wg) {
    items.resize(plain ? cs.size() : (cs.size() + W - 1) / W);
    for (size_t i = 0; i < items.size(); i++) items[i] = i;
  }
  const size_t nitems = items.size() + (deg2 ? 1 : 0);
  int K = (int)std::min<size_t>((size_t)concurrent_sweeps(g, prm), std::max<size_t>(1, nitems));
  if (K > 1) {
    // as many helper workspaces as the device holds (fewer sweeps at once otherwise)
    const int rc = ensure_helpers(g, ws, prm, K - 1);
    if (rc != BCMI_OK) {
      if (rc != BCMI_ERR_NOMEM || ws->helpers_graph != &g) return rc;
      K = 1 + (int)ws->helpers.size();
    }
    // the kernel timer sums the timed kernel's launches on every workspace
    for (int t = 1; t < K; t++) {
      Workspace* h = ws->helpers[t - 1];
      hipLaunchKernelGGL(k_kt_reset, 1, 1, 0, h->st, h->kt, h->kt_ws, h->kt_we);
    }
  }
  std::vector<RunPart> parts(K);
  std::atomic<size_t> next{0};
  std::atomic<bool> stop{false};
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  auto work_body = [&](int t) {
    RunPart& P = parts[t];
    if (t > 0 && hipSetDevice(dev) != hipSuccess) {
      P.rc = BCMI_ERR_HIP;
      P.msg = "hipSetDevice failed in a sweep thread";
      stop = true;
      return;
    }
    Workspace* w = t == 0 ? ws : ws->helpers[t - 1];
    if (deg2 && t == 0) {
      const int rc = deg2_run(g, w, prm, cs, cm, plan, P);
      if (rc != BCMI_OK) {
        P.rc = rc;
        P.msg = bcmi_last_error();
        stop = true;
        return;
      }
    }
    while (!stop) {
      const size_t q = next.fetch_add(1);
      if (q >= items.size()) break;
      const size_t it = items[q];
      int rc;
      if (plain) {
        rc = run_plain_one(g, w, prm, cs[it], cm[it], P);
      } else {
        size_t i0 = it * W;
        int k = (int)std::min<size_t>(W, cs.size() - i0);
        rc = run_batch(g, w, prm, cs.data() + i0, cm.data() + i0, k, P);
      }
      if (rc != BCMI_OK) {
        P.rc = rc;
        P.msg = bcmi_last_error();
        stop = true;
        break;
      }
    }
  };
  // an exception in a sweep thread (host memory) stops the call with an
  // error instead of ending the process
  auto work = [&](int t) {
    try {
      work_body(t);
    } catch (...) {
      parts[t].rc = BCMI_ERR_NOMEM;
      parts[t].msg = "out of host memory in a sweep thread";
      stop = true;
    }
  };
  {
    // a helper thread that cannot start leaves its share to the others (the
    // work items come from the shared counter)
    std::vector<std::thread> th;
    th.reserve(K);
    for (int t = 1; t < K; t++) {
      try {
        th.emplace_back(work, t);
      } catch (const std::system_error&) {
        break;
      }
    }
    work(0);
    for (auto& x : th) x.join();
  }
  for (RunPart& P : parts) {
    if (P.rc != BCMI_OK) return set_error(P.rc, P.msg);
    RunResult& Q = P.R;
    R.n_s.insert(R.n_s.end(), Q.n_s.begin(), Q.n_s.end());
    R.m_s.insert(R.m_s.end(), Q.m_s.begin(), Q.m_s.end());
    R.a_s.insert(R.a_s.end(), Q.a_s.begin(), Q.a_s.end());
    R.levels.insert(R.levels.end(), Q.levels.begin(), Q.levels.end());
    R.t_s.insert(R.t_s.end(), Q.t_s.begin(), Q.t_s.end());
    R.uni_n.insert(R.uni_n.end(), Q.uni_n.begin(), Q.uni_n.end());
    R.uni_m.insert(R.uni_m.end(), Q.uni_m.begin(), Q.uni_m.end());
    R.pairs += Q.pairs;
    R.n_td += Q.n_td; R.n_bu += Q.n_bu;
    R.sweeps += Q.sweeps; R.launches += Q.launches;
    R.scaled_batches += Q.scaled_batches;
    R.deg2_sources += Q.deg2_sources;
    R.max_levels = std::max(R.max_levels, Q.max_levels);
    ovf |= P.ovf;
  }
  // fold the helpers' BC into this workspace's
  for (int t = 1; t < K; t++) {
    Workspace* h = ws->helpers[t - 1];
    HIP_TRY(hipStreamSynchronize(h->st));
    hipLaunchKernelGGL(k_add_take_f64, grid_for(g.n), 256, 0, st, ws->bc, h->bc, g.n);
  }
  if (fin) {
    hipLaunchKernelGGL(k_finalize, grid_for(g.n), 256, 0, st, g.label, g.round, (const u64*)g.w, ws->cnt0, ws->cntS,
                       ws->sa, ws->srcK, g.comp_size, g.n, ws->bc);
  }
  HIP_TRY(hipEventRecord(ws->ev1, st));
  HIP_TRY(hipEventSynchronize(ws->ev1));
  HIP_TRY(hipGetLastError());
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
  R.gpu_ms = ms;
  R.wall_ms = now_ms() - t0;
  {
    // timed kernel: launch durations and launches summed over the workspaces
    // that ran sweeps (launches of concurrent sweeps overlap, so each one's
    // duration includes the device time the others took meanwhile)
    u64 total = 0, launches = 0;
    for (int t = 0; t < K; t++) {
      Workspace* w = t == 0 ? ws : ws->helpers[t - 1];
      HIP_TRY(hipMemcpyAsync(w->hkt, w->kt, sizeof(KT), hipMemcpyDeviceToHost, w->st));
      HIP_TRY(hipStreamSynchronize(w->st));
      total += w->hkt->total;
      launches += w->hkt->launches;
    }
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    R.ktime_ms = khz > 0 ? (double)total / (double)khz : 0.0;
    R.klaunches = (int64_t)launches;
  }
  for (size_t i = 0; i < R.n_s.size(); i++) {
    R.sum_n += R.n_s[i]; R.sum_m += R.m_s[i]; R.sum_a += R.a_s[i];
  }
  R.te = 2.0 * (double)R.sum_m;
  // Algorithmic HBM bytes (frozen model):
  //   plain   B(s) = 16 m_s + 24 a_s + 64 n_s
  //   batched B(S) = 24 m_S + 32 n_S + sum_s (24 a_s + 32 n_s), m_S/n_S = union reach
  R.bytes = 0;
  if (mode == BCMI_MODE_PLAIN) {
    for (size_t i = 0; i < R.n_s.size(); i++)
      R.bytes += 16.0 * R.m_s[i] + 24.0 * R.a_s[i] + 64.0 * R.n_s[i];
  } else {
    for (size_t i = 0; i < R.uni_n.size(); i++) R.bytes += 24.0 * R.uni_m[i] + 32.0 * R.uni_n[i];
    for (size_t i = 0; i < R.n_s.size(); i++) R.bytes += 24.0 * R.a_s[i] + 32.0 * R.n_s[i];
  }
  if (out) *out = R;
  if (ovf) return set_error(BCMI_ERR_OVERFLOW, "a shortest-path count exceeded 2^64-1");
  return BCMI_OK;
}

// ====================================== allocation-free plain sweeps
// Plain sweeps whose per-sweep state lives in the static device array of
// static_scratch.hip (one per GPU) instead of a workspace, for callers that
// may not allocate device memory (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The kernels are the plain
// sweep's, launched directly: instantiating a captured level loop would
// allocate.  The array keeps the last call's contents, so every call first
// clears the state it reads; calls on one device are serialised here.
static std::mutex g_scratch_mu[kMaxDevices];  // one call at a time per device

// Byte offsets of the plain sweep's arrays inside the scratch (256-aligned).
struct ScratchLayout {
  size_t ctl, src, mult, rec_lo, depth, sigma, delta, coef, rec_v, rec_j, end;
  int64_t cap;
};
static ScratchLayout scratch_layout(int64_t n, int64_t m, int chunk) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  ScratchLayout L;
  L.cap = n + m / chunk + 1024;  // a plain sweep appends each vertex once (record_capacity)
  size_t o = 0;
  L.ctl = o; o = up(o + sizeof(Ctl));
  L.src = o; o = up(o + sizeof(int32_t) * 64);
  L.mult = o; o = up(o + sizeof(double) * 64);
  L.rec_lo = o; o = up(o + sizeof(int64_t) * (size_t)(n + 3));
  L.depth = o; o = up(o + sizeof(int32_t) * (size_t)n);
  L.sigma = o; o = up(o + sizeof(u64) * (size_t)n);
  L.delta = o; o = up(o + sizeof(double) * (size_t)n);
  L.coef = o; o = up(o + sizeof(double) * (size_t)n);
  L.rec_v = o; o = up(o + sizeof(int32_t) * (size_t)L.cap);
  L.rec_j = o; o = up(o + sizeof(int32_t) * (size_t)L.cap);
  L.end = o;
  return L;
}

size_t static_scratch_need(const DevGraph& g) { return scratch_layout(g.n, g.m, g.chunk).end; }

int bc_run_static(const DevGraph& g, double* bc, hipStream_t st, int device, const Params& prm,
                  const int32_t* sources, int64_t nsrc, StaticRun* out) {
  *out = StaticRun();
  const double t0 = now_ms();
  if (device < 0 || device >= kMaxDevices) return set_error(BCMI_ERR_ARG, "device index out of range");
  TRY(static_scratch_init(device));  // the first call on this device looks the array up
  if (g.heuristic) return set_error(BCMI_ERR_STATE, "peeled graphs run through bcmi_bc_run");
  const ScratchLayout L = scratch_layout(g.n, g.m, g.chunk);
  if (L.end > kStaticScratchBytes)
    return set_error(BCMI_ERR_ARG, "graph needs " + std::to_string(L.end) + " bytes of sweep state; the static scratch holds " +
                                       std::to_string(kStaticScratchBytes) + " (about 1.5M vertices at 16 arcs each)");
  for (int64_t i = 0; i < nsrc; i++)
    if (sources[i] < 0 || sources[i] >= g.n) return set_error(BCMI_ERR_RANGE, "source id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu[device]);
  char* base = (char*)static_scratch_ptr(device);
  PA pa;
  pa.ctl = (Ctl*)(base + L.ctl);
  pa.depth = (int32_t*)(base + L.depth);
  pa.sigma = (u64*)(base + L.sigma);
  pa.delta = (double*)(base + L.delta);
  pa.coef = (double*)(base + L.coef);
  pa.done = nullptr;
  pa.rec_v = (int32_t*)(base + L.rec_v);
  pa.rec_j = (int32_t*)(base + L.rec_j);
  pa.cap = L.cap;
  pa.rec_lo = (int64_t*)(base + L.rec_lo);
  pa.srcbuf = (const int32_t*)(base + L.src);
  pa.multbuf = (const double*)(base + L.mult);
  pa.bc = bc;
  pa.alpha = prm.alpha_plain;
  pa.beta = prm.beta_plain;
  const GA ga = make_ga(g, prm);
  const int64_t n = g.n;
  const int G = grid_for(n);
  // the previous call (maybe on another graph) left its state behind
  hipLaunchKernelGGL(k_fill_i32, G, 256, 0, st, pa.depth, -1, n);
  hipLaunchKernelGGL(k_fill_u64, G, 256, 0, st, pa.sigma, 0ull, n);
  hipLaunchKernelGGL(k_fill_f64, G, 256, 0, st, pa.delta, 0.0, n);
  hipLaunchKernelGGL(k_fill_f64, G, 256, 0, st, pa.coef, 0.0, n);
  hipLaunchKernelGGL(k_fill_f64, 1, 64, 0, st, (double*)pa.multbuf, 1.0, (int64_t)1);
  const int Gk = prm.grid_plain, B = prm.block, steps = prm.steps;
  KT* const nokt = nullptr;
  int ovf = 0;
  for (int64_t i = 0; i < nsrc; i++) {
    const int32_t s = g.to_int(sources[i]);
    if (g.h_off[s + 1] == g.h_off[s]) continue;  // degree 0: skipped (bc2d.c:1026)
    hipLaunchKernelGGL(k_fill_i32, 1, 64, 0, st, (int32_t*)pa.srcbuf, s, (int64_t)1);
    hipLaunchKernelGGL(k_plain_init, 1, 1, 0, st, ga, pa);
    Ctl hc;
    for (;;) {  // forward levels, `steps` at a time, until the decide kernel closes the sweep
      for (int t = 0; t < steps; t++) {
        hipLaunchKernelGGL(k_plain_decide, 1, 1, 0, st, ga, pa, 1);
        hipLaunchKernelGGL(k_plain_td, Gk, B, 0, st, ga, pa, nokt);
        hipLaunchKernelGGL(k_plain_bu, Gk, B, 0, st, ga, pa, nokt);
      }
      HIP_TRY(hipMemcpyAsync(&hc, pa.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      if (hc.fdone) break;
    }
    if (hc.overflow) return set_error(BCMI_ERR_STATE, "plain sweep overflowed its work list");
    while (!hc.bdone) {  // backward levels, deepest first
      for (int t = 0; t < steps; t++) {
        hipLaunchKernelGGL(k_bdecide, 1, 1, 0, st, pa.ctl);
        hipLaunchKernelGGL(k_plain_back, Gk, B, 0, st, ga, pa, nokt);
        hipLaunchKernelGGL(k_plain_bfin, Gk, B, 0, st, ga, pa);
      }
      HIP_TRY(hipMemcpyAsync(&hc, pa.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
    }
    hipLaunchKernelGGL(k_plain_cleanup, grid_for(std::min<int64_t>(hc.rec_count, L.cap)), 256, 0, st, pa);
    if (hc.sigma_ovf) ovf = 1;
    out->sweeps++;
    out->sum_n += hc.tot_verts;
    out->sum_m += hc.tot_arcs;
    out->sum_a += hc.sum_a;
    out->bytes += 16.0 * hc.tot_arcs + 24.0 * hc.sum_a + 64.0 * hc.tot_verts;
    out->max_levels = std::max<int64_t>(out->max_levels, hc.last + 1);
  }
  HIP_TRY(hipStreamSynchronize(st));
  HIP_TRY(hipGetLastError());
  out->te = 2.0 * (double)out->sum_m;
  out->wall_ms = now_ms() - t0;
  if (ovf) return set_error(BCMI_ERR_OVERFLOW, "a shortest-path count exceeded 2^64-1");
  return BCMI_OK;
}

int sssp_debug(const DevGraph& g, Workspace* ws, const Params& prm, int32_t s, int32_t* depth, uint64_t* sigma,
               double* delta) {
  if (ws->mode != BCMI_MODE_PLAIN) return set_error(BCMI_ERR_STATE, "plain workspace required");
  if (s < 0 || s >= g.n) return set_error(BCMI_ERR_RANGE, "source out of range");
  hipStream_t st = ws->st;
  const int32_t s_caller = s;
  s = g.to_int(s);
  // Sweep into a scratch BC (bc unchanged): the multiplier is 0.
  ws->hsrc[0] = s;
  ws->hmult[0] = 0.0;
  HIP_TRY(hipMemcpyAsync(ws->srcbuf, ws->hsrc, sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ws->multbuf, ws->hmult, sizeof(double), hipMemcpyHostToDevice, st));
  int lv;
  TRY(sweep(g, ws, prm, 1, true, &lv));
  int64_t n = g.n;
  HIP_TRY(hipMemcpyAsync(depth, ws->depth, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(sigma, ws->sigma1, sizeof(u64) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(delta, ws->delta1, sizeof(double) * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int64_t v = 0; v < n; v++)
    if (depth[v] < 0) { sigma[v] = 0; delta[v] = 0.0; }
  // the source's delta is not computed by the sweep (depths D-2..1 only)
  delta[s] = 0.0;
  g.to_caller(depth);
  g.to_caller(sigma);
  g.to_caller(delta);
  (void)s_caller;
  int rc = sweep_cleanup(g, ws, prm);
  if (ws->hctl->sigma_ovf) return set_error(BCMI_ERR_OVERFLOW, "a shortest-path count exceeded 2^64-1");
  return rc;
}

int msbfs_debug(const DevGraph& g, Workspace* ws, const Params& prm, const int32_t* src, int k, int32_t* depth,
                uint64_t* sigma, double* delta) {
  if (ws->mode != BCMI_MODE_BATCHED) return set_error(BCMI_ERR_STATE, "batched workspace required");
  if (k < 1 || k > 64) return set_error(BCMI_ERR_ARG, "1..64 sources");
  for (int q = 0; q < k; q++)
    if (src[q] < 0 || src[q] >= g.n) return set_error(BCMI_ERR_RANGE, "source out of range");
  hipStream_t st = ws->st;
  int64_t n = g.n;
  if (!ws->dbg_delta) {
    ALLOC_LOCK;
    TRY(dalloc(&ws->dbg_delta, 64 * n));
    free_graphs(ws);
  }
  hipLaunchKernelGGL(k_fill_f64, grid_for(64 * n), 256, 0, st, ws->dbg_delta, 0.0, 64 * n);
  for (int q = 0; q < 64; q++) {
    ws->hsrc[q] = q < k ? g.to_int(src[q]) : 0;
    ws->hmult[q] = 0.0;
  }
  HIP_TRY(hipMemcpyAsync(ws->srcbuf, ws->hsrc, 64 * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(ws->multbuf, ws->hmult, 64 * sizeof(double), hipMemcpyHostToDevice, st));
  int lv;
  int narrow = std::min(prm.narrow_sigma, g.narrow_cap.load());
  int rc = sweep(g, ws, prm, k, true, &lv, 0, narrow);
  while (rc == BCMI_OK && narrow > 0 && ws->hctl->sigma_ovf) {
    // counts past the row's range: redo with the next wider rows
    narrow--;
    lower_narrow_cap(g, narrow);
    rc = sweep_cleanup(g, ws, prm);
    if (rc == BCMI_OK) rc = sweep(g, ws, prm, k, true, &lv, 0, narrow);
  }
  if (rc != BCMI_OK) return rc;
  const int nw = ws->hctl->narrow;
  int64_t nrec = std::min<int64_t>(ws->hctl->rec_count, ws->cap);
  int last = ws->hctl->last;
  std::vector<int32_t> rv(nrec), rj(nrec);
  std::vector<u64> rm(nrec);
  std::vector<int64_t> lo(last + 2);
  std::vector<u64> sg((size_t)64 * n);
  std::vector<double> dd((size_t)64 * n);
  HIP_TRY(hipMemcpyAsync(rv.data(), ws->rec_v, sizeof(int32_t) * nrec, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rj.data(), ws->rec_j, sizeof(int32_t) * nrec, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(rm.data(), ws->rec_mask, sizeof(u64) * nrec, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(lo.data(), ws->rec_lo, sizeof(int64_t) * (last + 2), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(sg.data(), ws->sigma, sizeof(u64) * 64 * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(dd.data(), ws->dbg_delta, sizeof(double) * 64 * n, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (nw > 0) {  // narrow rows: widen in place, from the end
    const char* sn = reinterpret_cast<const char*>(sg.data());
    const size_t wb = nw == 2 ? 2 : 4;
    for (size_t i = (size_t)64 * n; i-- > 0;) {
      uint32_t x = 0;
      memcpy(&x, sn + wb * i, wb);
      sg[i] = x;
    }
  }
  for (int q = 0; q < k; q++)
    for (int64_t v = 0; v < n; v++) { depth[(int64_t)q * n + v] = -1; sigma[(int64_t)q * n + v] = 0; delta[(int64_t)q * n + v] = 0.0; }
  for (int d = 0; d <= last; d++)
    for (int64_t r = lo[d]; r < lo[d + 1]; r++) {
      if (rj[r] != 0) continue;
      int32_t v = rv[r];
      u64 m = rm[r];
      for (int q = 0; q < k; q++)
        if ((m >> q) & 1) {
          depth[(int64_t)q * n + v] = d;
          sigma[(int64_t)q * n + v] = sg[(size_t)v * 64 + q];
          delta[(int64_t)q * n + v] = d >= 1 ? dd[(size_t)v * 64 + q] : 0.0;
        }
    }
  for (int q = 0; q < k; q++) {
    g.to_caller(depth + (int64_t)q * n);
    g.to_caller(sigma + (int64_t)q * n);
    g.to_caller(delta + (int64_t)q * n);
  }
  rc = sweep_cleanup(g, ws, prm);
  if (ws->hctl->sigma_ovf) return set_error(BCMI_ERR_OVERFLOW, "a shortest-path count exceeded 2^64-1");
  return rc;
}

}  // namespace bcmi
