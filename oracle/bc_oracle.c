/*
 * bc_oracle.c -- CPU restatement of the ck-bc `program/bc-cuda` Brandes
 * betweenness-centrality path.  TEST INFRASTRUCTURE ONLY: this file is the
 * checker that the HIP engine is compared against.  Only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg may load it; the engine never links
 * or calls it.
 *
 * Arithmetic: sigma (shortest-path counts) in uint64 with overflow detection,
 * delta and BC in IEEE binary64.  Every function cites the reference lines it
 * restates (paths relative to the reference repository root; "c" is
 * program/bc-cuda/bc2d.c, "cu" is program/bc-cuda/cuda_kernels.cu).
 *
 * Deliberate differences from the reference (each is a reference defect):
 *   - sigma is uint64 (reference: uint32 that wraps, cu:536);
 *   - delta/BC are FP64 (reference: float32 with atomics, cu:592-595);
 *   - unreached vertices receive no BC (reference adds delta=reach to them,
 *     cu:263-277 covers all N);
 *   - no 1000-level cap (reference MAX_LVL, bc2d.h:25);
 *   - degree-1 reduction is iterative (reference: a single pass, c:314-419),
 *     with exact BC for pruned vertices (reference leaves them at 0).
 *
 * Past 2^64, orc_brandes_scaled keeps path counts as binary64 values with a
 * power-of-two scale per level (not exact; pinned against the big-integer
 * Brandes of oracle.py and against orc_brandes where counts fit).
 *
 * Parity pins (tests/test_oracle.py): closed forms (paths, stars, cycles,
 * cliques, grids, hypercubes incl. sigma overflow), brute-force all-pairs path
 * counting on small random graphs, networkx 3.4.2 golden vectors (facebook
 * sources 0..999, R-MAT scale 10), peeling == no-peeling BC invariance, and
 * the single-pass reduction vs the iterative peel's first round on facebook.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <ctype.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_OVERFLOW 1
#define ORC_ERR_ARG 2
#define ORC_ERR_IO 3
#define ORC_ERR_RANGE 4
#define ORC_ERR_NOMEM 5

/* ------------------------------------------------------------------------ */
/* Graph ingest: restates read_graph (c:175-254), mirror (c:152-170),        */
/* norm_graph (c:430-453, cmpedge c:297-312 at a 1x1 mesh) and build_csc     */
/* (c:472-501).                                                              */
/* ------------------------------------------------------------------------ */

static int cmp_pair(const void *a, const void *b) {
    const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

/* Reads `u v` lines; '#' lines are comments (c:218-226).  Returns the number
 * of arcs after mirroring (c:152-170: reversed copy of every non-loop edge
 * unless directed), or -ORC_ERR_* on failure.  *bad_line gets the 1-based
 * count of data lines read when an id >= n is found (c:232-237). */
long long orc_read_edges(const char *path, uint64_t n, int directed,
                         uint64_t **arcs_out, uint64_t *bad_line) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -ORC_ERR_IO;
    size_t cap = 1 << 16, cnt = 0;
    uint64_t *ed = (uint64_t *)malloc(cap * 2 * sizeof(uint64_t));
    char line[1024];
    uint64_t lc = 0, pu = 0, pv = 0;
    int have_prev = 0;
    while (fgets(line, sizeof line, fp)) {
        if (strchr(line, '#')) continue;
        if (line[0] == '\0') continue;
        lc++;
        unsigned long long u, v;
        if (sscanf(line, "%llu %llu", &u, &v) != 2) {
            /* c:227-230: a failed sscanf leaves i,j unchanged, so the
             * previous edge is repeated; dedupe removes it later. */
            if (!have_prev) continue;
            u = pu; v = pv;
        }
        if (u >= n || v >= n) {
            free(ed); fclose(fp);
            if (bad_line) *bad_line = lc;
            return -ORC_ERR_RANGE;
        }
        if (cnt == cap) { cap *= 2; ed = (uint64_t *)realloc(ed, cap * 2 * sizeof(uint64_t)); }
        ed[2 * cnt] = u; ed[2 * cnt + 1] = v; cnt++;
        pu = u; pv = v; have_prev = 1;
    }
    fclose(fp);
    if (!directed) {
        size_t extra = 0;
        for (size_t i = 0; i < cnt; i++) extra += ed[2 * i] != ed[2 * i + 1];
        ed = (uint64_t *)realloc(ed, (cnt + extra + 1) * 2 * sizeof(uint64_t));
        size_t k = cnt;
        for (size_t i = 0; i < cnt; i++)
            if (ed[2 * i] != ed[2 * i + 1]) { ed[2 * k] = ed[2 * i + 1]; ed[2 * k + 1] = ed[2 * i]; k++; }
        cnt = k;
    }
    *arcs_out = ed;
    return (long long)cnt;
}

void orc_free(void *p) { free(p); }

/* norm_graph (c:430-453): sort by (u,v), drop duplicates and self loops. */
uint64_t orc_normalize(uint64_t *arcs, uint64_t na) {
    if (na == 0) return 0;
    qsort(arcs, na, 2 * sizeof(uint64_t), cmp_pair);
    uint64_t l = 0;
    for (uint64_t i = 0; i < na; i++) {
        if (arcs[2 * i] == arcs[2 * i + 1]) continue;
        if (l > 0 && arcs[2 * (l - 1)] == arcs[2 * i] && arcs[2 * (l - 1) + 1] == arcs[2 * i + 1]) continue;
        arcs[2 * l] = arcs[2 * i]; arcs[2 * l + 1] = arcs[2 * i + 1]; l++;
    }
    return l;
}

/* build_csc (c:472-501): column j lists every u with arc (u,j), ascending
 * because the arcs arrive sorted by (u,v).  For a mirrored graph this is the
 * CSR adjacency. */
void orc_build_csc(const uint64_t *arcs, uint64_t na, uint64_t n, int64_t *off, int32_t *adj) {
    memset(off, 0, (n + 1) * sizeof(int64_t));
    for (uint64_t i = 0; i < na; i++) off[arcs[2 * i + 1] + 1]++;
    for (uint64_t j = 0; j < n; j++) off[j + 1] += off[j];
    int64_t *cur = (int64_t *)malloc(n * sizeof(int64_t));
    memcpy(cur, off, n * sizeof(int64_t));
    for (uint64_t i = 0; i < na; i++) adj[cur[arcs[2 * i + 1]]++] = (int32_t)arcs[2 * i];
    free(cur);
}

/* ------------------------------------------------------------------------ */
/* Single-source Brandes: forward rule cu:526-537, backward rule cu:590-596, */
/* seed cu:689-703, reset cu:1089-1119 (delta starts at the pruned-subtree   */
/* weight w, cu:1095).                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t *depth;
    uint64_t *sigma;
    double *delta;
    int32_t *queue;
    int64_t *lvl_off; /* start of each level in queue; size n+2 */
} orc_ws;

static int ws_alloc(orc_ws *w, int64_t n) {
    w->depth = (int32_t *)malloc(n * sizeof(int32_t));
    w->sigma = (uint64_t *)malloc(n * sizeof(uint64_t));
    w->delta = (double *)malloc(n * sizeof(double));
    w->queue = (int32_t *)malloc((n + 1) * sizeof(int32_t));
    w->lvl_off = (int64_t *)malloc((n + 2) * sizeof(int64_t));
    if (!w->depth || !w->sigma || !w->delta || !w->queue || !w->lvl_off) return -1;
    for (int64_t i = 0; i < n; i++) w->depth[i] = -1;
    return 0;
}

static void ws_free(orc_ws *w) {
    free(w->depth); free(w->sigma); free(w->delta); free(w->queue); free(w->lvl_off);
}

/* Runs forward + backward from s.  w == NULL means all weights 0.  On return
 * depth/sigma/delta hold the results for the reached vertices (queue[0..nq)).
 * Returns number of levels (>=1) or -1 on sigma overflow.
 * Stats: *n_s reached vertices, *m_s sum of degrees of reached vertices,
 * *a_s shortest-path DAG arcs, *nw sum over reached of (1 + w). */
static int64_t sssp(int64_t n, const int64_t *off, const int32_t *adj, int32_t s,
                    const uint64_t *w, orc_ws *ws, int64_t *nq_out,
                    int64_t *n_s, int64_t *m_s, int64_t *a_s, uint64_t *nw) {
    int32_t *depth = ws->depth; uint64_t *sigma = ws->sigma; double *delta = ws->delta;
    int32_t *q = ws->queue; int64_t *lo = ws->lvl_off;
    int overflow = 0;
    int64_t nq = 0, arcs_dag = 0, msum = 0; uint64_t wsum = 0;
    q[nq++] = s; depth[s] = 0; sigma[s] = 1;
    lo[0] = 0; lo[1] = 1;
    int64_t L = 0;
    /* Forward, level-synchronous (c:535-542).  For each arc c->r out of the
     * frontier: first visit sets depth (cu:524-530); when depth[r] == level,
     * sigma[r] += sigma[c] (cu:532-537). */
    for (;;) {
        int64_t b = lo[L], e = lo[L + 1];
        if (b == e) break;
        int32_t d1 = (int32_t)(L + 1);
        for (int64_t i = b; i < e; i++) {
            int32_t c = q[i];
            for (int64_t k = off[c]; k < off[c + 1]; k++) {
                int32_t r = adj[k];
                if (depth[r] == -1) { depth[r] = d1; sigma[r] = 0; q[nq++] = r; }
                if (depth[r] == d1) {
                    uint64_t o = sigma[r];
                    sigma[r] = o + sigma[c];
                    if (sigma[r] < o) overflow = 1;
                    arcs_dag++;
                }
            }
        }
        L++;
        lo[L + 1] = nq;
    }
    /* levels 0..L-1 are non-empty */
    for (int64_t i = 0; i < nq; i++) {
        int32_t v = q[i];
        delta[v] = w ? (double)w[v] : 0.0;
        msum += off[v + 1] - off[v];
        wsum += 1 + (w ? w[v] : 0);
    }
    /* Backward (c:543-551, cu:590-596): for depth D-2 .. 1, for each vertex c
     * at that depth and each neighbour r one level deeper,
     * delta[c] += (delta[r] + 1) / sigma[r] * sigma[c].  Sequential push over
     * c's neighbour list, so each delta[c] sums in adjacency order. */
    for (int64_t d = L - 2; d >= 1; d--) {
        for (int64_t i = lo[d]; i < lo[d + 1]; i++) {
            int32_t c = q[i];
            double acc = delta[c];
            double sc = (double)sigma[c];
            for (int64_t k = off[c]; k < off[c + 1]; k++) {
                int32_t r = adj[k];
                if (depth[r] == d + 1) {
                    double t = (delta[r] + 1.0) / (double)sigma[r];
                    acc = acc + t * sc; /* built with -ffp-contract=off */
                }
            }
            delta[c] = acc;
        }
    }
    *nq_out = nq;
    if (n_s) *n_s = nq;
    if (m_s) *m_s = msum;
    if (a_s) *a_s = arcs_dag;
    if (nw) *nw = wsum;
    (void)n;
    return overflow ? -1 : L;
}

static void ws_reset(orc_ws *ws, int64_t nq) {
    for (int64_t i = 0; i < nq; i++) ws->depth[ws->queue[i]] = -1;
}

/* One source, full outputs (for bit-exact depth/sigma/delta checks).
 * depth_out/sigma_out/delta_out are n-sized; unreached: -1 / 0 / 0.
 * Returns levels, or -ORC_ERR_OVERFLOW. */
int64_t orc_sssp(int64_t n, const int64_t *off, const int32_t *adj, int32_t s,
                 const uint64_t *w, int32_t *depth_out, uint64_t *sigma_out, double *delta_out) {
    orc_ws ws;
    if (ws_alloc(&ws, n)) return -ORC_ERR_NOMEM;
    int64_t nq;
    int64_t L = sssp(n, off, adj, s, w, &ws, &nq, NULL, NULL, NULL, NULL);
    for (int64_t v = 0; v < n; v++) { depth_out[v] = -1; sigma_out[v] = 0; delta_out[v] = 0.0; }
    for (int64_t i = 0; i < nq; i++) {
        int32_t v = ws.queue[i];
        depth_out[v] = ws.depth[v]; sigma_out[v] = ws.sigma[v]; delta_out[v] = ws.delta[v];
    }
    ws_free(&ws);
    return L < 0 ? -ORC_ERR_OVERFLOW : L;
}

/* Weighted multi-source Brandes accumulation (c:1017-1050 source loop,
 * cu:263-277 update_bc with the source multiplier (reach_s + 1)):
 *   bc[v] += mult[i] * delta_{src[i]}(v) for every reached v != src[i].
 * mult == NULL means 1.  Sources with degree 0 contribute nothing
 * (c:1026-1035).  bc is ordered-pair (not halved).  Per-source statistics are
 * optional.  nw_out[i] = sum over reached of (1 + w).  OpenMP over sources
 * with thread-private BC partials summed in thread order.
 * Returns ORC_OK or ORC_ERR_OVERFLOW. */
int orc_brandes(int64_t n, const int64_t *off, const int32_t *adj,
                const int32_t *src, int64_t nsrc, const uint64_t *w, const uint64_t *mult,
                double *bc, int nthreads,
                int64_t *n_s, int64_t *m_s, int64_t *a_s, uint64_t *nw_out, int32_t *levels_out) {
    int nt = nthreads > 0 ? nthreads : 1;
    double *part = (double *)calloc((size_t)nt * n, sizeof(double));
    if (!part) return ORC_ERR_NOMEM;
    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *my = part + (size_t)tid * n;
        orc_ws ws;
        if (ws_alloc(&ws, n)) { bad = ORC_ERR_NOMEM; }
        else {
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t i = 0; i < nsrc; i++) {
                int32_t s = src[i];
                if (off[s + 1] == off[s]) {
                    if (n_s) n_s[i] = 1;
                    if (m_s) m_s[i] = 0;
                    if (a_s) a_s[i] = 0;
                    if (nw_out) nw_out[i] = 1 + (w ? w[s] : 0);
                    if (levels_out) levels_out[i] = 1;
                    continue;
                }
                int64_t nq;
                int64_t ns_, ms_, as_; uint64_t nw_;
                int64_t L = sssp(n, off, adj, s, w, &ws, &nq, &ns_, &ms_, &as_, &nw_);
                if (L < 0) bad = ORC_ERR_OVERFLOW;
                double m = mult ? (double)mult[i] : 1.0;
                for (int64_t k = 1; k < nq; k++) {
                    int32_t v = ws.queue[k];
                    my[v] += ws.delta[v] * m;
                }
                if (n_s) n_s[i] = ns_;
                if (m_s) m_s[i] = ms_;
                if (a_s) a_s[i] = as_;
                if (nw_out) nw_out[i] = nw_;
                if (levels_out) levels_out[i] = (int32_t)L;
                ws_reset(&ws, nq);
            }
            ws_free(&ws);
        }
    }
    for (int64_t v = 0; v < n; v++) {
        double acc = 0.0;
        for (int t = 0; t < nt; t++) acc += part[(size_t)t * n + v];
        bc[v] += acc;
    }
    free(part);
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Degree-1 reduction.                                                       */
/* ------------------------------------------------------------------------ */

/* Restatement of the reference's SINGLE pass degree_reduction (c:314-419)
 * followed by init_bc_1degree (c:455-465).  Input: arcs sorted by (u,v),
 * duplicate- and loop-free (norm_graph output), na >= 1.  Outputs: kept arcs
 * are compacted in place (return value = count); removed[2*k] = v,
 * removed[2*k+1] = u for each pendant u with single neighbour v (o_ed,
 * c:349-354), *nrem = pnod (c:413), reach[v] += 1 per consumed removed arc
 * (c:459-464, uses the first pnod entries of the sorted o_ed). */
uint64_t orc_ref_degree_reduction(uint64_t *arcs, uint64_t na, uint64_t *removed,
                                  uint64_t *nrem, uint64_t *ncouple_out, uint32_t *reach) {
    uint64_t *n_ed = (uint64_t *)malloc(2 * na * sizeof(uint64_t));
    uint64_t nod = 0, ne = 0, pnod = 0, ncouple = 0;
    uint64_t prev_u = (uint64_t)-1, u, v;
    for (uint64_t i = 0; i + 1 < na; i++) {
        u = arcs[2 * i]; v = arcs[2 * i + 1];
        uint64_t nu = arcs[2 * i + 2], nv = arcs[2 * i + 3];
        if (u == v || (u == nu && v == nv)) { prev_u = u; continue; }
        if (u != nu && u != prev_u) { removed[2 * nod] = v; removed[2 * nod + 1] = u; nod++; }
        else { n_ed[2 * ne] = u; n_ed[2 * ne + 1] = v; ne++; }
        prev_u = u;
    }
    u = arcs[2 * na - 2]; v = arcs[2 * na - 1];
    if (u == prev_u) { n_ed[2 * ne] = u; n_ed[2 * ne + 1] = v; ne++; }
    else if (u != v) { removed[2 * nod] = v; removed[2 * nod + 1] = u; nod++; }
    qsort(removed, nod, 2 * sizeof(uint64_t), cmp_pair);
    uint64_t out = 0;
    if (nod > 0) {
        for (uint64_t i = 0; i < ne; i++) {
            while (pnod < nod && n_ed[2 * i] > removed[2 * pnod]) { ncouple++; pnod++; }
            if (pnod < nod && n_ed[2 * i] == removed[2 * pnod] && n_ed[2 * i + 1] == removed[2 * pnod + 1]) {
                pnod++;
                continue;
            }
            arcs[2 * out] = n_ed[2 * i]; arcs[2 * out + 1] = n_ed[2 * i + 1]; out++;
        }
    } else {
        memcpy(arcs, n_ed, 2 * ne * sizeof(uint64_t));
        out = ne;
    }
    free(n_ed);
    for (uint64_t k = 0; k < pnod; k++) reach[removed[2 * k]]++;
    *nrem = pnod;
    if (ncouple_out) *ncouple_out = ncouple;
    return out;
}

/* Iterative, round-synchronous pendant peeling (the engine's extension of
 * c:314-419).  Round r: candidates are live vertices whose live degree is
 * exactly 1 at the start of the round.  A candidate u with live neighbour p
 * is removed in round r, unless p is also a candidate and p > u (an isolated
 * edge keeps its smaller endpoint).  On removal: parent[u] = p, and
 * w[p] += w[u] + 1 (w = number of vertices in the pruned trees hanging at
 * a vertex).  Outputs (n-sized): round_out[v] = removal round (1-based) or 0
 * for core vertices; parent_out (-1 for core); w_out; core_deg_out (live
 * degree after peeling).  Returns the number of rounds. */
int32_t orc_peel(int64_t n, const int64_t *off, const int32_t *adj,
                 int32_t *round_out, int32_t *parent_out, uint64_t *w_out, int32_t *core_deg_out) {
    int32_t *deg = core_deg_out;
    int32_t *cand = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t *list = (int32_t *)malloc(n * sizeof(int32_t));
    for (int64_t v = 0; v < n; v++) {
        deg[v] = (int32_t)(off[v + 1] - off[v]);
        round_out[v] = 0; parent_out[v] = -1; w_out[v] = 0; cand[v] = 0;
    }
    int32_t r = 0;
    for (;;) {
        int64_t nl = 0;
        for (int64_t v = 0; v < n; v++)
            if (round_out[v] == 0 && deg[v] == 1) list[nl++] = (int32_t)v;
        if (nl == 0) break;
        r++;
        for (int64_t i = 0; i < nl; i++) cand[list[i]] = r;
        /* Decide removals against start-of-round state. */
        int64_t nrm = 0;
        for (int64_t i = 0; i < nl; i++) {
            int32_t u = list[i], p = -1;
            for (int64_t k = off[u]; k < off[u + 1]; k++)
                if (round_out[adj[k]] == 0) { p = adj[k]; break; }
            if (cand[p] == r && p > u) continue;
            parent_out[u] = p;
            list[nrm++] = u;
        }
        for (int64_t i = 0; i < nrm; i++) round_out[list[i]] = r;
        for (int64_t i = 0; i < nrm; i++) {
            int32_t u = list[i], p = parent_out[u];
            w_out[p] += w_out[u] + 1;
            deg[p]--;
            deg[u] = 0;
        }
    }
    free(cand); free(list);
    return r;
}

/* Connected-component labels: label[v] = smallest vertex id in v's
 * component; comp_size[label] = component size (other entries 0). */
void orc_components(int64_t n, const int64_t *off, const int32_t *adj, int32_t *label, int64_t *comp_size) {
    int32_t *q = (int32_t *)malloc(n * sizeof(int32_t));
    for (int64_t v = 0; v < n; v++) { label[v] = -1; comp_size[v] = 0; }
    for (int64_t s = 0; s < n; s++) {
        if (label[s] != -1) continue;
        int64_t h = 0, t = 0;
        q[t++] = (int32_t)s; label[s] = (int32_t)s;
        while (h < t) {
            int32_t c = q[h++];
            for (int64_t k = off[c]; k < off[c + 1]; k++)
                if (label[adj[k]] == -1) { label[adj[k]] = (int32_t)s; q[t++] = adj[k]; }
        }
        comp_size[s] = t;
    }
    free(q);
}

/* Full BC pipeline over a multiset of sources on an undirected CSR graph.
 *   heuristic 0: plain Brandes, bc[v] = sum_{x in S} delta_x(v).
 *   heuristic 1: peel, run weighted Brandes on the core from each root c
 *                whose pruned tree holds sources (multiplicity cntS(c)), then
 *                add the closed-form terms for pruned vertices and roots.
 * expand != 0 (reference source semantics, c:1026 + cu:274): the source set
 * is first replaced by the union of T(c) over listed sources c that are core
 * vertices; listed pruned vertices are dropped.
 * The result (ordered pairs, not halved) equals plain Brandes over the same
 * (possibly expanded) source set on the original graph.  bc must be zeroed by
 * the caller.  Returns ORC_OK / error. */
int orc_bc_pipeline(int64_t n, const int64_t *off, const int32_t *adj,
                    const int32_t *src, int64_t nsrc, int heuristic, int expand,
                    double *bc, int nthreads) {
    if (heuristic == 0 && !expand)
        return orc_brandes(n, off, adj, src, nsrc, NULL, NULL, bc, nthreads, NULL, NULL, NULL, NULL, NULL);
    int32_t *round = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t *parent = (int32_t *)malloc(n * sizeof(int32_t));
    uint64_t *w = (uint64_t *)calloc(n, sizeof(uint64_t));
    int32_t *cdeg = (int32_t *)malloc(n * sizeof(int32_t));
    int32_t nr = 0;
    if (heuristic) nr = orc_peel(n, off, adj, round, parent, w, cdeg);
    else for (int64_t v = 0; v < n; v++) { round[v] = 0; parent[v] = -1; }
    /* core CSR */
    int64_t *coff = (int64_t *)malloc((n + 1) * sizeof(int64_t));
    int64_t tot = 0;
    coff[0] = 0;
    for (int64_t v = 0; v < n; v++) {
        if (round[v] == 0)
            for (int64_t k = off[v]; k < off[v + 1]; k++) tot += round[adj[k]] == 0;
        coff[v + 1] = tot;
    }
    int32_t *cadj = (int32_t *)malloc((tot + 1) * sizeof(int32_t));
    for (int64_t v = 0, p = 0; v < n; v++)
        if (round[v] == 0)
            for (int64_t k = off[v]; k < off[v + 1]; k++)
                if (round[adj[k]] == 0) cadj[p++] = adj[k];
    /* order of removal: rounds ascending */
    int32_t *order = (int32_t *)malloc(n * sizeof(int32_t));
    int64_t no = 0;
    {
        int64_t *cnt = (int64_t *)calloc((size_t)nr + 2, sizeof(int64_t));
        for (int64_t v = 0; v < n; v++) if (round[v] > 0) cnt[round[v]]++;
        int64_t acc = 0;
        for (int32_t r = 1; r <= nr; r++) { int64_t c = cnt[r]; cnt[r] = acc; acc += c; }
        for (int64_t v = 0; v < n; v++) if (round[v] > 0) order[cnt[round[v]]++] = (int32_t)v;
        no = acc;
        free(cnt);
    }
    /* root of every vertex (core: itself): process removals in descending round */
    int32_t *root = (int32_t *)malloc(n * sizeof(int32_t));
    for (int64_t v = 0; v < n; v++) root[v] = (int32_t)v;
    for (int64_t i = no - 1; i >= 0; i--) {
        int32_t u = order[i];
        root[u] = root[parent[u]];
    }
    /* source multiset counts */
    uint64_t *cnt0 = (uint64_t *)calloc(n, sizeof(uint64_t));
    if (expand) {
        uint8_t *listed = (uint8_t *)calloc(n, 1);
        for (int64_t i = 0; i < nsrc; i++) if (round[src[i]] == 0) listed[src[i]] = 1;
        for (int64_t v = 0; v < n; v++) if (listed[root[v]]) cnt0[v] = 1;
        free(listed);
    } else {
        for (int64_t i = 0; i < nsrc; i++) cnt0[src[i]]++;
    }
    /* bottom-up tree accumulation in removal order:
     * cntS(u) = cnt0(u) + sum children cntS; sa(u) = sum children cntS*(1+w) */
    uint64_t *cntS = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *sa = (uint64_t *)calloc(n, sizeof(uint64_t));
    for (int64_t v = 0; v < n; v++) cntS[v] = cnt0[v];
    for (int64_t i = 0; i < no; i++) {
        int32_t u = order[i], p = parent[u];
        cntS[p] += cntS[u];
        sa[p] += cntS[u] * (1 + w[u]);
    }
    /* components on the original graph */
    int32_t *label = (int32_t *)malloc(n * sizeof(int32_t));
    int64_t *csize = (int64_t *)malloc(n * sizeof(int64_t));
    orc_components(n, off, adj, label, csize);
    uint64_t *srcK = (uint64_t *)calloc(n, sizeof(uint64_t));
    for (int64_t v = 0; v < n; v++) srcK[label[v]] += cnt0[v];
    /* core BFS sources: roots with cntS > 0, in increasing id order */
    int64_t ncs = 0;
    for (int64_t v = 0; v < n; v++) if (round[v] == 0 && cntS[v] > 0) ncs++;
    int32_t *cs = (int32_t *)malloc((ncs + 1) * sizeof(int32_t));
    uint64_t *cm = (uint64_t *)malloc((ncs + 1) * sizeof(uint64_t));
    ncs = 0;
    for (int64_t v = 0; v < n; v++)
        if (round[v] == 0 && cntS[v] > 0) { cs[ncs] = (int32_t)v; cm[ncs] = cntS[v]; ncs++; }
    int rc = orc_brandes(n, coff, cadj, cs, ncs, w, cm, bc, nthreads, NULL, NULL, NULL, NULL, NULL);
    /* closed-form terms (integer exact, then converted) */
    for (int64_t v = 0; v < n; v++) {
        uint64_t nK = (uint64_t)csize[label[v]];
        uint64_t desc = cntS[v] - cnt0[v];
        int64_t t = (int64_t)(desc * (nK - 1)) - (int64_t)sa[v];
        if (round[v] > 0) t += (int64_t)(w[v] * (srcK[label[v]] - cntS[v]));
        bc[v] += (double)t;
    }
    free(round); free(parent); free(w); free(cdeg); free(coff); free(cadj); free(order);
    free(root); free(cnt0); free(cntS); free(sa); free(label); free(csize); free(srcK);
    free(cs); free(cm);
    return rc;
}

/* Algorithmic byte model for the roofline (per source / per batch).
 * plain:   B(s) = 16 m_s + 24 a_s + 64 n_s
 * batched: B(S) = 24 m_S + 32 n_S + sum_s (24 a_s + 32 n_s)
 * (m_S, n_S: union reach).  Provided here so tests can check the engine's
 * own statistics against an independent count. */
double orc_bytes_plain(int64_t n_s, int64_t m_s, int64_t a_s) {
    return 16.0 * (double)m_s + 24.0 * (double)a_s + 64.0 * (double)n_s;
}

/* Distance checksum per source: out[i] = sum over vertices t reached from
 * src[i], t != src[i], of (d(src[i], t) - 1).  It equals the sum over v of
 * the dependency delta_s(v) (every shortest s-t path has d(s,t) - 1 inner
 * vertices and the path fractions of each t sum to one), so it checks a BC
 * run without path counts - also where they exceed 2^64.  Plain BFS,
 * OpenMP over sources. */
int orc_depth_sums(int64_t n, const int64_t *off, const int32_t *adj, const int32_t *src, int64_t nsrc,
                   int nthreads, int64_t *out) {
    int err = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        int32_t *depth = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        if (!depth || !q) {
#pragma omp atomic write
            err = 1;
        } else {
            for (int64_t v = 0; v < n; v++) depth[v] = -1;
#pragma omp for schedule(dynamic, 1)
            for (int64_t i = 0; i < nsrc; i++) {
                int32_t s = src[i];
                int64_t nq = 0, sum = 0;
                q[nq++] = s;
                depth[s] = 0;
                for (int64_t h = 0; h < nq; h++) {
                    int32_t c = q[h];
                    for (int64_t k = off[c]; k < off[c + 1]; k++) {
                        int32_t r = adj[k];
                        if (depth[r] == -1) { depth[r] = depth[c] + 1; q[nq++] = r; sum += depth[r] - 1; }
                    }
                }
                for (int64_t h = 0; h < nq; h++) depth[q[h]] = -1;
                out[i] = sum;
            }
        }
        free(depth);
        free(q);
    }
    return err ? -ORC_ERR_NOMEM : 0;
}

/* Brandes with path counts past 2^64 (and past binary64's range): sigma is
 * held as a binary64 value per vertex times a power of two per level,
 * sigma(v) = sig[v] * 2^E[depth v].  After each forward level the level's
 * counts are rescaled by a power of two (exact) so that the largest lies in
 * [1, 2).  Same recurrences as sssp (cu:524-537, cu:590-596):
 * sigma[r] += sigma[c] over arcs to the next level, and
 * delta[c] += (delta[r] + 1) / sigma[r] * sigma[c] in adjacency order, where
 * sigma[c] / sigma[r] = sig[c] / sig[r] * 2^(E[d] - E[d+1]).  Counts above
 * 2^53 round, so the result carries binary64 rounding over the levels rather
 * than being exact; used on long bridged-block chains, where orc_brandes
 * reports overflow.  Ordered-pair BC added into bc (not halved); OpenMP over
 * sources with thread-private partials summed in thread order. */
int orc_brandes_scaled(int64_t n, const int64_t *off, const int32_t *adj, const int32_t *src, int64_t nsrc,
                       double *bc, int nthreads) {
    int nt = nthreads > 0 ? nthreads : 1;
    double *part = (double *)calloc((size_t)nt * (size_t)(n > 0 ? n : 1), sizeof(double));
    if (!part) return ORC_ERR_NOMEM;
    int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double *my = part + (size_t)tid * n;
        size_t nn = (size_t)(n > 0 ? n : 1);
        int32_t *depth = (int32_t *)malloc(nn * sizeof(int32_t));
        double *sig = (double *)malloc(nn * sizeof(double));
        double *delta = (double *)malloc(nn * sizeof(double));
        int32_t *q = (int32_t *)malloc((nn + 1) * sizeof(int32_t));
        int64_t *lo = (int64_t *)malloc((nn + 2) * sizeof(int64_t));
        int *shift = (int *)malloc((nn + 2) * sizeof(int));  /* E[d] - E[d-1] */
        if (!depth || !sig || !delta || !q || !lo || !shift) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
            bad = ORC_ERR_NOMEM;
        } else {
            for (int64_t v = 0; v < n; v++) depth[v] = -1;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
            for (int64_t i = 0; i < nsrc; i++) {
                int32_t s = src[i];
                if (off[s + 1] == off[s]) continue;
                int64_t nq = 0, L = 0;
                q[nq++] = s; depth[s] = 0; sig[s] = 1.0;
                lo[0] = 0; lo[1] = 1; shift[0] = 0;
                for (;;) {
                    int64_t b = lo[L], e = lo[L + 1];
                    if (b == e) break;
                    int32_t d1 = (int32_t)(L + 1);
                    for (int64_t k = b; k < e; k++) {
                        int32_t c = q[k];
                        for (int64_t a = off[c]; a < off[c + 1]; a++) {
                            int32_t r = adj[a];
                            if (depth[r] == -1) { depth[r] = d1; sig[r] = 0.0; q[nq++] = r; }
                            if (depth[r] == d1) sig[r] = sig[r] + sig[c];
                        }
                    }
                    L++;
                    lo[L + 1] = nq;
                    /* rescale the new level so that its largest count is in [1, 2) */
                    double mx = 0.0;
                    for (int64_t k = lo[L]; k < nq; k++) mx = sig[q[k]] > mx ? sig[q[k]] : mx;
                    int ex = 0;
                    if (mx > 0.0) { (void)frexp(mx, &ex); ex -= 1; }
                    shift[L] = ex;
                    for (int64_t k = lo[L]; k < nq; k++) sig[q[k]] = ldexp(sig[q[k]], -ex);
                }
                for (int64_t k = 0; k < nq; k++) delta[q[k]] = 0.0;
                for (int64_t d = L - 2; d >= 1; d--) {
                    for (int64_t k = lo[d]; k < lo[d + 1]; k++) {
                        int32_t c = q[k];
                        double acc = delta[c];
                        double sc = ldexp(sig[c], -shift[d + 1]);  /* sigma[c] in level d+1's units */
                        for (int64_t a = off[c]; a < off[c + 1]; a++) {
                            int32_t r = adj[a];
                            if (depth[r] == d + 1) {
                                double t = (delta[r] + 1.0) / sig[r];
                                acc = acc + t * sc;
                            }
                        }
                        delta[c] = acc;
                    }
                }
                for (int64_t k = 1; k < nq; k++) my[q[k]] += delta[q[k]];
                for (int64_t k = 0; k < nq; k++) depth[q[k]] = -1;
            }
        }
        free(depth); free(sig); free(delta); free(q); free(lo); free(shift);
    }
    for (int64_t v = 0; v < n; v++) {
        double acc = 0.0;
        for (int t = 0; t < nt; t++) acc += part[(size_t)t * n + v];
        bc[v] += acc;
    }
    free(part);
    return bad;
}

int orc_version(void) { return 1; }
