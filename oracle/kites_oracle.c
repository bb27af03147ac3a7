/*
 * kites_oracle.c -- CPU oracle for the vertex-centric iteration engine.
 *
 * TEST INFRASTRUCTURE ONLY (see kites_oracle.h for who may use it and
 * what pins each function).  Build: see oracle/Makefile.
 */
#define _GNU_SOURCE
#include "kites_oracle.h"

#include <sched.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define INF32 0xFFFFFFFFu

/* ---- functor sets: one C function per lambda of the engine's sets ---- */

/* nbrCompF: partial value contributed by in-neighbour value s over weight w */
static inline uint32_t nbr_comp(int algo, uint32_t s, uint32_t w) {
  switch (algo) {
    case KITES_ORACLE_SSSP_SAT: {
      if (s == INF32) return INF32;
      uint32_t t = s + w;
      return (t < s) ? INF32 : t;
    }
    case KITES_ORACLE_BFS:
      return (s == INF32) ? INF32 : s + 1u;
    case KITES_ORACLE_CC:
      return s;
    case KITES_ORACLE_WIDEST:
      return s < w ? s : w;
    default: /* KITES_ORACLE_SSSP_WRAP: examples/sssp_basic_single_device.cu:42-43 */
      return s + w;
  }
}
/* redF: if( lhs > rhs ) lhs = rhs   (sssp_basic_single_device.cu:44-45);
 * the WIDEST set uses max instead. */
static int g_widest = 0;
static inline void red_min(uint32_t* lhs, uint32_t rhs) {
  if (g_widest) { if (*lhs < rhs) *lhs = rhs; }
  else if (*lhs > rhs) *lhs = rhs;
}
/* updF: return new < old            (sssp_basic_single_device.cu:46-47);
 * the WIDEST set uses new > old. */
static inline int upd_lt(uint32_t nv, uint32_t ov) { return g_widest ? nv > ov : nv < ov; }

/* ---- Gauss-Seidel sweep: restates kites/graph/sanity_check.cuh:21-47 ---- */
uint32_t oracle_fixpoint_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                             const uint32_t* E, uint32_t* V, int algo,
                             uint32_t max_sweeps) {
  uint32_t sweeps = 0;
  int any;
  g_widest = (algo == KITES_ORACLE_WIDEST);
  do {
    any = 0;
    for (uint32_t v = 0; v < n; ++v) {
      uint32_t pre = V[v];
      uint32_t acc = pre;                           /* initF: loc = glob   */
      /* sanity_check.cuh:30 masks R with 0x7FFFFFFF (multi-GPU boundary
       * bit); single-device graphs never set it, so R is used as is.     */
      for (uint32_t e = R[v]; e < R[v + 1]; ++e) {
        uint32_t tmp = nbr_comp(algo, V[C[e]], E ? E[e] : 1u);
        red_min(&acc, tmp);
      }
      if (upd_lt(acc, pre)) {
        V[v] = acc;
        any = 1;
      }
    }
    ++sweeps;
    if (max_sweeps && sweeps >= max_sweeps && any) return 0;
  } while (any);
  return sweeps;
}

/* ---- synchronous activation-rule rounds (reference GPU semantics) ---- */
uint32_t oracle_frontier_rounds_u32(uint32_t n, const uint32_t* R,
                                    const uint32_t* C, const uint32_t* E,
                                    uint32_t* V, const uint8_t* frontier0,
                                    int algo, uint32_t max_rounds) {
  uint8_t* act = (uint8_t*)malloc(n ? n : 1);
  uint8_t* nxt = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t* newv = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint8_t* upd = (uint8_t*)calloc(n ? n : 1, 1);
  memcpy(act, frontier0, n);
  /* out-lists for the activation rule: transpose of the in-CSR */
  uint32_t* oR = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  uint64_t m = R[n];
  uint32_t* oC = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  for (uint64_t e = 0; e < m; ++e) oR[C[e] + 1]++;
  for (uint32_t v = 0; v < n; ++v) oR[v + 1] += oR[v];
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  memcpy(cur, oR, sizeof(uint32_t) * ((size_t)n + 1));
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t e = R[v]; e < R[v + 1]; ++e) oC[cur[C[e]]++] = v;

  uint32_t rounds = 0;
  g_widest = (algo == KITES_ORACLE_WIDEST);
  for (;;) {
    int any = 0;
    /* Jacobi round: every active vertex reads the values of the previous
     * round (one deterministic schedule among those the GPU may take). */
    for (uint32_t v = 0; v < n; ++v) {
      upd[v] = 0;
      if (!act[v]) continue;
      uint32_t acc = V[v];
      for (uint32_t e = R[v]; e < R[v + 1]; ++e)
        red_min(&acc, nbr_comp(algo, V[C[e]], E ? E[e] : 1u));
      if (upd_lt(acc, V[v])) { newv[v] = acc; upd[v] = 1; any = 1; }
    }
    memset(nxt, 0, n);
    for (uint32_t v = 0; v < n; ++v) {
      if (!upd[v]) continue;
      V[v] = newv[v];
      for (uint32_t e = oR[v]; e < oR[v + 1]; ++e) nxt[oC[e]] = 1;
    }
    ++rounds;
    if (!any) break;
    if (max_rounds && rounds >= max_rounds) { rounds = 0; break; }
    uint8_t* t = act; act = nxt; nxt = t;
  }
  free(act); free(nxt); free(newv); free(upd); free(oR); free(oC); free(cur);
  return rounds;
}

/* ---- textbook Dijkstra (binary heap), independent of the fixpoint ---- */
typedef struct { uint32_t d, v; } heap_item;

static void heap_push(heap_item* h, size_t* sz, heap_item it) {
  size_t i = (*sz)++;
  h[i] = it;
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    heap_item t = h[p]; h[p] = h[i]; h[i] = t;
    i = p;
  }
}
static heap_item heap_pop(heap_item* h, size_t* sz) {
  heap_item top = h[0];
  h[0] = h[--(*sz)];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, s = i;
    if (l < *sz && h[l].d < h[s].d) s = l;
    if (r < *sz && h[r].d < h[s].d) s = r;
    if (s == i) break;
    heap_item t = h[s]; h[s] = h[i]; h[i] = t;
    i = s;
  }
  return top;
}

static void transpose(uint32_t n, const uint32_t* R, const uint32_t* C,
                      const uint32_t* E, uint32_t** oR_, uint32_t** oC_,
                      uint32_t** oE_) {
  uint64_t m = R[n];
  uint32_t* oR = (uint32_t*)calloc((size_t)n + 1, sizeof(uint32_t));
  uint32_t* oC = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t* oE = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
  for (uint64_t e = 0; e < m; ++e) oR[C[e] + 1]++;
  for (uint32_t v = 0; v < n; ++v) oR[v + 1] += oR[v];
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  memcpy(cur, oR, sizeof(uint32_t) * ((size_t)n + 1));
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t e = R[v]; e < R[v + 1]; ++e) {
      uint32_t k = cur[C[e]]++;
      oC[k] = v;
      oE[k] = E ? E[e] : 1u;
    }
  free(cur);
  *oR_ = oR; *oC_ = oC; *oE_ = oE;
}

void oracle_dijkstra_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                         const uint32_t* E, uint32_t src, uint32_t* dist) {
  uint32_t *oR, *oC, *oE;
  transpose(n, R, C, E, &oR, &oC, &oE);
  uint64_t m = R[n];
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (m + 2));
  size_t sz = 0;
  uint8_t* done = (uint8_t*)calloc(n ? n : 1, 1);
  for (uint32_t v = 0; v < n; ++v) dist[v] = INF32;
  dist[src] = 0;
  heap_push(h, &sz, (heap_item){0, src});
  while (sz) {
    heap_item it = heap_pop(h, &sz);
    if (done[it.v]) continue;
    done[it.v] = 1;
    for (uint32_t e = oR[it.v]; e < oR[it.v + 1]; ++e) {
      uint32_t w = oE[e], u = oC[e];
      uint64_t nd = (uint64_t)it.d + w;
      uint32_t d32 = nd >= INF32 ? INF32 : (uint32_t)nd;
      if (d32 < dist[u]) {
        dist[u] = d32;
        heap_push(h, &sz, (heap_item){d32, u});
      }
    }
  }
  free(h); free(done); free(oR); free(oC); free(oE);
}

void oracle_widest_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                       const uint32_t* E, uint32_t src, uint32_t* width) {
  /* Max-bottleneck Dijkstra: settle vertices in decreasing width order;
   * the heap stores ~width so the min-heap pops the widest first. */
  uint32_t *oR, *oC, *oE;
  transpose(n, R, C, E, &oR, &oC, &oE);
  uint64_t m = R[n];
  heap_item* h = (heap_item*)malloc(sizeof(heap_item) * (m + 2));
  size_t sz = 0;
  uint8_t* done = (uint8_t*)calloc(n ? n : 1, 1);
  for (uint32_t v = 0; v < n; ++v) width[v] = 0;
  width[src] = INF32;
  heap_push(h, &sz, (heap_item){~INF32, src});
  while (sz) {
    heap_item it = heap_pop(h, &sz);
    if (done[it.v]) continue;
    done[it.v] = 1;
    uint32_t wv = width[it.v];
    for (uint32_t e = oR[it.v]; e < oR[it.v + 1]; ++e) {
      uint32_t u = oC[e], b = wv < oE[e] ? wv : oE[e];
      if (b > width[u]) { width[u] = b; heap_push(h, &sz, (heap_item){~b, u}); }
    }
  }
  free(h); free(done); free(oR); free(oC); free(oE);
}

void oracle_bfs_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                    uint32_t src, uint32_t* level) {
  uint32_t *oR, *oC, *oE;
  transpose(n, R, C, NULL, &oR, &oC, &oE);
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t v = 0; v < n; ++v) level[v] = INF32;
  size_t head = 0, tail = 0;
  level[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint32_t e = oR[u]; e < oR[u + 1]; ++e) {
      uint32_t v = oC[e];
      if (level[v] == INF32) { level[v] = level[u] + 1; q[tail++] = v; }
    }
  }
  free(q); free(oR); free(oC); free(oE);
}

static uint32_t uf_find(uint32_t* p, uint32_t x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

void oracle_cc_min_label(uint32_t n, const uint32_t* R, const uint32_t* C,
                         uint32_t* label) {
  uint32_t* p = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  for (uint32_t v = 0; v < n; ++v) p[v] = v;
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t e = R[v]; e < R[v + 1]; ++e) {
      uint32_t a = uf_find(p, v), b = uf_find(p, C[e]);
      if (a == b) continue;
      /* keep the smaller id as root so the root is the component minimum */
      if (a < b) p[b] = a; else p[a] = b;
    }
  for (uint32_t v = 0; v < n; ++v) label[v] = uf_find(p, v);
  free(p);
}

/* ---- ordered (Jacobi) iterations of a non-idempotent functor set ----
 * The combine order the engine defines for options::ordered
 * (kites_amd/csrc/kites/detail/ordered_kernel.hpp), restated for the
 * float PageRank functor set of examples/pagerank_ordered.hip:
 *   initF: acc = base;  nbrCompF: x = (damp * V[u]) * w;  redF: a = a + b;
 *   updF: update when |acc - old| > tol.
 * Per active vertex: acc = base; for each 64-edge chunk of its in-list in
 * order, x[0..k) as above, then for s = 1, 2, ..., 32: x[l] = x[l] + x[l+s]
 * for l divisible by 2s with l+s < k, then acc = acc + x[0].  Jacobi: reads
 * the previous iteration's V; active vertices follow the reference
 * (kites/device/kernels_single_device.cuh:119-165: the out-neighbours of
 * the vertices that changed).  Returns the iterations run (the last one
 * changes nothing) or 0 when max_iter is reached first. */
uint32_t oracle_ordered_pagerank_f32(uint32_t n, const uint32_t* R, const uint32_t* C,
                                     const float* W, const uint32_t* oR,
                                     const uint32_t* oC, float* V, const uint8_t* active0,
                                     float base, float damp, float tol, uint32_t max_iter) {
  float* next = (float*)malloc(sizeof(float) * (n ? n : 1));
  uint8_t* act = (uint8_t*)malloc(n ? n : 1);
  uint8_t* nact = (uint8_t*)malloc(n ? n : 1);
  memcpy(act, active0, n);
  uint32_t it = 0, done = 0;
  while (it < max_iter && !done) {
    memcpy(next, V, sizeof(float) * n);
    memset(nact, 0, n);
    int changed = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (!act[v]) continue;
      float acc = base;
      for (uint32_t c0 = R[v]; c0 < R[v + 1]; c0 += 64) {
        uint32_t const k = (R[v + 1] - c0) < 64 ? (R[v + 1] - c0) : 64;
        float x[64];
        for (uint32_t l = 0; l < k; ++l) x[l] = (damp * V[C[c0 + l]]) * W[c0 + l];
        for (uint32_t s = 1; s < 64; s <<= 1)
          for (uint32_t l = 0; l + s < k; l += 2 * s) x[l] = x[l] + x[l + s];
        acc = acc + x[0];
      }
      float const d = acc - V[v];
      if (d > tol || d < -tol) {
        next[v] = acc;
        changed = 1;
        for (uint32_t e = oR[v]; e < oR[v + 1]; ++e) nact[oC[e]] = 1;
      }
    }
    memcpy(V, next, sizeof(float) * n);
    uint8_t* t = act;
    act = nact;
    nact = t;
    ++it;
    if (!changed) done = 1;
  }
  free(next);
  free(act);
  free(nact);
  return done ? it : 0;
}

/* ---- multi-core frontier Bellman-Ford: the all-core CPU baseline ---- */
static inline int atomic_relax(uint32_t* slot, uint32_t cand) {
  uint32_t old = __atomic_load_n(slot, __ATOMIC_RELAXED);
  while (cand < old) {
    if (__atomic_compare_exchange_n(slot, &old, cand, 1, __ATOMIC_RELAXED,
                                    __ATOMIC_RELAXED))
      return 1;
  }
  return 0;
}

uint32_t oracle_frontier_omp_u32(uint32_t n, const uint32_t* R,
                                 const uint32_t* C, const uint32_t* E,
                                 uint32_t* V, const uint8_t* frontier0,
                                 int algo, int threads, int* threads_used) {
  uint32_t *oR, *oC, *oE;
  transpose(n, R, C, E, &oR, &oC, &oE);
  uint32_t const rounds = oracle_frontier_omp_csr_u32(n, R, C, E, oR, oC, oE, V, frontier0,
                                                      algo, threads, threads_used);
  free(oR); free(oC); free(oE);
  return rounds;
}

uint32_t oracle_frontier_omp_csr_u32(uint32_t n, const uint32_t* R,
                                     const uint32_t* C, const uint32_t* E,
                                     const uint32_t* oR, const uint32_t* oC,
                                     const uint32_t* oE, uint32_t* V,
                                     const uint8_t* frontier0, int algo,
                                     int threads, int* threads_used) {
  g_widest = (algo == KITES_ORACLE_WIDEST);
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  else omp_set_num_threads(oracle_affinity_cores());
  int team = 1;
#pragma omp parallel
  {
#pragma omp single
    team = omp_get_num_threads();
  }
#else
  int team = 1;
#endif
  if (threads_used) *threads_used = team;

  uint8_t* inq = (uint8_t*)calloc(n ? n : 1, 1);
  uint32_t* q = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* q2 = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint64_t qn = 0;
  uint32_t rounds = 1;

  /* Pass 0: recompute the initial frontier over in-edges (the reference's
   * first kernel pass), owner writes only. */
  {
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : cnt)
    for (int64_t vv = 0; vv < (int64_t)n; ++vv) {
      uint32_t v = (uint32_t)vv;
      if (!frontier0[v]) continue;
      uint32_t pre = __atomic_load_n(&V[v], __ATOMIC_RELAXED), acc = pre;
      for (uint32_t e = R[v]; e < R[v + 1]; ++e)
        red_min(&acc, nbr_comp(algo, __atomic_load_n(&V[C[e]], __ATOMIC_RELAXED),
                               E ? E[e] : 1u));
      if (upd_lt(acc, pre)) {
        __atomic_store_n(&V[v], acc, __ATOMIC_RELAXED);
        inq[v] = 1;
        cnt++;
      }
    }
    (void)cnt;
    for (uint32_t v = 0; v < n; ++v)
      if (inq[v]) { q[qn++] = v; inq[v] = 0; }
  }

  while (qn) {
    uint64_t qn2 = 0;
#pragma omp parallel
    {
      uint32_t local[256];
      int ln = 0;
#pragma omp for schedule(dynamic, 64)
      for (int64_t k = 0; k < (int64_t)qn; ++k) {
        uint32_t u = q[k];
        uint32_t val = __atomic_load_n(&V[u], __ATOMIC_RELAXED);
        for (uint32_t e = oR[u]; e < oR[u + 1]; ++e) {
          uint32_t v = oC[e];
          if (atomic_relax(&V[v], nbr_comp(algo, val, oE ? oE[e] : 1u))) {
            if (!__atomic_exchange_n(&inq[v], 1, __ATOMIC_RELAXED)) {
              local[ln++] = v;
              if (ln == 256) {
                uint64_t at = __atomic_fetch_add(&qn2, (uint64_t)ln, __ATOMIC_RELAXED);
                memcpy(q2 + at, local, sizeof(uint32_t) * ln);
                ln = 0;
              }
            }
          }
        }
      }
      if (ln) {
        uint64_t at = __atomic_fetch_add(&qn2, (uint64_t)ln, __ATOMIC_RELAXED);
        memcpy(q2 + at, local, sizeof(uint32_t) * ln);
      }
    }
    for (uint64_t k = 0; k < qn2; ++k) inq[q2[k]] = 0;
    uint32_t* t = q; q = q2; q2 = t;
    qn = qn2;
    ++rounds;
  }
  free(inq); free(q); free(q2);
  return rounds;
}

/* ---- certificate check of many single-source results at once ----
 * D: n x K row-major, column k claimed to be the SSSP_SAT (E given) or BFS
 * (E NULL: unit weights) distances from sources[k] over the in-edge CSR.
 * With positive weights the following hold exactly for the true distances
 * and for nothing else: D[s_k][k] = 0; D[v][k] <= sat(D[u][k] + w) for
 * every edge u -> v; every finite D[v][k] with v != s_k has a tight
 * in-edge (D[u][k] + w == D[v][k]), so a chain of tight edges with
 * strictly falling values leads back to s_k.  Counts into bad[0..2]:
 * wrong source entries, violated edges, finite entries without a tight
 * in-edge (or a zero entry off the source).  Returns their sum. */
uint64_t oracle_check_cols_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                               const uint32_t* E, uint32_t K, const uint32_t* D,
                               const uint32_t* sources, int threads, uint64_t* bad) {
  uint64_t b0 = 0, b1 = 0, b2 = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  else omp_set_num_threads(oracle_affinity_cores());
#endif
  for (uint32_t k = 0; k < K; ++k)
    if (sources[k] >= n || D[(size_t)sources[k] * K + k] != 0) b0++;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : b1, b2)
  for (int64_t vv = 0; vv < (int64_t)n; ++vv) {
    uint32_t const v = (uint32_t)vv;
    const uint32_t* dv = D + (size_t)v * K;
    uint8_t tight[256];
    uint32_t const kk = K < 256 ? K : 256;
    for (uint32_t k0 = 0; k0 < K; k0 += kk) {
      uint32_t const kn = (K - k0) < kk ? (K - k0) : kk;
      memset(tight, 0, kn);
      for (uint32_t e = R[v]; e < R[v + 1]; ++e) {
        const uint32_t* du = D + (size_t)C[e] * K + k0;
        uint32_t const w = E ? E[e] : 1u;
        for (uint32_t k = 0; k < kn; ++k) {
          uint32_t const c = nbr_comp(KITES_ORACLE_SSSP_SAT, du[k], w);
          if (dv[k0 + k] > c) b1++;
          else if (dv[k0 + k] == c) tight[k] = 1;
        }
      }
      for (uint32_t k = 0; k < kn; ++k) {
        uint32_t const d = dv[k0 + k];
        int const is_src = sources[k0 + k] == v;
        if (!is_src && d != INF32 && (!tight[k] || d == 0)) b2++;
      }
    }
  }
  if (bad) { bad[0] = b0; bad[1] = b1; bad[2] = b2; }
  return b0 + b1 + b2;
}

/* ---- per-graph Dijkstra over a disjoint union of graphs ----
 * Graph i owns vertices [voff[i], voff[i+1]) and all their in-edges come
 * from the same range; dist gets graph i's textbook Dijkstra from
 * sources[i] inside its range (INF elsewhere in the range).  The graphs
 * run in parallel, one heap each. */
int oracle_dijkstra_ranges_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                               const uint32_t* E, uint32_t ngraph, const uint32_t* voff,
                               const uint32_t* sources, int threads, uint32_t* dist) {
  int err = 0;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
  else omp_set_num_threads(oracle_affinity_cores());
#endif
  if (voff[0] != 0 || voff[ngraph] > n) return 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
  for (int64_t gi = 0; gi < (int64_t)ngraph; ++gi) {
    uint32_t const lo = voff[gi], hi = voff[gi + 1];
    if (hi < lo || sources[gi] < lo || sources[gi] >= hi) { err |= 1; continue; }
    uint32_t const gn = hi - lo;
    uint64_t const e0 = R[lo], e1 = R[hi];
    /* local in-CSR with ids relative to lo */
    uint32_t* lR = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)gn + 1));
    uint32_t* lC = (uint32_t*)malloc(sizeof(uint32_t) * ((e1 - e0) ? (e1 - e0) : 1));
    for (uint32_t v = 0; v <= gn; ++v) lR[v] = (uint32_t)(R[lo + v] - e0);
    int bad = 0;
    for (uint64_t e = e0; e < e1; ++e) {
      if (C[e] < lo || C[e] >= hi) bad = 1;
      lC[e - e0] = C[e] - lo;
    }
    if (bad) { err |= 1; free(lR); free(lC); continue; }
    oracle_dijkstra_u32(gn, lR, lC, E + e0, sources[gi] - lo, dist + lo);
    free(lR); free(lC);
  }
  return err;
}

int oracle_affinity_cores(void) {
  cpu_set_t set;
  CPU_ZERO(&set);
  if (sched_getaffinity(0, sizeof(set), &set) != 0) return 1;
  return CPU_COUNT(&set);
}
