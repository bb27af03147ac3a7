/*
 * kites_oracle.h -- CPU oracle for the vertex-centric iteration engine.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (kites_amd/, include/)
 * may include, link or call this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, and only as the checker or as the
 * timed CPU baseline.
 *
 * Contents and what pins each part (see the tests named in brackets):
 *   oracle_fixpoint_u32   restates the reference CPU path, a single-thread
 *                         Gauss-Seidel sweep over all vertices until no
 *                         update (kites/graph/sanity_check.cuh:21-47), with
 *                         the functor sets below.  Pinned against textbook
 *                         Dijkstra / BFS / union-find, exhaustive
 *                         Bellman-Ford on tiny graphs and closed forms on
 *                         grids, paths, stars  [tests/test_oracle.py].
 *   oracle_dijkstra_u32   textbook binary-heap Dijkstra (independent).
 *   oracle_bfs_u32        textbook queue BFS levels (independent).
 *   oracle_cc_min_label   union-find, min vertex id per component.
 *   oracle_frontier_omp_u32 / oracle_frontier_omp_csr_u32
 *                         multi-core frontier Bellman-Ford (OpenMP) with the
 *                         same functor sets: the all-core CPU baseline (the
 *                         _csr form takes the out-CSR the caller already
 *                         holds).  Pinned by equality with
 *                         oracle_fixpoint_u32.
 *   oracle_check_cols_u32 shortest-path certificate of many single-source
 *                         results at once (full-size checks of batches).
 *                         Pinned by Dijkstra columns passing and by planted
 *                         single-entry mistakes failing.
 *   oracle_dijkstra_ranges_u32
 *                         per-graph Dijkstra over a disjoint union of
 *                         graphs.  Pinned by equality with one
 *                         oracle_dijkstra_u32 call per graph.
 *
 * Functor sets (each mirrors a __host__ __device__ lambda set of the
 * engine; init = copy, red = min, upd = new < old as in
 * examples/sssp_basic_single_device.cu:40-47):
 *   KITES_ORACLE_SSSP_SAT   partial = src==INF || src+w overflows ? INF : src+w
 *   KITES_ORACLE_BFS        partial = src==INF ? INF : src+1
 *   KITES_ORACLE_CC         partial = src
 *   KITES_ORACLE_SSSP_WRAP  partial = src + w (u32 wrap-around), the
 *                           reference example's functor verbatim
 *                           (examples/sssp_basic_single_device.cu:42-43).
 */
#ifndef KITES_ORACLE_H_
#define KITES_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  KITES_ORACLE_SSSP_SAT = 0,
  KITES_ORACLE_BFS = 1,
  KITES_ORACLE_CC = 2,
  KITES_ORACLE_SSSP_WRAP = 3,
  KITES_ORACLE_WIDEST = 4   /* partial = min(src, w), red = max, upd = new > old */
};

/* Gauss-Seidel fixpoint over the in-edge CSR (R: n+1 offsets, C: source ids,
 * E: weights).  V is updated in place.  Returns the number of sweeps, or 0
 * if max_sweeps was reached without convergence. */
uint32_t oracle_fixpoint_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                             const uint32_t* E, uint32_t* V, int algo,
                             uint32_t max_sweeps);

/* Same fixpoint, one synchronous round per step, over the in-edge CSR,
 * computing only the vertices whose in-neighbours changed in the previous
 * round (the reference GPU's activation rule, kernels_single_device.cuh:
 * 122-167).  frontier0: n bytes, 1 = initially active.  Returns rounds. */
uint32_t oracle_frontier_rounds_u32(uint32_t n, const uint32_t* R,
                                    const uint32_t* C, const uint32_t* E,
                                    uint32_t* V, const uint8_t* frontier0,
                                    int algo, uint32_t max_rounds);

/* Textbook Dijkstra from src over the graph whose in-edge CSR is (R,C,E).
 * dist[v] = INF (0xFFFFFFFF) if unreachable. Saturating path sums. */
void oracle_dijkstra_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                         const uint32_t* E, uint32_t src, uint32_t* dist);

/* Textbook max-bottleneck (widest path) Dijkstra from src; width[src] =
 * INF, unreachable = 0. */
void oracle_widest_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                       const uint32_t* E, uint32_t src, uint32_t* width);

/* Textbook BFS levels from src over the in-edge CSR (R,C). */
void oracle_bfs_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                    uint32_t src, uint32_t* level);

/* Union-find over all edges (direction ignored): label[v] = min id in v's
 * weakly connected component. */
void oracle_cc_min_label(uint32_t n, const uint32_t* R, const uint32_t* C,
                         uint32_t* label);

/* Multi-core frontier Bellman-Ford (push over the out-CSR built
 * internally), same functor sets.  frontier0 as above.  threads <= 0 means
 * all cores available to the process.  Returns rounds; *threads_used gets
 * the OpenMP team size. */
uint32_t oracle_frontier_omp_u32(uint32_t n, const uint32_t* R,
                                 const uint32_t* C, const uint32_t* E,
                                 uint32_t* V, const uint8_t* frontier0,
                                 int algo, int threads, int* threads_used);

/* The same over a prepared out-edge CSR (oR: n+1 offsets, oC: destinations,
 * oE: weights, NULL = unit weights), so a caller that already holds the
 * out-CSR times only the relaxation rounds. */
uint32_t oracle_frontier_omp_csr_u32(uint32_t n, const uint32_t* R,
                                     const uint32_t* C, const uint32_t* E,
                                     const uint32_t* oR, const uint32_t* oC,
                                     const uint32_t* oE, uint32_t* V,
                                     const uint8_t* frontier0, int algo,
                                     int threads, int* threads_used);

/* Certificate check of K single-source SSSP_SAT (E given) or BFS (E NULL)
 * results D (n x K row-major) from sources[0..K): source entries zero,
 * no edge violated, a tight in-edge for every other finite entry.  These
 * hold for the exact distances only (positive weights).  bad[0..2] get the
 * three counts (may be NULL); returns their sum. */
uint64_t oracle_check_cols_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                               const uint32_t* E, uint32_t K, const uint32_t* D,
                               const uint32_t* sources, int threads, uint64_t* bad);

/* Textbook Dijkstra of every graph of a disjoint union (graph i = vertices
 * [voff[i], voff[i+1]), all of whose in-edges stay inside the range) from
 * sources[i], graphs in parallel.  Returns nonzero if a range, source or
 * edge breaks that layout. */
int oracle_dijkstra_ranges_u32(uint32_t n, const uint32_t* R, const uint32_t* C,
                               const uint32_t* E, uint32_t ngraph, const uint32_t* voff,
                               const uint32_t* sources, int threads, uint32_t* dist);

/* Ordered (Jacobi) iterations of the float PageRank functor set with the
 * engine's defined combine order (kites_oracle.c).  V in/out, active0: n
 * bytes (nonzero = active).  Returns iterations, 0 if max_iter is reached. */
uint32_t oracle_ordered_pagerank_f32(uint32_t n, const uint32_t* R, const uint32_t* C,
                                     const float* W, const uint32_t* oR,
                                     const uint32_t* oC, float* V, const uint8_t* active0,
                                     float base, float damp, float tol, uint32_t max_iter);

/* Number of CPUs in this process's affinity mask. */
int oracle_affinity_cores(void);

#ifdef __cplusplus
}
#endif

#endif /* KITES_ORACLE_H_ */
