/*
 * kites_amd.h -- C ABI of the MI355X vertex-centric iteration engine.
 *
 * The engine's native interface is the header-only C++ template API under
 * kites_amd/csrc/kites/ (kites::graph_csr, kites::make_graph_csr_for_device,
 * kites::process<launch::sync>(devGraph, dev, initF, nbrCompF, redF, updF)),
 * a drop-in for the reference's kites/ headers; user lambdas are template
 * arguments, so that API cannot cross a binary boundary.  This C ABI is the
 * binary boundary for callers that cannot compile C++ (ctypes, other
 * languages): it exposes the same pipeline with fixed functor sets
 * instantiated inside the library, each written as the reference's four
 * lambdas (init, nbrComp, red, upd).  Every entry point returns 0 on success
 * or a KITES_ERR_* code; kites_last_error() describes the last failure of the
 * calling thread.  Nothing here allocates device memory outside the calls
 * that create or upload a graph (and the first run of a graph, which
 * allocates that graph's per-run workspace).
 */
#ifndef KITES_AMD_H_
#define KITES_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  KITES_OK = 0,
  KITES_ERR_ARG = 1,        /* invalid argument */
  KITES_ERR_RUNTIME = 2,    /* HIP runtime error */
  KITES_ERR_NOCONV = 3,     /* iteration limit reached */
  KITES_ERR_OTHER = 4
};

/* Functor sets.  Values u32; INF = 0xFFFFFFFF.  init copies, red is min
 * ("if (lhs > rhs) lhs = rhs"), upd is "new < old", except WIDEST. */
enum {
  KITES_ALGO_SSSP = 0,       /* partial = src==INF or overflow ? INF : src + w */
  KITES_ALGO_BFS = 1,        /* partial = src==INF ? INF : src + 1; runs from kites_graph_init /
                                kites_graph_init_sources pull bottom-up: reached vertices are
                                skipped and a pull stops at the first frontier in-neighbour */
  KITES_ALGO_CC = 2,         /* partial = src (min label) */
  KITES_ALGO_SSSP_WRAP = 3,  /* partial = src + w (u32 wrap): the reference
                                example's functor, examples/sssp_basic_single_device.cu:40-47 */
  KITES_ALGO_WIDEST = 4      /* partial = min(src, w), red = max, upd = new > old */
};

enum {
  KITES_SCHED_AUTO = 0,      /* plain when the functor set allows delta steps, else classic */
  KITES_SCHED_CLASSIC = 1,   /* the reference algorithm every iteration */
  KITES_SCHED_PLAIN = 2,     /* fixed push/pull threshold and granularity */
  KITES_SCHED_FORCED = 3,    /* every delta iteration runs forced_variant */
  KITES_SCHED_LEARNED = 4    /* the loaded MLP picks each iteration's variant */
};

/* Variant ids (kites::detail::variant_id). */
enum {
  KITES_VAR_CLASSIC = 0, KITES_VAR_PUSH_T = 1, KITES_VAR_PUSH_W = 2, KITES_VAR_PUSH_B = 3,
  KITES_VAR_PUSH_LB = 4, KITES_VAR_PULL_T = 5, KITES_VAR_PULL_W = 6, KITES_VAR_PULL_B = 7,
  KITES_VAR_PULL_LB = 8, KITES_VAR_INIT_LB = 9,
  KITES_VAR_SPLIT = 10,     /* near-far: far-pile entries below the new threshold to the queue */
  KITES_VAR_PERSIST = 11    /* a push iteration of the one-workgroup loop (statistics only) */
};

typedef struct kites_run_opts {
  int32_t sched;            /* KITES_SCHED_* */
  int32_t gran;             /* plain schedule granularity: 0 thread, 1 wave, 2 workgroup,
                               3 edge-balanced */
  int32_t forced_variant;   /* KITES_VAR_* for KITES_SCHED_FORCED */
  float alpha;              /* plain: pull when out-edges(frontier) * alpha > m */
  int32_t use_graph;        /* replay iteration chunks from a captured hipGraph */
  int32_t chunk;            /* iterations per chunk */
  int32_t grid;             /* workgroups per launch, 0 = occupancy-derived */
  uint32_t max_iterations;
  int32_t keep_stats;       /* keep per-iteration statistics (kites_graph_stats) */
  float margin;             /* learned: keep the plain choice unless another variant is
                               predicted faster by more than this (log2 units) */
  uint32_t delta;           /* near-far push bucket width, 0 = off (SSSP, BFS, SSSP_WRAP):
                               relaxations to values >= the threshold go to a far pile that
                               is not expanded until the queue runs dry; then a split
                               iteration raises the threshold to the least far value + delta.
                               Not with the recorder or kites_sssp_u32_noalloc. */
  int32_t split_pull;       /* edge-balanced pull over the source-split in-CSR (one part of
                               the sources per XCD; one more copy of C and E, built on the
                               first run that sets it); delta schedules only */
  int32_t persist;          /* 1 (default): after a small push frontier, iterations whose
                               queue fits one workgroup (up to 2,048 out-edges) run many per
                               launch in one workgroup (near-far runs too: far-pile appends
                               and splits of up to twice the limit); plain and learned
                               schedules of SSSP, BFS, CC and SSSP_WRAP.  0: off.  Above 1:
                               that out-edge limit (at most 16,384). */
  int32_t small_grid;       /* workgroups of the launches that follow a small push frontier
                               (entries + out-edges <= 2^18; 0 = one per compute unit) */
  int32_t tiny_grid;        /* ... and a tiny one (<= 2^16; 0 = a quarter of small_grid's
                               default) */
  int32_t pipeline;         /* 1 (default): after the first chunk of iterations, queue the next
                               chunk before reading the last one's state; 0: one at a time */
} kites_run_opts;

typedef struct kites_run_info {
  uint32_t iterations;
  float ms;                 /* device time of the run (hipEvents on the run's stream) */
  int32_t delta;            /* delta variants were allowed */
  int32_t min_fast;         /* native-min path was used */
  uint32_t graph_builds;    /* iteration-chunk hipGraphs captured by this graph's runs so far
                               (a run that repeats an earlier run's launches replays them) */
} kites_run_info;

/* Per-iteration statistics (layout of kites::detail::step_stats). */
typedef struct kites_iter_stats {
  uint32_t count;           /* vertices updated */
  uint32_t variant;         /* KITES_VAR_* that ran */
  uint32_t max_outdeg;
  uint32_t in_count;        /* frontier queue length consumed */
  uint64_t sum_outdeg;      /* out-edges of the updated vertices */
  uint64_t work_edges;      /* edges examined */
  uint64_t gathers;         /* source values gathered */
  uint64_t active;          /* vertices / queue entries processed */
  uint64_t qe;              /* packed append counter (internal) */
} kites_iter_stats;

typedef struct kites_hgraph kites_hgraph;  /* host CSR graph */
typedef struct kites_graph kites_graph;    /* device-resident graph */

const char* kites_last_error(void);
int kites_device_count(int32_t* count);
void kites_run_opts_default(kites_run_opts* o);

/* ---- host graphs ------------------------------------------------------- */
/* R-MAT (a, b, c; d = 1-a-b-c), self-loops and duplicates removed; n_limit
 * > 0 rejects endpoints >= n_limit; exact_edges > 0 draws until exactly that
 * many distinct edges exist, else draws edge_factor * 2^scale edges.  The
 * vertex count is padded to a multiple of 128.  threads <= 0: all cores. */
int kites_hgraph_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b,
                      double c, uint32_t wmin, uint32_t wmax, uint32_t n_limit,
                      uint64_t exact_edges, int32_t threads, kites_hgraph** out);
/* `count` independent R-MAT graphs (a=.57, b=.19, c=.19) with scale and
 * edge factor drawn uniformly from [scale_lo, scale_hi] x [ef_lo, ef_hi]
 * (all seeded), as one disjoint union: graph i owns vertices
 * [voff[i], voff[i+1]) (voff: count+1 entries) and a contiguous edge range;
 * sources[i] (count entries) = its max-out-degree vertex (lowest id on
 * ties) as a union id.  No counterpart in the reference. */
int kites_hgraph_rmat_batch(uint32_t count, uint64_t seed, uint32_t scale_lo, uint32_t scale_hi,
                            uint32_t ef_lo, uint32_t ef_hi, uint32_t wmin, uint32_t wmax,
                            int32_t threads, kites_hgraph** out, uint32_t* voff,
                            uint32_t* sources);
/* G(n, draws) with self-loops and duplicates removed. */
int kites_hgraph_er(uint32_t n, uint64_t draws, uint64_t seed, uint32_t wmin, uint32_t wmax,
                    int32_t threads, kites_hgraph** out);
/* rows x cols 4-neighbour grid, directed edges both ways, own weight each. */
int kites_hgraph_grid(uint32_t rows, uint32_t cols, uint64_t seed, uint32_t wmin, uint32_t wmax,
                      int32_t threads, kites_hgraph** out);
/* Union of the edges and their reverses, duplicates removed. */
int kites_hgraph_symmetrize(const kites_hgraph* in, int32_t threads, kites_hgraph** out);
/* From an in-edge CSR (R: n+1, C/E: R[n]); n is used as given. */
int kites_hgraph_from_csr(uint32_t n, const uint32_t* R, const uint32_t* C, const uint32_t* E,
                          int32_t threads, kites_hgraph** out);
/* Host build from m edges src[k] -> dst[k] with weight w[k] (ids < n), the
 * same arrays as kites_graph_from_edges builds on the GPU: in-lists sorted
 * by source with duplicates in input order (graph_raw.h:102,131-133), n
 * padded to a multiple of 128 (graph_raw.h:173-182); dedup != 0 drops
 * self-loops and later copies of a pair.  threads <= 0: all cores. */
int kites_hgraph_from_edges(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* w, int32_t dedup, int32_t threads,
                            kites_hgraph** out);
/* Prepared-graph files.  Text: the reference's dump, written as
 * kites::io::save_to_file writes it (kites/io/graph_output.cuh:29-35) and
 * read as graph_csr(file) reads it (kites/graph/graph_csr.h:112-166; one
 * vertex per group).  Binary (no counterpart in the reference): the same
 * arrays plus the out-edge CSR, raw little-endian, loaded without parsing
 * or transpose.  threads <= 0: all cores (the transpose of a text load). */
int kites_hgraph_save_text(const kites_hgraph* g, const char* path);
int kites_hgraph_load_text(const char* path, int32_t threads, kites_hgraph** out);
int kites_hgraph_save_binary(const kites_hgraph* g, const char* path);
int kites_hgraph_load_binary(const char* path, kites_hgraph** out);
/* Edge-list text parsed as the reference's graph_raw does
 * (kites/graph/graph_raw.h:58-190) with the reference example's
 * pre-processing callback (examples/sssp_basic_single_device.cu:10-26)
 * for the given source. */
int kites_hgraph_load_edgelist(const char* path, uint32_t vpg, int32_t s_d, int32_t directed,
                               int32_t complete, uint32_t source, kites_hgraph** out);
int kites_hgraph_info(const kites_hgraph* g, uint32_t* n, uint32_t* n_real, uint64_t* m);
int kites_hgraph_arrays(const kites_hgraph* g, const uint32_t** R, const uint32_t** C,
                        const uint32_t** E, const uint32_t** oR, const uint32_t** oC,
                        const uint32_t** oE);
/* Edge-list loads only: initial values (n), frontier bits (one byte per
 * bit, nbits), group out-lists dir_R (ndir_r) / dir_C (ndir_c). */
int kites_hgraph_load_arrays(const kites_hgraph* g, const uint32_t** V0, const uint8_t** frontier,
                             uint64_t* nbits, const uint32_t** dir_R, uint64_t* ndir_r,
                             const uint32_t** dir_C, uint64_t* ndir_c);
void kites_hgraph_destroy(kites_hgraph* g);

/* ---- device graphs: replaces make_graph_csr_for_device
 *      (kites/graph/graph_csr_device.cuh:146-152) ------------------------ */
int kites_graph_upload(int32_t device, const kites_hgraph* h, kites_graph** out);
void kites_graph_destroy(kites_graph* g);
int kites_graph_info(const kites_graph* g, uint32_t* n, uint64_t* m);
/* Builds the device graph on the GPU from m edges src[k] -> dst[k] with
 * weight w[k] (host arrays, ids < n): replaces the host build of
 * graph_raw.h:58-190 and graph_csr.h:45-110 (vertex-per-group 1, directed)
 * with two stable radix sorts on the device.  Same arrays as the host path:
 * in-lists sorted by source with duplicates kept in input order, the
 * weighted out-lists sorted by destination, n padded to a multiple of 128.
 * dedup != 0 drops self-loops and later copies of a (src, dst) pair (the
 * generators' rule).  build_ms (may be NULL): device time from the edges in
 * HBM to the finished graph. */
int kites_graph_from_edges(int32_t device, uint32_t n, uint64_t m, const uint32_t* src,
                           const uint32_t* dst, const uint32_t* w, int32_t dedup,
                           kites_graph** out, float* build_ms);
/* Copies the device graph's arrays to the host (any pointer may be NULL):
 * R, oR n+1 entries; C, E, oC, oE m; nzi nnz; nzR nnz+1 (the vertices with
 * in-edges and their in-list offsets).  No counterpart in the reference,
 * whose device graph is read only through V. */
int kites_graph_arrays(const kites_graph* g, uint32_t* R, uint32_t* C, uint32_t* E, uint32_t* oR,
                       uint32_t* oC, uint32_t* oE, uint32_t* nzi, uint32_t* nzR, uint32_t* nnz);
/* Initial values (n u32) and initial frontier (n bytes, nonzero = active;
 * NULL = empty), as the reference's pre-processing callback produces them
 * (graph_raw.h:103-110). */
int kites_graph_set_state(kites_graph* g, const uint32_t* V, const uint8_t* frontier);
/* Standard initial state on the device: SSSP/BFS from `source` (V = INF
 * but V[source] = 0, frontier = out-neighbours of source); SSSP_WRAP the
 * same with the reference example's INF = 1073741823; CC
 * (V[v] = v, every vertex active); WIDEST (V = 0 but V[source] = INF,
 * frontier = out-neighbours of source). */
int kites_graph_init(kites_graph* g, int32_t algo, uint32_t source);
/* The same for a set of sources (SSSP, BFS, SSSP_WRAP): every source starts
 * at 0 and every other vertex at INF, so each vertex ends at its value from
 * the nearest source; on a disjoint union of graphs with one source each
 * (kites_hgraph_rmat_batch) every graph gets its own single-source result
 * in one run.  Duplicates are ignored. */
int kites_graph_init_sources(kites_graph* g, int32_t algo, const uint32_t* sources,
                             uint32_t count);

/* ---- the fixpoint: replaces kites::process<launch::sync>(devGrph, dev,
 *      initF, nbrCompF, redF, updF) (kites/device/compute.cuh:29-137) ---- */
int kites_graph_run(kites_graph* g, int32_t algo, const kites_run_opts* o, kites_run_info* info);
/* Copies the vertex values to the host (n u32): replaces
 * host_pinned_buffer(devGrph.V) (kites/memory/host_pinned_buffer.cuh:52-57). */
int kites_graph_values(kites_graph* g, uint32_t* out);
/* Statistics of the last run made with keep_stats. */
int kites_graph_stats(kites_graph* g, kites_iter_stats* out, uint32_t cap, uint32_t* count);

/* Float PageRank, V(v) = (1 - damping) + damping * sum over in-edges (u, v)
 * of V(u) / outdeg(u), every vertex starting at 1 and active, a vertex
 * updated while its value moves by more than tol; run in the ordered mode
 * (Jacobi iterations, a fixed combine order of each vertex's in-edges), so
 * the float result is defined bit for bit.  No counterpart in the reference
 * (its in-place sum would depend on timing, kernels_single_device.cuh:92-132).
 * out: n floats (may be NULL); max_iterations 0 = 10000. */
int kites_pagerank_f32(kites_graph* g, float damping, float tol, uint32_t max_iterations,
                       float* out, kites_run_info* info);

/* Convenience: init + run + (optional) copy-out. */
int kites_sssp_u32(kites_graph* g, uint32_t source, const kites_run_opts* o, kites_run_info* info,
                   uint32_t* dist_out);
int kites_bfs_u32(kites_graph* g, uint32_t source, const kites_run_opts* o, kites_run_info* info,
                  uint32_t* level_out);
int kites_cc_u32(kites_graph* g, const kites_run_opts* o, kites_run_info* info,
                 uint32_t* label_out);

/* kites_sssp_u32 without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The engine's run state (frontier bitmaps, queues, statistics, control
 * block) lives in a device array of about 72 MB that the library holds on
 * every GPU it runs on until the process ends; the caller passes the graph
 * (uploaded beforehand) and the output.  Limits: at most 4,194,304 vertices
 * and max_iterations <= 65,536; split_pull and keep_stats are refused
 * (KITES_ERR_ARG), and so is delta (the near-far far pile lives in the
 * graph's workspace).  Iteration chunks are launched directly (use_graph is
 * ignored) and info->ms is host-clock time.  Calls on one GPU share the
 * array and are serialized by the library (overlapping calls would race on
 * it); the array keeps the last call's contents, which every run resets
 * before reading.  Results equal kites_sssp_u32's. */
int kites_sssp_u32_noalloc(kites_graph* g, uint32_t source, const kites_run_opts* o,
                           kites_run_info* info, uint32_t* dist_out);

/* ---- multi-source runs ---------------------------------------------------- */
/* No counterpart in the reference, whose process() runs one frontier
 * (kites/device/compute.cuh:29-137): up to 64 sources of SSSP or BFS run
 * together, column k equal to a single-source run from sources[k]. */
typedef struct kites_batch_opts {
  uint32_t edges_per_item;  /* work item = this many in-edges: 128, 256, 512 (64 x items per
                               lane; bit-parallel BFS: any multiple of 64 up to 4096) */
  uint32_t block;           /* threads per workgroup: 256, 512 (bit-parallel BFS: also 128) */
  uint32_t unroll;          /* neighbour rows in flight per wave: 8, 16 */
  int32_t grid;             /* workgroups per launch, 0 = occupancy-derived */
  int32_t use_graph;        /* replay iteration chunks from a captured hipGraph */
  int32_t chunk;            /* iterations per chunk */
  uint32_t max_iterations;
  int32_t keep_stats;       /* keep per-iteration statistics (kites_batch_stats) */
  int32_t per_source_masks; /* K-wide runs: a 64-bit mask per vertex of the sources whose
                               value moved, besides the one-bit frontier (same results; the
                               rows are read whole either way); default shape only */
  int32_t top_down;         /* bit-parallel BFS: 1 (default): a level whose frontier has fewer
                               than m / 4 out-edges (none above 4,096) runs top-down over them;
                               0: every level pulls over all in-edges; above 1: that divisor.
                               K-wide runs always pull */
} kites_batch_opts;

typedef struct kites_batch_info {
  uint32_t iterations;
  float ms;                 /* device time: value init + iterations */
  uint32_t graph_builds;    /* iteration-chunk hipGraphs captured by this graph's batch runs */
} kites_batch_info;

typedef struct kites_batch_iter_stats {
  uint32_t updated;         /* vertices whose row moved */
  uint32_t flushes;         /* rows compared against their update */
  uint64_t active_edges;    /* in-edges whose source mask was non-zero */
  uint64_t gathered;        /* neighbour values loaded (mask bits over those edges) */
  uint64_t moved;           /* (vertex, source) values that moved */
} kites_batch_iter_stats;

typedef struct kites_tune_entry {
  uint32_t edges_per_item, block, unroll;
  float ms;
} kites_tune_entry;

void kites_batch_opts_default(kites_batch_opts* o);
/* algo: KITES_ALGO_SSSP or KITES_ALGO_BFS; sources: K (<= 64) vertex ids. */
int kites_batch_run(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t K,
                    const kites_batch_opts* o, kites_batch_info* info);
/* Result of the last multi-source run: n x K values, row-major. */
int kites_batch_values(kites_graph* g, uint32_t* out);
int kites_batch_stats(kites_graph* g, kites_batch_iter_stats* out, uint32_t cap,
                      uint32_t* count);
/* Times every (edges_per_item, block, unroll) shape on this graph and
 * sources (best of `repeats`), checks that all give identical rows (a device
 * digest of the n x K values) and writes the fastest into *best (other
 * fields taken from *best on entry). */
int kites_batch_autotune(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t K,
                         uint32_t repeats, kites_batch_opts* best, float* best_ms,
                         kites_tune_entry* table, uint32_t cap, uint32_t* count);

/* Bit-parallel BFS from up to 256 sources in one run (no counterpart in
 * the reference): levels as u32, 0xFFFFFFFF = unreached; column k equals a
 * single-source BFS from sources[k].  Uses edges_per_item / block /
 * use_graph / chunk / max_iterations / keep_stats of the options. */
typedef struct kites_bfs256_iter_stats {
  uint32_t updated;         /* vertices that gained sources */
  uint32_t push;            /* 1: the level ran top-down */
  uint64_t active_edges;    /* edges from a frontier vertex into a vertex not yet reached by
                               every source */
  uint64_t new_pairs;       /* (vertex, source) pairs reached */
  uint64_t out_edges;       /* out-edges of the vertices updated (the next frontier) */
  uint32_t max_out;         /* largest out-degree among them */
  uint32_t pad;
} kites_bfs256_iter_stats;

int kites_bfs256_run(kites_graph* g, const uint32_t* sources, uint32_t K,
                     const kites_batch_opts* o, kites_batch_info* info);
/* Levels of the last run: n x K, row-major. */
int kites_bfs256_levels(kites_graph* g, uint32_t* out);
int kites_bfs256_stats(kites_graph* g, kites_bfs256_iter_stats* out, uint32_t cap,
                       uint32_t* count);
/* Times every (edges_per_item, block) shape of the bit-parallel BFS on this
 * graph and sources (best of `repeats`), checks that all give identical
 * levels (a device digest of the n x K levels) and writes the fastest into
 * *best (other fields taken from *best on entry); table entries have unroll
 * 0.  The levels of the last shape stay for kites_bfs256_levels.  No
 * counterpart in the reference. */
int kites_bfs256_autotune(kites_graph* g, const uint32_t* sources, uint32_t K, uint32_t repeats,
                          kites_batch_opts* best, float* best_ms, kites_tune_entry* table,
                          uint32_t cap, uint32_t* count);

/* Single-source (or seed-list) runs: times every (workgroups per CU, chunk)
 * shape of `algo` from the standard start (one source: kites_graph_init;
 * several: kites_graph_init_sources; CC: no source), best of 2, checks that
 * all shapes give identical values and writes the fastest grid / chunk into
 * *best (other fields taken from *best on entry).  Table entries:
 * (workgroups per CU, 256, chunk, ms).  No counterpart in the reference. */
int kites_graph_autotune(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t nsrc,
                         kites_run_opts* best, float* best_ms, kites_tune_entry* table,
                         uint32_t cap, uint32_t* count);

/* ---- learned schedule ---------------------------------------------------- */
/* No counterpart in the reference (its schedule is one fixed kernel per
 * iteration, kites/device/compute.cuh:69-117). */

/* Scheduler network: MLP_IN = 10 features -> 32 ReLU -> 8 outputs (predicted
 * relative log2 time of variants KITES_VAR_PUSH_T .. KITES_VAR_PULL_LB). */
uint32_t kites_mlp_nparam(void);
/* Weights used by KITES_SCHED_LEARNED runs of this graph. */
int kites_graph_set_model(kites_graph* g, const float* weights, uint32_t nparam);

/* One recorded state: features of the iteration and the device time (ms)
 * of every delta variant run from that state. */
typedef struct kites_variant_sample {
  float feat[10];
  float ms[8];
  uint32_t iter;
  uint32_t plain_variant;   /* what the run's own schedule chose */
} kites_variant_sample;

/* Runs `algo` from the state set by kites_graph_init / set_state with the
 * delta schedule in `o`, forking every iteration to time all 8 delta
 * variants; writes up to cap samples.  rollout_variant = 0 times the forked
 * iteration alone; a KITES_VAR_* id times it plus the rest of the run with
 * that variant forced.  The final values equal a normal run's. */
int kites_graph_record(kites_graph* g, int32_t algo, const kites_run_opts* o,
                       uint32_t rollout_variant, kites_variant_sample* out, uint32_t cap,
                       uint32_t* count);

/* Batch inference, loss gradient and Adam training of the network on GPU
 * `device` with the project's HIP kernels.  X: B x 9, T / mask / Y: B x 8
 * (row-major, host memory); mask may be NULL (all ones).  Loss = mean over
 * samples and outputs of mask * (y - t)^2. */
int kites_mlp_forward(int32_t device, const float* w, const float* X, uint32_t B, float* Y);
int kites_mlp_loss_grad(int32_t device, const float* w, const float* X, const float* T,
                        const float* mask, uint32_t B, float* grad, float* loss);
int kites_mlp_train(int32_t device, float* w, const float* X, const float* T, const float* mask,
                    uint32_t B, uint32_t epochs, float lr, float* loss_first, float* loss_last);
const char* kites_mlp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* KITES_AMD_H_ */
