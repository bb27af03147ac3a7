// capi_internal.hpp -- handle types and error plumbing shared by the C-ABI
// translation units (capi.hip, batch_capi.hip).
#ifndef KITES_AMD_CAPI_INTERNAL_HPP_
#define KITES_AMD_CAPI_INTERNAL_HPP_

#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kites_amd.h"
#include "hgraph.hpp"
#include "kites/kites.hpp"

struct kites_hgraph {
  kites_amd::hgraph g;
};

struct kites_graph {
  std::unique_ptr<kites::gpu> dev;
  std::unique_ptr<kites::graph_csr_device<unsigned, unsigned>> g;
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<kites::detail::step_stats> stats;
  kites::device_buffer<float> model;
  // Start recorded by kites_graph_init for the next delta run.
  kites::options::seeding seed = kites::options::seeding::none;
  uint32_t seed_source = 0;
  kites::device_buffer<unsigned> seed_list;
  uint32_t seed_count = 0;
  // The recorded start holds 0 at its sources and 0xFFFFFFFF elsewhere
  // (kites_graph_init / init_sources for SSSP or BFS): a BFS run from it
  // reaches every vertex first at its final level.
  bool unit_start = false;
  // Multi-source runs: value rows (n x 64) and the last run's width/stats.
  kites::device_buffer<unsigned> rows;
  uint32_t batch_k = 0;
  std::vector<kites::detail::batch_stats> batch_stats;
  // Bit-parallel BFS: levels (n x 256) and the last run's width/stats.
  kites::device_buffer<unsigned> levels;
  uint32_t bfs_k = 0;
  std::vector<kites::detail::bitbfs_stats> bfs_stats;
};

namespace kites_capi {

constexpr unsigned kInf = 0xFFFFFFFFu;

// Message of the calling thread's last failed call (kites_last_error).
inline std::string& last_error() {
  static thread_local std::string e;
  return e;
}

inline int fail(std::exception const& e) {
  std::string& s = last_error();
  s = e.what();
  if (s.rfind("HIPAssert", 0) == 0) return KITES_ERR_RUNTIME;
  if (s.find("iteration limit") != std::string::npos) return KITES_ERR_NOCONV;
  if (dynamic_cast<std::invalid_argument const*>(&e)) return KITES_ERR_ARG;
  return KITES_ERR_OTHER;
}

inline void need(bool c, const char* what) {
  if (!c) throw std::invalid_argument(what);
}

inline kites::options to_options(kites_graph* g, const kites_run_opts* o) {
  kites::options opt;
  kites_run_opts d;
  kites_run_opts_default(&d);
  if (!o) o = &d;
  switch (o->sched) {
    case KITES_SCHED_AUTO: opt.sched = kites::schedule::automatic; break;
    case KITES_SCHED_CLASSIC: opt.sched = kites::schedule::classic; break;
    case KITES_SCHED_PLAIN: opt.sched = kites::schedule::plain; break;
    case KITES_SCHED_FORCED: opt.sched = kites::schedule::forced; break;
    case KITES_SCHED_LEARNED: opt.sched = kites::schedule::learned; break;
    default: throw std::invalid_argument("unknown schedule");
  }
  need(o->gran >= 0 && o->gran <= 3, "granularity must be 0..3");
  need(o->forced_variant >= 1 && o->forced_variant <= 8, "forced variant must be 1..8");
  opt.gran = static_cast<kites::granularity>(o->gran);
  opt.forced_variant = static_cast<unsigned>(o->forced_variant);
  opt.alpha = o->alpha;
  opt.margin = o->margin;
  opt.delta = o->delta;
  opt.split_pull = o->split_pull != 0;
  need(o->persist >= 0, "persist must be >= 0");
  opt.persist = o->persist != 0;
  if (o->persist > 1) opt.persist_edges = static_cast<unsigned>(o->persist);
  need(o->small_grid >= 0, "small_grid must be >= 0");
  opt.small_grid = static_cast<unsigned>(o->small_grid);
  need(o->tiny_grid >= 0, "tiny_grid must be >= 0");
  opt.tiny_grid = static_cast<unsigned>(o->tiny_grid);
  opt.pipeline = o->pipeline != 0;
  opt.use_graph = o->use_graph != 0;
  opt.chunk = o->chunk > 0 ? static_cast<unsigned>(o->chunk) : 16u;
  opt.grid = o->grid > 0 ? static_cast<unsigned>(o->grid) : 0u;
  opt.max_iterations = o->max_iterations ? o->max_iterations : (1u << 16);
  opt.keep_stats = o->keep_stats != 0;
  opt.mlp_weights = g->model.size() ? g->model.get_ptr() : nullptr;
  opt.verbose = false;
  if (opt.sched != kites::schedule::classic) {
    opt.seed = g->seed;
    opt.seed_source = g->seed_source;
    opt.seed_list = g->seed_list.size() ? g->seed_list.get_ptr() : nullptr;
    opt.seed_count = g->seed_count;
  }
  return opt;
}

// BFS from a unit start (sources at 0, the rest 0xFFFFFFFF): levels below
// 0xFFFFFFFF are final once written, and a vertex is offered one level in
// the iteration that reaches it, so pulls skip reached vertices and stop at
// the first updated in-neighbour (options::settled_below).
inline void settle_bfs(kites_graph* g, int32_t algo, kites::options& opt) {
  if (algo == KITES_ALGO_BFS && g->unit_start && g->seed != kites::options::seeding::none)
    opt.settled_below = kInf;
}

}  // namespace kites_capi

#define KITES_API_BEGIN try {
#define KITES_API_END                                 \
  }                                                   \
  catch (std::exception const& e) {                   \
    return kites_capi::fail(e);                       \
  }                                                   \
  catch (...) {                                       \
    kites_capi::last_error() = "unknown exception";   \
    return KITES_ERR_OTHER;                           \
  }                                                   \
  return KITES_OK;

#endif  // KITES_AMD_CAPI_INTERNAL_HPP_
