// batch_capi.hip -- C ABI of multi-source runs and their autotuner
// (declarations in include/kites_amd.h).  Runs kites::process_batch with
// the reference-style lambdas of the SSSP and BFS functor sets.
#include <chrono>
#include <cstdio>

#include "capi_internal.hpp"

namespace {

using kites_capi::kInf;
using kites_capi::need;

template <class Body>
void with_min_functors(int algo, Body&& body) {
  using vT = unsigned;
  using eT = unsigned;
  auto initF = [] __host__ __device__(volatile vT& locV, vT& globV) { locV = globV; };
  auto redMin = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
    if (lhsV > rhsV) lhsV = rhsV;
  };
  auto updLess = [] __host__ __device__(volatile vT& newV, vT oldV) { return newV < oldV; };
  switch (algo) {
    case KITES_ALGO_SSSP: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
        vT const t = srcV + *connE;
        partialV = (srcV == 0xFFFFFFFFu || t < srcV) ? 0xFFFFFFFFu : t;
      };
      body(initF, nbr, redMin, updLess);
      break;
    }
    case KITES_ALGO_BFS: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT*) {
        partialV = (srcV == 0xFFFFFFFFu) ? 0xFFFFFFFFu : srcV + 1u;
      };
      body(initF, nbr, redMin, updLess);
      break;
    }
    default:
      throw std::invalid_argument("multi-source runs support KITES_ALGO_SSSP and KITES_ALGO_BFS");
  }
}

kites::batch_options to_batch_options(const kites_batch_opts* o) {
  kites_batch_opts d;
  kites_batch_opts_default(&d);
  if (!o) o = &d;
  kites::batch_options b;
  b.edges_per_item = o->edges_per_item;
  b.block = o->block;
  b.unroll = o->unroll;
  b.grid = o->grid > 0 ? static_cast<unsigned>(o->grid) : 0u;
  b.use_graph = o->use_graph != 0;
  b.chunk = o->chunk > 0 ? static_cast<unsigned>(o->chunk) : 8u;
  b.max_iterations = o->max_iterations ? o->max_iterations : (1u << 16);
  b.keep_stats = o->keep_stats != 0;
  b.per_source_masks = o->per_source_masks != 0;
  need(o->top_down >= 0, "top_down must be >= 0");
  b.top_down = o->top_down == 1 ? kites::detail::kBitPushDiv : static_cast<unsigned>(o->top_down);
  return b;
}

// Digest of the first K columns of n device rows `stride` values apart
// (kites::detail::rows_digest_kernel).
unsigned long long rows_digest(kites_graph* g, const unsigned* rows, unsigned stride, unsigned K) {
  g->dev->setAsActive();
  kites::device_buffer<unsigned long long> dig(1, g->dev->getDevIdx());
  KITES_HIP_CHECK(hipMemset(dig.get_ptr(), 0, sizeof(unsigned long long)));
  hipLaunchKernelGGL(kites::detail::rows_digest_kernel, dim3(1024), dim3(256), 0, 0, rows,
                     static_cast<unsigned>(g->n), stride, K, dig.get_ptr());
  KITES_HIP_CHECK(hipPeekAtLastError());
  unsigned long long h = 0;
  KITES_HIP_CHECK(hipMemcpy(&h, dig.get_ptr(), sizeof(h), hipMemcpyDeviceToHost));
  return h;
}

float run_batch(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t K,
                kites::batch_options b, unsigned* iters, unsigned* builds = nullptr) {
  std::vector<unsigned> src(sources, sources + K);
  kites::batch_report rep;
  b.report = &rep;
  std::vector<kites::detail::batch_stats> st;
  if (b.keep_stats) b.stats = &st;
  with_min_functors(algo, [&](auto fi, auto fn, auto fr, auto fu) {
    kites::process_batch<kites::launch::sync>(*g->g, *g->dev, src, 0u, kInf, g->rows, fi, fn, fr,
                                              fu, b);
  });
  g->batch_k = K;
  if (b.keep_stats) g->batch_stats = st;
  if (iters) *iters = rep.iterations;
  if (builds) *builds = rep.graph_builds;
  return rep.ms;
}

}  // namespace

extern "C" {

void kites_batch_opts_default(kites_batch_opts* o) {
  if (!o) return;
  o->edges_per_item = 256;
  o->block = 512;
  o->unroll = 8;
  o->grid = 0;
  o->use_graph = 1;
  o->chunk = 8;
  o->max_iterations = 1u << 16;
  o->keep_stats = 0;
  o->per_source_masks = 0;
  o->top_down = 1;
}

int kites_batch_run(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t K,
                    const kites_batch_opts* o, kites_batch_info* info) {
  KITES_API_BEGIN
  need(g != nullptr && (sources != nullptr || K == 0), "null argument");
  need(K <= 64, "at most 64 sources per run");
  unsigned it = 0, builds = 0;
  float const ms = run_batch(g, algo, sources, K, to_batch_options(o), &it, &builds);
  if (info) {
    info->iterations = it;
    info->ms = ms;
    info->graph_builds = builds;
  }
  KITES_API_END
}

int kites_batch_values(kites_graph* g, uint32_t* out) {
  KITES_API_BEGIN
  need(g != nullptr && out != nullptr, "null argument");
  need(g->rows.size() >= static_cast<std::size_t>(g->n) * 64, "no multi-source run yet");
  g->dev->setAsActive();
  if (g->batch_k == 64)   // whole rows: one contiguous copy
    KITES_HIP_CHECK(hipMemcpy(out, g->rows.get_ptr(),
                              static_cast<std::size_t>(g->n) * 64 * sizeof(unsigned),
                              hipMemcpyDeviceToHost));
  else if (g->batch_k)
    KITES_HIP_CHECK(hipMemcpy2D(out, g->batch_k * sizeof(unsigned), g->rows.get_ptr(),
                                64 * sizeof(unsigned), g->batch_k * sizeof(unsigned), g->n,
                                hipMemcpyDeviceToHost));
  KITES_API_END
}

int kites_batch_stats(kites_graph* g, kites_batch_iter_stats* out, uint32_t cap, uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  static_assert(sizeof(kites_batch_iter_stats) == sizeof(kites::detail::batch_stats),
                "stats layout");
  uint32_t const k = static_cast<uint32_t>(g->batch_stats.size());
  if (count) *count = k;
  if (out) std::memcpy(out, g->batch_stats.data(), std::min(cap, k) * sizeof(kites_batch_iter_stats));
  KITES_API_END
}

int kites_batch_autotune(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t K,
                         uint32_t repeats, kites_batch_opts* best, float* best_ms,
                         kites_tune_entry* table, uint32_t cap, uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr && sources != nullptr && best != nullptr, "null argument");
  need(K >= 1 && K <= 64, "1..64 sources");
  static const unsigned kEpi[] = {128, 256, 512};
  static const unsigned kBlock[] = {256, 512};
  static const unsigned kUnroll[] = {8, 16};
  kites_batch_opts base = *best;
  float bms = 0.f;
  bool have = false;
  uint32_t k = 0;
  unsigned long long ref = 0;
  for (unsigned epi : kEpi)
    for (unsigned blk : kBlock)
      for (unsigned un : kUnroll) {
        kites_batch_opts o = base;
        o.edges_per_item = epi;
        o.block = blk;
        o.unroll = un;
        o.keep_stats = 0;
        float t = 0.f;
        for (uint32_t r = 0; r < std::max(1u, repeats); ++r) {
          float const ms = run_batch(g, algo, sources, K, to_batch_options(&o), nullptr);
          t = (r == 0 || ms < t) ? ms : t;
        }
        // Every shape must give the same rows (a device digest against the
        // first shape's).
        unsigned long long const h = rows_digest(g, g->rows.get_ptr(), 64u, K);
        if (k == 0) ref = h;
        else if (h != ref)
          throw std::runtime_error("autotune: a shape gave different results");
        if (table && k < cap) table[k] = kites_tune_entry{epi, blk, un, t};
        ++k;
        if (!have || t < bms) {
          bms = t;
          *best = o;
          have = true;
        }
      }
  if (count) *count = k;
  if (best_ms) *best_ms = bms;
  KITES_API_END
}

int kites_bfs256_run(kites_graph* g, const uint32_t* sources, uint32_t K,
                     const kites_batch_opts* o, kites_batch_info* info) {
  KITES_API_BEGIN
  need(g != nullptr && (sources != nullptr || K == 0), "null argument");
  need(K <= 256, "at most 256 sources per run");
  std::vector<unsigned> src(sources, sources + K);
  kites::batch_options b = to_batch_options(o);
  kites::batch_report rep;
  b.report = &rep;
  std::vector<kites::detail::bitbfs_stats> st;
  kites::process_bfs_bitparallel(*g->g, *g->dev, src, g->levels, b, b.keep_stats ? &st : nullptr);
  g->bfs_k = K;
  if (b.keep_stats) g->bfs_stats = st;
  if (info) {
    info->iterations = rep.iterations;
    info->ms = rep.ms;
    info->graph_builds = rep.graph_builds;
  }
  KITES_API_END
}

int kites_bfs256_autotune(kites_graph* g, const uint32_t* sources, uint32_t K, uint32_t repeats,
                          kites_batch_opts* best, float* best_ms, kites_tune_entry* table,
                          uint32_t cap, uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr && sources != nullptr && best != nullptr, "null argument");
  need(K >= 1 && K <= 256, "1..256 sources");
  static const unsigned kEpi[] = {256, 512, 1024, 2048};
  static const unsigned kBlock[] = {256, 512};
  std::vector<unsigned> src(sources, sources + K);
  kites_batch_opts const base = *best;
  unsigned long long ref = 0;
  float bms = 0.f;
  bool have = false;
  uint32_t k = 0;
  for (unsigned epi : kEpi)
    for (unsigned blk : kBlock) {
      kites_batch_opts o = base;
      o.edges_per_item = epi;
      o.block = blk;
      o.keep_stats = 0;
      kites::batch_options b = to_batch_options(&o);
      float t = 0.f;
      for (uint32_t r = 0; r < std::max(1u, repeats); ++r) {
        kites::batch_report rep;
        b.report = &rep;
        kites::process_bfs_bitparallel(*g->g, *g->dev, src, g->levels, b, nullptr);
        t = (r == 0 || rep.ms < t) ? rep.ms : t;
      }
      g->bfs_k = K;
      // Every shape must give the same levels (digest against the first).
      unsigned long long const h = rows_digest(g, g->levels.get_ptr(), 256u, K);
      if (k == 0) ref = h;
      else if (h != ref)
        throw std::runtime_error("autotune: a shape gave different levels");
      if (table && k < cap) table[k] = kites_tune_entry{epi, blk, 0u, t};
      ++k;
      if (!have || t < bms) {
        bms = t;
        *best = o;
        have = true;
      }
    }
  if (count) *count = k;
  if (best_ms) *best_ms = bms;
  KITES_API_END
}

int kites_bfs256_levels(kites_graph* g, uint32_t* out) {
  KITES_API_BEGIN
  need(g != nullptr && out != nullptr, "null argument");
  need(g->levels.size() >= static_cast<std::size_t>(g->n) * 256, "no bit-parallel BFS run yet");
  g->dev->setAsActive();
  if (g->bfs_k == 256)   // whole rows: one contiguous copy
    KITES_HIP_CHECK(hipMemcpy(out, g->levels.get_ptr(),
                              static_cast<std::size_t>(g->n) * 256 * sizeof(unsigned),
                              hipMemcpyDeviceToHost));
  else if (g->bfs_k)
    KITES_HIP_CHECK(hipMemcpy2D(out, g->bfs_k * sizeof(unsigned), g->levels.get_ptr(),
                                256 * sizeof(unsigned), g->bfs_k * sizeof(unsigned), g->n,
                                hipMemcpyDeviceToHost));
  KITES_API_END
}

int kites_bfs256_stats(kites_graph* g, kites_bfs256_iter_stats* out, uint32_t cap,
                       uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  static_assert(sizeof(kites_bfs256_iter_stats) == sizeof(kites::detail::bitbfs_stats),
                "stats layout");
  uint32_t const k = static_cast<uint32_t>(g->bfs_stats.size());
  if (count) *count = k;
  if (out) std::memcpy(out, g->bfs_stats.data(), std::min(cap, k) * sizeof(kites_bfs256_iter_stats));
  KITES_API_END
}

}  // extern "C"
