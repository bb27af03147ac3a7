// capi.hip -- C ABI of the engine (declarations and contract in
// include/kites_amd.h).  Each functor set is written as the reference's
// four __host__ __device__ lambdas and run through the same
// kites::process() template a C++ user calls, so the C ABI measures the
// drop-in path itself.
#include <algorithm>
#include <fstream>
#include <string>

#include "capi_internal.hpp"
#include "functor_sets.hpp"

namespace {

using kites_capi::kInf;
using kites_capi::need;
using kites_capi::settle_bfs;
using kites_capi::to_options;
using kites_capi::with_functors;

__global__ void fill_values_kernel(unsigned* V, unsigned n, unsigned fill, unsigned src,
                                   unsigned src_val, int identity, unsigned long long* bm,
                                   unsigned nwords, int all_active) {
  unsigned const tid = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned const nt = gridDim.x * blockDim.x;
  for (unsigned i = tid; i < n; i += nt) V[i] = identity ? i : (i == src ? src_val : fill);
  for (unsigned w = tid; w < nwords; w += nt) {
    unsigned long long word = 0;
    if (all_active) {
      unsigned const lo = w * 64u;
      if (lo < n) word = (n - lo >= 64) ? ~0ull : ((1ull << (n - lo)) - 1ull);
    }
    bm[w] = word;
  }
}

__global__ void mark_out_kernel(const unsigned* oR, const unsigned* oC, unsigned src,
                                unsigned long long* bm) {
  unsigned const b = oR[src], e = oR[src + 1];
  for (unsigned i = b + blockIdx.x * blockDim.x + threadIdx.x; i < e; i += gridDim.x * blockDim.x) {
    unsigned const v = oC[i];
    atomicOr(bm + (v >> 6), 1ull << (v & 63));
  }
}

__global__ void seed_values_kernel(unsigned* V, const unsigned* list, unsigned count,
                                   unsigned value, unsigned long long* bm, const unsigned* oR,
                                   const unsigned* oC) {
  // one workgroup per source: its value, then its out-neighbours' frontier bits
  unsigned const i = blockIdx.x;
  if (i >= count) return;
  unsigned const s = list[i];
  if (threadIdx.x == 0) V[s] = value;
  for (unsigned e = oR[s] + threadIdx.x; e < oR[s + 1]; e += blockDim.x) {
    unsigned const v = oC[e];
    atomicOr(bm + (v >> 6), 1ull << (v & 63));
  }
}


}  // namespace

extern "C" {

const char* kites_last_error(void) { return kites_capi::last_error().c_str(); }

int kites_device_count(int32_t* count) {
  KITES_API_BEGIN
  need(count != nullptr, "count is null");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  KITES_API_END
}

void kites_run_opts_default(kites_run_opts* o) {
  if (!o) return;
  o->sched = KITES_SCHED_AUTO;
  o->gran = 3;
  o->forced_variant = KITES_VAR_PULL_LB;
  o->alpha = 14.f;
  o->margin = 0.f;
  o->delta = 0;
  o->use_graph = 1;
  o->chunk = 16;
  o->grid = 0;
  o->max_iterations = 1u << 16;
  o->keep_stats = 0;
  o->split_pull = 0;
  o->persist = 1;
  o->small_grid = 0;
  o->tiny_grid = 0;
  o->pipeline = 1;
}

// ---- host graphs -----------------------------------------------------------

int kites_hgraph_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b,
                      double c, uint32_t wmin, uint32_t wmax, uint32_t n_limit,
                      uint64_t exact_edges, int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(out != nullptr, "out is null");
  need(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, "R-MAT probabilities out of range");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::gen_rmat(scale, edge_factor, seed, a, b, c, wmin, wmax, n_limit, exact_edges,
                      threads, h->g);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_rmat_batch(uint32_t count, uint64_t seed, uint32_t scale_lo, uint32_t scale_hi,
                            uint32_t ef_lo, uint32_t ef_hi, uint32_t wmin, uint32_t wmax,
                            int32_t threads, kites_hgraph** out, uint32_t* voff,
                            uint32_t* sources) {
  KITES_API_BEGIN
  need(out != nullptr && voff != nullptr && sources != nullptr, "null argument");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  std::vector<uint32_t> vo, src;
  kites_amd::gen_rmat_batch(count, seed, scale_lo, scale_hi, ef_lo, ef_hi, wmin, wmax, threads,
                            h->g, vo, src);
  std::memcpy(voff, vo.data(), vo.size() * 4);
  std::memcpy(sources, src.data(), src.size() * 4);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_er(uint32_t n, uint64_t draws, uint64_t seed, uint32_t wmin, uint32_t wmax,
                    int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(out != nullptr, "out is null");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::gen_er(n, draws, seed, wmin, wmax, threads, h->g);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_grid(uint32_t rows, uint32_t cols, uint64_t seed, uint32_t wmin, uint32_t wmax,
                      int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(out != nullptr, "out is null");
  need(wmin <= wmax, "weight range empty");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::gen_grid(rows, cols, seed, wmin, wmax, threads, h->g);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_symmetrize(const kites_hgraph* in, int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(in != nullptr && out != nullptr, "null graph");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::symmetrize(in->g, threads, h->g);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_from_csr(uint32_t n, const uint32_t* R, const uint32_t* C, const uint32_t* E,
                          int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(out != nullptr && R != nullptr, "null argument");
  uint64_t const m = R[n];
  need(m == 0 || (C != nullptr && E != nullptr), "null edge arrays");
  for (uint32_t v = 0; v < n; ++v) need(R[v] <= R[v + 1], "R must be non-decreasing");
  for (uint64_t e = 0; e < m; ++e) need(C[e] < n, "edge source out of range");
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  auto& g = h->g;
  g.n = n;
  g.n_real = n;
  g.m = m;
  g.R.assign(R, R + n + 1);
  g.C.assign(C, C + m);
  g.E.assign(E, E + m);
  kites_amd::transpose(g, threads);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_from_edges(uint32_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                            const uint32_t* w, int32_t dedup, int32_t threads,
                            kites_hgraph** out) {
  KITES_API_BEGIN
  need(out != nullptr, "null argument");
  need(m == 0 || (src != nullptr && dst != nullptr && w != nullptr), "null edge array");
  need(n > 0, "a graph needs at least one vertex");
  std::vector<uint32_t> s(src, src + m), d(dst, dst + m), ww(w, w + m);
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  for (uint64_t k = 0; k < m; ++k) need(s[k] < n && d[k] < n, "edge endpoint out of range");
  kites_amd::build_csr(kites_amd::pad_vertices(n), s, d, ww, dedup != 0, threads, h->g);
  h->g.n_real = n;
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_save_text(const kites_hgraph* h, const char* path) {
  KITES_API_BEGIN
  need(h != nullptr && path != nullptr, "null argument");
  std::ofstream out(path, std::ios::binary);
  need(static_cast<bool>(out), (std::string("cannot open ") + path).c_str());
  kites_amd::save_text(h->g, out);
  KITES_API_END
}

int kites_hgraph_load_text(const char* path, int32_t threads, kites_hgraph** out) {
  KITES_API_BEGIN
  need(path != nullptr && out != nullptr, "null argument");
  std::ifstream in(path, std::ios::binary);
  need(static_cast<bool>(in), (std::string("Failed to open specified file: ") + path).c_str());
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::load_text(in, h->g, threads);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_save_binary(const kites_hgraph* h, const char* path) {
  KITES_API_BEGIN
  need(h != nullptr && path != nullptr, "null argument");
  std::ofstream out(path, std::ios::binary);
  need(static_cast<bool>(out), (std::string("cannot open ") + path).c_str());
  kites_amd::save_binary(h->g, out);
  KITES_API_END
}

int kites_hgraph_load_binary(const char* path, kites_hgraph** out) {
  KITES_API_BEGIN
  need(path != nullptr && out != nullptr, "null argument");
  std::ifstream in(path, std::ios::binary);
  need(static_cast<bool>(in), (std::string("Failed to open specified file: ") + path).c_str());
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  kites_amd::load_binary(in, h->g);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_load_edgelist(const char* path, uint32_t vpg, int32_t s_d, int32_t directed,
                               int32_t complete, uint32_t source, kites_hgraph** out) {
  KITES_API_BEGIN
  need(path != nullptr && out != nullptr, "null argument");
  need(vpg >= 1, "vertex-per-group must be >= 1");
  // The reference example's pre-processing function
  // (examples/sssp_basic_single_device.cu:10-26), parameterised by source.
  auto cb = [source](unsigned argc, char** argv, unsigned s, unsigned d, unsigned& sv,
                     unsigned& dv, unsigned& w) {
    sv = (s != source) ? 1073741823u : 0u;
    dv = (d != source) ? 1073741823u : 0u;
    w = (argc > 0) ? static_cast<unsigned>(std::atoi(argv[0])) : ((s + 1) * (d + 2) % 127);
    return s == source;
  };
  kites::graph_raw<unsigned, unsigned> raw(
      std::string(path), cb, vpg,
      s_d ? kites::input_graph_form::edge_list_s_d : kites::input_graph_form::edge_list_d_s,
      directed ? kites::graph_property::directed : kites::graph_property::undirected,
      complete ? kites::edge_list_expression::complete : kites::edge_list_expression::incomplete);
  std::unique_ptr<kites_hgraph> h(new kites_hgraph());
  auto& g = h->g;
  g.n = static_cast<uint32_t>(raw.get_num_vertices());
  g.n_real = g.n;
  g.m = raw.get_num_edges();
  g.vpg = vpg;
  g.directed = directed != 0;
  g.V0.assign(raw.rawVVec.begin(), raw.rawVVec.end());
  g.R.assign(static_cast<std::size_t>(g.n) + 1, 0);
  g.C.reserve(g.m);
  g.E.reserve(g.m);
  for (uint32_t v = 0; v < g.n; ++v) {
    for (auto& nb : raw.rawNbrsVec[v]) {
      g.C.push_back(nb.nbrIdx);
      g.E.push_back(nb.EdgeVal);
    }
    g.R[v + 1] = static_cast<uint32_t>(g.C.size());
  }
  g.frontier0.resize(raw.vBitmap.size());
  for (std::size_t i = 0; i < raw.vBitmap.size(); ++i) g.frontier0[i] = raw.vBitmap[i] ? 1 : 0;
  if (directed) {
    g.dir_R.assign(raw.rawIVecOutgoing.size() + 1, 0);
    for (std::size_t gi = 0; gi < raw.rawIVecOutgoing.size(); ++gi) {
      for (auto x : raw.rawIVecOutgoing[gi]) g.dir_C.push_back(x);
      g.dir_R[gi + 1] = static_cast<uint32_t>(g.dir_C.size());
    }
  } else {
    g.dir_R.assign(1, 0);
  }
  kites_amd::transpose(g, 0);
  *out = h.release();
  KITES_API_END
}

int kites_hgraph_info(const kites_hgraph* g, uint32_t* n, uint32_t* n_real, uint64_t* m) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  if (n) *n = g->g.n;
  if (n_real) *n_real = g->g.n_real;
  if (m) *m = g->g.m;
  KITES_API_END
}

int kites_hgraph_arrays(const kites_hgraph* g, const uint32_t** R, const uint32_t** C,
                        const uint32_t** E, const uint32_t** oR, const uint32_t** oC,
                        const uint32_t** oE) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  if (R) *R = g->g.R.data();
  if (C) *C = g->g.C.data();
  if (E) *E = g->g.E.data();
  if (oR) *oR = g->g.oR.data();
  if (oC) *oC = g->g.oC.data();
  if (oE) *oE = g->g.oE.data();
  KITES_API_END
}

int kites_hgraph_load_arrays(const kites_hgraph* g, const uint32_t** V0, const uint8_t** frontier,
                             uint64_t* nbits, const uint32_t** dir_R, uint64_t* ndir_r,
                             const uint32_t** dir_C, uint64_t* ndir_c) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  if (V0) *V0 = g->g.V0.data();
  if (frontier) *frontier = g->g.frontier0.data();
  if (nbits) *nbits = g->g.frontier0.size();
  if (dir_R) *dir_R = g->g.dir_R.data();
  if (ndir_r) *ndir_r = g->g.dir_R.size();
  if (dir_C) *dir_C = g->g.dir_C.data();
  if (ndir_c) *ndir_c = g->g.dir_C.size();
  KITES_API_END
}

void kites_hgraph_destroy(kites_hgraph* g) { delete g; }

// ---- device graphs ------------------------------------------------------------

int kites_graph_upload(int32_t device, const kites_hgraph* h, kites_graph** out) {
  KITES_API_BEGIN
  need(h != nullptr && out != nullptr, "null argument");
  auto const& hg = h->g;
  need(hg.oR.size() == static_cast<std::size_t>(hg.n) + 1, "host graph lacks its out-edge CSR");
  std::unique_ptr<kites_graph> g(new kites_graph());
  g->dev.reset(new kites::gpu(device));
  g->g.reset(new kites::graph_csr_device<unsigned, unsigned>(
      hg.n, hg.m, hg.R.data(), hg.C.data(), hg.E.data(), hg.oR.data(), hg.oC.data(),
      hg.oE.data(),
      hg.directed ? kites::graph_property::directed : kites::graph_property::undirected,
      *g->dev));
  if (hg.vpg > 1) {
    auto& dg = *g->g;
    dg.vertexPerGroup = hg.vpg;
    dg.dir_R.alloc(hg.dir_R.size(), device);
    dg.dir_C.alloc(hg.dir_C.size() ? hg.dir_C.size() : 1, device);
    KITES_HIP_CHECK(hipMemcpy(dg.dir_R.get_ptr(), hg.dir_R.data(), hg.dir_R.size() * 4,
                              hipMemcpyHostToDevice));
    if (!hg.dir_C.empty())
      KITES_HIP_CHECK(hipMemcpy(dg.dir_C.get_ptr(), hg.dir_C.data(), hg.dir_C.size() * 4,
                                hipMemcpyHostToDevice));
  }
  g->n = hg.n;
  g->m = hg.m;
  *out = g.release();
  KITES_API_END
}

void kites_graph_destroy(kites_graph* g) {
  try {
    delete g;
  } catch (...) {
  }
}

int kites_graph_info(const kites_graph* g, uint32_t* n, uint64_t* m) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  KITES_API_END
}

int kites_graph_set_state(kites_graph* g, const uint32_t* V, const uint8_t* frontier) {
  KITES_API_BEGIN
  need(g != nullptr && V != nullptr, "null argument");
  auto& dg = *g->g;
  g->dev->setAsActive();
  KITES_HIP_CHECK(hipMemcpy(dg.V.get_ptr(), V, g->n * sizeof(unsigned), hipMemcpyHostToDevice));
  std::vector<unsigned long long> words(dg.vBitmapEven.size(), 0ull);
  if (frontier) {
    // Bits per group: the caller passes one byte per bit (vertex or group).
    std::size_t const nbits = std::min<std::size_t>(dg.vBitmapEven.getnBits(), words.size() * 64);
    for (std::size_t i = 0; i < nbits && i < static_cast<std::size_t>(g->n); ++i)
      if (frontier[i]) words[i >> 6] |= 1ull << (i & 63);
  }
  KITES_HIP_CHECK(hipMemcpy(dg.vBitmapEven.get_ptr(), words.data(), words.size() * 8,
                            hipMemcpyHostToDevice));
  g->seed = kites::options::seeding::none;
  g->unit_start = false;
  KITES_API_END
}

int kites_graph_init(kites_graph* g, int32_t algo, uint32_t source) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  auto& dg = *g->g;
  g->dev->setAsActive();
  hipStream_t s = g->dev->getSsEven().get();
  unsigned const nwords = static_cast<unsigned>(dg.vBitmapEven.size());
  // The frontier written here is the reference protocol's (used by classic
  // runs); delta runs start from the equivalent seed instead.
  if (algo == KITES_ALGO_CC) {
    hipLaunchKernelGGL(fill_values_kernel, dim3(1024), dim3(256), 0, s, dg.V.get_ptr(), g->n, 0u,
                       0u, 0u, 1, dg.vBitmapEven.get_ptr(), nwords, 1);
    g->seed = kites::options::seeding::all;
    g->unit_start = false;
  } else {
    need(source < g->n, "source out of range");
    bool const widest = (algo == KITES_ALGO_WIDEST);
    need(algo == KITES_ALGO_SSSP || algo == KITES_ALGO_BFS || algo == KITES_ALGO_SSSP_WRAP ||
             widest,
         "unknown functor set");
    // SSSP_WRAP adds without saturation, so it starts from the reference
    // example's INF = 2^30 - 1 (sssp_basic_single_device.cu:20-21), which
    // cannot wrap; the other sets start from 0xFFFFFFFF.
    unsigned const fill = widest ? 0u : (algo == KITES_ALGO_SSSP_WRAP ? 1073741823u : kInf);
    hipLaunchKernelGGL(fill_values_kernel, dim3(1024), dim3(256), 0, s, dg.V.get_ptr(), g->n,
                       fill, source, widest ? kInf : 0u, 0,
                       dg.vBitmapEven.get_ptr(), nwords, 0);
    hipLaunchKernelGGL(mark_out_kernel, dim3(64), dim3(256), 0, s, dg.oR.get_ptr(),
                       dg.oC.get_ptr(), source, dg.vBitmapEven.get_ptr());
    g->seed = kites::options::seeding::source;
    g->seed_source = source;
    g->unit_start = fill == kInf && !widest;
  }
  KITES_HIP_CHECK(hipPeekAtLastError());
  KITES_API_END
}

int kites_graph_init_sources(kites_graph* g, int32_t algo, const uint32_t* sources,
                             uint32_t count) {
  KITES_API_BEGIN
  need(g != nullptr && (sources != nullptr || count == 0), "null argument");
  need(algo == KITES_ALGO_SSSP || algo == KITES_ALGO_BFS || algo == KITES_ALGO_SSSP_WRAP,
       "multi-source start supports SSSP, BFS and SSSP_WRAP");
  std::vector<unsigned> list(sources, sources + count);
  std::sort(list.begin(), list.end());
  list.erase(std::unique(list.begin(), list.end()), list.end());
  for (unsigned v : list) need(v < g->n, "source out of range");
  auto& dg = *g->g;
  g->dev->setAsActive();
  hipStream_t s = g->dev->getSsEven().get();
  unsigned const nwords = static_cast<unsigned>(dg.vBitmapEven.size());
  unsigned const fill = algo == KITES_ALGO_SSSP_WRAP ? 1073741823u : kInf;
  if (g->seed_list.size() < list.size()) {
    g->seed_list.free();
    g->seed_list.alloc(list.size(), g->dev->getDevIdx());
  }
  if (!list.empty())
    KITES_HIP_CHECK(hipMemcpyAsync(g->seed_list.get_ptr(), list.data(), list.size() * 4,
                                   hipMemcpyHostToDevice, s));
  // values = fill everywhere (no source: the fill kernel's source index is n)
  hipLaunchKernelGGL(fill_values_kernel, dim3(1024), dim3(256), 0, s, dg.V.get_ptr(), g->n, fill,
                     g->n, 0u, 0, dg.vBitmapEven.get_ptr(), nwords, 0);
  if (!list.empty())
    hipLaunchKernelGGL(seed_values_kernel, dim3(static_cast<unsigned>(list.size())), dim3(256), 0,
                       s, dg.V.get_ptr(), g->seed_list.get_ptr(),
                       static_cast<unsigned>(list.size()), 0u, dg.vBitmapEven.get_ptr(),
                       dg.oR.get_ptr(), dg.oC.get_ptr());
  KITES_HIP_CHECK(hipPeekAtLastError());
  g->seed = kites::options::seeding::list;
  g->seed_count = static_cast<uint32_t>(list.size());
  g->unit_start = fill == kInf;
  KITES_API_END
}

int kites_graph_run(kites_graph* g, int32_t algo, const kites_run_opts* o, kites_run_info* info) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  kites::run_report rep;
  kites::options opt = to_options(g, o);
  opt.report = &rep;
  settle_bfs(g, algo, opt);
  // The recorded start is consumed by this run, as the frontier is.
  g->seed = kites::options::seeding::none;
  g->unit_start = false;
  with_functors(algo, [&](auto fi, auto fn, auto fr, auto fu) {
    kites::process<kites::launch::sync>(*g->g, *g->dev, fi, fn, fr, fu, opt);
  });
  if (opt.keep_stats) g->stats = rep.stats;
  if (info) {
    info->iterations = rep.iterations;
    info->ms = static_cast<float>(rep.ms);
    info->delta = rep.delta;
    info->min_fast = rep.min_fast;
    info->graph_builds = rep.graph_builds;
  }
  KITES_API_END
}

int kites_graph_values(kites_graph* g, uint32_t* out) {
  KITES_API_BEGIN
  need(g != nullptr && out != nullptr, "null argument");
  g->dev->setAsActive();
  KITES_HIP_CHECK(hipMemcpy(out, g->g->V.get_ptr(), g->n * sizeof(unsigned),
                            hipMemcpyDeviceToHost));
  KITES_API_END
}

int kites_graph_stats(kites_graph* g, kites_iter_stats* out, uint32_t cap, uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  static_assert(sizeof(kites_iter_stats) == sizeof(kites::detail::step_stats), "stats layout");
  uint32_t const k = static_cast<uint32_t>(g->stats.size());
  if (count) *count = k;
  if (out) std::memcpy(out, g->stats.data(), std::min(cap, k) * sizeof(kites_iter_stats));
  KITES_API_END
}

int kites_sssp_u32(kites_graph* g, uint32_t source, const kites_run_opts* o, kites_run_info* info,
                   uint32_t* dist_out) {
  int rc = kites_graph_init(g, KITES_ALGO_SSSP, source);
  if (rc) return rc;
  rc = kites_graph_run(g, KITES_ALGO_SSSP, o, info);
  if (rc || !dist_out) return rc;
  return kites_graph_values(g, dist_out);
}

int kites_bfs_u32(kites_graph* g, uint32_t source, const kites_run_opts* o, kites_run_info* info,
                  uint32_t* level_out) {
  int rc = kites_graph_init(g, KITES_ALGO_BFS, source);
  if (rc) return rc;
  rc = kites_graph_run(g, KITES_ALGO_BFS, o, info);
  if (rc || !level_out) return rc;
  return kites_graph_values(g, level_out);
}

int kites_cc_u32(kites_graph* g, const kites_run_opts* o, kites_run_info* info,
                 uint32_t* label_out) {
  int rc = kites_graph_init(g, KITES_ALGO_CC, 0);
  if (rc) return rc;
  rc = kites_graph_run(g, KITES_ALGO_CC, o, info);
  if (rc || !label_out) return rc;
  return kites_graph_values(g, label_out);
}

int kites_graph_record(kites_graph* g, int32_t algo, const kites_run_opts* o,
                       uint32_t rollout_variant, kites_variant_sample* out, uint32_t cap,
                       uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  need(rollout_variant <= 8, "rollout variant must be 0..8");
  static_assert(sizeof(kites_variant_sample) == sizeof(kites::variant_sample), "sample layout");
  std::vector<kites::variant_sample> rec;
  kites::options opt = to_options(g, o);
  need(opt.sched != kites::schedule::classic, "the recorder needs a delta schedule");
  opt.record = &rec;
  opt.record_rollout = rollout_variant;
  settle_bfs(g, algo, opt);
  g->seed = kites::options::seeding::none;
  g->unit_start = false;
  with_functors(algo, [&](auto fi, auto fn, auto fr, auto fu) {
    kites::process<kites::launch::sync>(*g->g, *g->dev, fi, fn, fr, fu, opt);
  });
  if (count) *count = static_cast<uint32_t>(rec.size());
  if (out) std::memcpy(out, rec.data(), std::min<std::size_t>(cap, rec.size()) * sizeof(kites_variant_sample));
  KITES_API_END
}

int kites_graph_autotune(kites_graph* g, int32_t algo, const uint32_t* sources, uint32_t nsrc,
                         kites_run_opts* best, float* best_ms, kites_tune_entry* table,
                         uint32_t cap, uint32_t* count) {
  KITES_API_BEGIN
  need(g != nullptr && best != nullptr, "null argument");
  need(algo == KITES_ALGO_CC || (sources != nullptr && nsrc >= 1), "sources needed");
  int cus = 0;
  KITES_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                        g->dev->getDevIdx()));
  auto start = [&]() {
    int rc;
    if (algo == KITES_ALGO_CC) rc = kites_graph_init(g, algo, 0);
    else if (nsrc == 1) rc = kites_graph_init(g, algo, sources[0]);
    else rc = kites_graph_init_sources(g, algo, sources, nsrc);
    if (rc) throw std::runtime_error(kites_capi::last_error());
  };
  // workgroups per compute unit x iterations per replayed chunk; every
  // shape must give the same values (checked against the first).  Up to six
  // rounds of the 7 resident workgroups per CU: SSSP on R-MAT-22 is fastest
  // at 35 per CU, BFS at the default 21.
  static const unsigned kPerCu[] = {2, 4, 8, 14, 21, 28, 35, 42};
  static const unsigned kChunk[] = {4, 8, 16};
  kites_run_opts const base = *best;
  std::vector<unsigned> ref, cur(g->n);
  float bms = 0.f;
  bool have = false;
  uint32_t k = 0;
  for (unsigned pc : kPerCu)
    for (unsigned ch : kChunk) {
      kites_run_opts o = base;
      o.grid = static_cast<int32_t>(pc * static_cast<unsigned>(cus));
      o.chunk = static_cast<int32_t>(ch);
      o.keep_stats = 0;
      float t = 0.f;
      for (int r = 0; r < 2; ++r) {
        start();
        kites_run_info info{};
        if (kites_graph_run(g, algo, &o, &info)) throw std::runtime_error(kites_capi::last_error());
        t = (r == 0 || info.ms < t) ? info.ms : t;
      }
      g->dev->setAsActive();
      KITES_HIP_CHECK(hipMemcpy(cur.data(), g->g->V.get_ptr(), g->n * 4, hipMemcpyDeviceToHost));
      if (ref.empty()) ref = cur;
      else if (cur != ref) throw std::runtime_error("autotune: a shape gave different values");
      if (table && k < cap) table[k] = kites_tune_entry{pc, 256u, ch, t};
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

uint32_t kites_mlp_nparam(void) { return static_cast<uint32_t>(kites::detail::MLP_NPARAM); }

int kites_graph_set_model(kites_graph* g, const float* weights, uint32_t nparam) {
  KITES_API_BEGIN
  need(g != nullptr && weights != nullptr, "null argument");
  need(nparam == static_cast<uint32_t>(kites::detail::MLP_NPARAM), "wrong parameter count");
  g->dev->setAsActive();
  if (g->model.size() == 0) g->model.alloc(nparam, g->dev->getDevIdx());
  KITES_HIP_CHECK(hipMemcpy(g->model.get_ptr(), weights, nparam * sizeof(float),
                            hipMemcpyHostToDevice));
  KITES_API_END
}

}  // extern "C"
