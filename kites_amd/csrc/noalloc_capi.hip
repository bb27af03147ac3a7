// noalloc_capi.hip -- kites_sssp_u32_noalloc: single-source SSSP on a
// device graph that allocates nothing (declaration and contract in
// include/kites_amd.h).
//
// The run state of the engine (frontier bitmap ring, queues with their
// out-edge offsets, per-iteration statistics, control block, group
// statistics) lives in `scratch`, a file-scope device array sized for
// graphs of up to kMaxVertices vertices and runs of up to kMaxIterations
// iterations.  Every GPU holds its own copy of the array from the moment
// this code object is loaded until the process ends.  Its device address
// is taken once per device with hipGetSymbolAddress, with that device
// current; host code only ever passes that address to the kernels, never
// the array's name (which names a host-side stand-in).
//
// The array keeps the last call's contents; the engine's reset kernel
// clears everything a run reads before it is written.  Calls on one device
// share the array, so they are serialized by a per-device mutex here:
// calls from several host threads to one device wait for each other
// (calls that overlapped on one device would race on the array).
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "functor_sets.hpp"

namespace {

using kites_capi::need;
using kites::detail::group_stats;
using kites::detail::kTicketGroups;
using kites::detail::step_ctl;
using kites::detail::step_stats;

constexpr std::size_t kMaxVertices = std::size_t(1) << 22;   // 4,194,304 (R-MAT scale 22)
constexpr unsigned kMaxIterations = 1u << 16;               // the default iteration limit
constexpr std::size_t kMaxDevices = 64;

constexpr std::size_t align256(std::size_t b) { return (b + 255) & ~std::size_t(255); }

// Byte offsets of the regions inside the array (each 256-byte aligned).
constexpr std::size_t kMaxWords = kMaxVertices / 64 + 1;
constexpr std::size_t kOffBm = 0;
constexpr std::size_t kOffQ = kOffBm + align256(3 * kMaxWords * sizeof(unsigned long long));
constexpr std::size_t kOffQe = kOffQ + align256(2 * kMaxVertices * sizeof(unsigned));
constexpr std::size_t kOffStats = kOffQe + align256(2 * (kMaxVertices + 1) * sizeof(unsigned));
constexpr std::size_t kOffCtl = kOffStats + align256(std::size_t(kMaxIterations) * sizeof(step_stats));
constexpr std::size_t kOffGst = kOffCtl + align256(sizeof(step_ctl));
constexpr std::size_t kOffWord = kOffGst + align256(kTicketGroups * sizeof(group_stats));
constexpr std::size_t kScratchBytes = kOffWord + 256;
static_assert(kScratchBytes < (std::size_t(80) << 20), "the array stays below 80 MB");
static_assert(alignof(step_ctl) <= 256 && alignof(group_stats) <= 256, "region alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu[kMaxDevices];
std::once_flag g_once[kMaxDevices];
unsigned char* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device`, which must be current.
unsigned char* scratch_on(int device) {
  need(device >= 0 && static_cast<std::size_t>(device) < kMaxDevices, "device index out of range");
  std::call_once(g_once[device], [device] {
    void* p = nullptr;
    hipError_t const e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("HIPAssert: hipGetSymbolAddress(scratch): ") +
                               hipGetErrorString(e));
    g_ptr[device] = static_cast<unsigned char*>(p);
  });
  return g_ptr[device];
}

}  // namespace

extern "C" {

int kites_sssp_u32_noalloc(kites_graph* g, uint32_t source, const kites_run_opts* o,
                           kites_run_info* info, uint32_t* dist_out) {
  KITES_API_BEGIN
  need(g != nullptr, "null graph");
  need(g->g->get_num_vertices() <= kMaxVertices,
       "kites_sssp_u32_noalloc: the graph has more than 4,194,304 vertices (the scratch limit)");
  need(!o || o->max_iterations <= kMaxIterations,
       "kites_sssp_u32_noalloc: max_iterations above 65,536 (the scratch limit)");
  need(!o || !o->split_pull, "kites_sssp_u32_noalloc: the split pull allocates");
  need(!o || !o->keep_stats, "kites_sssp_u32_noalloc: per-iteration statistics allocate");
  need(!o || !o->delta, "kites_sssp_u32_noalloc: near-far runs need the far pile of the graph's workspace");
  int const device = g->dev->getDevIdx();
  need(device >= 0 && static_cast<std::size_t>(device) < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu[device]);
  g->dev->setAsActive();
  unsigned char* const base = scratch_on(device);

  int rc = kites_graph_init(g, KITES_ALGO_SSSP, source);
  if (rc) return rc;
  kites::options opt = kites_capi::to_options(g, o);
  g->seed = kites::options::seeding::none;   // the recorded start is consumed

  kites::detail::run_state st;
  st.bm = reinterpret_cast<unsigned long long*>(base + kOffBm);
  st.q = reinterpret_cast<unsigned*>(base + kOffQ);
  st.qe = reinterpret_cast<unsigned*>(base + kOffQe);
  st.stats = reinterpret_cast<step_stats*>(base + kOffStats);
  st.ctl = reinterpret_cast<step_ctl*>(base + kOffCtl);
  st.gst = reinterpret_cast<group_stats*>(base + kOffGst);
  st.word = reinterpret_cast<unsigned*>(base + kOffWord);
  st.nwords = static_cast<unsigned>((g->g->get_num_vertices() + 63) / 64 + 1);
  st.max_iter = kMaxIterations;
  step_ctl ctl_host{};
  st.ctl_host = &ctl_host;

  kites::run_report rep;
  opt.report = &rep;
  kites_capi::with_sssp([&](auto fi, auto fn, auto fr, auto fu) {
    kites::process_noalloc<kites::launch::sync>(*g->g, *g->dev, fi, fn, fr, fu, opt, st);
  });
  if (info) {
    info->iterations = rep.iterations;
    info->ms = static_cast<float>(rep.ms);
    info->delta = rep.delta;
    info->min_fast = rep.min_fast;
    info->graph_builds = 0;
  }
  if (dist_out)
    KITES_HIP_CHECK(hipMemcpy(dist_out, g->g->V.get_ptr(), g->n * sizeof(unsigned),
                              hipMemcpyDeviceToHost));
  KITES_API_END
}

}  // extern "C"
