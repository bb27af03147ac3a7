// ordered_capi.hip -- C ABI of the ordered mode (declarations in
// include/kites_amd.h): float PageRank, a functor set whose reduction is a
// floating-point sum, run with options::ordered so the result is defined bit
// for bit.
#include "capi_internal.hpp"

namespace {

using kites_capi::need;

// in-edge weight 1 / outdeg(source), as float
static __global__ void inv_outdeg_kernel(const unsigned* __restrict__ oR,
                                         const unsigned* __restrict__ C, unsigned m,
                                         float* __restrict__ W) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    unsigned const u = C[e];
    W[e] = 1.0f / static_cast<float>(oR[u + 1] - oR[u]);
  }
}

static __global__ void fill_kernel(float* __restrict__ V, unsigned n, float x,
                                   unsigned long long* __restrict__ bits, unsigned words) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) V[i] = x;
  for (unsigned w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    unsigned const lo = w * 64u;
    bits[w] = lo >= n ? 0ull : (n - lo >= 64 ? ~0ull : ((1ull << (n - lo)) - 1ull));
  }
}

}  // namespace

extern "C" {

int kites_pagerank_f32(kites_graph* g, float damping, float tol, uint32_t max_iterations,
                       float* out, kites_run_info* info) {
  KITES_API_BEGIN
  need(g != nullptr && g->g, "null graph");
  need(damping >= 0.f && damping <= 1.f, "damping must be in [0, 1]");
  need(tol >= 0.f, "tolerance must be non-negative");
  auto& src = *g->g;
  kites::gpu& dev = *g->dev;
  int const did = dev.getDevIdx();
  std::size_t const n = g->n;
  std::size_t const m = g->m;
  // A float view of the graph: the same in-CSR and out-lists, weights
  // 1 / outdeg(source).
  kites::graph_csr_device<float, float> f(n, dev);
  kites::stream& s = dev.getSsEven();
  auto d2d = [&](void* d, const void* h, std::size_t bytes) {
    if (bytes) KITES_HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyDeviceToDevice, s.get()));
  };
  f.C.alloc(m ? m : 1, did);
  f.E.alloc(m ? m : 1, did);
  f.oC.alloc(m ? m : 1, did);
  d2d(f.R.get_ptr(), src.R.get_ptr(), (n + 1) * 4);
  d2d(f.C.get_ptr(), src.C.get_ptr(), m * 4);
  d2d(f.oR.get_ptr(), src.oR.get_ptr(), (n + 1) * 4);
  d2d(f.oC.get_ptr(), src.oC.get_ptr(), m * 4);
  f.set_num_edges(m);
  unsigned const cus = static_cast<unsigned>(dev.numComputeUnits());
  if (m)
    hipLaunchKernelGGL(inv_outdeg_kernel, dim3(cus * 8), dim3(256), 0, s.get(),
                       reinterpret_cast<const unsigned*>(src.oR.get_ptr()),
                       reinterpret_cast<const unsigned*>(src.C.get_ptr()),
                       static_cast<unsigned>(m), f.E.get_ptr());
  // every vertex starts at 1 and active
  hipLaunchKernelGGL(fill_kernel, dim3(cus * 8), dim3(256), 0, s.get(), f.V.get_ptr(),
                     static_cast<unsigned>(n), 1.0f, f.vBitmapEven.get_ptr(),
                     static_cast<unsigned>(f.vBitmapEven.size()));
  KITES_HIP_CHECK(hipPeekAtLastError());
  s.sync();

  float const base = 1.0f - damping;
  auto initF = [base] __host__ __device__(volatile float& locV, float&) { locV = base; };
  auto nbrF = [damping] __host__ __device__(volatile float& partialV, float srcV, float* connE) {
    partialV = damping * srcV * (*connE);
  };
  auto redF = [] __host__ __device__(volatile float& lhsV, volatile float& rhsV) {
    lhsV = lhsV + rhsV;
  };
  auto updF = [tol] __host__ __device__(volatile float& newV, float oldV) {
    float const d = newV - oldV;
    return d > tol || d < -tol;
  };
  kites::run_report rep;
  kites::options o;
  o.ordered = true;
  o.max_iterations = max_iterations ? max_iterations : 10000u;
  o.report = &rep;
  kites::process<kites::launch::sync>(f, dev, initF, nbrF, redF, updF, o);
  if (out && n)
    KITES_HIP_CHECK(hipMemcpy(out, f.V.get_ptr(), n * sizeof(float), hipMemcpyDeviceToHost));
  if (info) {
    info->iterations = rep.iterations;
    info->ms = static_cast<float>(rep.ms);
    info->delta = 0;
    info->min_fast = 0;
    info->graph_builds = 0;
  }
  KITES_API_END
}

}  // extern "C"
