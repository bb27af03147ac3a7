// ingest_capi.hip -- C ABI of the GPU graph build (declarations in
// include/kites_amd.h).  A translation unit of its own: hipCUB's sort and
// scan headers are compiled only here.
#include "capi_internal.hpp"
#include "kites/ingest.hpp"

namespace {

using kites_capi::need;

template <class T>
void download(T* dst, const kites::device_buffer<T>& src, std::size_t count) {
  if (dst && count)
    KITES_HIP_CHECK(hipMemcpy(dst, src.get_ptr(), count * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace

extern "C" {

int kites_graph_from_edges(int32_t device, uint32_t n, uint64_t m, const uint32_t* src,
                           const uint32_t* dst, const uint32_t* w, int32_t dedup,
                           kites_graph** out, float* build_ms) {
  KITES_API_BEGIN
  need(out != nullptr, "null argument");
  need(m == 0 || (src != nullptr && dst != nullptr && w != nullptr), "null edge array");
  need(n > 0, "a graph needs at least one vertex");
  std::unique_ptr<kites_graph> g(new kites_graph());
  g->dev.reset(new kites::gpu(device));
  int const did = g->dev->getDevIdx();
  kites::device_buffer<unsigned> ds(m ? m : 1, did), dd(m ? m : 1, did), dw(m ? m : 1, did);
  if (m) {
    KITES_HIP_CHECK(hipMemcpy(ds.get_ptr(), src, m * 4, hipMemcpyHostToDevice));
    KITES_HIP_CHECK(hipMemcpy(dd.get_ptr(), dst, m * 4, hipMemcpyHostToDevice));
    KITES_HIP_CHECK(hipMemcpy(dw.get_ptr(), w, m * 4, hipMemcpyHostToDevice));
  }
  kites::ingest_report rep;
  g->g.reset(new kites::graph_csr_device<unsigned, unsigned>(
      kites::make_graph_csr_from_edges<unsigned, unsigned, unsigned>(
          *g->dev, n, m, ds.get_ptr(), dd.get_ptr(), dw.get_ptr(), dedup != 0, &rep)));
  g->n = static_cast<uint32_t>(rep.n);
  g->m = rep.m;
  if (build_ms) *build_ms = rep.build_ms;
  *out = g.release();
  KITES_API_END
}

int kites_graph_arrays(const kites_graph* g, uint32_t* R, uint32_t* C, uint32_t* E, uint32_t* oR,
                       uint32_t* oC, uint32_t* oE, uint32_t* nzi, uint32_t* nzR, uint32_t* nnz) {
  KITES_API_BEGIN
  need(g != nullptr && g->g, "null graph");
  auto const& d = *g->g;
  std::size_t const n = g->n;
  std::size_t const m = g->m;
  download(R, d.R, n + 1);
  download(C, d.C, m);
  download(E, d.E, m);
  download(oR, d.oR, n + 1);
  download(oC, d.oC, m);
  download(oE, d.oE, m);
  download(nzi, d.nzi, d.nnz);
  download(nzR, d.nzR, static_cast<std::size_t>(d.nnz) + 1);
  if (nnz) *nnz = d.nnz;
  KITES_API_END
}

}  // extern "C"
