// graph_device.hpp -- the device-resident graph.
//
// graph_csr_device keeps the reference's public members
// (kites/graph/graph_csr_device.cuh:21-52: V, R, C, E, vBitmapEven,
// vBitmapOdd, updateFlag, updateFlagHost, gProp, dir_R, dir_C,
// vertexPerGroup) and make_graph_csr_for_device (:146-152).  It adds the
// weighted out-edge CSR (oR, oC, oE), the transpose of the in-edge CSR,
// which the push variants and the frontier expansion read; it is built on
// the host at upload time so its order is deterministic (per source,
// destinations ascending, as the reference's dir_C for vertex-per-group 1,
// graph_raw.h:135-139).  The engine's per-run workspace is allocated on the
// first process() call and kept with the graph.
//
// HBM layout (u32 everywhere, n padded to 128): V n | R n+1 | C m | E m |
// oR n+1 | oC m | oE m | dir_R/dir_C (group lists, vertex-per-group > 1
// only) | workspace: 3 frontier bitmaps of ceil(n/64)+1 u64 words, 2
// queues of n u32 with 2 out-edge offset arrays of n+1 u32,
// per-iteration statistics, control block.
#ifndef KITES_GRAPH_DEVICE_HPP_
#define KITES_GRAPH_DEVICE_HPP_

#include <algorithm>
#include <chrono>
#include <iostream>
#include <memory>
#include <vector>

#include "common.hpp"
#include "detail/batch_kernel.hpp"
#include "detail/bitbfs_kernel.hpp"
#include "detail/launch_cache.hpp"
#include "detail/step_kernel.hpp"
#include "device.hpp"
#include "graph.hpp"
#include "memory.hpp"

namespace kites {
namespace detail {

// Scratch of multi-source runs (kites/batch.hpp): the 3-slot frontier
// mask ring, work-item start positions, statistics and control block.
struct batch_workspace {
  int dev = 0;
  std::size_t n = 0;
  unsigned epi = 0, nitems = 0, max_iter = 0;
  device_buffer<unsigned long long> F;      // K-wide frontier ring (3 x one bit per vertex)
  device_buffer<unsigned long long> M;      // K-wide per-source masks (3 x n words, opt-in)
  device_buffer<unsigned> cstart;          // non-empty-vertex position per 64-edge chunk
  std::size_t cstart_m = 0;
  device_buffer<unsigned> src;
  device_buffer<batch_stats> stats;
  device_buffer<batch_ctl> ctl;
  host_pinned_buffer<batch_ctl> ctl_host;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // bit-parallel BFS (256 sources): visited masks, 3-slot frontier masks,
  // 3-slot one-bit summaries, statistics
  std::size_t bn = 0;
  device_buffer<unsigned long long> VIS, F4, A;
  device_buffer<unsigned> L8w;             // byte levels, n x 256 B (as u32 words)
  device_buffer<bitbfs_stats> bstats;
  device_buffer<unsigned> bsrc;
  launch_cache graphs;                     // captured iteration chunks
  ~batch_workspace() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

// Per-run scratch of the engine (one per device graph).
struct workspace {
  int dev = 0;
  std::size_t n = 0;
  unsigned nwords = 0;
  unsigned max_iter = 0;
  device_buffer<unsigned long long> bm;   // 3 bitmaps of nwords words
  device_buffer<unsigned> q;              // 2 queues of n entries
  device_buffer<unsigned> qe;             // 2 offset arrays of n + 1 entries
  device_buffer<unsigned> fq;             // near-far: 2 far piles of n entries (on first use)
  device_buffer<unsigned long long> fbm;  // near-far: far bitmap of nwords words
  device_buffer<step_stats> stats;
  device_buffer<step_ctl> ctl;
  host_pinned_buffer<step_ctl> ctl_host;
  device_buffer<group_stats> gst;         // per-group partial statistics
  launch_cache graphs;                     // captured iteration chunks
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evc[2] = {nullptr, nullptr};  // completion of the two control-line copies in flight
  unsigned max_outdeg = 0;                 // the graph's largest out-degree, once computed
  bool have_max_outdeg = false;

  // Largest out-degree of the graph whose out-offsets are oR (n vertices),
  // computed on the device the first time and kept.
  unsigned max_outdeg_of(const unsigned* oR, std::size_t nv, stream& s) {
    if (have_max_outdeg) return max_outdeg;
    device_buffer<unsigned> out(1, dev);
    KITES_HIP_CHECK(hipMemsetAsync(out.get_ptr(), 0, sizeof(unsigned), s.get()));
    if (nv) {
      unsigned const blocks = static_cast<unsigned>(std::min<std::size_t>((nv + 255) / 256, 4096));
      hipLaunchKernelGGL(max_outdeg_kernel, dim3(blocks), dim3(256), 0, s.get(), oR,
                         static_cast<unsigned>(nv), out.get_ptr());
      KITES_HIP_CHECK(hipPeekAtLastError());
    }
    KITES_HIP_CHECK(hipMemcpyAsync(&max_outdeg, out.get_ptr(), sizeof(unsigned),
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    have_max_outdeg = true;
    return max_outdeg;
  }

  void ensure(std::size_t nv, unsigned iters, int device) {
    dev = device;
    unsigned const words = static_cast<unsigned>((nv + 63) / 64 + 1);
    if (nv > n || words > nwords) {
      bm.free();
      bm.alloc(3ull * words, dev);
      q.free();
      q.alloc(2 * (nv ? nv : 1), dev);
      qe.free();
      qe.alloc(2 * (nv + 1), dev);
      fq.free();
      fbm.free();
      n = nv;
      nwords = words;
    }
    if (iters > max_iter) {
      stats.free();
      stats.alloc(iters, dev);
      max_iter = iters;
    }
    if (ctl.size() == 0) {
      ctl.alloc(1, dev);
      ctl_host.alloc(2);
      gst.alloc(kTicketGroups, dev);
      KITES_HIP_CHECK(hipMemset(ctl.get_ptr(), 0, sizeof(step_ctl)));
      KITES_HIP_CHECK(hipEventCreate(&ev0));
      KITES_HIP_CHECK(hipEventCreate(&ev1));
      KITES_HIP_CHECK(hipEventCreateWithFlags(&evc[0], hipEventDisableTiming));
      KITES_HIP_CHECK(hipEventCreateWithFlags(&evc[1], hipEventDisableTiming));
    }
  }
  // The near-far far piles and bitmap, allocated by the first run that
  // needs them (after ensure()).
  void ensure_far() {
    if (fq.size() == 0) fq.alloc(2 * (n ? n : 1), dev);
    if (fbm.size() == 0) fbm.alloc(nwords, dev);
  }
  ~workspace() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (hipEvent_t e : evc)
      if (e) (void)hipEventDestroy(e);
  }
};

// Host transpose of an in-edge CSR into the weighted out-edge CSR.
template <class eT>
inline void transpose_host(std::size_t n, const unsigned* R, const unsigned* C, const eT* E,
                           std::vector<unsigned>& oR, std::vector<unsigned>& oC,
                           std::vector<eT>& oE) {
  std::size_t const m = n ? R[n] : 0;
  oR.assign(n + 1, 0);
  oC.resize(m);
  oE.resize(m);
  for (std::size_t e = 0; e < m; ++e) oR[C[e] + 1]++;
  for (std::size_t v = 0; v < n; ++v) oR[v + 1] += oR[v];
  std::vector<unsigned> cur(oR.begin(), oR.end() - 1);
  for (std::size_t v = 0; v < n; ++v)
    for (unsigned e = R[v]; e < R[v + 1]; ++e) {
      unsigned const k = cur[C[e]]++;
      oC[k] = static_cast<unsigned>(v);
      oE[k] = E[e];
    }
}

// Vertices with in-edges and their in-edge offsets (the work list of the
// edge-balanced pull).
inline void nonempty_host(std::size_t n, const unsigned* R, std::vector<unsigned>& nzi,
                          std::vector<unsigned>& nzR) {
  nzi.clear();
  nzR.clear();
  for (std::size_t v = 0; v < n; ++v)
    if (R[v + 1] != R[v]) {
      nzi.push_back(static_cast<unsigned>(v));
      nzR.push_back(R[v]);
    }
  nzR.push_back(n ? R[n] : 0u);
}

}  // namespace detail

template <class vT, class eT, class vIdxT = unsigned int>
class graph_csr_device : public graph_base {
  static_assert(sizeof(vIdxT) == 4, "the engine indexes vertices and edges with 32 bits");
  using eIdxT = vIdxT;

 public:
  int residing_device_id;
  device_buffer<vT> V;
  device_buffer<eIdxT> R;
  device_buffer<vIdxT> C;
  device_buffer<eT> E;
  device_bitmap vBitmapEven;
  device_bitmap vBitmapOdd;
  device_buffer<unsigned> updateFlag;
  host_pinned_buffer<unsigned> updateFlagHost;
  graph_property gProp;
  device_buffer<eIdxT> dir_R;
  device_buffer<vIdxT> dir_C;
  unsigned vertexPerGroup;
  // Weighted out-edge CSR (transpose of R, C, E).
  device_buffer<eIdxT> oR;
  device_buffer<vIdxT> oC;
  device_buffer<eT> oE;
  // Vertices with in-edges and their offsets (edge-balanced pull).
  device_buffer<vIdxT> nzi;
  device_buffer<eIdxT> nzR;
  std::unique_ptr<detail::workspace> ws;
  std::unique_ptr<detail::batch_workspace> bws;   // multi-source runs (batch.hpp)
  // Source-split in-CSR of the edge-balanced pull (split.hpp; built on the
  // first run that asks for it).
  device_buffer<vIdxT> xC;
  device_buffer<eT> xE;
  device_buffer<vIdxT> xnzi;
  device_buffer<eIdxT> xnzR;
  unsigned xseg[detail::kSplitParts + 1] = {};
  std::size_t xm = 0;
  bool xbuilt = false;

  void upload_nonempty(const eIdxT* Rh, std::size_t n, stream& s) {
    std::vector<unsigned> a, b;
    detail::nonempty_host(n, Rh, a, b);
    nzi.alloc(a.size() ? a.size() : 1, residing_device_id);
    nzR.alloc(b.size(), residing_device_id);
    if (!a.empty())
      KITES_HIP_CHECK(hipMemcpyAsync(nzi.get_ptr(), a.data(), a.size() * 4, hipMemcpyHostToDevice,
                                     s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(nzR.get_ptr(), b.data(), b.size() * 4, hipMemcpyHostToDevice,
                                   s.get()));
    s.sync();
    nnz = static_cast<unsigned>(a.size());
  }
  unsigned nnz = 0;
  // bit v set iff v has out-edges (push variants test it per updated edge)
  device_buffer<unsigned long long> outbits;
  void build_outbits(std::size_t n, stream& s) {
    unsigned const nw = static_cast<unsigned>((n + 63) / 64 + 1);
    outbits.alloc(nw, residing_device_id);
    hipLaunchKernelGGL(detail::outbits_kernel, dim3((nw + 255) / 256), dim3(256), 0, s.get(),
                       oR.get_ptr(), static_cast<unsigned>(n), outbits.get_ptr(), nw);
    KITES_HIP_CHECK(hipPeekAtLastError());
  }

  int getResidingDeviceID() const { return residing_device_id; }
  unsigned getVertexPerGroup() const { return vertexPerGroup; }

  graph_csr_device() = default;
  graph_csr_device(graph_csr_device&&) = default;
  graph_csr_device& operator=(graph_csr_device&&) = default;

  graph_csr_device(graph_csr<vT, eT, vIdxT>& h, gpu& dev)
      : residing_device_id{[&] {
          dev.setAsActive();
          return dev.getDevIdx();
        }()},
        V{h.V.size(), residing_device_id},
        R{h.R.size(), residing_device_id},
        C{h.C.size(), residing_device_id},
        E{h.E.size(), residing_device_id},
        vBitmapEven{std::max(h.vBitmap.getnBits(), h.get_num_vertices()), residing_device_id},
        vBitmapOdd{std::max(h.vBitmap.getnBits(), h.get_num_vertices()), residing_device_id},
        updateFlag{1, residing_device_id},
        updateFlagHost{1},
        gProp{h.gProp},
        vertexPerGroup{h.getVertexPerGroup() ? h.getVertexPerGroup() : 1u} {
    auto t1 = std::chrono::steady_clock::now();
    stream& s = dev.getSsEven();
    std::size_t const n = h.get_num_vertices();
    copyMemOnStream(V, h.V, s);
    copyMemOnStream(R, h.R, s);
    copyMemOnStream(C, h.C, s);
    copyMemOnStream(E, h.E, s);
    copyMemOnStream(vBitmapEven, h.vBitmap, s);
    KITES_HIP_CHECK(hipMemsetAsync(vBitmapOdd.get_ptr(), 0, vBitmapOdd.sizeInBytes(), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(updateFlag.get_ptr(), 0, sizeof(unsigned), s.get()));
    updateFlagHost[0] = 0;
    if (vertexPerGroup != 1 && gProp == graph_property::directed) {
      dir_R.alloc(h.dir_R.size(), residing_device_id);
      dir_C.alloc(h.dir_C.size(), residing_device_id);
      copyMemOnStream(dir_R, h.dir_R, s);
      copyMemOnStream(dir_C, h.dir_C, s);
    }
    std::vector<eIdxT> hoR;
    std::vector<vIdxT> hoC;
    std::vector<eT> hoE;
    detail::transpose_host(n, h.R.get_ptr(), h.C.get_ptr(), h.E.get_ptr(), hoR, hoC, hoE);
    oR.alloc(hoR.size(), residing_device_id);
    oC.alloc(hoC.size(), residing_device_id);
    oE.alloc(hoE.size(), residing_device_id);
    if (!hoR.empty())
      KITES_HIP_CHECK(hipMemcpyAsync(oR.get_ptr(), hoR.data(), hoR.size() * sizeof(eIdxT),
                                     hipMemcpyHostToDevice, s.get()));
    if (!hoC.empty()) {
      KITES_HIP_CHECK(hipMemcpyAsync(oC.get_ptr(), hoC.data(), hoC.size() * sizeof(vIdxT),
                                     hipMemcpyHostToDevice, s.get()));
      KITES_HIP_CHECK(hipMemcpyAsync(oE.get_ptr(), hoE.data(), hoE.size() * sizeof(eT),
                                     hipMemcpyHostToDevice, s.get()));
    }
    set_num_vertices(n);
    set_num_edges(h.get_num_edges());
    s.sync();
    upload_nonempty(h.R.get_ptr(), n, s);
    build_outbits(n, s);
    s.sync();
    auto t2 = std::chrono::steady_clock::now();
    std::cout << "H2D copies took " << std::chrono::duration<double, std::milli>(t2 - t1).count()
              << " (ms) in total.\n";
  }

  // Extension: the vertex arrays of an n-vertex directed graph (vertex-per-
  // group 1) with zero values and an empty frontier; the edge arrays are
  // filled by make_graph_csr_from_edges (ingest.hpp).
  graph_csr_device(std::size_t n, gpu& dev)
      : residing_device_id{[&] {
          dev.setAsActive();
          return dev.getDevIdx();
        }()},
        V{n, residing_device_id},
        R{n + 1, residing_device_id},
        vBitmapEven{n, residing_device_id},
        vBitmapOdd{n, residing_device_id},
        updateFlag{1, residing_device_id},
        updateFlagHost{1},
        gProp{graph_property::directed},
        vertexPerGroup{1},
        oR{n + 1, residing_device_id} {
    stream& s = dev.getSsEven();
    KITES_HIP_CHECK(hipMemsetAsync(V.get_ptr(), 0, n * sizeof(vT), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(vBitmapEven.get_ptr(), 0, vBitmapEven.sizeInBytes(), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(vBitmapOdd.get_ptr(), 0, vBitmapOdd.sizeInBytes(), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(updateFlag.get_ptr(), 0, sizeof(unsigned), s.get()));
    updateFlagHost[0] = 0;
    set_num_vertices(n);
    set_num_edges(0);
    s.sync();
  }

  // Extension: from host in-edge CSR arrays already holding the out-edge
  // CSR (no host transpose).  All pointers are host memory.
  graph_csr_device(std::size_t n, std::size_t m, const eIdxT* Rh, const vIdxT* Ch, const eT* Eh,
                   const eIdxT* oRh, const vIdxT* oCh, const eT* oEh, graph_property gp, gpu& dev)
      : residing_device_id{[&] {
          dev.setAsActive();
          return dev.getDevIdx();
        }()},
        V{n, residing_device_id},
        R{n + 1, residing_device_id},
        C{m, residing_device_id},
        E{m, residing_device_id},
        vBitmapEven{n, residing_device_id},
        vBitmapOdd{n, residing_device_id},
        updateFlag{1, residing_device_id},
        updateFlagHost{1},
        gProp{gp},
        vertexPerGroup{1},
        oR{n + 1, residing_device_id},
        oC{m, residing_device_id},
        oE{m, residing_device_id} {
    stream& s = dev.getSsEven();
    auto up = [&](void* d, const void* hsrc, std::size_t bytes) {
      if (bytes) KITES_HIP_CHECK(hipMemcpyAsync(d, hsrc, bytes, hipMemcpyHostToDevice, s.get()));
    };
    up(R.get_ptr(), Rh, (n + 1) * sizeof(eIdxT));
    up(C.get_ptr(), Ch, m * sizeof(vIdxT));
    up(E.get_ptr(), Eh, m * sizeof(eT));
    up(oR.get_ptr(), oRh, (n + 1) * sizeof(eIdxT));
    up(oC.get_ptr(), oCh, m * sizeof(vIdxT));
    up(oE.get_ptr(), oEh, m * sizeof(eT));
    KITES_HIP_CHECK(hipMemsetAsync(vBitmapEven.get_ptr(), 0, vBitmapEven.sizeInBytes(), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(vBitmapOdd.get_ptr(), 0, vBitmapOdd.sizeInBytes(), s.get()));
    KITES_HIP_CHECK(hipMemsetAsync(updateFlag.get_ptr(), 0, sizeof(unsigned), s.get()));
    updateFlagHost[0] = 0;
    set_num_vertices(n);
    set_num_edges(m);
    s.sync();
    upload_nonempty(Rh, n, s);
    build_outbits(n, s);
    s.sync();
  }
};

template <class vT, class eT, class vIdxT>
graph_csr_device<vT, eT, vIdxT> make_graph_csr_for_device(graph_csr<vT, eT, vIdxT>& h, gpu& dev) {
  return graph_csr_device<vT, eT, vIdxT>(h, dev);
}

}  // namespace kites

#endif  // KITES_GRAPH_DEVICE_HPP_
