// split.hpp -- the source-split in-CSR of the edge-balanced pull.
//
// Each XCD of the MI355X has an L2 of its own (4 MB); V of an R-MAT-22
// graph is 16 MB, so a pull whose waves gather V[source] from anywhere
// misses L2 on most gathers.  The split groups the in-edges by the part of
// their source (8 parts; a part owns whole 32-vertex lines of V, chosen by a
// hash of the line, detail::split_part) and, within a part, by destination
// and then source, the in-CSR's own order.  Workgroup w of a pull takes part
// w % 8, and consecutive workgroups are placed on consecutive XCDs, so each
// XCD gathers from its own eighth of V (2 MB at R-MAT-22).  A destination's
// partial minima from different parts meet in the flush's atomic relax, as
// partials of one destination from different waves already do, so results
// do not depend on where a workgroup runs; only locality does.
//
// Built on the GPU on first use: one stable radix sort of 32-bit keys
// (part, destination) with the edge position as value, a gather of C and E,
// and a flag scan for the segments.  It holds one more copy of C and E.
#ifndef KITES_SPLIT_HPP_
#define KITES_SPLIT_HPP_

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "common.hpp"
#include "detail/step_kernel.hpp"
#include "device.hpp"
#include "graph_device.hpp"
#include "ingest.hpp"
#include "memory.hpp"

namespace kites {
namespace detail {

// key = (part of the source) << nb | destination; value = edge position.
static __global__ void split_keys_kernel(const unsigned* __restrict__ R, unsigned n,
                                         const unsigned* __restrict__ C, unsigned m, unsigned nb,
                                         unsigned* __restrict__ key, unsigned* __restrict__ val) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += stride) {
    // destination: the largest v with R[v] <= e (then R[v + 1] > e)
    unsigned lo = 0, hi = n - 1;
    while (lo < hi) {
      unsigned const mid = lo + (hi - lo + 1) / 2;
      if (R[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    key[e] = (split_part(C[e]) << nb) | lo;
    val[e] = e;
  }
}

template <class eT>
__global__ void split_gather_kernel(const unsigned* __restrict__ key, const unsigned* __restrict__ val,
                                    unsigned m, const unsigned* __restrict__ C,
                                    const eT* __restrict__ E, unsigned* __restrict__ xC,
                                    eT* __restrict__ xE, unsigned* __restrict__ head) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    unsigned const e = val[i];
    xC[i] = C[e];
    xE[i] = E[e];
    head[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
  }
}

static __global__ void split_segments_kernel(const unsigned* __restrict__ key,
                                             const unsigned* __restrict__ head,
                                             const unsigned* __restrict__ pos, unsigned m,
                                             unsigned nb, unsigned nseg,
                                             unsigned* __restrict__ xnzi,
                                             unsigned* __restrict__ xnzR,
                                             unsigned* __restrict__ segpart) {
  unsigned const stride = gridDim.x * blockDim.x;
  unsigned const mask = (1u << nb) - 1u;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    if (head[i]) {
      unsigned const s = pos[i];
      xnzi[s] = key[i] & mask;
      xnzR[s] = i;
      segpart[s] = key[i] >> nb;
    }
    if (i == m - 1) xnzR[nseg] = m;
  }
}

// xseg[b] = first segment whose part is >= b, b = 0..kSplitParts.
static __global__ void split_bounds_kernel(const unsigned* __restrict__ segpart, unsigned nseg,
                                           unsigned* __restrict__ xseg) {
  unsigned const b = threadIdx.x;
  if (b > kSplitParts) return;
  unsigned lo = 0, hi = nseg;
  while (lo < hi) {
    unsigned const mid = lo + (hi - lo) / 2;
    if (segpart[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  xseg[b] = lo;
}

// Builds (or keeps) the split of g's in-CSR.  Needs vertex-per-group 1 and
// at most 2^29 vertices (3 part bits + the destination in a 32-bit key).
template <class vT, class eT, class vIdxT>
void ensure_split(graph_csr_device<vT, eT, vIdxT>& g, stream& s) {
  std::size_t const n = g.get_num_vertices();
  std::size_t const m = g.get_num_edges();
  if (g.xbuilt && g.xm == m) return;
  if (n == 0 || n > (std::size_t{1} << 29))
    throw std::invalid_argument("kites: the split pull needs 1 to 2^29 vertices");
  int const did = g.getResidingDeviceID();
  g.xbuilt = false;
  g.xC.free();
  g.xE.free();
  g.xnzi.free();
  g.xnzR.free();
  for (unsigned b = 0; b <= kSplitParts; ++b) g.xseg[b] = 0;
  if (m == 0) {
    g.xC.alloc(1, did);
    g.xE.alloc(1, did);
    g.xnzi.alloc(1, did);
    g.xnzR.alloc(1, did);
    KITES_HIP_CHECK(hipMemsetAsync(g.xnzR.get_ptr(), 0, sizeof(unsigned), s.get()));
    s.sync();
    g.xm = 0;
    g.xbuilt = true;
    return;
  }
  unsigned const nb = id_bits(n);
  unsigned const mm = static_cast<unsigned>(m);
  device_buffer<unsigned> keyA(m, did), keyB(m, did), valA(m, did), valB(m, did);
  scratch_bytes tmp;
  hipLaunchKernelGGL(split_keys_kernel, dim3(ingest_grid(m)), dim3(kIngestBlock), 0, s.get(),
                     reinterpret_cast<const unsigned*>(g.R.get_ptr()), static_cast<unsigned>(n),
                     reinterpret_cast<const unsigned*>(g.C.get_ptr()), mm, nb, keyA.get_ptr(),
                     valA.get_ptr());
  KITES_HIP_CHECK(hipPeekAtLastError());
  std::size_t tb = 0;
  int const bits = static_cast<int>(nb + 3);
  KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keyA.get_ptr(), keyB.get_ptr(),
                                                     valA.get_ptr(), valB.get_ptr(), m, 0, bits,
                                                     s.get()));
  tmp.ensure(tb);
  KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keyA.get_ptr(), keyB.get_ptr(),
                                                     valA.get_ptr(), valB.get_ptr(), m, 0, bits,
                                                     s.get()));
  g.xC.alloc(m, did);
  g.xE.alloc(m, did);
  unsigned* const head = keyA.get_ptr();
  unsigned* const pos = valA.get_ptr();
  hipLaunchKernelGGL(split_gather_kernel<eT>, dim3(ingest_grid(m)), dim3(kIngestBlock), 0,
                     s.get(), keyB.get_ptr(), valB.get_ptr(), mm,
                     reinterpret_cast<const unsigned*>(g.C.get_ptr()), g.E.get_ptr(),
                     reinterpret_cast<unsigned*>(g.xC.get_ptr()), g.xE.get_ptr(), head);
  KITES_HIP_CHECK(hipPeekAtLastError());
  KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, head, pos, m, s.get()));
  tmp.ensure(tb);
  KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, head, pos, m, s.get()));
  unsigned last[2] = {0, 0};
  KITES_HIP_CHECK(hipMemcpyAsync(&last[0], pos + (m - 1), 4, hipMemcpyDeviceToHost, s.get()));
  KITES_HIP_CHECK(hipMemcpyAsync(&last[1], head + (m - 1), 4, hipMemcpyDeviceToHost, s.get()));
  s.sync();
  unsigned const nseg = last[0] + last[1];
  g.xnzi.alloc(nseg, did);
  g.xnzR.alloc(nseg + 1, did);
  hipLaunchKernelGGL(split_segments_kernel, dim3(ingest_grid(m)), dim3(kIngestBlock), 0, s.get(),
                     keyB.get_ptr(), head, pos, mm, nb, nseg,
                     reinterpret_cast<unsigned*>(g.xnzi.get_ptr()),
                     reinterpret_cast<unsigned*>(g.xnzR.get_ptr()), valB.get_ptr());
  KITES_HIP_CHECK(hipPeekAtLastError());
  device_buffer<unsigned> bounds(kSplitParts + 1, did);
  hipLaunchKernelGGL(split_bounds_kernel, dim3(1), dim3(64), 0, s.get(), valB.get_ptr(), nseg,
                     bounds.get_ptr());
  KITES_HIP_CHECK(hipPeekAtLastError());
  KITES_HIP_CHECK(hipMemcpyAsync(g.xseg, bounds.get_ptr(), sizeof(unsigned) * (kSplitParts + 1),
                                 hipMemcpyDeviceToHost, s.get()));
  s.sync();
  g.xm = m;
  g.xbuilt = true;
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_SPLIT_HPP_
