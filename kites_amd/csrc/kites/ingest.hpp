// ingest.hpp -- a device graph built on the GPU from an edge list in HBM.
//
// Extension of the reference, which builds its graph on the host in one
// thread (kites/graph/graph_raw.h:58-190 and graph_csr.h:45-110; the
// startup bottleneck at R-MAT scale 22).  The arrays are the ones the host
// path produces for vertex-per-group 1:
//   * in-edge CSR R, C, E: in-list of v = the edges whose destination is v,
//     sorted by source id, duplicates kept in input order (graph_raw.h:102,
//     131-133; one of the orders its std::sort may produce, the one the
//     host build here produces);
//   * weighted out-edge CSR oR, oC, oE: out-list of u sorted by destination,
//     duplicates in in-list order (the host transpose);
//   * the non-empty vertex list (nzi, nzR) and the out-edge bitmap.
// With `dedup`, self-loops and every later copy of a (source, destination)
// pair are dropped (the generators' rule: the first copy keeps its weight).
// The vertex count is padded to a multiple of 128 (graph_raw.h:173-182).
//
// Two stable LSD radix sorts do the work: edges by the key (destination,
// source) with the input position as value, then the sorted edges by
// source.  Offsets are lower bounds over the sorted row ids; the non-empty
// list is a flag scan.  Sorting and scans use hipCUB (rocPRIM).
#ifndef KITES_INGEST_HPP_
#define KITES_INGEST_HPP_

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "common.hpp"
#include "device.hpp"
#include "graph_device.hpp"
#include "memory.hpp"

namespace kites {
namespace detail {

constexpr unsigned kIngestBlock = 256;

inline unsigned ingest_grid(std::size_t items) {
  std::size_t const g = (items + kIngestBlock - 1) / kIngestBlock;
  return static_cast<unsigned>(g < 65535u * 16u ? (g ? g : 1) : 65535u * 16u);
}

// key[k] = (dst << nb) | src, val[k] = k; bad[0] |= 1 for an endpoint >= n.
static __global__ void ingest_keys_kernel(const unsigned* __restrict__ src,
                                          const unsigned* __restrict__ dst, std::uint64_t m,
                                          unsigned n, unsigned nb, std::uint64_t* __restrict__ key,
                                          unsigned* __restrict__ val, unsigned* __restrict__ bad) {
  std::uint64_t const stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
  bool oob = false;
  for (std::uint64_t k = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < m;
       k += stride) {
    unsigned const s = src[k], d = dst[k];
    oob |= (s >= n) || (d >= n);
    key[k] = (static_cast<std::uint64_t>(d) << nb) | s;
    val[k] = static_cast<unsigned>(k);
  }
  if (oob) atomicOr(bad, 1u);
}

// keep[i] = 1 unless dedup drops edge i of the sorted list (a self-loop, or
// the same key as edge i - 1).
static __global__ void ingest_keep_kernel(const std::uint64_t* __restrict__ key, std::uint64_t m,
                                          unsigned nb, int dedup, unsigned* __restrict__ keep) {
  std::uint64_t const stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
  std::uint64_t const mask = (std::uint64_t{1} << nb) - 1;
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += stride) {
    std::uint64_t const k = key[i];
    bool drop = false;
    if (dedup) drop = ((k >> nb) == (k & mask)) || (i > 0 && key[i - 1] == k);
    keep[i] = drop ? 0u : 1u;
  }
}

// Kept edge i goes to position pos[i]: C (source), E (weight of its input
// edge), D (destination, the row id the offsets are searched in).
template <class eT>
__global__ void ingest_scatter_kernel(const std::uint64_t* __restrict__ key,
                                      const unsigned* __restrict__ val,
                                      const unsigned* __restrict__ keep,
                                      const unsigned* __restrict__ pos, std::uint64_t m,
                                      unsigned nb, const eT* __restrict__ w,
                                      unsigned* __restrict__ C, eT* __restrict__ E,
                                      unsigned* __restrict__ D) {
  std::uint64_t const stride = static_cast<std::uint64_t>(gridDim.x) * blockDim.x;
  std::uint64_t const mask = (std::uint64_t{1} << nb) - 1;
  for (std::uint64_t i = static_cast<std::uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
       i += stride) {
    if (!keep[i]) continue;
    unsigned const p = pos[i];
    std::uint64_t const k = key[i];
    C[p] = static_cast<unsigned>(k & mask);
    D[p] = static_cast<unsigned>(k >> nb);
    E[p] = w[val[i]];
  }
}

// off[v] = number of entries of the sorted `rows` below v, v = 0..n.
static __global__ void ingest_offsets_kernel(const unsigned* __restrict__ rows, unsigned m,
                                             unsigned n, unsigned* __restrict__ off) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += stride) {
    unsigned lo = 0, hi = m;
    while (lo < hi) {
      unsigned const mid = lo + (hi - lo) / 2;
      if (rows[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    off[v] = lo;
  }
}

// Out-CSR entry j = the in-CSR edge perm[j]: destination D, weight E.
template <class eT>
__global__ void ingest_out_kernel(const unsigned* __restrict__ perm, unsigned m,
                                  const unsigned* __restrict__ D, const eT* __restrict__ E,
                                  unsigned* __restrict__ oC, eT* __restrict__ oE) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned j = blockIdx.x * blockDim.x + threadIdx.x; j < m; j += stride) {
    unsigned const e = perm[j];
    oC[j] = D[e];
    oE[j] = E[e];
  }
}

static __global__ void ingest_iota_kernel(unsigned* __restrict__ a, unsigned m) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) a[i] = i;
}

static __global__ void ingest_nz_flags_kernel(const unsigned* __restrict__ R, unsigned n,
                                              unsigned* __restrict__ flag) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
    flag[v] = R[v + 1] > R[v] ? 1u : 0u;
}

// nzi[pos[v]] = v and nzR[pos[v]] = R[v] for every non-empty v; the last
// thread also writes nzR[nnz] = m.
static __global__ void ingest_nz_scatter_kernel(const unsigned* __restrict__ R, unsigned n,
                                                const unsigned* __restrict__ flag,
                                                const unsigned* __restrict__ pos,
                                                unsigned* __restrict__ nzi,
                                                unsigned* __restrict__ nzR) {
  unsigned const stride = gridDim.x * blockDim.x;
  for (unsigned v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    if (flag[v]) {
      nzi[pos[v]] = v;
      nzR[pos[v]] = R[v];
    }
    if (v == n - 1) nzR[pos[v] + flag[v]] = R[n];
  }
}

inline unsigned id_bits(std::size_t n) {
  unsigned b = 1;
  while (b < 32 && (std::size_t{1} << b) < n) ++b;
  return b;
}

// Device scratch that frees itself.
struct scratch_bytes {
  void* p = nullptr;
  std::size_t bytes = 0;
  void ensure(std::size_t b) {
    if (b <= bytes) return;
    if (p) KITES_HIP_CHECK(hipFree(p));
    p = nullptr;
    KITES_HIP_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
  ~scratch_bytes() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace detail

struct ingest_report {
  std::size_t n = 0;          // padded vertex count
  std::size_t m = 0;          // edges kept
  float build_ms = 0.f;       // device time from the edge list in HBM to the finished graph
};

// Builds the device graph from m edges (src[k] -> dst[k], weight w[k]) held in
// device memory on `dev`.  n: vertex ids are < n; the graph has n rounded up
// to a multiple of 128 vertices.  Throws std::invalid_argument for an
// endpoint >= n or more than 2^32 - 1 edges.
template <class vT, class eT, class vIdxT>
graph_csr_device<vT, eT, vIdxT> make_graph_csr_from_edges(gpu& dev, std::size_t n, std::size_t m,
                                                           const vIdxT* src, const vIdxT* dst,
                                                           const eT* w, bool dedup,
                                                           ingest_report* rep = nullptr) {
  using namespace detail;
  static_assert(sizeof(vIdxT) == 4, "32-bit vertex ids");
  if (m >= 0xffffffffull) throw std::invalid_argument("ingest: at most 2^32 - 2 edges");
  if (n == 0) throw std::invalid_argument("ingest: a graph needs at least one vertex");
  if (n > 0xffffffffull - 256) throw std::invalid_argument("ingest: too many vertices");
  dev.setAsActive();
  int const did = dev.getDevIdx();
  stream& s = dev.getSsEven();
  std::size_t const np = n % VERTEX_PADDING ? (n / VERTEX_PADDING + 1) * VERTEX_PADDING : n;
  unsigned const nb = id_bits(np);
  hipEvent_t e0, e1;
  KITES_HIP_CHECK(hipEventCreate(&e0));
  KITES_HIP_CHECK(hipEventCreate(&e1));
  KITES_HIP_CHECK(hipEventRecord(e0, s.get()));

  device_buffer<std::uint64_t> keyA(m ? m : 1, did), keyB(m ? m : 1, did);
  device_buffer<unsigned> valA(m ? m : 1, did), valB(m ? m : 1, did);
  device_buffer<unsigned> bad(1, did);
  KITES_HIP_CHECK(hipMemsetAsync(bad.get_ptr(), 0, 4, s.get()));
  scratch_bytes tmp;
  std::size_t mk = 0;   // edges kept
  device_buffer<unsigned> D;
  graph_csr_device<vT, eT, vIdxT> g(np, dev);
  if (m) {
    hipLaunchKernelGGL(ingest_keys_kernel, dim3(ingest_grid(m)), dim3(kIngestBlock), 0, s.get(),
                       reinterpret_cast<const unsigned*>(src),
                       reinterpret_cast<const unsigned*>(dst), static_cast<std::uint64_t>(m),
                       static_cast<unsigned>(n), nb, keyA.get_ptr(), valA.get_ptr(), bad.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
    unsigned hbad = 0;
    KITES_HIP_CHECK(hipMemcpyAsync(&hbad, bad.get_ptr(), 4, hipMemcpyDeviceToHost, s.get()));
    s.sync();
    if (hbad) throw std::invalid_argument("ingest: edge endpoint out of range");
    // 1. edges by (destination, source), stable: duplicates keep input order
    std::size_t tb = 0;
    KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
        nullptr, tb, keyA.get_ptr(), keyB.get_ptr(), valA.get_ptr(), valB.get_ptr(), m, 0,
        static_cast<int>(2 * nb), s.get()));
    tmp.ensure(tb);
    KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
        tmp.p, tb, keyA.get_ptr(), keyB.get_ptr(), valA.get_ptr(), valB.get_ptr(), m, 0,
        static_cast<int>(2 * nb), s.get()));
    // 2. kept edges and their positions (keep flags in valA, positions in keyA)
    unsigned* const keep = valA.get_ptr();
    unsigned* const pos = reinterpret_cast<unsigned*>(keyA.get_ptr());
    hipLaunchKernelGGL(ingest_keep_kernel, dim3(ingest_grid(m)), dim3(kIngestBlock), 0, s.get(),
                       keyB.get_ptr(), static_cast<std::uint64_t>(m), nb, dedup ? 1 : 0, keep);
    KITES_HIP_CHECK(hipPeekAtLastError());
    KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, keep, pos, m, s.get()));
    tmp.ensure(tb);
    KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, keep, pos, m, s.get()));
    unsigned last[2] = {0, 0};
    KITES_HIP_CHECK(hipMemcpyAsync(&last[0], pos + (m - 1), 4, hipMemcpyDeviceToHost, s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(&last[1], keep + (m - 1), 4, hipMemcpyDeviceToHost, s.get()));
    s.sync();
    mk = static_cast<std::size_t>(last[0]) + last[1];
    g.C.alloc(mk ? mk : 1, did);
    g.E.alloc(mk ? mk : 1, did);
    D.alloc(mk ? mk : 1, did);
    hipLaunchKernelGGL(ingest_scatter_kernel<eT>, dim3(ingest_grid(m)), dim3(kIngestBlock), 0,
                       s.get(), keyB.get_ptr(), valB.get_ptr(), keep, pos,
                       static_cast<std::uint64_t>(m), nb, w,
                       reinterpret_cast<unsigned*>(g.C.get_ptr()), g.E.get_ptr(), D.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
  } else {
    g.C.alloc(1, did);
    g.E.alloc(1, did);
  }
  unsigned const m32 = static_cast<unsigned>(mk);
  hipLaunchKernelGGL(ingest_offsets_kernel, dim3(ingest_grid(np + 1)), dim3(kIngestBlock), 0,
                     s.get(), D.size() ? D.get_ptr() : nullptr, m32, static_cast<unsigned>(np),
                     reinterpret_cast<unsigned*>(g.R.get_ptr()));
  KITES_HIP_CHECK(hipPeekAtLastError());
  // 3. out-CSR: the in-CSR order sorted by source, stable
  g.oC.alloc(mk ? mk : 1, did);
  g.oE.alloc(mk ? mk : 1, did);
  if (mk) {
    unsigned* const skey = valA.get_ptr();               // sorted sources
    unsigned* const perm = reinterpret_cast<unsigned*>(keyA.get_ptr());
    hipLaunchKernelGGL(ingest_iota_kernel, dim3(ingest_grid(mk)), dim3(kIngestBlock), 0, s.get(),
                       valB.get_ptr(), m32);
    KITES_HIP_CHECK(hipPeekAtLastError());
    std::size_t tb = 0;
    const unsigned* const ckeys = reinterpret_cast<const unsigned*>(g.C.get_ptr());
    KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, ckeys, skey, valB.get_ptr(),
                                                       perm, mk, 0, static_cast<int>(nb),
                                                       s.get()));
    tmp.ensure(tb);
    KITES_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, ckeys, skey, valB.get_ptr(),
                                                       perm, mk, 0, static_cast<int>(nb),
                                                       s.get()));
    hipLaunchKernelGGL(ingest_out_kernel<eT>, dim3(ingest_grid(mk)), dim3(kIngestBlock), 0,
                       s.get(), perm, m32, D.get_ptr(), g.E.get_ptr(),
                       reinterpret_cast<unsigned*>(g.oC.get_ptr()), g.oE.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
    hipLaunchKernelGGL(ingest_offsets_kernel, dim3(ingest_grid(np + 1)), dim3(kIngestBlock), 0,
                       s.get(), skey, m32, static_cast<unsigned>(np),
                       reinterpret_cast<unsigned*>(g.oR.get_ptr()));
    KITES_HIP_CHECK(hipPeekAtLastError());
  } else {
    hipLaunchKernelGGL(ingest_offsets_kernel, dim3(ingest_grid(np + 1)), dim3(kIngestBlock), 0,
                       s.get(), nullptr, 0u, static_cast<unsigned>(np),
                       reinterpret_cast<unsigned*>(g.oR.get_ptr()));
    KITES_HIP_CHECK(hipPeekAtLastError());
  }
  // 4. vertices with in-edges
  {
    device_buffer<unsigned> flag(np, did), fpos(np, did);
    const unsigned* const R = reinterpret_cast<const unsigned*>(g.R.get_ptr());
    hipLaunchKernelGGL(ingest_nz_flags_kernel, dim3(ingest_grid(np)), dim3(kIngestBlock), 0,
                       s.get(), R, static_cast<unsigned>(np), flag.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
    std::size_t tb = 0;
    KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, flag.get_ptr(), fpos.get_ptr(),
                                                     np, s.get()));
    tmp.ensure(tb);
    KITES_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, flag.get_ptr(), fpos.get_ptr(),
                                                     np, s.get()));
    unsigned last[2] = {0, 0};
    KITES_HIP_CHECK(hipMemcpyAsync(&last[0], fpos.get_ptr() + (np - 1), 4,
                                   hipMemcpyDeviceToHost, s.get()));
    KITES_HIP_CHECK(hipMemcpyAsync(&last[1], flag.get_ptr() + (np - 1), 4,
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    unsigned const nnz = last[0] + last[1];
    g.nzi.alloc(nnz ? nnz : 1, did);
    g.nzR.alloc(static_cast<std::size_t>(nnz) + 1, did);
    hipLaunchKernelGGL(ingest_nz_scatter_kernel, dim3(ingest_grid(np)), dim3(kIngestBlock), 0,
                       s.get(), R, static_cast<unsigned>(np), flag.get_ptr(), fpos.get_ptr(),
                       reinterpret_cast<unsigned*>(g.nzi.get_ptr()),
                       reinterpret_cast<unsigned*>(g.nzR.get_ptr()));
    KITES_HIP_CHECK(hipPeekAtLastError());
    g.nnz = nnz;
    s.sync();
  }
  g.set_num_edges(mk);
  g.build_outbits(np, s);
  KITES_HIP_CHECK(hipEventRecord(e1, s.get()));
  s.sync();
  float ms = 0.f;
  KITES_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rep) {
    rep->n = np;
    rep->m = mk;
    rep->build_ms = ms;
  }
  return g;
}

}  // namespace kites

#endif  // KITES_INGEST_HPP_
