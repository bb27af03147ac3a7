// batch.hpp -- many sources in one run (extension of the reference API).
//
// kites::process_batch runs up to 64 independent single-source problems of
// one min-semilattice functor set (red = min, upd = "<", as SSSP and BFS)
// on one device graph together.  Result rows: rows[v * 64 + k] is the value
// of vertex v for source k (256 B per row for 4-byte values), so every
// neighbour value a wave gathers is one coalesced row access.  The kernel
// and its data layout are described in detail/batch_kernel.hpp.
//
// The reference has no batch entry point; its single-source process()
// (kites/device/compute.cuh:29-137) is what each column reproduces: column
// k equals a single-source run from sources[k].
#ifndef KITES_BATCH_HPP_
#define KITES_BATCH_HPP_

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <stdexcept>
#include <vector>

#include "detail/batch_kernel.hpp"
#include "engine.hpp"

namespace kites {

struct batch_report {
  unsigned iterations = 0;
  float ms = 0.f;   // device time of init + iterations
  unsigned graph_builds = 0;  // iteration-chunk graphs captured by this graph's batch runs so far
};

struct batch_options {
  unsigned edges_per_item = 256;   // 64 * items-per-lane: 128, 256 or 512 (bit-parallel
                                   // BFS: any multiple of 64 up to 4096)
  unsigned block = 512;            // threads per workgroup: 256 or 512 (bit-parallel BFS:
                                   // also 128)
  unsigned unroll = 8;             // neighbour rows in flight per wave: 8 or 16
  unsigned grid = 0;               // workgroups per launch, 0 = occupancy-derived
  bool use_graph = true;           // replay iteration chunks from a captured hipGraph
  unsigned chunk = 8;
  unsigned max_iterations = 1u << 16;
  bool keep_stats = false;
  // K-wide runs: besides the one-bit frontier, a 64-bit mask per vertex of
  // the sources whose value moved (detail/batch_kernel.hpp); same results.
  bool per_source_masks = false;
  // Bit-parallel BFS: a level whose frontier's out-edges are below
  // m / top_down runs top-down (detail/bitbfs_kernel.hpp); 0: every level
  // pulls.  (K-wide runs always pull.)
  unsigned top_down = detail::kBitPushDiv;
  batch_report* report = nullptr;
  std::vector<detail::batch_stats>* stats = nullptr;
};

namespace detail {

template <int BLOCK, int UNROLL, int IPL, bool PS, class P>
unsigned batch_grid(gpu& dev, unsigned requested) {
  if (requested) return requested;
  int per_cu = 0;
  KITES_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(&batch_step_kernel<BLOCK, UNROLL, IPL, PS, P>), BLOCK,
      0));
  per_cu = std::max(1, std::min(per_cu, 2048 / BLOCK));
  return static_cast<unsigned>(dev.numComputeUnits() * per_cu);
}

template <int BLOCK, int UNROLL, int IPL, bool PS, class P>
unsigned batch_loop(P const& prm, gpu& dev, stream& s, batch_options const& o,
                    batch_workspace& ws) {
  auto kern = &batch_step_kernel<BLOCK, UNROLL, IPL, PS, P>;
  unsigned const grid = batch_grid<BLOCK, UNROLL, IPL, PS, P>(dev, o.grid);
  unsigned const chunk = std::max(1u, o.chunk);
  hipGraphExec_t const exec =
      o.use_graph ? ws.graphs.chunk(kern, grid, BLOCK, chunk, prm, s.get()) : nullptr;
  unsigned launched = 0;
  for (;;) {
    if (exec) {
      KITES_HIP_CHECK(hipGraphLaunch(exec, s.get()));
    } else {
      for (unsigned k = 0; k < chunk; ++k)
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, s.get(), prm);
      KITES_HIP_CHECK(hipPeekAtLastError());
    }
    launched += chunk;
    KITES_HIP_CHECK(hipMemcpyAsync(ws.ctl_host.get_ptr(), prm.ctl, offsetof(batch_ctl, board),
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    if (ws.ctl_host[0].done || launched > ws.max_iter + chunk) break;
  }
  return launched;
}

template <int BLOCK, int UNROLL, class P>
void batch_dispatch_ipl(P const& prm, gpu& dev, stream& s, batch_options const& o,
                        batch_workspace& ws) {
  if (o.per_source_masks) {
    // the per-source-mask variant is built for the default shape only
    if constexpr (BLOCK == 512 && UNROLL == 8) {
      if (o.edges_per_item == 256) {
        batch_loop<BLOCK, UNROLL, 4, true>(prm, dev, s, o, ws);
        return;
      }
    }
    throw std::invalid_argument(
        "process_batch: per-source masks run at the default shape (256 edges per item, block "
        "512, unroll 8)");
  }
  switch (o.edges_per_item) {
    case 128: batch_loop<BLOCK, UNROLL, 2, false>(prm, dev, s, o, ws); break;
    case 256: batch_loop<BLOCK, UNROLL, 4, false>(prm, dev, s, o, ws); break;
    default: batch_loop<BLOCK, UNROLL, 8, false>(prm, dev, s, o, ws); break;
  }
}

template <int BLOCK, class P>
void batch_dispatch_unroll(P const& prm, gpu& dev, stream& s, batch_options const& o,
                           batch_workspace& ws) {
  switch (o.unroll) {
    case 8: batch_dispatch_ipl<BLOCK, 8>(prm, dev, s, o, ws); break;
    case 16: batch_dispatch_ipl<BLOCK, 16>(prm, dev, s, o, ws); break;
    default: throw std::invalid_argument("process_batch: unroll must be 8 or 16");
  }
}

}  // namespace detail

// Runs sources.size() (<= 64) single-source fixpoints together.  Every
// vertex starts at `other_value`, source k's own vertex at `source_value`
// in column k, and only the sources are active at the start (the frontier
// start of the reference example, kites/graph/graph_raw.h:103-110).  `rows`
// is resized to n * 64 values when smaller.
template <launch launchMode, class vT, class eT, class vIdxT, class funcInitT,
          class funcCompNbrT, class funcRedT, class funcUpdT>
void process_batch(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev,
                   std::vector<vIdxT> const& sources, vT source_value, vT other_value,
                   device_buffer<vT>& rows, funcInitT funcInit, funcCompNbrT funcCompNbr,
                   funcRedT funcRed, funcUpdT funcUpd, batch_options const& o = batch_options()) {
  using namespace detail;
  static_assert(sizeof(vIdxT) == 4, "process_batch: 32-bit vertex ids");
  dev.setAsActive();
  stream& s = dev.getSsEven();
  std::size_t const n = grph.get_num_vertices();
  std::size_t const m = grph.get_num_edges();
  unsigned const K = static_cast<unsigned>(sources.size());
  if (K > kBatchLanes) throw std::invalid_argument("process_batch: at most 64 sources");
  for (auto v : sources)
    if (static_cast<std::size_t>(v) >= n)
      throw std::invalid_argument("process_batch: source out of range");
  probe_result const pr = probe_functors<vT, eT>(funcInit, funcCompNbr, funcRed, funcUpd);
  if (!pr.min_fast)
    throw std::invalid_argument(
        "process_batch: the functor set must be a min semilattice (red = min, upd = <)");
  if (o.block != 256 && o.block != 512)
    throw std::invalid_argument("process_batch: block must be 256 or 512");
  if (o.edges_per_item != 128 && o.edges_per_item != 256 && o.edges_per_item != 512)
    throw std::invalid_argument("process_batch: edges_per_item must be 128, 256 or 512");
  if (m >= (1ull << 32) - 4096) throw std::invalid_argument("process_batch: too many edges");
  if (grph.nnz == 0 && m != 0) throw std::runtime_error("process_batch: graph has no in-edge list");

  if (!grph.bws) grph.bws.reset(new batch_workspace());
  batch_workspace& ws = *grph.bws;
  int const did = grph.getResidingDeviceID();
  ws.dev = did;
  unsigned const aw = static_cast<unsigned>(std::max<std::size_t>(n, 1) / 64 + 1);
  if (ws.n < n || ws.F.size() < 3ull * aw) {
    ws.F.free();
    ws.F.alloc(3ull * aw, did);
    ws.M.free();
    ws.n = n;
  }
  if (o.per_source_masks && ws.M.size() < 3ull * std::max<std::size_t>(n, 1))
    ws.M.alloc(3ull * std::max<std::size_t>(n, 1), did);
  // chunk start positions: one per 64 in-edges, shared by every item size
  unsigned const epi = o.edges_per_item;
  unsigned const nitems = static_cast<unsigned>((m + epi - 1) / epi);
  unsigned const nchunks = static_cast<unsigned>((m + 63) / 64);
  if (ws.cstart.size() < std::max(1u, nchunks) || ws.cstart_m != m) {
    ws.cstart.free();
    ws.cstart.alloc(std::max(1u, nchunks), did);
    if (nchunks)
      hipLaunchKernelGGL(batch_istart_kernel, dim3((nchunks + 255) / 256), dim3(256), 0, s.get(),
                         grph.nzR.get_ptr(), grph.nnz, 64u, nchunks, ws.cstart.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
    ws.cstart_m = m;
  }
  unsigned const iters = std::max(2u, o.max_iterations);
  if (ws.stats.size() < iters) {
    ws.stats.free();
    ws.stats.alloc(iters, did);
    ws.max_iter = std::max(ws.max_iter, iters);
  }
  if (ws.ctl.size() == 0) {
    ws.ctl.alloc(1, did);
    ws.ctl_host.alloc(1);
    ws.src.alloc(kBatchLanes, did);
    KITES_HIP_CHECK(hipEventCreate(&ws.ev0));
    KITES_HIP_CHECK(hipEventCreate(&ws.ev1));
  }
  if (rows.size() < n * kBatchLanes) {
    rows.free();
    rows.alloc(std::max<std::size_t>(n, 1) * kBatchLanes, did);
  }
  std::vector<unsigned> hs(kBatchLanes, 0u);
  for (unsigned k = 0; k < K; ++k) hs[k] = static_cast<unsigned>(sources[k]);
  KITES_HIP_CHECK(hipMemcpyAsync(ws.src.get_ptr(), hs.data(), kBatchLanes * 4,
                                 hipMemcpyHostToDevice, s.get()));
  KITES_HIP_CHECK(hipMemsetAsync(ws.stats.get_ptr(), 0, iters * sizeof(batch_stats), s.get()));

  using P = batch_params<vT, eT, funcCompNbrT>;
  param_block<P> pb(funcCompNbr);
  P& prm = pb.get();
  prm.n = static_cast<unsigned>(n);
  prm.m = static_cast<unsigned>(m);
  prm.C = grph.C.get_ptr();
  prm.E = grph.E.get_ptr();
  prm.nzi = grph.nzi.get_ptr();
  prm.nzR = grph.nzR.get_ptr();
  prm.nnz = grph.nnz;
  prm.istart = ws.cstart.get_ptr();
  prm.nitems = nitems;
  prm.epi = epi;
  prm.V = rows.get_ptr();
  prm.A = ws.F.get_ptr();
  prm.aw = aw;
  prm.M = o.per_source_masks ? ws.M.get_ptr() : nullptr;
  prm.ctl = ws.ctl.get_ptr();
  prm.stats = ws.stats.get_ptr();
  prm.collect = o.keep_stats || o.stats != nullptr;

  if (nitems && !prm.istart) throw std::logic_error("process_batch: no chunk starts");
  KITES_HIP_CHECK(hipEventRecord(ws.ev0, s.get()));
  {
    int const cus = dev.numComputeUnits();
    hipLaunchKernelGGL(batch_fill_kernel<vT>, dim3(cus * 8), dim3(256), 0, s.get(),
                       rows.get_ptr(), static_cast<unsigned long long>(n), other_value,
                       ws.F.get_ptr(), 3ull * aw);
    if (prm.M)
      hipLaunchKernelGGL(batch_fill_kernel<vT>, dim3(cus * 8), dim3(256), 0, s.get(),
                         rows.get_ptr(), 0ull, other_value, prm.M, 3ull * n);
    hipLaunchKernelGGL(batch_seed_kernel<vT>, dim3(1), dim3(kBatchLanes), 0, s.get(),
                       rows.get_ptr(), ws.F.get_ptr(), ws.src.get_ptr(), K, source_value,
                       ws.ctl.get_ptr(), iters, prm.M);
    KITES_HIP_CHECK(hipPeekAtLastError());
  }
  switch (o.block) {
    case 256: batch_dispatch_unroll<256>(prm, dev, s, o, ws); break;
    default: batch_dispatch_unroll<512>(prm, dev, s, o, ws); break;
  }
  KITES_HIP_CHECK(hipEventRecord(ws.ev1, s.get()));
  s.sync();
  float ms = 0.f;
  KITES_HIP_CHECK(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
  batch_ctl const ctl = ws.ctl_host[0];
  if (!ctl.done || ctl.overflow)
    throw std::runtime_error("kites::process_batch: iteration limit reached before convergence");
  if (o.report) {
    o.report->iterations = ctl.iter;
    o.report->ms = ms;
    o.report->graph_builds = ws.graphs.builds();
  }
  if (o.stats) {
    o.stats->resize(ctl.iter);
    if (ctl.iter)
      KITES_HIP_CHECK(hipMemcpy(o.stats->data(), ws.stats.get_ptr(),
                                ctl.iter * sizeof(batch_stats), hipMemcpyDeviceToHost));
  }
}

namespace detail {

template <int BLOCK>
unsigned bitbfs_loop(bitbfs_params const& prm, gpu& dev, stream& s, batch_options const& o,
                     batch_workspace& ws) {
  auto kern = &bitbfs_step_kernel<BLOCK>;
  unsigned grid = o.grid;
  if (!grid) {
    int per_cu = 0;
    KITES_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(kern), BLOCK, 0));
    per_cu = std::max(1, std::min(per_cu, 2048 / BLOCK));
    grid = static_cast<unsigned>(dev.numComputeUnits() * per_cu);
  }
  unsigned const chunk = std::max(1u, o.chunk);
  hipGraphExec_t const exec =
      o.use_graph ? ws.graphs.chunk(kern, grid, BLOCK, chunk, prm, s.get()) : nullptr;
  unsigned launched = 0;
  for (;;) {
    if (exec) {
      KITES_HIP_CHECK(hipGraphLaunch(exec, s.get()));
    } else {
      for (unsigned k = 0; k < chunk; ++k)
        hipLaunchKernelGGL(kern, dim3(grid), dim3(BLOCK), 0, s.get(), prm);
      KITES_HIP_CHECK(hipPeekAtLastError());
    }
    launched += chunk;
    KITES_HIP_CHECK(hipMemcpyAsync(ws.ctl_host.get_ptr(), prm.ctl, offsetof(batch_ctl, board),
                                   hipMemcpyDeviceToHost, s.get()));
    s.sync();
    if (ws.ctl_host[0].done || launched > ws.max_iter + chunk) break;
  }
  return launched;
}

}  // namespace detail

// Bit-parallel BFS from up to 256 sources (levels as u32, unreached =
// 0xFFFFFFFF): levels[v * 256 + k] is the BFS level of v from sources[k].
// Column k equals a single-source BFS run (the reference example's
// functors with unit weights).  `levels` is resized to n * 256 when smaller.
template <class vT, class eT, class vIdxT>
void process_bfs_bitparallel(graph_csr_device<vT, eT, vIdxT>& grph, gpu& dev,
                             std::vector<vIdxT> const& sources, device_buffer<unsigned>& levels,
                             batch_options const& o = batch_options(),
                             std::vector<detail::bitbfs_stats>* stats = nullptr) {
  using namespace detail;
  static_assert(sizeof(vIdxT) == 4, "process_bfs_bitparallel: 32-bit vertex ids");
  dev.setAsActive();
  stream& s = dev.getSsEven();
  std::size_t const n = grph.get_num_vertices();
  std::size_t const m = grph.get_num_edges();
  unsigned const K = static_cast<unsigned>(sources.size());
  if (K > kBitSources) throw std::invalid_argument("process_bfs_bitparallel: at most 256 sources");
  for (auto v : sources)
    if (static_cast<std::size_t>(v) >= n)
      throw std::invalid_argument("process_bfs_bitparallel: source out of range");
  if (o.block != 128 && o.block != 256 && o.block != 512)
    throw std::invalid_argument("process_bfs_bitparallel: block must be 128, 256 or 512");
  if (o.edges_per_item < 64 || (o.edges_per_item & 63) || o.edges_per_item > 4096)
    throw std::invalid_argument(
        "process_bfs_bitparallel: edges_per_item must be a multiple of 64 in [64, 4096]");
  if (m >= (1ull << 32) - 4096) throw std::invalid_argument("process_bfs_bitparallel: too many edges");
  if (grph.nnz == 0 && m != 0) throw std::runtime_error("process_bfs_bitparallel: no in-edge list");

  if (!grph.bws) grph.bws.reset(new batch_workspace());
  batch_workspace& ws = *grph.bws;
  int const did = grph.getResidingDeviceID();
  std::size_t const nn = std::max<std::size_t>(n, 1);
  unsigned const aw = static_cast<unsigned>(nn / 64 + 1);
  if (ws.bn < n || ws.VIS.size() == 0) {
    ws.VIS.free();
    ws.F4.free();
    ws.A.free();
    ws.L8w.free();
    ws.VIS.alloc(nn * kBitWords, did);
    ws.F4.alloc(3 * nn * kBitWords, did);
    ws.A.alloc(4ull * aw, did);   // 3 summary slots + the done bits
    ws.L8w.alloc(nn * (kBitSources / 4), did);
    ws.bn = n;
  }
  // chunk start positions: one per 64 in-edges, shared by every item size
  unsigned const epi = o.edges_per_item;
  unsigned const nitems = static_cast<unsigned>((m + epi - 1) / epi);
  unsigned const nchunks = static_cast<unsigned>((m + 63) / 64);
  if (ws.cstart.size() < std::max(1u, nchunks) || ws.cstart_m != m) {
    ws.cstart.free();
    ws.cstart.alloc(std::max(1u, nchunks), did);
    if (nchunks)
      hipLaunchKernelGGL(batch_istart_kernel, dim3((nchunks + 255) / 256), dim3(256), 0, s.get(),
                         grph.nzR.get_ptr(), grph.nnz, 64u, nchunks, ws.cstart.get_ptr());
    KITES_HIP_CHECK(hipPeekAtLastError());
    ws.cstart_m = m;
  }
  unsigned const iters = std::max(2u, o.max_iterations);
  if (ws.max_iter < iters || ws.bstats.size() < iters) {
    ws.bstats.free();
    ws.bstats.alloc(iters, did);
    ws.max_iter = std::max(ws.max_iter, iters);
  }
  if (ws.ctl.size() == 0) {
    ws.ctl.alloc(1, did);
    ws.ctl_host.alloc(1);
    ws.src.alloc(kBatchLanes, did);
    KITES_HIP_CHECK(hipEventCreate(&ws.ev0));
    KITES_HIP_CHECK(hipEventCreate(&ws.ev1));
  }
  if (ws.bsrc.size() == 0) ws.bsrc.alloc(kBitSources, did);
  if (levels.size() < nn * kBitSources) {
    levels.free();
    levels.alloc(nn * kBitSources, did);
  }
  std::vector<unsigned> hs(kBitSources, 0u);
  for (unsigned k = 0; k < K; ++k) hs[k] = static_cast<unsigned>(sources[k]);
  KITES_HIP_CHECK(hipMemcpyAsync(ws.bsrc.get_ptr(), hs.data(), kBitSources * 4,
                                 hipMemcpyHostToDevice, s.get()));
  KITES_HIP_CHECK(hipMemsetAsync(ws.bstats.get_ptr(), 0, iters * sizeof(bitbfs_stats), s.get()));

  bitbfs_params prm;
  zero_params(prm);
  prm.n = static_cast<unsigned>(n);
  prm.m = static_cast<unsigned>(m);
  prm.C = grph.C.get_ptr();
  prm.nzi = grph.nzi.get_ptr();
  prm.nzR = grph.nzR.get_ptr();
  prm.nnz = grph.nnz;
  prm.istart = ws.cstart.get_ptr();
  prm.nitems = nitems;
  prm.epi = epi;
  prm.VIS = ws.VIS.get_ptr();
  prm.F = ws.F4.get_ptr();
  prm.A = ws.A.get_ptr();
  prm.aw = aw;
  prm.DONE = ws.A.get_ptr() + 3ull * aw;
  for (unsigned w = 0; w < kBitWords; ++w) {
    unsigned const lo = w * 64u;
    prm.full[w] = K >= lo + 64u ? ~0ull : (K > lo ? (1ull << (K - lo)) - 1ull : 0ull);
  }
  prm.L = levels.get_ptr();
  prm.L8 = reinterpret_cast<unsigned char*>(ws.L8w.get_ptr());
  prm.ctl = ws.ctl.get_ptr();
  prm.stats = ws.bstats.get_ptr();
  prm.oR = grph.oR.get_ptr();
  prm.oC = grph.oC.get_ptr();
  prm.push_div = o.top_down;
  // the seed reads the sources' out-degrees whatever the direction rule;
  // top-down levels also walk the out-lists (an edgeless graph has none)
  if (!prm.oR || (prm.push_div && m && !prm.oC))
    throw std::logic_error("process_bfs_bitparallel: no out-edge CSR");

  if (nitems && !prm.istart) throw std::logic_error("process_bfs_bitparallel: no chunk starts");
  KITES_HIP_CHECK(hipEventRecord(ws.ev0, s.get()));
  int const cus = dev.numComputeUnits();
  // (the byte levels need no clearing: the expansion reads a pair's byte
  // only when its VIS bit says this run reached it)
  hipLaunchKernelGGL(bitbfs_init_kernel, dim3(cus * 8), dim3(256), 0, s.get(), ws.VIS.get_ptr(),
                     static_cast<unsigned long long>(n) * kBitWords, ws.F4.get_ptr(),
                     3ull * n * kBitWords, ws.A.get_ptr(), 4ull * aw);
  hipLaunchKernelGGL(bitbfs_seed_kernel, dim3(1), dim3(kBitSources), 0, s.get(), prm,
                     ws.bsrc.get_ptr(), K, iters);
  KITES_HIP_CHECK(hipPeekAtLastError());
  switch (o.block) {
    case 128: bitbfs_loop<128>(prm, dev, s, o, ws); break;
    case 256: bitbfs_loop<256>(prm, dev, s, o, ws); break;
    default: bitbfs_loop<512>(prm, dev, s, o, ws); break;
  }
  // the loop ended on a synchronized control block: its iteration count
  // says whether any level went past the byte range
  batch_ctl const ctl = ws.ctl_host[0];
  if (n) {
    unsigned const eb = static_cast<unsigned>(std::min<std::size_t>((n + 3) / 4, cus * 16));
    hipLaunchKernelGGL(bitbfs_expand_kernel, dim3(eb), dim3(256), 0, s.get(), levels.get_ptr(),
                       ws.L8w.get_ptr(), ws.VIS.get_ptr(), static_cast<unsigned>(n),
                       ctl.iter >= 255u ? 1 : 0);
    KITES_HIP_CHECK(hipPeekAtLastError());
  }
  KITES_HIP_CHECK(hipEventRecord(ws.ev1, s.get()));
  s.sync();
  float ms = 0.f;
  KITES_HIP_CHECK(hipEventElapsedTime(&ms, ws.ev0, ws.ev1));
  if (!ctl.done || ctl.overflow)
    throw std::runtime_error("kites::process_bfs_bitparallel: iteration limit reached before convergence");
  if (o.report) {
    o.report->iterations = ctl.iter;
    o.report->ms = ms;
    o.report->graph_builds = ws.graphs.builds();
  }
  if (stats) {
    stats->resize(ctl.iter);
    if (ctl.iter)
      KITES_HIP_CHECK(hipMemcpy(stats->data(), ws.bstats.get_ptr(), ctl.iter * sizeof(bitbfs_stats),
                                hipMemcpyDeviceToHost));
  }
}

}  // namespace kites

#endif  // KITES_BATCH_HPP_
