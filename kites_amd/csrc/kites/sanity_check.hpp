// sanity_check.hpp -- the reference's host-side self-check of a GPU run.
//
// sanity_check(hostGraph, devGraph, F...) runs the same four functors on the
// host as a single-thread in-place sweep over all vertices until no vertex
// changes, then compares every slot with the device result
// (kites/graph/sanity_check.cuh:13-68).  It prints as the reference does and
// also returns whether the results match.  It modifies the host graph's V,
// as the reference does.
#ifndef KITES_SANITY_CHECK_HPP_
#define KITES_SANITY_CHECK_HPP_

#include <iostream>

#include "graph.hpp"
#include "graph_device.hpp"

namespace kites {

template <class vT, class eT, class vIdxT, class funcInitT, class funcCompNbrT, class funcRedT,
          class funcUpdT>
bool sanity_check(graph_csr<vT, eT, vIdxT>& h, graph_csr_device<vT, eT, vIdxT>& d,
                  funcInitT funcInit, funcCompNbrT funcCompNbr, funcRedT funcRed,
                  funcUpdT funcUpd) {
  std::cout << "### SANITY CHECK ...\n";
  std::size_t const n = h.get_num_vertices();
  unsigned sweeps = 0;
  bool any;
  do {
    any = false;
    for (std::size_t v = 0; v < n; ++v) {
      vT pre = h.V[v];
      volatile vT inHand;
      funcInit(inHand, pre);
      for (vIdxT e = h.R[v]; e < h.R[v + 1]; ++e) {
        volatile vT tmp;
        funcCompNbr(tmp, h.V.at(h.C.at(e)), h.E.get_ptr() + e);
        funcRed(inHand, tmp);
      }
      if (funcUpd(inHand, pre)) {
        h.V[v] = inHand;
        any = true;
      }
    }
    ++sweeps;
  } while (any);
  std::cout << "### SANITY CHECK: Single CPU thread performed " << sweeps
            << " iterations in total.\n";
  host_pinned_buffer<vT> dv(d.V);
  for (std::size_t i = 0; i < dv.size(); ++i)
    if (!(dv[i] == h.V[i])) {
      std::cout << "### SANITY CHECK: MISMATCH at index " << i << "! Expected " << h.V[i]
                << "\tbut got " << dv[i] << "\n";
      return false;
    }
  std::cout << "### SANITY CHECK: PASSED!\n";
  return true;
}

}  // namespace kites

#endif  // KITES_SANITY_CHECK_HPP_
