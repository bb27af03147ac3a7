// functor_sets.hpp -- the C ABI's fixed functor sets (KITES_ALGO_*), each
// written as the reference's four __host__ __device__ lambdas
// (README.md:16-23), shared by the C-ABI translation units.
#ifndef KITES_AMD_FUNCTOR_SETS_HPP_
#define KITES_AMD_FUNCTOR_SETS_HPP_

#include <stdexcept>

#include "../../include/kites_amd.h"

namespace kites_capi {

// Runs `body(initF, nbrCompF, redF, updF)` with the lambdas of the
// saturating SSSP set (KITES_ALGO_SSSP), written exactly as a user of the
// reference API writes them (README.md:16-23).
template <class Body>
void with_sssp(Body&& body) {
  using vT = unsigned;
  using eT = unsigned;
  auto initF = [] __host__ __device__(volatile vT& locV, vT& globV) { locV = globV; };
  auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
    vT const t = srcV + *connE;
    partialV = (srcV == 0xFFFFFFFFu || t < srcV) ? 0xFFFFFFFFu : t;
  };
  auto redMin = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
    if (lhsV > rhsV) lhsV = rhsV;
  };
  auto updLess = [] __host__ __device__(volatile vT& newV, vT oldV) { return newV < oldV; };
  body(initF, nbr, redMin, updLess);
}

// Runs `body(initF, nbrCompF, redF, updF)` with the lambdas of functor set
// `algo`.
template <class Body>
void with_functors(int algo, Body&& body) {
  using vT = unsigned;
  using eT = unsigned;
  auto initF = [] __host__ __device__(volatile vT& locV, vT& globV) { locV = globV; };
  auto redMin = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
    if (lhsV > rhsV) lhsV = rhsV;
  };
  auto updLess = [] __host__ __device__(volatile vT& newV, vT oldV) { return newV < oldV; };
  switch (algo) {
    case KITES_ALGO_SSSP:
      with_sssp(body);
      break;
    case KITES_ALGO_BFS: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT*) {
        partialV = (srcV == 0xFFFFFFFFu) ? 0xFFFFFFFFu : srcV + 1u;
      };
      body(initF, nbr, redMin, updLess);
      break;
    }
    case KITES_ALGO_CC: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT*) { partialV = srcV; };
      body(initF, nbr, redMin, updLess);
      break;
    }
    case KITES_ALGO_SSSP_WRAP: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
        partialV = srcV + (*connE);
      };
      body(initF, nbr, redMin, updLess);
      break;
    }
    case KITES_ALGO_WIDEST: {
      auto nbr = [] __host__ __device__(volatile vT& partialV, vT srcV, eT* connE) {
        vT const w = *connE;
        partialV = srcV < w ? srcV : w;
      };
      auto redMax = [] __host__ __device__(volatile vT& lhsV, volatile vT& rhsV) {
        if (lhsV < rhsV) lhsV = rhsV;
      };
      auto updMore = [] __host__ __device__(volatile vT& newV, vT oldV) { return newV > oldV; };
      body(initF, nbr, redMax, updMore);
      break;
    }
    default:
      throw std::invalid_argument("unknown functor set");
  }
}

}  // namespace kites_capi

#endif  // KITES_AMD_FUNCTOR_SETS_HPP_
