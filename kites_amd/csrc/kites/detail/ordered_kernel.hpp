// ordered_kernel.hpp -- ordered (Jacobi) iterations for functor sets whose
// reduction is not idempotent (sums, PageRank-style updates).
//
// The reference runs every functor set in place with chaotic relaxation and
// combines a vertex's in-edges in an undefined tree order
// (kites/device/kernels_single_device.cuh:92-132), so a sum's result depends
// on timing and launch shape.  This mode defines both:
//   * Jacobi: iteration i reads Vin (the values after iteration i-1) and
//     writes Vout; vertices not recomputed keep their value.
//   * Combine order: for an active vertex v,
//       acc = initF(Vin[v]);
//       for each 64-edge chunk of v's in-list, in order:
//         x[l] = nbrCompF(Vin[C[e_l]], &E[e_l]) for the chunk's edges e_0..e_k-1;
//         a pairwise tree: for s = 1, 2, 4, ..., 32, x[l] = redF(x[l], x[l+s])
//           for every l divisible by 2s with l+s < k (lhs = the left operand);
//         acc = redF(acc, x[0]);
//       if updF(acc, Vin[v]): Vout[v] = acc and v's out-neighbours become
//       active for the next iteration.
//   Activeness follows the reference: the initial frontier, then the
//   out-neighbours of the vertices that changed.
// redF therefore needs neither associativity nor commutativity here, and the
// result is the same for every grid and run.
#ifndef KITES_DETAIL_ORDERED_KERNEL_HPP_
#define KITES_DETAIL_ORDERED_KERNEL_HPP_

#include <hip/hip_runtime.h>

#include <type_traits>

#include "wave.hpp"

namespace kites {
namespace detail {

template <class vT, class eT, class FI, class FN, class FR, class FU>
struct ordered_params {
  unsigned n;
  const unsigned* R;   // in-edge CSR
  const unsigned* C;
  eT* E;
  const unsigned* oR;  // out-lists (activation of out-neighbours)
  const unsigned* oC;
  const vT* Vin;
  vT* Vout;
  const unsigned long long* in_bm;   // active vertices of this iteration
  unsigned long long* out_bm;        // active vertices of the next one
  unsigned nwords;
  unsigned* flag;                    // set to 1 when a value changes
  FI fi;
  FN fn;
  FR fr;
  FU fu;
  ordered_params(FI a, FN b, FR r, FU u) : fi(a), fn(b), fr(r), fu(u) {}
};

template <class T>
__device__ __forceinline__ volatile T& ovs(T& x) {
  return *reinterpret_cast<volatile T*>(&x);
}

// One wave per active vertex: the waves walk the active bitmap a 64-bit word
// at a time and take the word's vertices in order.
template <int BLOCK, class P>
__global__ __launch_bounds__(BLOCK) void ordered_kernel(P p) {
  using vT = typename std::remove_const<typename std::remove_reference<decltype(*p.Vin)>::type>::type;
  constexpr int kW = BLOCK / 64;
  __shared__ vT slot[BLOCK];
  __shared__ vT acc[kW];
  unsigned const lane = lane_id(), wib = threadIdx.x >> 6, t = threadIdx.x;
  unsigned const nwaves = gridDim.x * kW;
  for (unsigned w = blockIdx.x * kW + wib; w < p.nwords; w += nwaves) {
    unsigned long long bits = p.in_bm[w];
    while (bits) {
      unsigned const v = w * 64u + static_cast<unsigned>(__builtin_ctzll(bits));
      bits &= bits - 1;
      if (v >= p.n) break;
      unsigned const beg = p.R[v], end = p.R[v + 1];
      vT const old = p.Vin[v];
      if (lane == 0) {
        vT g = old;
        p.fi(ovs(acc[wib]), g);
      }
      for (unsigned c0 = beg; c0 < end; c0 += 64) {
        unsigned const k = end - c0 < 64u ? end - c0 : 64u;
        if (lane < k) {
          unsigned const e = c0 + lane;
          p.fn(ovs(slot[t]), p.Vin[p.C[e]], p.E + e);
        }
        wave_lds_sync();
#pragma unroll
        for (unsigned s = 1; s < 64; s <<= 1) {
          if ((lane & (2 * s - 1)) == 0 && lane + s < k) p.fr(ovs(slot[t]), ovs(slot[t + s]));
          wave_lds_sync();
        }
        if (lane == 0) p.fr(ovs(acc[wib]), ovs(slot[t]));
        wave_lds_sync();
      }
      unsigned moved = 0;
      if (lane == 0) {
        if (p.fu(ovs(acc[wib]), old)) {
          p.Vout[v] = acc[wib];
          *p.flag = 1u;
          moved = 1;
        }
      }
      if (ballot64(moved != 0)) {
        unsigned const ob = p.oR[v], oe = p.oR[v + 1];
        for (unsigned e = ob + lane; e < oe; e += 64) {
          unsigned const u = p.oC[e];
          atomicOr(p.out_bm + (u >> 6), 1ull << (u & 63));
        }
      }
      wave_lds_sync();
    }
  }
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_ORDERED_KERNEL_HPP_
