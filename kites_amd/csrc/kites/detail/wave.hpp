// wave.hpp -- wave64 primitives for gfx950.
//
// Replace the reference's 32-lane warp helpers
// (kites/device/utility_device_functions.cuh:30-149: shfl.up scan, bfe,
// %laneid, %lanemask_lt, 5-step LDS binary search) with CDNA forms: DPP row
// shifts and row broadcasts for scans and sums, mbcnt for lane ranks inside
// a ballot, 64-bit ballots, and a 6-step binary search over a wave's 64
// prefix entries in LDS.
//
// Collectives that read other lanes' registers (dpp_*, wave_*_scan,
// wave_sum, wave_max) must be called with all 64 lanes active.
#ifndef KITES_DETAIL_WAVE_HPP_
#define KITES_DETAIL_WAVE_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kites {
namespace detail {

__device__ __forceinline__ unsigned lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ unsigned long long ballot64(bool p) {
  return __ballot(p);
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ unsigned rank_in(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u));
}

__device__ __forceinline__ unsigned readlane(unsigned x, int lane) {
  return __builtin_amdgcn_readlane(x, lane);
}

__device__ __forceinline__ unsigned readfirstlane(unsigned x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// DPP controls (GFX9 encoding): row_shr:n = 0x110+n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143.  Lanes whose DPP source is invalid or masked read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ unsigned dpp(unsigned x) {
  return __builtin_amdgcn_update_dpp(0u, x, CTRL, ROW_MASK, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (u32).
__device__ __forceinline__ unsigned wave_incl_scan(unsigned x) {
  x += dpp<0x111>(x);        // row_shr:1
  x += dpp<0x112>(x);        // row_shr:2
  x += dpp<0x114>(x);        // row_shr:4
  x += dpp<0x118>(x);        // row_shr:8  -> inclusive within each row of 16
  x += dpp<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
  x += dpp<0x143, 0xc>(x);   // row_bcast:31 into rows 2 and 3
  return x;
}

__device__ __forceinline__ unsigned wave_sum(unsigned x) {
  return readlane(wave_incl_scan(x), 63);
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) x += __shfl_xor(x, s, 64);
  return x;
}

__device__ __forceinline__ unsigned wave_max(unsigned x) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    unsigned y = __shfl_xor(x, s, 64);
    x = x > y ? x : y;
  }
  return x;
}

// Largest j in [0, 64) with pre[j] <= e, for a non-decreasing pre[] with
// pre[0] <= e.
__device__ __forceinline__ unsigned owner_search(const unsigned* pre, unsigned e) {
  unsigned lo = 0;
#pragma unroll
  for (unsigned step = 32; step > 0; step >>= 1)
    if (pre[lo + step] <= e) lo += step;
  return lo;
}

// Orders LDS accesses of the lanes of one wave (no cross-wave effect).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Shuffle of an arbitrary trivially copyable value made of 32-bit words.
template <class T>
__device__ __forceinline__ T shfl_down_any(T v, unsigned delta) {
  static_assert(sizeof(T) % 4 == 0, "value type must be a multiple of 4 bytes");
  union U {
    T t;
    unsigned w[sizeof(T) / 4];
  } in, out;
  in.t = v;
#pragma unroll
  for (unsigned i = 0; i < sizeof(T) / 4; ++i) out.w[i] = __shfl_down(in.w[i], delta, 64);
  return out.t;
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_WAVE_HPP_
