// mlp_device.hpp -- the scheduler's network: features and inference.
//
// A 2-layer perceptron (MLP_IN inputs -> MLP_HID ReLU units -> MLP_OUT
// linear outputs), FP32.  Output k is the predicted log2 of the duration of
// delta variant k+1 (PUSH_T, PUSH_W, PUSH_B, PUSH_LB, PULL_T, PULL_W, PULL_B,
// PULL_LB) for the coming iteration; the scheduler runs the variant with the smallest
// prediction.  Weights are one flat float array (layout below) written by
// the training kernels or loaded from a file.
#ifndef KITES_DETAIL_MLP_DEVICE_HPP_
#define KITES_DETAIL_MLP_DEVICE_HPP_

#include <hip/hip_runtime.h>

namespace kites {
namespace detail {

constexpr int MLP_IN = 10;
constexpr int MLP_HID = 32;
constexpr int MLP_OUT = 8;
// layout: W1[HID][IN] | b1[HID] | W2[OUT][HID] | b2[OUT]
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_HID * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_OUT * MLP_HID;
constexpr int MLP_NPARAM = MLP_OFF_B2 + MLP_OUT;

// Features of the coming iteration, computed from the statistics of the
// iteration that just ended (its updated set is the next frontier).
__host__ __device__ inline void make_features(float* f, unsigned n, unsigned long long m,
                                              unsigned count, unsigned long long sum_outdeg,
                                              unsigned max_outdeg, unsigned iter, bool settled) {
  float const fn = n ? static_cast<float>(n) : 1.f;
  float const fm = m ? static_cast<float>(m) : 1.f;
  float const fc = static_cast<float>(count);
  float const fs = static_cast<float>(sum_outdeg);
  f[0] = log2f(1.f + fc) / 24.f;
  f[1] = log2f(1.f + fs) / 32.f;
  f[2] = log2f(1.f + static_cast<float>(max_outdeg)) / 24.f;
  f[3] = fc / fn;
  f[4] = fs / fm;
  f[5] = log2f(1.f + (count ? fs / fc : 0.f)) / 16.f;
  f[6] = log2f(1.f + static_cast<float>(iter)) / 16.f;
  f[7] = log2f(1.f + fm / fn) / 8.f;
  f[8] = log2f(fn) / 24.f;
  f[9] = settled ? 1.f : 0.f;   // the run's values settle (options::settled_below)
}

// Forward pass for one sample; out[MLP_OUT].
__host__ __device__ inline void mlp_forward1(const float* w, const float* x, float* out) {
  float h[MLP_HID];
  for (int j = 0; j < MLP_HID; ++j) {
    float a = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_IN; ++i) a += w[MLP_OFF_W1 + j * MLP_IN + i] * x[i];
    h[j] = a > 0.f ? a : 0.f;
  }
  for (int k = 0; k < MLP_OUT; ++k) {
    float a = w[MLP_OFF_B2 + k];
    for (int j = 0; j < MLP_HID; ++j) a += w[MLP_OFF_W2 + k * MLP_HID + j] * h[j];
    out[k] = a;
  }
}

// Variant with the smallest predicted time; `fallback` if the network is
// absent or produces a non-finite value.
__host__ __device__ inline unsigned mlp_choose(const float* w, const float* x, unsigned fallback) {
  if (!w) return fallback;
  float out[MLP_OUT];
  mlp_forward1(w, x, out);
  int best = 0;
  for (int k = 0; k < MLP_OUT; ++k) {
    if (!(out[k] == out[k]) || out[k] > 1e30f || out[k] < -1e30f) return fallback;
    if (out[k] < out[best]) best = k;
  }
  return static_cast<unsigned>(best + 1);
}

// One lane's share of the network for the wave-parallel inference: lane l
// computes hidden unit j = l mod 32 (both halves of the wave hold the same
// unit) and the products of that unit with outputs 4h..4h+3 of its half
// h = l / 32.
struct mlp_lane_weights {
  float w1[MLP_IN];
  float b1;
  float w2[4];
  float b2[4];
};
static_assert(MLP_HID == 32 && MLP_OUT == 8, "the lane mapping assumes 32 hidden units, 8 outputs");

// Issues the lane's weight loads (19 independent loads, all in flight
// together; nothing waits for them here), so a caller can start them beside
// its other loads and take one round trip for all.  Call with the 64 lanes
// of one wave; `w` must not be null.
__device__ inline void mlp_load_lane(const float* w, mlp_lane_weights& r) {
  unsigned const lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  unsigned const j = lane & 31u, h = lane >> 5;
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) r.w1[i] = w[MLP_OFF_W1 + j * MLP_IN + i];
  r.b1 = w[MLP_OFF_B1 + j];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    r.w2[t] = w[MLP_OFF_W2 + (4 * h + t) * MLP_HID + j];
    r.b2[t] = w[MLP_OFF_B2 + 4 * h + t];
  }
}

// Hidden unit (lane mod 32) of the wave-parallel inference.
__device__ inline float mlp_lane_hidden(float const* w1, float b1, const float* x) {
  float a = b1;
#pragma unroll
  for (int i = 0; i < MLP_IN; ++i) a += w1[i] * x[i];
  return a > 0.f ? a : 0.f;
}

// Outputs from the lanes' hidden units: each lane's 4 products, a 5-step
// butterfly within each 32-lane half, then outputs 0-3 from lane 0 and 4-7
// from lane 32; the chosen variant (the same in every lane).
__device__ inline unsigned mlp_lane_output(float hu, float const* w2, float const* b2,
                                           unsigned fallback, float margin) {
  float p[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) p[t] = w2[t] * hu;
#pragma unroll
  for (int off = 16; off >= 1; off >>= 1) {
#pragma unroll
    for (int t = 0; t < 4; ++t) p[t] += __shfl_xor(p[t], off);
  }
  float out[MLP_OUT];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    float const o = p[t] + b2[t];
    out[t] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(o), 0));
    out[4 + t] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(o), 32));
  }
  // Straight-line argmin (no exit inside the loop, so it unrolls and the
  // outputs stay in registers).
  bool bad = false;
  int best = 0;
  float bv = out[0];
  float of = out[0];
#pragma unroll
  for (int k = 0; k < MLP_OUT; ++k) {
    float const o = out[k];
    bad = bad || !(o == o) || o > 1e30f || o < -1e30f;
    if (o < bv) {
      bv = o;
      best = k;
    }
    if (static_cast<unsigned>(k + 1) == fallback) of = o;
  }
  if (bad) return fallback;
  if (fallback >= 1 && fallback <= static_cast<unsigned>(MLP_OUT) && !(of - bv > margin))
    return fallback;
  return static_cast<unsigned>(best + 1);
}

// Wave-parallel inference for the device scheduler from weights loaded by
// mlp_load_lane (call with the same 64 lanes; x is the same in every lane):
// registers and cross-lane sums only.  `fallback` (a delta variant 1..8) is
// kept unless the best prediction is lower than the fallback's by more than
// `margin`; it is also returned if the network is absent or produces a
// non-finite value.
__device__ inline unsigned mlp_choose_lane(const float* w, mlp_lane_weights const& r,
                                           const float* x, unsigned fallback,
                                           float margin = 0.f) {
  if (!w) return fallback;
  return mlp_lane_output(mlp_lane_hidden(r.w1, r.b1, x), r.w2, r.b2, fallback, margin);
}


// LDS the wave-parallel inference needs: the staged weights, then the
// hidden units and the outputs (floats).
constexpr int MLP_SCRATCH = MLP_NPARAM + MLP_HID + MLP_OUT;

// The same inference through LDS, for callers with little register room
// (the one-workgroup loop): the weights are first staged in `scratch` (LDS,
// MLP_SCRATCH floats) with every lane's loads in flight at once, then lane
// j computes hidden unit j and lane k output k from LDS, one product at a
// time.
// `fallback` (a delta variant 1..8) is kept unless the best prediction is
// lower than the fallback's by more than `margin`.
__device__ inline unsigned mlp_choose_lds(const float* w, const float* x, unsigned fallback,
                                          float* scratch, float margin = 0.f) {
  if (!w) return fallback;
  unsigned const lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  {
    constexpr int kRounds = (MLP_NPARAM + 63) / 64;
    float t[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      int const i = r * 64 + static_cast<int>(lane);
      t[r] = i < MLP_NPARAM ? w[i] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      int const i = r * 64 + static_cast<int>(lane);
      if (i < MLP_NPARAM) scratch[i] = t[r];
    }
  }
  float* const hid = scratch + MLP_NPARAM;
  float* const out = hid + MLP_HID;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < static_cast<unsigned>(MLP_HID)) {
    float a = scratch[MLP_OFF_B1 + lane];
#pragma unroll 1
    for (int i = 0; i < MLP_IN; ++i) a += scratch[MLP_OFF_W1 + lane * MLP_IN + i] * x[i];
    hid[lane] = a > 0.f ? a : 0.f;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < static_cast<unsigned>(MLP_OUT)) {
    float a = scratch[MLP_OFF_B2 + lane];
#pragma unroll 1
    for (int j = 0; j < MLP_HID; ++j) a += scratch[MLP_OFF_W2 + lane * MLP_HID + j] * hid[j];
    out[lane] = a;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  int best = 0;
  float bv = out[0];
#pragma unroll 1
  for (int k = 0; k < MLP_OUT; ++k) {
    float const o = out[k];
    if (!(o == o) || o > 1e30f || o < -1e30f) return fallback;
    if (o < bv) {
      bv = o;
      best = k;
    }
  }
  if (fallback >= 1 && fallback <= static_cast<unsigned>(MLP_OUT) &&
      !(out[fallback - 1] - bv > margin))
    return fallback;
  return static_cast<unsigned>(best + 1);
}

}  // namespace detail
}  // namespace kites

#endif  // KITES_DETAIL_MLP_DEVICE_HPP_
