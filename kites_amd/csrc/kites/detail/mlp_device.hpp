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

// LDS the wave-parallel inference needs: the staged weights, then the
// hidden units and the outputs (floats).
constexpr int MLP_SCRATCH = MLP_NPARAM + MLP_HID + MLP_OUT;

// Wave-parallel inference for the device scheduler (call with the 64 lanes
// of one wave; x is the same in every lane).  The weights are first staged
// in `scratch` (LDS, MLP_SCRATCH floats) with every lane's loads in flight
// at once -- one round trip instead of one per weight a lane reads -- then
// lane j computes hidden unit j and lane k output k from LDS.  Keeps the
// scheduler's register footprint small inside the iteration kernel.
// `fallback` (a delta variant 1..8) is kept unless the best prediction is
// lower than the fallback's by more than `margin`.
__device__ inline unsigned mlp_choose_wave(const float* w, const float* x, unsigned fallback,
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
