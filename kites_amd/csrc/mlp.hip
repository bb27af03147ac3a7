// mlp.hip -- training and batch inference of the scheduler network on the
// GPU (declarations in include/kites_amd.h).
//
// Network: x (MLP_IN) -> h = relu(W1 x + b1) (MLP_HID) -> y = W2 h + b2
// (MLP_OUT), FP32, one flat parameter array (layout in
// kites/detail/mlp_device.hpp).  Loss: mean over samples and outputs of
// mask * (y - t)^2 (mask 0 drops a target).  Gradients are summed over the
// batch by one thread per parameter in sample order, so they are
// deterministic run to run.  Optimiser: Adam (Kingma & Ba, ICLR 2015).
#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kites_amd.h"
#include "kites/detail/mlp_device.hpp"
#include "kites/memory.hpp"

namespace {

using namespace kites::detail;

constexpr int kTpb = 256;

__global__ void mlp_forward_kernel(const float* __restrict__ w, const float* __restrict__ X,
                                   unsigned B, float* __restrict__ H, float* __restrict__ Y) {
  unsigned const b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float x[MLP_IN];
  for (int i = 0; i < MLP_IN; ++i) x[i] = X[static_cast<size_t>(b) * MLP_IN + i];
  float h[MLP_HID];
  for (int j = 0; j < MLP_HID; ++j) {
    float a = w[MLP_OFF_B1 + j];
    for (int i = 0; i < MLP_IN; ++i) a += w[MLP_OFF_W1 + j * MLP_IN + i] * x[i];
    h[j] = a > 0.f ? a : 0.f;
    H[static_cast<size_t>(b) * MLP_HID + j] = h[j];
  }
  for (int k = 0; k < MLP_OUT; ++k) {
    float a = w[MLP_OFF_B2 + k];
    for (int j = 0; j < MLP_HID; ++j) a += w[MLP_OFF_W2 + k * MLP_HID + j] * h[j];
    Y[static_cast<size_t>(b) * MLP_OUT + k] = a;
  }
}

// dY = 2 * mask * (y - t) / (B * OUT); per-sample squared error into sq[b].
__global__ void mlp_output_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                       const float* __restrict__ M, unsigned B,
                                       float* __restrict__ dY, float* __restrict__ sq) {
  unsigned const b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float const scale = 2.f / (static_cast<float>(B) * MLP_OUT);
  float acc = 0.f;
  for (int k = 0; k < MLP_OUT; ++k) {
    size_t const idx = static_cast<size_t>(b) * MLP_OUT + k;
    float const m = M ? M[idx] : 1.f;
    float const d = (Y[idx] - T[idx]) * m;
    dY[idx] = scale * d;
    acc += d * d;
  }
  sq[b] = acc;
}

// dH = relu'(h) * (W2^T dY), per sample.
__global__ void mlp_hidden_grad_kernel(const float* __restrict__ w, const float* __restrict__ H,
                                       const float* __restrict__ dY, unsigned B,
                                       float* __restrict__ dH) {
  unsigned const b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  for (int j = 0; j < MLP_HID; ++j) {
    float a = 0.f;
    if (H[static_cast<size_t>(b) * MLP_HID + j] > 0.f)
      for (int k = 0; k < MLP_OUT; ++k)
        a += w[MLP_OFF_W2 + k * MLP_HID + j] * dY[static_cast<size_t>(b) * MLP_OUT + k];
    dH[static_cast<size_t>(b) * MLP_HID + j] = a;
  }
}

// One thread per parameter; sums over samples in order.
__global__ void mlp_param_grad_kernel(const float* __restrict__ X, const float* __restrict__ H,
                                      const float* __restrict__ dY,
                                      const float* __restrict__ dH, unsigned B,
                                      float* __restrict__ g) {
  int const p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float acc = 0.f;
  if (p < MLP_OFF_B1) {
    int const j = p / MLP_IN, i = p % MLP_IN;
    for (unsigned b = 0; b < B; ++b)
      acc += dH[static_cast<size_t>(b) * MLP_HID + j] * X[static_cast<size_t>(b) * MLP_IN + i];
  } else if (p < MLP_OFF_W2) {
    int const j = p - MLP_OFF_B1;
    for (unsigned b = 0; b < B; ++b) acc += dH[static_cast<size_t>(b) * MLP_HID + j];
  } else if (p < MLP_OFF_B2) {
    int const q = p - MLP_OFF_W2, k = q / MLP_HID, j = q % MLP_HID;
    for (unsigned b = 0; b < B; ++b)
      acc += dY[static_cast<size_t>(b) * MLP_OUT + k] * H[static_cast<size_t>(b) * MLP_HID + j];
  } else {
    int const k = p - MLP_OFF_B2;
    for (unsigned b = 0; b < B; ++b) acc += dY[static_cast<size_t>(b) * MLP_OUT + k];
  }
  g[p] = acc;
}

__global__ void mlp_adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                float* __restrict__ m1, float* __restrict__ m2, float lr,
                                float b1, float b2, float eps, float bc1, float bc2) {
  int const p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float const gp = g[p];
  float const a = b1 * m1[p] + (1.f - b1) * gp;
  float const v = b2 * m2[p] + (1.f - b2) * gp * gp;
  m1[p] = a;
  m2[p] = v;
  w[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

__global__ void sum_kernel(const float* __restrict__ x, unsigned n, float* __restrict__ out) {
  // One block: deterministic tree sum.
  __shared__ float s[kTpb];
  float acc = 0.f;
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) acc += x[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (unsigned st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

thread_local std::string g_mlp_err;

struct mlp_session {
  unsigned B;
  kites::device_buffer<float> w, X, T, M, H, Y, dY, dH, g, sq, loss, m1, m2;
  mlp_session(int dev, unsigned b, const float* wh, const float* Xh, const float* Th,
              const float* Mh)
      : B(b), w(MLP_NPARAM, dev), X(static_cast<size_t>(b) * MLP_IN, dev),
        T(static_cast<size_t>(b) * MLP_OUT, dev), H(static_cast<size_t>(b) * MLP_HID, dev),
        Y(static_cast<size_t>(b) * MLP_OUT, dev), dY(static_cast<size_t>(b) * MLP_OUT, dev),
        dH(static_cast<size_t>(b) * MLP_HID, dev), g(MLP_NPARAM, dev), sq(b, dev), loss(1, dev),
        m1(MLP_NPARAM, dev), m2(MLP_NPARAM, dev) {
    KITES_HIP_CHECK(hipMemcpy(w.get_ptr(), wh, MLP_NPARAM * 4, hipMemcpyHostToDevice));
    KITES_HIP_CHECK(hipMemcpy(X.get_ptr(), Xh, X.sizeInBytes(), hipMemcpyHostToDevice));
    if (Th) KITES_HIP_CHECK(hipMemcpy(T.get_ptr(), Th, T.sizeInBytes(), hipMemcpyHostToDevice));
    if (Mh) {
      M.alloc(static_cast<size_t>(b) * MLP_OUT, dev);
      KITES_HIP_CHECK(hipMemcpy(M.get_ptr(), Mh, M.sizeInBytes(), hipMemcpyHostToDevice));
    }
    KITES_HIP_CHECK(hipMemset(m1.get_ptr(), 0, m1.sizeInBytes()));
    KITES_HIP_CHECK(hipMemset(m2.get_ptr(), 0, m2.sizeInBytes()));
  }
  dim3 sgrid() const { return dim3((B + kTpb - 1) / kTpb); }
  void forward() {
    hipLaunchKernelGGL(mlp_forward_kernel, sgrid(), dim3(kTpb), 0, 0, w.get_ptr(), X.get_ptr(), B,
                       H.get_ptr(), Y.get_ptr());
  }
  void backward() {
    hipLaunchKernelGGL(mlp_output_grad_kernel, sgrid(), dim3(kTpb), 0, 0, Y.get_ptr(), T.get_ptr(),
                       M.size() ? M.get_ptr() : nullptr, B, dY.get_ptr(), sq.get_ptr());
    hipLaunchKernelGGL(mlp_hidden_grad_kernel, sgrid(), dim3(kTpb), 0, 0, w.get_ptr(), H.get_ptr(),
                       dY.get_ptr(), B, dH.get_ptr());
    hipLaunchKernelGGL(mlp_param_grad_kernel, dim3((MLP_NPARAM + kTpb - 1) / kTpb), dim3(kTpb), 0,
                       0, X.get_ptr(), H.get_ptr(), dY.get_ptr(), dH.get_ptr(), B, g.get_ptr());
    hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(kTpb), 0, 0, sq.get_ptr(), B, loss.get_ptr());
  }
  float read_loss() {
    float l = 0.f;
    KITES_HIP_CHECK(hipMemcpy(&l, loss.get_ptr(), 4, hipMemcpyDeviceToHost));
    return l / (static_cast<float>(B) * MLP_OUT);
  }
};

int mlp_fail(std::exception const& e) {
  g_mlp_err = e.what();
  return dynamic_cast<std::invalid_argument const*>(&e) ? KITES_ERR_ARG : KITES_ERR_RUNTIME;
}

}  // namespace

extern "C" {

const char* kites_mlp_last_error(void) { return g_mlp_err.c_str(); }

int kites_mlp_forward(int32_t device, const float* w, const float* X, uint32_t B, float* Y) {
  try {
    if (!w || !X || !Y || !B) throw std::invalid_argument("null argument or empty batch");
    KITES_HIP_CHECK(hipSetDevice(device));
    mlp_session s(device, B, w, X, nullptr, nullptr);
    s.forward();
    KITES_HIP_CHECK(hipMemcpy(Y, s.Y.get_ptr(), s.Y.sizeInBytes(), hipMemcpyDeviceToHost));
  } catch (std::exception const& e) {
    return mlp_fail(e);
  }
  return KITES_OK;
}

int kites_mlp_loss_grad(int32_t device, const float* w, const float* X, const float* T,
                        const float* mask, uint32_t B, float* grad, float* loss) {
  try {
    if (!w || !X || !T || !grad || !B) throw std::invalid_argument("null argument or empty batch");
    KITES_HIP_CHECK(hipSetDevice(device));
    mlp_session s(device, B, w, X, T, mask);
    s.forward();
    s.backward();
    KITES_HIP_CHECK(hipMemcpy(grad, s.g.get_ptr(), MLP_NPARAM * 4, hipMemcpyDeviceToHost));
    if (loss) *loss = s.read_loss();
  } catch (std::exception const& e) {
    return mlp_fail(e);
  }
  return KITES_OK;
}

int kites_mlp_train(int32_t device, float* w, const float* X, const float* T, const float* mask,
                    uint32_t B, uint32_t epochs, float lr, float* loss_first, float* loss_last) {
  try {
    if (!w || !X || !T || !B) throw std::invalid_argument("null argument or empty batch");
    KITES_HIP_CHECK(hipSetDevice(device));
    mlp_session s(device, B, w, X, T, mask);
    float const b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.f, bc2 = 1.f;
    for (uint32_t e = 0; e < epochs; ++e) {
      s.forward();
      s.backward();
      if (e == 0 && loss_first) *loss_first = s.read_loss();
      bc1 *= b1;
      bc2 *= b2;
      hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + kTpb - 1) / kTpb), dim3(kTpb), 0, 0,
                         s.w.get_ptr(), s.g.get_ptr(), s.m1.get_ptr(), s.m2.get_ptr(), lr, b1, b2,
                         eps, 1.f - bc1, 1.f - bc2);
    }
    s.forward();
    s.backward();
    if (loss_last) *loss_last = s.read_loss();
    KITES_HIP_CHECK(hipMemcpy(w, s.w.get_ptr(), MLP_NPARAM * 4, hipMemcpyDeviceToHost));
  } catch (std::exception const& e) {
    return mlp_fail(e);
  }
  return KITES_OK;
}

}  // extern "C"
