// Training kernels of the friction warm-start MLP (6-32-32-1, tanh, FP32):
// mean-squared-error loss and gradient over a batch, Adam update, forward.
// A block of kThreads (128) threads takes 128 samples: forward and backward per sample
// in registers, the per-sample hidden activations and deltas in LDS, then
// each thread reduces a slice of the weight gradients over the block's
// samples and adds it to the global gradient with one atomic per entry.
#include <hip/hip_runtime.h>

#include "engine.h"
#include "friction_common.h"

namespace sdg {
namespace friction {
namespace {

constexpr int kThreads = 128;
constexpr int kStride = kMlpInputs + 1;  // samples: features then target

__global__ void kZero(float* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

__global__ __launch_bounds__(kThreads) void kLossGrad(const float* __restrict__ prm,
                                                      const float* __restrict__ samples, int64_t n,
                                                      float* __restrict__ grads, float* __restrict__ loss) {
  __shared__ float sZ[kThreads][kMlpInputs];
  __shared__ float sH1[kThreads][kMlpHidden + 1];
  __shared__ float sD1[kThreads][kMlpHidden + 1];
  __shared__ float sH2[kThreads][kMlpHidden + 1];
  __shared__ float sD2[kThreads][kMlpHidden + 1];
  __shared__ float sDy[kThreads];
  const int t = threadIdx.x;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + t;
  const bool ok = i < n;
  const float inv = 1.0f / static_cast<float>(n);
  float z[kMlpInputs];
  float target = 0.0f;
#pragma unroll
  for (int j = 0; j < kMlpInputs; ++j) {
    const float x = ok ? samples[i * kStride + j] : 0.0f;
    z[j] = (x - prm[kOffInMean + j]) / prm[kOffInScale + j];
    sZ[t][j] = ok ? z[j] : 0.0f;
  }
  if (ok) target = (samples[i * kStride + kMlpInputs] - prm[kOffOutMean]) / prm[kOffOutScale];
  float h1[kMlpHidden];
  for (int k = 0; k < kMlpHidden; ++k) {
    float s = prm[kOffB1 + k];
#pragma unroll
    for (int j = 0; j < kMlpInputs; ++j) s += prm[kOffW1 + k * kMlpInputs + j] * z[j];
    h1[k] = tanhf(s);
    sH1[t][k] = ok ? h1[k] : 0.0f;
  }
  float h2[kMlpHidden];
  float y = prm[kOffB3];
  for (int k = 0; k < kMlpHidden; ++k) {
    float s = prm[kOffB2 + k];
    for (int j = 0; j < kMlpHidden; ++j) s += prm[kOffW2 + k * kMlpHidden + j] * h1[j];
    h2[k] = tanhf(s);
    sH2[t][k] = ok ? h2[k] : 0.0f;
    y += prm[kOffW3 + k] * h2[k];
  }
  const float r = y - target;
  const float dy = ok ? 2.0f * r * inv : 0.0f;
  sDy[t] = dy;
  float d2[kMlpHidden];
  for (int k = 0; k < kMlpHidden; ++k) {
    d2[k] = dy * prm[kOffW3 + k] * (1.0f - h2[k] * h2[k]);
    sD2[t][k] = d2[k];
  }
  for (int j = 0; j < kMlpHidden; ++j) {
    float s = 0.0f;
    for (int k = 0; k < kMlpHidden; ++k) s += prm[kOffW2 + k * kMlpHidden + j] * d2[k];
    sD1[t][j] = s * (1.0f - h1[j] * h1[j]);
  }
  // block sum of the loss
  float l = ok ? r * r * inv : 0.0f;
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);  // wave64 reduction
  if ((t & 63) == 0) atomicAdd(loss, l);
  __syncthreads();
  // weight gradients: entry e handled by thread e % kThreads, summed over the block's samples
  for (int e = t; e < kMlpParams - kOffW1; e += kThreads) {
    const int idx = kOffW1 + e;
    float acc = 0.0f;
    if (idx < kOffB1) {
      const int k = (idx - kOffW1) / kMlpInputs, j = (idx - kOffW1) % kMlpInputs;
      for (int s = 0; s < kThreads; ++s) acc += sD1[s][k] * sZ[s][j];
    } else if (idx < kOffW2) {
      const int k = idx - kOffB1;
      for (int s = 0; s < kThreads; ++s) acc += sD1[s][k];
    } else if (idx < kOffB2) {
      const int k = (idx - kOffW2) / kMlpHidden, j = (idx - kOffW2) % kMlpHidden;
      for (int s = 0; s < kThreads; ++s) acc += sD2[s][k] * sH1[s][j];
    } else if (idx < kOffW3) {
      const int k = idx - kOffB2;
      for (int s = 0; s < kThreads; ++s) acc += sD2[s][k];
    } else if (idx < kOffB3) {
      const int k = idx - kOffW3;
      for (int s = 0; s < kThreads; ++s) acc += sDy[s] * sH2[s][k];
    } else {
      for (int s = 0; s < kThreads; ++s) acc += sDy[s];
    }
    atomicAdd(&grads[idx], acc);
  }
}

__global__ void kAdam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                      float* __restrict__ v, float lr, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kOffW1 || i >= kMlpParams) return;  // normalisation constants are not trained
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void kForward(const float* __restrict__ prm, const float* __restrict__ x, int64_t n, int stride,
                         float* __restrict__ out) {
  __shared__ float sP[kMlpParams];
  for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = prm[i];
  __syncthreads();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[kMlpInputs];
  for (int j = 0; j < kMlpInputs; ++j) f[j] = x[i * stride + j];
  out[i] = mlpEvaluate(sP, f);
}

}  // namespace
}  // namespace friction
}  // namespace sdg

using namespace sdg;

extern "C" {

int sdg_mlp_param_count(void) { return friction::kMlpParams; }

int sdg_mlp_loss_grad(const float* params, const float* samples, int64_t n, float* grads, float* loss, void* stream) {
  if (params == nullptr || samples == nullptr || grads == nullptr || loss == nullptr || n <= 0) {
    setError("sdg_mlp_loss_grad: invalid arguments");
    return SDG_ERR_INVALID;
  }
  auto st = static_cast<hipStream_t>(stream);
  friction::kZero<<<(friction::kMlpParams + 255) / 256, 256, 0, st>>>(grads, friction::kMlpParams);
  friction::kZero<<<1, 64, 0, st>>>(loss, 1);
  const unsigned grid = static_cast<unsigned>((n + friction::kThreads - 1) / friction::kThreads);
  friction::kLossGrad<<<grid, friction::kThreads, 0, st>>>(params, samples, n, grads, loss);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

int sdg_mlp_adam(float* params, const float* grads, float* m, float* v, int step, float lr, void* stream) {
  if (params == nullptr || grads == nullptr || m == nullptr || v == nullptr || step < 1) {
    setError("sdg_mlp_adam: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const float c1 = 1.0f - powf(0.9f, static_cast<float>(step));
  const float c2 = 1.0f - powf(0.999f, static_cast<float>(step));
  friction::kAdam<<<(friction::kMlpParams + 255) / 256, 256, 0, static_cast<hipStream_t>(stream)>>>(
      params, grads, m, v, lr, c1, c2);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

int sdg_mlp_forward(const float* params, const float* features, int64_t n, int stride, float* out, void* stream) {
  if (params == nullptr || features == nullptr || out == nullptr || n < 0 || stride < friction::kMlpInputs) {
    setError("sdg_mlp_forward: invalid arguments");
    return SDG_ERR_INVALID;
  }
  if (n == 0) return SDG_OK;
  friction::kForward<<<static_cast<unsigned>((n + 255) / 256), 256, 0, static_cast<hipStream_t>(stream)>>>(
      params, features, n, stride, out);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

}  // extern "C"
