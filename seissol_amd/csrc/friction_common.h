// Warm-start MLP shared by the friction kernel and the training kernels.
//
// Parameter layout (FP32): in_mean[6], in_scale[6], out_mean, out_scale,
// W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3.
// The network predicts the root of the first inner slip-rate solve of a time
// point, g(V) = S - N asinh(V exp(psi / a) / (2 V0)) - V = 0 with
// S = invEtaS tau (the root lies in (0, S]), N = invEtaS |sigma_n| a and psi
// the state after the first state update.  The closed-form creep root
// V_c = 2 V0 exp(-psi / a) sinh(S / N) solves g(V) + V = 0 and bounds the root
// from above; it is the root itself up to V_c where V_c is small.  The network
// serves the points where it is not, learning the correction to
//   base = min(log10 V_c, log10 S)
// from the features
//   x = (log10 V_c, log10 S, log10 N, psi / a, log10 V_prev, 100 a)
// and outputs log10(root) - base.
#pragma once
#include <hip/hip_runtime.h>

namespace sdg {
namespace friction {

constexpr int kMlpInputs = 6;
constexpr int kMlpHidden = 32;
constexpr int kOffInMean = 0;
constexpr int kOffInScale = 6;
constexpr int kOffOutMean = 12;
constexpr int kOffOutScale = 13;
constexpr int kOffW1 = 14;
constexpr int kOffB1 = kOffW1 + kMlpHidden * kMlpInputs;
constexpr int kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden;
constexpr int kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpHidden;
constexpr int kMlpParams = kOffB3 + 1;

// features of one solve from the natural log of V_c; returns base (FP64)
__device__ inline double mlpFeatures(double logVc, double S, double N, double psiOverA, double vPrev, double a,
                                     float (&x)[kMlpInputs]) {
  const double ls = log10(fmax(S, 1e-300));
  const double creep = fmin(fmax(logVc * 0.43429448190325182765, -300.0), 300.0);
  x[0] = static_cast<float>(creep);
  x[1] = static_cast<float>(ls);
  x[2] = static_cast<float>(log10(fmax(N, 1e-300)));
  x[3] = static_cast<float>(psiOverA);
  x[4] = static_cast<float>(log10(fmax(vPrev, 1e-300)));
  x[5] = static_cast<float>(100.0 * a);
  return fmin(creep, ls);
}

// hidden activations of the normalised input z
__device__ inline float mlpEvaluate(const float* prm, const float (&x)[kMlpInputs]) {
  float z[kMlpInputs];
#pragma unroll
  for (int i = 0; i < kMlpInputs; ++i) z[i] = (x[i] - prm[kOffInMean + i]) / prm[kOffInScale + i];
  float h1[kMlpHidden];
#pragma unroll
  for (int k = 0; k < kMlpHidden; ++k) {
    float s = prm[kOffB1 + k];
#pragma unroll
    for (int i = 0; i < kMlpInputs; ++i) s += prm[kOffW1 + k * kMlpInputs + i] * z[i];
    h1[k] = tanhf(s);
  }
  float y = prm[kOffB3];
  for (int k = 0; k < kMlpHidden; ++k) {
    float s = prm[kOffB2 + k];
#pragma unroll
    for (int i = 0; i < kMlpHidden; ++i) s += prm[kOffW2 + k * kMlpHidden + i] * h1[i];
    y += prm[kOffW3 + k] * tanhf(s);
  }
  return y * prm[kOffOutScale] + prm[kOffOutMean];
}

}  // namespace friction
}  // namespace sdg
