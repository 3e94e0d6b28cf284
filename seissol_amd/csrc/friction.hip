// Rate-and-state fast-velocity-weakening friction (law 103) on the GPU, FP64.
//
// One lane per (face, Gauss point); the lanes of a face form a contiguous
// power-of-two group inside a wave so the face-wide convergence decisions of
// the CPU implementation (CpuImpl/RateAndState.h:209-222, 375-386) are wave
// ballots restricted to the face's lanes.  Per time point: stresses from the
// interpolated DOFs, nucleation, fixed-point state update around a Newton
// solve for the slip rate, slip/traction update and outputs; afterwards the
// state resampling and the imposed states (FrictionSolverCommon.h:146-631,
// FastVelocityWeakeningLaw.h:43-165, RateAndStateCommon.h:22-139).
//
// Optional warm start of the first slip-rate solve of each time point, before
// its first residual test: a point whose closed-form creep root V_c (the root
// of g(V) + V = 0, so g(V_c) = -V_c) lies below half the tolerance starts at
// V_c and passes that test at once; with a 6-32-32-1 MLP (FP32) every other
// point starts at the network's predicted root when it lies more than a decade
// from the previous slip rate.  No evaluation is added and the residual test
// is unchanged; g is strictly decreasing in V, so a converged solve returns
// the plain solver's root up to the tolerance.
#include <hip/hip_runtime.h>

#include <cmath>

#include "engine.h"
#include "friction_common.h"

namespace sdg {
namespace friction {
namespace {

constexpr double kAlmostZero = 1e-45;
constexpr int kMaxOrder = 8;
constexpr double kMovedDecades = 0.30103;  // log10(2)
constexpr double kJumpDecades = 1.0;
constexpr double kLn10 = 2.30258509299404568402;

// The network's forward pass, kept out of line: inlined into kFriction, its
// unrolled products and LDS weight loads pushed the kernel past its registers
// (916 B of spills per lane); as a call it costs the warm-start branch only.
__device__ __noinline__ float mlpEvaluateCall(const float* prm, const float (&x)[kMlpInputs]) {
  return mlpEvaluate(prm, x);
}

// natural log of the closed-form creep root V_c = 2 V0 exp(-psi/a) sinh(S/N),
// which solves N asinh(V exp(psi/a) / (2 V0)) = S; -inf when S or N is not positive
__device__ inline double creepLogRoot(double S, double N, double psiOverA, double sr0) {
  if (!(S > 0.0) || !(N > 0.0)) return -INFINITY;
  const double c = S / N;
  return log(2.0 * sr0) - psiOverA + c + log(-0.5 * expm1(-2.0 * c));
}

__device__ inline double arsinhexp(double x, double expLog, double ex) {
  const double sw = 10.0, threshold = 50.0, log2 = 0.69314718055994530943;
  int xexp = 0;
  (void)frexp(x, &xexp);
  if (expLog + (xexp > 0 ? xexp : 0) * log2 > sw || expLog >= threshold) {
    if (expLog <= 0) ex = 1.0 / ex;
    const double xa = fabs(x);
    const double xs = x >= 0 ? 1.0 : -1.0;
    return xs * (expLog + log(xa + sqrt(xa * xa + ex * ex)));
  }
  if (expLog > 0) ex = 1.0 / ex;
  return asinh(ex * x);
}

__device__ inline double arsinhexpDerivative(double x, double expLog, double ex) {
  const double sw = 10.0, threshold = 50.0, log2 = 0.69314718055994530943;
  int xexp = 0;
  (void)frexp(x, &xexp);
  if (expLog + (xexp > 0 ? xexp : 0) * log2 > sw || expLog >= threshold) {
    if (expLog <= 0) ex = 1.0 / ex;
    return 1.0 / sqrt(x * x + ex * ex);
  }
  if (expLog > 0) ex = 1.0 / ex;
  const double v = ex * x;
  return ex / sqrt(1.0 + v * v);
}

__device__ inline double computeCExp(double c) { return c > 0 ? exp(-c) : exp(c); }

__device__ inline double logsinh(double x, double c) {
  const double sign = c >= 0 ? 1.0 : -1.0;
  const double ac = fabs(c);
  return ac + log(x / 2 * -sign * expm1(-2 * ac));
}

__device__ inline double smoothStep(double t, double t0) {
  if (t <= 0) return 0.0;
  if (t < t0) {
    const double tau = t - t0;
    return exp(tau * tau / (t * (t - 2.0 * t0)));
  }
  return 1.0;
}

struct PointParams {
  double a, sl0, f0, muw, b, srw;
};

__device__ inline double updateStateVariable(const PointParams& pp, double sr0, double svRef, double dt, double V) {
  const double lowV = fmax(0.0, pp.f0 - (pp.b - pp.a) * log(V / sr0));
  const double r = V / pp.srw;
  const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
  const double muSS = pp.muw + (lowV - pp.muw) / pow(1.0 + r8, 0.125);
  const double psiSS = pp.a * logsinh(sr0 / V * 2, muSS / pp.a);
  const double pre = -V * (dt / pp.sl0);
  return psiSS * (-expm1(pre)) + exp(pre) * svRef;
}

__device__ inline double muOf(double a, double sr0, double V, double psi) {
  const double cExpLog = psi / a;
  return a * arsinhexp(0.5 / sr0 * V, cExpLog, computeCExp(cExpLog));
}

__device__ inline double dmuOf(double a, double sr0, double V, double psi) {
  const double cExpLog = psi / a;
  return a * (0.5 / sr0) * arsinhexpDerivative(0.5 / sr0 * V, cExpLog, computeCExp(cExpLog));
}

// true if pred holds on every real point of this lane's face
template <int TPF>
__device__ inline bool faceAll(bool pred, int numPoints) {
  const unsigned long long b = __ballot(pred);
  const int lane = threadIdx.x & 63;
  const int seg = (lane / TPF) * TPF;
  const unsigned long long full = numPoints >= 64 ? ~0ULL : ((1ULL << numPoints) - 1ULL);
  return ((b >> seg) & full) == full;
}

struct StepArgs {
  double deltaT[kMaxOrder];
  double timeWeights[kMaxOrder];
};

template <int TPF>
__global__ __launch_bounds__(256) void kFriction(sdg_friction_params prm, sdg_fault_layer L, double fullUpdateTime,
                                                 StepArgs args, const double* __restrict__ resample,
                                                 sdg_friction_options opt) {
  constexpr int FPB = 256 / TPF;
  __shared__ double sDelta[256];
  __shared__ float sMlp[kMlpParams];
  const bool useMlp = opt.mlp_params != nullptr;
  if (useMlp)
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sMlp[i] = opt.mlp_params[i];
  __syncthreads();

  const int pnt = threadIdx.x % TPF;
  const int64_t face = static_cast<int64_t>(blockIdx.x) * FPB + threadIdx.x / TPF;
  const int NP = L.num_points, P = L.points, O = L.order;
  const bool active = face < L.num_faces && pnt < NP;
  const int64_t f = active ? face : 0;
  const int p = active ? pnt : 0;
  const int64_t k = f * P + p;

  const double* ie = L.imp_and_eta + f * 11;
  const double etaS = ie[5], invEtaS = ie[6], invZp = ie[7], invZs = ie[8], invZpN = ie[9], invZsN = ie[10];
  const double etaP = ie[4] * (prm.eta_damp_end > fullUpdateTime ? prm.eta_damp : 1.0);
  const PointParams pp{L.rs_a[k], L.rs_sl0[k], L.rs_f0[k], L.rs_muw[k], L.rs_b[k], L.rs_srw[k]};
  auto qp = [&](int o, int q) { return L.q_interpolated_plus[((f * O + o) * 9 + q) * P + p]; };
  auto qm = [&](int o, int q) { return L.q_interpolated_minus[((f * O + o) * 9 + q) * P + p]; };

  double fsN[kMaxOrder], fsT1[kMaxOrder], fsT2[kMaxOrder], trT1[kMaxOrder], trT2[kMaxOrder];
  for (int o = 0; o < O; ++o) {
    fsN[o] = etaP * (qm(o, 6) - qp(o, 6) + qp(o, 0) * invZp + qm(o, 0) * invZpN);
    fsT1[o] = etaS * (qm(o, 7) - qp(o, 7) + qp(o, 3) * invZs + qm(o, 3) * invZsN);
    fsT2[o] = etaS * (qm(o, 8) - qp(o, 8) + qp(o, 5) * invZs + qm(o, 5) * invZsN);
  }
  double* is = L.initial_stress + f * 6 * P + p;  // component c at is[c * P]
  double sv = L.state_variable[k];
  double slipRate1 = L.slip_rate1[k], slipRate2 = L.slip_rate2[k];
  double srm = L.slip_rate_magnitude[k];
  double muv = L.mu[k];
  uint8_t convInner = L.convergence_inner[k], convOuter = L.convergence_outer[k];
  // the per-point outputs accumulated over the time points, read once here and
  // written once after the time loop (a read-modify-write per time point
  // waited on each load in turn)
  double accSlip = L.accumulated_slip[k], slip1 = L.slip1[k], slip2 = L.slip2[k];
  double peakSr = L.peak_slip_rate[k], ruptureTime = L.rupture_time[k], dynStressTime = L.dyn_stress_time[k];
  uint8_t rupturePending = L.rupture_time_pending[k], dynStressPending = L.dyn_stress_time_pending[k];
  int newtonEvals = 0, newtonSteps = 0;
  const bool closedForm = opt.closed_form_start != 0;

  double updateTime = fullUpdateTime;
  for (int o = 0; o < O; ++o) {
    const double dt = args.deltaT[o];
    const double startTime = updateTime;
    updateTime += dt;
    if (prm.nucleation_count > 0 && L.nucleation_stress != nullptr && updateTime <= prm.t0 + prm.s0 &&
        updateTime >= prm.s0) {
      const double gNuc = smoothStep(updateTime - prm.s0, prm.t0) - smoothStep(updateTime - prm.s0 - dt, prm.t0);
      if (active) {
        for (int c = 0; c < 6; ++c) is[c * P] += L.nucleation_stress[f * 6 * P + c * P + p] * gNuc;
        if (L.nucleation_pressure != nullptr) L.initial_pressure[k] += L.nucleation_pressure[k] * gNuc;
      }
    }
    const double pressure = L.initial_pressure[k];
    const double normal = fmin(0.0, fsN[o] + is[0] + pressure);
    const double svRef = sv;
    const double absShear = sqrt((is[3 * P] + fsT1[o]) * (is[3 * P] + fsT1[o]) + (is[5 * P] + fsT2[o]) * (is[5 * P] + fsT2[o]));
    srm = fmax(kAlmostZero, sqrt(slipRate1 * slipRate1 + slipRate2 * slipRate2));
    const double vPrev = srm;
    const double upper = fmax(kAlmostZero, invEtaS * absShear);  // g(upper) <= 0: the root lies below
    const double nScale = invEtaS * fabs(normal) * pp.a;
    double localSR = srm;
    bool convOuterPre = true;
    for (int j = 0; j < prm.rs_state_variable_updates; ++j) {
      sv = active ? updateStateVariable(pp, prm.rs_sr0, svRef, dt, localSR) : 0.0;
      double test = srm;
      // warm start of the first solve; lvc, base, feat also feed the sample recording
      double lvc = -INFINITY, base = 0.0;
      float feat[kMlpInputs];
      if (j == 0 && active && (closedForm || useMlp || opt.samples != nullptr)) {
        lvc = creepLogRoot(invEtaS * absShear, nScale, sv / pp.a, prm.rs_sr0);
        const bool creeping = lvc < log(0.5 * prm.rs_slip_rate_tolerance);
        if (closedForm && creeping) {
          test = fmax(kAlmostZero, exp(lvc));
        } else if (useMlp && !creeping) {
          base = mlpFeatures(lvc, upper, nScale, sv / pp.a, vPrev, pp.a, feat);
          // the root lies in (0, min(V_c, S)]
          const double lg = fmin(base + static_cast<double>(mlpEvaluateCall(sMlp, feat)), fmin(lvc / kLn10, log10(upper)));
          if (isfinite(lg) && fabs(lg - log10(vPrev)) > kJumpDecades) test = fmax(kAlmostZero, pow(10.0, lg));
        }
      }
      double g = 0.0, muF = muv;
      bool converged = false;
      // Newton on g(V) = -invEtaS (|sigma_n| mu(V) - tau) - V from the
      // previous slip rate (RateAndState.h:339-416)
      for (int it = 0; it < prm.rs_max_slip_rate_updates; ++it) {
        if (active) {
          muF = muOf(pp.a, prm.rs_sr0, test, sv);
          g = -invEtaS * (fabs(normal) * muF - absShear) - test;
          ++newtonEvals;
        }
        converged = faceAll<TPF>(!active || fabs(g) < prm.rs_slip_rate_tolerance, NP);
        if (converged) break;
        if (active) {
          const double dG = -invEtaS * (fabs(normal) * dmuOf(pp.a, prm.rs_sr0, test, sv)) - 1.0;
          test = fmax(kAlmostZero, test - g / dG);
          ++newtonSteps;
        }
      }
      if (converged) muv = muF;
      else convInner &= (fabs(g) < prm.rs_slip_rate_tolerance) ? 1 : 0;
      if (opt.samples != nullptr && !useMlp && j == 0 && active && converged &&
          !(lvc < log(0.5 * prm.rs_slip_rate_tolerance))) {
        // (features, log10 root - base) of the first solve of the time point
        // where the closed form does not apply; solves that move the slip rate
        // by more than a factor of two are all kept, the others thinned by a
        // hash to 1 in sample_stride
        base = mlpFeatures(lvc, upper, nScale, sv / pp.a, vPrev, pp.a, feat);
        const uint64_t h = (static_cast<uint64_t>(face) * 2654435761ULL) ^ (static_cast<uint64_t>(p) * 40503ULL) ^
                           (static_cast<uint64_t>(o) * 97ULL) ^ static_cast<uint64_t>(fullUpdateTime * 1e9);
        const bool moved = fabs(log10(test) - log10(vPrev)) > kMovedDecades;
        if (moved || opt.sample_stride <= 1 || (h % static_cast<uint64_t>(opt.sample_stride)) == 0) {
          const unsigned long long idx = atomicAdd(opt.sample_count, 1ULL);
          if (static_cast<int64_t>(idx) < opt.capacity) {
            for (int i = 0; i < kMlpInputs; ++i) opt.samples[idx * (kMlpInputs + 1) + i] = feat[i];
            opt.samples[idx * (kMlpInputs + 1) + kMlpInputs] = static_cast<float>(log10(test) - base);
          }
        }
      }
      const double old = srm;
      localSR = 0.5 * (old + test);
      convOuterPre = fabs(test - old) < prm.rs_state_tolerance;
      srm = test;
      if (faceAll<TPF>(!active || convOuterPre, NP)) break;
    }
    convOuter &= convOuterPre ? 1 : 0;
    // final state, friction, slip rates and tractions
    sv = active ? updateStateVariable(pp, prm.rs_sr0, svRef, dt, localSR) : 0.0;
    if (active) {
      muv = muOf(pp.a, prm.rs_sr0, srm, sv);
      const double strength = -muv * normal;
      const double t1 = is[3 * P] + fsT1[o], t2 = is[5 * P] + fsT2[o];
      const double divisor = strength + etaS * srm;
      slipRate1 = srm * t1 / divisor;
      slipRate2 = srm * t2 / divisor;
      trT1[o] = fsT1[o] - etaS * slipRate1;
      trT2[o] = fsT2[o] - etaS * slipRate2;
      accSlip += srm * dt;
      slip1 += slipRate1 * dt;
      slip2 += slipRate2 * dt;
      if (rupturePending && srm > 0.001) {
        ruptureTime = startTime;
        rupturePending = 0;
      }
      if (ruptureTime > 0.0 && ruptureTime <= startTime && dynStressPending &&
          muv <= pp.muw + 0.05 * (pp.f0 - pp.muw)) {
        dynStressTime = startTime;
        dynStressPending = 0;
      }
      peakSr = fmax(peakSr, srm);
    }
  }
  if (active) {
    L.accumulated_slip[k] = accSlip;
    L.slip1[k] = slip1;
    L.slip2[k] = slip2;
    L.peak_slip_rate[k] = peakSr;
    L.rupture_time[k] = ruptureTime;
    L.rupture_time_pending[k] = rupturePending;
    L.dyn_stress_time[k] = dynStressTime;
    L.dyn_stress_time_pending[k] = dynStressPending;
  }
  // state resampling over the face's points (postHook)
  __syncthreads();
  const double delta = active ? sv - L.state_variable[k] : 0.0;
  sDelta[threadIdx.x] = delta;
  __syncthreads();
  if (active) {
    const int base = (threadIdx.x / TPF) * TPF;
    double acc = 0.0;
    for (int q = 0; q < NP; ++q) acc += resample[p * NP + q] * sDelta[base + q];
    L.state_variable[k] += acc;
    L.slip_rate1[k] = slipRate1;
    L.slip_rate2[k] = slipRate2;
    L.slip_rate_magnitude[k] = srm;
    L.mu[k] = muv;
    L.traction1[k] = trT1[O - 1];
    L.traction2[k] = trT2[O - 1];
    L.convergence_inner[k] = convInner;
    L.convergence_outer[k] = convOuter;
    if (L.newton_iterations != nullptr) L.newton_iterations[k] += newtonEvals;
    if (L.slip_rate_updates != nullptr) L.slip_rate_updates[k] += newtonSteps;
    // imposed states (FrictionSolverCommon.h:333-401)
    double m[9] = {0}, pl[9] = {0};
    for (int o = 0; o < O; ++o) {
      const double w = args.timeWeights[o];
      const double ns = fsN[o], t1 = trT1[o], t2 = trT2[o];
      m[0] += w * ns; m[3] += w * t1; m[5] += w * t2;
      m[6] += w * (qm(o, 6) - invZpN * (ns - qm(o, 0)));
      m[7] += w * (qm(o, 7) - invZsN * (t1 - qm(o, 3)));
      m[8] += w * (qm(o, 8) - invZsN * (t2 - qm(o, 5)));
      pl[0] += w * ns; pl[3] += w * t1; pl[5] += w * t2;
      pl[6] += w * (qp(o, 6) + invZp * (ns - qp(o, 0)));
      pl[7] += w * (qp(o, 7) + invZs * (t1 - qp(o, 3)));
      pl[8] += w * (qp(o, 8) + invZs * (t2 - qp(o, 5)));
    }
    for (int q = 0; q < 9; ++q) {
      L.imposed_state_minus[(f * 9 + q) * P + p] = m[q];
      L.imposed_state_plus[(f * 9 + q) * P + p] = pl[q];
    }
  }
}

}  // namespace
}  // namespace friction
}  // namespace sdg

using namespace sdg;

extern "C" int sdg_friction_evaluate(const sdg_friction_params* params, const sdg_fault_layer* layer,
                                     double full_update_time, const double* delta_t, const double* time_weights,
                                     const double* resample, const sdg_friction_options* options, void* stream) {
  if (params == nullptr || layer == nullptr || delta_t == nullptr || time_weights == nullptr) {
    setError("sdg_friction_evaluate: NULL argument");
    return SDG_ERR_INVALID;
  }
  const sdg_fault_layer& L = *layer;
  if (L.num_faces < 0 || L.num_points < 1 || L.num_points > 64 || L.points < L.num_points || L.order < 1 ||
      L.order > friction::kMaxOrder) {
    setError("sdg_friction_evaluate: invalid layer sizes (num_points in [1, 64], order in [1, 8])");
    return SDG_ERR_INVALID;
  }
  if (L.num_faces == 0) return SDG_OK;
  if (resample == nullptr || L.imp_and_eta == nullptr || L.q_interpolated_plus == nullptr ||
      L.q_interpolated_minus == nullptr || L.initial_stress == nullptr || L.initial_pressure == nullptr ||
      L.rs_a == nullptr || L.rs_sl0 == nullptr || L.rs_f0 == nullptr || L.rs_muw == nullptr || L.rs_b == nullptr ||
      L.rs_srw == nullptr || L.state_variable == nullptr || L.slip_rate1 == nullptr || L.slip_rate2 == nullptr ||
      L.slip_rate_magnitude == nullptr || L.mu == nullptr || L.traction1 == nullptr || L.traction2 == nullptr ||
      L.slip1 == nullptr || L.slip2 == nullptr || L.accumulated_slip == nullptr || L.rupture_time == nullptr ||
      L.peak_slip_rate == nullptr || L.dyn_stress_time == nullptr || L.rupture_time_pending == nullptr ||
      L.dyn_stress_time_pending == nullptr || L.convergence_inner == nullptr || L.convergence_outer == nullptr ||
      L.imposed_state_plus == nullptr || L.imposed_state_minus == nullptr) {
    setError("sdg_friction_evaluate: missing fault-layer array");
    return SDG_ERR_INVALID;
  }
  if (params->nucleation_count > 1) {
    setError("sdg_friction_evaluate: at most one nucleation");
    return SDG_ERR_UNSUPPORTED;
  }
  sdg_friction_options opt{};
  if (options != nullptr) opt = *options;
  if (opt.samples != nullptr && opt.sample_count == nullptr) {
    setError("sdg_friction_evaluate: sample recording needs sample_count");
    return SDG_ERR_INVALID;
  }
  friction::StepArgs args{};
  for (int o = 0; o < L.order; ++o) {
    args.deltaT[o] = delta_t[o];
    args.timeWeights[o] = time_weights[o];
  }
  auto st = static_cast<hipStream_t>(stream);
  if (L.num_points <= 32) {
    const unsigned grid = static_cast<unsigned>((L.num_faces + 7) / 8);
    friction::kFriction<32><<<grid, 256, 0, st>>>(*params, L, full_update_time, args, resample, opt);
  } else {
    const unsigned grid = static_cast<unsigned>((L.num_faces + 3) / 4);
    friction::kFriction<64><<<grid, 256, 0, st>>>(*params, L, full_update_time, args, resample, opt);
  }
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}
