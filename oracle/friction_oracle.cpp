// ORACLE -- test infrastructure only.
//
// CPU FP64 restatement of the rate-and-state fast-velocity-weakening friction
// solve (friction law 103) of SeisSol's CPU implementation:
//   driver              src/DynamicRupture/FrictionLaws/CpuImpl/BaseFrictionLaw.h:49-189
//   rate-and-state      src/DynamicRupture/FrictionLaws/CpuImpl/RateAndState.h:40-416
//   FVW law             src/DynamicRupture/FrictionLaws/CpuImpl/FastVelocityWeakeningLaw.h:43-165
//   overflow-safe math  src/DynamicRupture/FrictionLaws/RateAndStateCommon.h:22-139
//   pre/post compute    src/DynamicRupture/FrictionLaws/FrictionSolverCommon.h:146-631
//   nucleation          src/Numerical/GaussianNucleationFunction.h
//   initial state       src/DynamicRupture/Initializer/RateAndStateInitializer.cpp:154-180
// Convergence checks are face-wide over the real Gauss points, as in the CPU
// implementation (RateAndState.h:209-222, 375-386).  No thermal
// pressurization (NoTP), elastic material.  Counts residual evaluations and
// Newton steps per point.  Option closed_form_start: a point whose closed-form
// creep root lies below half the slip-rate tolerance starts the first solve of
// a time point there (the library's warm-start option without the network).
//
// Also: a FP64 forward/backward of the 6-32-32-1 warm-start MLP used to check
// the GPU training kernels' gradients.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

typedef struct {
  int64_t num_faces;
  int32_t points;
  int32_t num_points;
  int32_t order;
  int32_t pad_;
  const double* imp_and_eta;
  const double* q_plus;
  const double* q_minus;
  double* initial_stress;
  double* initial_pressure;
  const double* nucleation_stress;
  const double* nucleation_pressure;
  const double* rs_a;
  const double* rs_sl0;
  const double* rs_f0;
  const double* rs_muw;
  const double* rs_b;
  const double* rs_srw;
  double* state_variable;
  double* slip_rate1;
  double* slip_rate2;
  double* slip_rate_magnitude;
  double* mu;
  double* traction1;
  double* traction2;
  double* slip1;
  double* slip2;
  double* accumulated_slip;
  double* rupture_time;
  double* peak_slip_rate;
  double* dyn_stress_time;
  uint8_t* rupture_time_pending;
  uint8_t* dyn_stress_time_pending;
  uint8_t* convergence_inner;
  uint8_t* convergence_outer;
  double* imposed_state_plus;
  double* imposed_state_minus;
  int32_t* newton_iterations;  // residual evaluations
  int32_t* slip_rate_updates;  // Newton steps
} OracleFaultLayer;

typedef struct {
  double rs_sr0;
  double rs_slip_rate_tolerance;
  double rs_state_tolerance;
  int32_t rs_max_slip_rate_updates;
  int32_t rs_state_variable_updates;
  double eta_damp;
  double eta_damp_end;
  int32_t nucleation_count;
  int32_t closed_form_start;  // creeping points start at the closed-form creep root
  double t0;
  double s0;
} OracleFrictionParams;

}  // extern "C"

namespace {

constexpr double kAlmostZero = 1e-45;  // RateAndStateCommon.h:22-31 (double)

// asinh(x * exp(c)) without overflow (RateAndStateCommon.h:46-79)
double arsinhexp(double x, double expLog, double ex) {
  const double sw = 10.0, threshold = 50.0, log2 = 0.69314718055994530943;
  int xexp = 0;
  (void)std::frexp(x, &xexp);
  if (expLog + std::max(xexp, 0) * log2 > sw || expLog >= threshold) {
    if (expLog <= 0) ex = 1 / ex;
    const double xa = std::fabs(x);
    const double xs = x >= 0 ? 1.0 : -1.0;
    return xs * (expLog + std::log(xa + std::sqrt(xa * xa + ex * ex)));
  }
  if (expLog > 0) ex = 1 / ex;
  return std::asinh(ex * x);
}

double computeCExp(double c) { return c > 0 ? std::exp(-c) : std::exp(c); }

double arsinhexpDerivative(double x, double expLog, double ex) {
  const double sw = 10.0, threshold = 50.0, log2 = 0.69314718055994530943;
  int xexp = 0;
  (void)std::frexp(x, &xexp);
  if (expLog + std::max(xexp, 0) * log2 > sw || expLog >= threshold) {
    if (expLog <= 0) ex = 1 / ex;
    return 1 / std::sqrt(x * x + ex * ex);
  }
  if (expLog > 0) ex = 1 / ex;
  const double v = ex * x;
  return ex / std::sqrt(1 + v * v);
}

// log(x * sinh(c)) (RateAndStateCommon.h:120-139)
double logsinh(double x, double c) {
  const double sign = c >= 0 ? 1.0 : -1.0;
  const double ac = std::fabs(c);
  return ac + std::log(x / 2 * -sign * std::expm1(-2 * ac));
}

double smoothStep(double t, double t0) {
  if (t <= 0) return 0.0;
  if (t < t0) {
    const double tau = t - t0;
    return std::exp(tau * tau / (t * (t - 2.0 * t0)));
  }
  return 1.0;
}

struct Face {
  const OracleFaultLayer* L;
  const OracleFrictionParams* prm;
  int64_t f;
  int P, NP, O;
  size_t at(int p) const { return static_cast<size_t>(f) * P + p; }
  double a(int p) const { return L->rs_a[at(p)]; }
  double is(int c, int p) const { return L->initial_stress[(static_cast<size_t>(f) * 6 + c) * P + p]; }

  // FastVelocityWeakeningLaw.h:43-77
  double updateStateVariable(int p, double svRef, double dt, double V) const {
    const double muW = L->rs_muw[at(p)], srW = L->rs_srw[at(p)], A = a(p), sl0 = L->rs_sl0[at(p)];
    const double lowV = std::max(0.0, L->rs_f0[at(p)] - (L->rs_b[at(p)] - A) * std::log(V / prm->rs_sr0));
    const double r = V / srW;
    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    const double muSS = muW + (lowV - muW) / std::pow(1.0 + r8, 1.0 / 8.0);
    const double psiSS = A * logsinh(prm->rs_sr0 / V * 2, muSS / A);
    const double pre = -V * (dt / sl0);
    return psiSS * (-std::expm1(pre)) + std::exp(pre) * svRef;
  }
  double mu(int p, double V, double psi) const {
    const double A = a(p), cLin = 0.5 / prm->rs_sr0, cExpLog = psi / A;
    return A * arsinhexp(cLin * V, cExpLog, computeCExp(cExpLog));
  }
  double dmu(int p, double V, double psi) const {
    const double A = a(p), cLin = 0.5 / prm->rs_sr0, cExpLog = psi / A;
    return A * cLin * arsinhexpDerivative(cLin * V, cExpLog, computeCExp(cExpLog));
  }
};

inline double magnitude(double x, double y) { return std::sqrt(x * x + y * y); }

void evaluateFace(const OracleFaultLayer* L, const OracleFrictionParams* prm, int64_t f, double fullUpdateTime,
                  const double* deltaT, const double* timeWeights, const double* resample) {
  Face F{L, prm, f, L->points, L->num_points, L->order};
  const int P = F.P, NP = F.NP, O = F.O;
  const double* ie = L->imp_and_eta + f * 11;
  const double etaS = ie[5], invEtaS = ie[6], invZp = ie[7], invZs = ie[8], invZpN = ie[9], invZsN = ie[10];
  const double etaPDamp = prm->eta_damp_end > fullUpdateTime ? prm->eta_damp : 1.0;
  const double etaP = ie[4] * etaPDamp;
  auto qp = [&](int o, int q, int p) { return L->q_plus[((static_cast<size_t>(f) * O + o) * 9 + q) * P + p]; };
  auto qm = [&](int o, int q, int p) { return L->q_minus[((static_cast<size_t>(f) * O + o) * 9 + q) * P + p]; };

  // precomputeStressFromQInterpolated (FrictionSolverCommon.h:146-195)
  std::vector<double> fsN(O * NP), fsT1(O * NP), fsT2(O * NP), trT1(O * NP), trT2(O * NP);
  for (int o = 0; o < O; ++o)
    for (int p = 0; p < NP; ++p) {
      fsN[o * NP + p] = etaP * (qm(o, 6, p) - qp(o, 6, p) + qp(o, 0, p) * invZp + qm(o, 0, p) * invZpN);
      fsT1[o * NP + p] = etaS * (qm(o, 7, p) - qp(o, 7, p) + qp(o, 3, p) * invZs + qm(o, 3, p) * invZsN);
      fsT2[o * NP + p] = etaS * (qm(o, 8, p) - qp(o, 8, p) + qp(o, 5, p) * invZs + qm(o, 5, p) * invZsN);
    }
  std::vector<double> sv(NP);
  for (int p = 0; p < NP; ++p) sv[p] = L->state_variable[F.at(p)];  // preHook

  double updateTime = fullUpdateTime;
  std::vector<double> normal(NP), absShear(NP), localSR(NP), svRef(NP), test(NP), g(NP), muF(NP);
  std::vector<uint8_t> convOuterPre(NP);
  for (int o = 0; o < O; ++o) {
    const double dt = deltaT[o];
    const double startTime = updateTime;
    updateTime += dt;
    // adjustInitialStress (FrictionSolverCommon.h:503-531)
    if (prm->nucleation_count > 0 && L->nucleation_stress != nullptr) {
      if (updateTime <= prm->t0 + prm->s0 && updateTime >= prm->s0) {
        const double gNuc = smoothStep(updateTime - prm->s0, prm->t0) - smoothStep(updateTime - prm->s0 - dt, prm->t0);
        for (int p = 0; p < NP; ++p) {
          for (int c = 0; c < 6; ++c)
            L->initial_stress[(static_cast<size_t>(f) * 6 + c) * P + p] +=
                L->nucleation_stress[(static_cast<size_t>(f) * 6 + c) * P + p] * gNuc;
          if (L->nucleation_pressure) L->initial_pressure[F.at(p)] += L->nucleation_pressure[F.at(p)] * gNuc;
        }
      }
    }
    // calcInitialVariables (RateAndState.h:120-156)
    for (int p = 0; p < NP; ++p) {
      normal[p] = std::min(0.0, fsN[o * NP + p] + F.is(0, p) + L->initial_pressure[F.at(p)]);
      svRef[p] = sv[p];
      absShear[p] = magnitude(F.is(3, p) + fsT1[o * NP + p], F.is(5, p) + fsT2[o * NP + p]);
      double srm = magnitude(L->slip_rate1[F.at(p)], L->slip_rate2[F.at(p)]);
      srm = std::max(kAlmostZero, srm);
      L->slip_rate_magnitude[F.at(p)] = srm;
      localSR[p] = srm;
    }
    // updateStateVariableIterative (RateAndState.h:159-242)
    for (int j = 0; j < prm->rs_state_variable_updates; ++j) {
      for (int p = 0; p < NP; ++p) sv[p] = F.updateStateVariable(p, svRef[p], dt, localSR[p]);
      for (int p = 0; p < NP; ++p)
        normal[p] = std::min(0.0, fsN[o * NP + p] + F.is(0, p) + L->initial_pressure[F.at(p)]);
      // invertSlipRateIterative (RateAndState.h:339-416): face-wide convergence
      for (int p = 0; p < NP; ++p) test[p] = L->slip_rate_magnitude[F.at(p)];
      if (prm->closed_form_start && j == 0)
        for (int p = 0; p < NP; ++p) {
          // V_c = 2 V0 exp(-psi/a) sinh(S/N) solves N asinh(V exp(psi/a) / (2 V0)) = S,
          // i.e. g(V_c) = -V_c: below half the tolerance the first test passes there
          const double S = invEtaS * absShear[p], N = invEtaS * std::fabs(normal[p]) * F.a(p);
          if (S > 0.0 && N > 0.0) {
            const double c = S / N;
            const double vc = std::exp(std::log(2.0 * prm->rs_sr0) - sv[p] / F.a(p) + c + std::log(-0.5 * std::expm1(-2.0 * c)));
            if (vc < 0.5 * prm->rs_slip_rate_tolerance) test[p] = std::max(kAlmostZero, vc);
          }
        }
      bool converged = false;
      auto evalAll = [&]() {
        for (int p = 0; p < NP; ++p) {
          muF[p] = F.mu(p, test[p], sv[p]);
          g[p] = -invEtaS * (std::fabs(normal[p]) * muF[p] - absShear[p]) - test[p];
        }
        if (L->newton_iterations)
          for (int p = 0; p < NP; ++p) L->newton_iterations[F.at(p)] += 1;
        bool all = true;
        for (int p = 0; p < NP; ++p) all = all && std::fabs(g[p]) < prm->rs_slip_rate_tolerance;
        return all;
      };
      for (int it = 0; it < prm->rs_max_slip_rate_updates; ++it) {
        converged = evalAll();
        if (converged) break;
        if (L->slip_rate_updates)
          for (int p = 0; p < NP; ++p) L->slip_rate_updates[F.at(p)] += 1;
        for (int p = 0; p < NP; ++p) {
          const double dG = -invEtaS * (std::fabs(normal[p]) * F.dmu(p, test[p], sv[p])) - 1.0;
          test[p] = std::max(kAlmostZero, test[p] - g[p] / dG);
        }
      }
      if (converged)
        for (int p = 0; p < NP; ++p) L->mu[F.at(p)] = muF[p];
      else
        for (int p = 0; p < NP; ++p)
          L->convergence_inner[F.at(p)] &= std::fabs(g[p]) < prm->rs_slip_rate_tolerance ? 1 : 0;
      bool all = true;
      for (int p = 0; p < NP; ++p) {
        const double old = L->slip_rate_magnitude[F.at(p)];
        localSR[p] = 0.5 * (old + test[p]);
        const bool pc = std::fabs(test[p] - old) < prm->rs_state_tolerance;
        all = all && pc;
        convOuterPre[p] = pc;
        L->slip_rate_magnitude[F.at(p)] = test[p];
      }
      if (all) break;
    }
    for (int p = 0; p < NP; ++p) L->convergence_outer[F.at(p)] &= convOuterPre[p];
    for (int p = 0; p < NP; ++p)
      normal[p] = std::min(0.0, fsN[o * NP + p] + F.is(0, p) + L->initial_pressure[F.at(p)]);
    // calcSlipRateAndTraction (RateAndState.h:244-307)
    for (int p = 0; p < NP; ++p) {
      const size_t k = F.at(p);
      sv[p] = F.updateStateVariable(p, svRef[p], dt, localSR[p]);
      const double srm = L->slip_rate_magnitude[k];
      L->mu[k] = F.mu(p, srm, sv[p]);
      const double strength = -L->mu[k] * normal[p];
      const double t1 = F.is(3, p) + fsT1[o * NP + p];
      const double t2 = F.is(5, p) + fsT2[o * NP + p];
      const double divisor = strength + etaS * srm;
      L->slip_rate1[k] = srm * t1 / divisor;
      L->slip_rate2[k] = srm * t2 / divisor;
      trT1[o * NP + p] = fsT1[o * NP + p] - etaS * L->slip_rate1[k];
      trT2[o * NP + p] = fsT2[o * NP + p] - etaS * L->slip_rate2[k];
      L->traction1[k] = trT1[o * NP + p];
      L->traction2[k] = trT2[o * NP + p];
      L->accumulated_slip[k] += srm * dt;
      L->slip1[k] += L->slip_rate1[k] * dt;
      L->slip2[k] += L->slip_rate2[k] * dt;
    }
    // outputs (FrictionSolverCommon.h:546-598, RateAndState.h:309-324)
    for (int p = 0; p < NP; ++p) {
      const size_t k = F.at(p);
      if (L->rupture_time_pending[k] && L->slip_rate_magnitude[k] > 0.001) {
        L->rupture_time[k] = startTime;
        L->rupture_time_pending[k] = 0;
      }
      if (L->rupture_time[k] > 0.0 && L->rupture_time[k] <= startTime && L->dyn_stress_time_pending[k] &&
          L->mu[k] <= L->rs_muw[k] + 0.05 * (L->rs_f0[k] - L->rs_muw[k])) {
        L->dyn_stress_time[k] = startTime;
        L->dyn_stress_time_pending[k] = 0;
      }
      L->peak_slip_rate[k] = std::max(L->peak_slip_rate[k], L->slip_rate_magnitude[k]);
    }
  }
  // postHook: resampleStateVar (FastVelocityWeakeningLaw.h:141-165)
  std::vector<double> delta(NP);
  for (int p = 0; p < NP; ++p) delta[p] = sv[p] - L->state_variable[F.at(p)];
  for (int p = 0; p < NP; ++p) {
    double acc = 0.0;
    for (int q = 0; q < NP; ++q) acc += resample[p * NP + q] * delta[q];
    L->state_variable[F.at(p)] += acc;
  }
  // postcomputeImposedStateFromNewStress (FrictionSolverCommon.h:333-401)
  for (int p = 0; p < NP; ++p) {
    double m[9] = {0}, pl[9] = {0};
    for (int o = 0; o < O; ++o) {
      const double w = timeWeights[o];
      const double ns = fsN[o * NP + p], t1 = trT1[o * NP + p], t2 = trT2[o * NP + p];
      m[0] += w * ns; m[3] += w * t1; m[5] += w * t2;
      m[6] += w * (qm(o, 6, p) - invZpN * (ns - qm(o, 0, p)));
      m[7] += w * (qm(o, 7, p) - invZsN * (t1 - qm(o, 3, p)));
      m[8] += w * (qm(o, 8, p) - invZsN * (t2 - qm(o, 5, p)));
      pl[0] += w * ns; pl[3] += w * t1; pl[5] += w * t2;
      pl[6] += w * (qp(o, 6, p) + invZp * (ns - qp(o, 0, p)));
      pl[7] += w * (qp(o, 7, p) + invZs * (t1 - qp(o, 3, p)));
      pl[8] += w * (qp(o, 8, p) + invZs * (t2 - qp(o, 5, p)));
    }
    for (int q = 0; q < 9; ++q) {
      L->imposed_state_minus[(static_cast<size_t>(f) * 9 + q) * P + p] = m[q];
      L->imposed_state_plus[(static_cast<size_t>(f) * 9 + q) * P + p] = pl[q];
    }
  }
}

}  // namespace

extern "C" {

int oracle_friction_evaluate(const OracleFaultLayer* L, const OracleFrictionParams* prm, double fullUpdateTime,
                             const double* deltaT, const double* timeWeights, const double* resample) {
#pragma omp parallel for schedule(static)
  for (int64_t f = 0; f < L->num_faces; ++f) evaluateFace(L, prm, f, fullUpdateTime, deltaT, timeWeights, resample);
  return 0;
}

double oracle_arsinhexp(double x, double expLog) { return arsinhexp(x, expLog, computeCExp(expLog)); }
double oracle_arsinhexp_derivative(double x, double expLog) {
  return arsinhexpDerivative(x, expLog, computeCExp(expLog));
}
double oracle_logsinh(double x, double c) { return logsinh(x, c); }

// RateAndStateFastVelocityInitializer::computeInitialStateAndFriction
// (RateAndStateInitializer.cpp:154-180): state and friction for a point
// sliding steadily at the initial slip rate under the given tractions.
int oracle_rs_initial_state(double traction1, double traction2, double pressure, double rsA, double rsSr0,
                            double initialSlipRate, double* stateVariable, double* frictionCoefficient) {
  const double absTraction = magnitude(traction1, traction2);
  const double tmp = std::fabs(absTraction / (rsA * pressure));
  *stateVariable = rsA * logsinh(2.0 * rsSr0 / initialSlipRate, tmp);
  const double explog = *stateVariable / rsA;
  *frictionCoefficient = rsA * arsinhexp(initialSlipRate * 0.5 / rsSr0, explog, computeCExp(explog));
  return 0;
}

// FrictionSolver::computeDeltaT (FrictionSolver.cpp:20-35)
int oracle_compute_delta_t(int n, const double* timePoints, double* deltaT) {
  deltaT[0] = timePoints[0];
  for (int i = 1; i < n; ++i) deltaT[i] = timePoints[i] - timePoints[i - 1];
  deltaT[n - 1] += timePoints[0];
  return 0;
}

// Warm-start MLP 6-32-32-1 (tanh hidden layers, linear output), FP64.
// params: W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3[1].
// loss = mean (y - t)^2 over the batch; grads has the same layout as params.
double oracle_mlp_loss_grad(int64_t n, const double* x, const double* t, const double* params, double* grads) {
  const double* W1 = params;
  const double* b1 = W1 + 32 * 6;
  const double* W2 = b1 + 32;
  const double* b2 = W2 + 32 * 32;
  const double* W3 = b2 + 32;
  const double* b3 = W3 + 32;
  const int total = 32 * 6 + 32 + 32 * 32 + 32 + 32 + 1;
  std::fill(grads, grads + total, 0.0);
  double* gW1 = grads;
  double* gb1 = gW1 + 32 * 6;
  double* gW2 = gb1 + 32;
  double* gb2 = gW2 + 32 * 32;
  double* gW3 = gb2 + 32;
  double* gb3 = gW3 + 32;
  double loss = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double* xi = x + 6 * i;
    double h1[32], h2[32];
    for (int k = 0; k < 32; ++k) {
      double s = b1[k];
      for (int j = 0; j < 6; ++j) s += W1[k * 6 + j] * xi[j];
      h1[k] = std::tanh(s);
    }
    for (int k = 0; k < 32; ++k) {
      double s = b2[k];
      for (int j = 0; j < 32; ++j) s += W2[k * 32 + j] * h1[j];
      h2[k] = std::tanh(s);
    }
    double y = b3[0];
    for (int j = 0; j < 32; ++j) y += W3[j] * h2[j];
    const double r = y - t[i];
    loss += r * r;
    const double dy = 2.0 * r / n;
    double d2[32], d1[32];
    gb3[0] += dy;
    for (int j = 0; j < 32; ++j) {
      gW3[j] += dy * h2[j];
      d2[j] = dy * W3[j] * (1.0 - h2[j] * h2[j]);
    }
    for (int k = 0; k < 32; ++k) {
      gb2[k] += d2[k];
      for (int j = 0; j < 32; ++j) gW2[k * 32 + j] += d2[k] * h1[j];
    }
    for (int j = 0; j < 32; ++j) {
      double s = 0.0;
      for (int k = 0; k < 32; ++k) s += W2[k * 32 + j] * d2[k];
      d1[j] = s * (1.0 - h1[j] * h1[j]);
    }
    for (int k = 0; k < 32; ++k) {
      gb1[k] += d1[k];
      for (int j = 0; j < 6; ++j) gW1[k * 6 + j] += d1[k] * xi[j];
    }
  }
  return loss / n;
}

}  // extern "C"

extern "C" {

// precomputeStressFromQInterpolated (FrictionSolverCommon.h:146-195) for one
// face: q [O][9][P]; imp: zp zs zpNeig zsNeig etaP etaS invEtaS invZp invZs
// invZpNeig invZsNeig; outputs [O][P].
int oracle_precompute_stress(int O, int P, const double* imp, const double* qPlus, const double* qMinus,
                             double etaPDamp, double* normal, double* t1, double* t2) {
  const double etaP = imp[4] * etaPDamp, etaS = imp[5];
  for (int o = 0; o < O; ++o)
    for (int p = 0; p < P; ++p) {
      auto qp = [&](int q) { return qPlus[(o * 9 + q) * P + p]; };
      auto qm = [&](int q) { return qMinus[(o * 9 + q) * P + p]; };
      normal[o * P + p] = etaP * (qm(6) - qp(6) + qp(0) * imp[7] + qm(0) * imp[9]);
      t1[o * P + p] = etaS * (qm(7) - qp(7) + qp(3) * imp[8] + qm(3) * imp[10]);
      t2[o * P + p] = etaS * (qm(8) - qp(8) + qp(5) * imp[8] + qm(5) * imp[10]);
    }
  return 0;
}

// postcomputeImposedStateFromNewStress (FrictionSolverCommon.h:333-401) for
// one face: normal, traction1, traction2 [O][P]; outputs [9][P].
int oracle_postcompute_imposed(int O, int P, const double* imp, const double* normal, const double* t1,
                               const double* t2, const double* qPlus, const double* qMinus, const double* w,
                               double* plus, double* minus) {
  for (int i = 0; i < 9 * P; ++i) plus[i] = minus[i] = 0.0;
  for (int o = 0; o < O; ++o)
    for (int p = 0; p < P; ++p) {
      auto qp = [&](int q) { return qPlus[(o * 9 + q) * P + p]; };
      auto qm = [&](int q) { return qMinus[(o * 9 + q) * P + p]; };
      const double ns = normal[o * P + p], a = t1[o * P + p], b = t2[o * P + p];
      minus[0 * P + p] += w[o] * ns;
      minus[3 * P + p] += w[o] * a;
      minus[5 * P + p] += w[o] * b;
      minus[6 * P + p] += w[o] * (qm(6) - imp[9] * (ns - qm(0)));
      minus[7 * P + p] += w[o] * (qm(7) - imp[10] * (a - qm(3)));
      minus[8 * P + p] += w[o] * (qm(8) - imp[10] * (b - qm(5)));
      plus[0 * P + p] += w[o] * ns;
      plus[3 * P + p] += w[o] * a;
      plus[5 * P + p] += w[o] * b;
      plus[6 * P + p] += w[o] * (qp(6) + imp[7] * (ns - qp(0)));
      plus[7 * P + p] += w[o] * (qp(7) + imp[8] * (a - qp(3)));
      plus[8 * P + p] += w[o] * (qp(8) + imp[8] * (b - qp(5)));
    }
  return 0;
}

}  // extern "C"
