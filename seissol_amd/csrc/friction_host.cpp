// Host-side helpers of the friction solver: time points of a DR step, the
// Stroud rule on the reference triangle with its resampling matrix, and the
// initial rate-and-state state.
#include <cmath>
#include <vector>

#include "../../include/seissol_amd.h"
#include "global_matrices.h"

namespace sdg {
void setError(const std::string& msg);
namespace {

// Jacobi polynomial P_n^{(a,b)}(x) and its derivative by the three-term recurrence
void jacobi(int n, double a, double b, double x, double* p, double* dp) {
  double p0 = 1.0, p1 = 0.5 * (a - b + (a + b + 2.0) * x);
  if (n == 0) { *p = 1.0; *dp = 0.0; return; }
  double d0 = 0.0, d1 = 0.5 * (a + b + 2.0);
  for (int m = 2; m <= n; ++m) {
    const double c = 2.0 * m + a + b;
    const double a1 = 2.0 * m * (m + a + b) * (c - 2.0);
    const double a2 = (c - 1.0) * (a * a - b * b);
    const double a3 = (c - 2.0) * (c - 1.0) * c;
    const double a4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * c;
    const double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
    const double d2 = ((a2 + a3 * x) * d1 + a3 * p1 - a4 * d0) / a1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  *p = p1;
  *dp = d1;
}

// roots of P_n^{(a,b)} on (-1, 1), ascending (Newton with deflation)
std::vector<double> jacobiRoots(int n, double a, double b) {
  std::vector<double> r(n);
  for (int k = 0; k < n; ++k) {
    double x = -std::cos(M_PI * (k + 0.5) / n);
    if (k > 0) x = 0.5 * (x + r[k - 1]) > r[k - 1] ? x : r[k - 1] + 1e-3;
    for (int it = 0; it < 200; ++it) {
      double p, dp;
      jacobi(n, a, b, x, &p, &dp);
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += 1.0 / (x - r[j]);
      const double dx = p / (dp - p * s);
      x -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    r[k] = x;
  }
  std::sort(r.begin(), r.end());
  return r;
}

}  // namespace
}  // namespace sdg

#include <algorithm>
#include <string>

using namespace sdg;

extern "C" {

int sdg_friction_delta_t(int order, const double* time_points, double* delta_t) {
  if (order < 1 || time_points == nullptr || delta_t == nullptr) {
    setError("sdg_friction_delta_t: invalid arguments");
    return SDG_ERR_INVALID;
  }
  delta_t[0] = time_points[0];
  for (int i = 1; i < order; ++i) delta_t[i] = time_points[i] - time_points[i - 1];
  delta_t[order - 1] += time_points[0];  // the segment [last point, dt] goes to the last point
  return SDG_OK;
}

int sdg_gauss_legendre(int n, double t0, double t1, double* points, double* weights) {
  if (n < 1 || points == nullptr || weights == nullptr) {
    setError("sdg_gauss_legendre: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const auto x = jacobiRoots(n, 0.0, 0.0);
  for (int k = 0; k < n; ++k) {
    double p, dp;
    jacobi(n, 0.0, 0.0, x[k], &p, &dp);
    points[k] = t0 + 0.5 * (x[k] + 1.0) * (t1 - t0);
    weights[k] = 0.5 * (t1 - t0) * 2.0 / ((1.0 - x[k] * x[k]) * dp * dp);
  }
  return SDG_OK;
}

int sdg_dr_quadrature(int order, double* points, double* weights, double* resample) {
  if (order < 1 || order > 8 || points == nullptr || weights == nullptr) {
    setError("sdg_dr_quadrature: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const int n = order + 1;
  // collapsed rule: chi ~ Gauss-Jacobi (1, 0) with weight (1 - chi), tau = (1 - chi) s, s ~ Gauss-Legendre
  const auto u = jacobiRoots(n, 1.0, 0.0);
  const auto g = jacobiRoots(n, 0.0, 0.0);
  std::vector<double> chi(n), wc(n), s(n), ws(n);
  for (int k = 0; k < n; ++k) {
    double p, dp;
    jacobi(n, 1.0, 0.0, u[k], &p, &dp);
    chi[k] = 0.5 * (1.0 + u[k]);
    wc[k] = 0.25 * 4.0 / ((1.0 - u[k] * u[k]) * dp * dp);
    jacobi(n, 0.0, 0.0, g[k], &p, &dp);
    s[k] = 0.5 * (1.0 + g[k]);
    ws[k] = 0.5 * 2.0 / ((1.0 - g[k] * g[k]) * dp * dp);
  }
  // order of the reference data: chi descending, then s descending
  int idx = 0;
  for (int i = n - 1; i >= 0; --i)
    for (int j = n - 1; j >= 0; --j, ++idx) {
      points[2 * idx] = chi[i];
      points[2 * idx + 1] = (1.0 - chi[i]) * s[j];
      weights[idx] = wc[i] * ws[j];
    }
  if (resample != nullptr) {
    // projection onto the face polynomials of degree < order, evaluated at the points:
    // R[i][j] = sum_m psi_m(x_i) psi_m(x_j) w_j / M2_m
    const int np = n * n, b2 = numBasis2(order);
    std::vector<double> psi(static_cast<size_t>(np) * b2), m2(b2, 0.0);
    for (int q = 0; q < np; ++q)
      for (int m = 0; m < b2; ++m) psi[q * b2 + m] = evalTriBasis(order, m, points[2 * q], points[2 * q + 1]);
    for (int m = 0; m < b2; ++m)
      for (int q = 0; q < np; ++q) m2[m] += weights[q] * psi[q * b2 + m] * psi[q * b2 + m];
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < np; ++j) {
        double acc = 0.0;
        for (int m = 0; m < b2; ++m) acc += psi[i * b2 + m] * psi[j * b2 + m] / m2[m];
        resample[i * np + j] = acc * weights[j];
      }
  }
  return SDG_OK;
}

int sdg_rs_initial_state(double traction1, double traction2, double pressure, double rs_a, double rs_sr0,
                         double initial_slip_rate, double* state_variable, double* friction_coefficient) {
  if (state_variable == nullptr || friction_coefficient == nullptr || rs_a <= 0 || initial_slip_rate <= 0) {
    setError("sdg_rs_initial_state: invalid arguments");
    return SDG_ERR_INVALID;
  }
  // steady sliding at the initial slip rate under the given tractions:
  // psi = a log(2 V0 / V sinh(|tau| / (a |sigma_n|))), mu = a asinh(V / (2 V0) exp(psi / a))
  const double absTraction = std::sqrt(traction1 * traction1 + traction2 * traction2);
  const double c = std::fabs(absTraction / (rs_a * pressure));
  const double x = 2.0 * rs_sr0 / initial_slip_rate;
  const double psi = rs_a * (std::fabs(c) + std::log(x / 2 * -(c >= 0 ? 1.0 : -1.0) * std::expm1(-2 * std::fabs(c))));
  *state_variable = psi;
  const double cExpLog = psi / rs_a;
  const double v = initial_slip_rate * 0.5 / rs_sr0;
  // asinh(v exp(cExpLog)) in the overflow-safe form
  if (cExpLog > 10.0) {
    *friction_coefficient = rs_a * (cExpLog + std::log(v + std::sqrt(v * v + std::exp(-2.0 * cExpLog))));
  } else {
    *friction_coefficient = rs_a * std::asinh(v * std::exp(cExpLog));
  }
  return SDG_OK;
}

}  // extern "C"

// Face evaluation matrices of dynamic rupture (codegen/kernels/dynamic_rupture.py:
// V3mTo2n(i, h), V3mTo2nTWDivM(i, h); data of dr_stroud_matrices_{order}.json):
// v3m [4][4][n][B]: the 3D basis at the n Stroud points of face i, seen with
// orientation h - 1 (h = 0: the face's own parameterisation, the plus side);
// v3mtw [4][4][B][n] = v3m^T diag(w) / M (the doubled surface area of the
// physical face enters through the flux solver).
extern "C" int sdg_dr_face_matrices(int order, double* v3m, double* v3mtw) {
  if (order < 1 || order > 8 || v3m == nullptr) {
    setError("sdg_dr_face_matrices: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const int n = (order + 1) * (order + 1), b = numBasis3(order);
  std::vector<double> pts(2 * n), w(n);
  if (sdg_dr_quadrature(order, pts.data(), w.data(), nullptr) != SDG_OK) return SDG_ERR_INVALID;
  std::vector<double> gm(globalMatrixLayout(order).total);
  generateGlobalMatrices(order, gm.data());
  const double* mass = gm.data() + globalMatrixLayout(order).mass;
  for (int i = 0; i < 4; ++i)
    for (int h = 0; h < 4; ++h)
      for (int k = 0; k < n; ++k) {
        double x[3];
        faceToReference(i, pts[2 * k], pts[2 * k + 1], h - 1, x);
        for (int l = 0; l < b; ++l) {
          const double v = evalTetBasis(order, l, x[0], x[1], x[2]);
          v3m[((i * 4 + h) * n + k) * b + l] = v;
          if (v3mtw != nullptr) v3mtw[((i * 4 + h) * b + l) * n + k] = v * w[k] / mass[l];
        }
      }
  return SDG_OK;
}

