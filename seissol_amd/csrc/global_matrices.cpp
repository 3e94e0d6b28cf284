// Global (element-independent) ADER-DG matrices, generated from the modal
// Dubiner basis on the reference tetrahedron.
//
// The reference ships these matrices as data (codegen/matrices/matrices_{B}.xml,
// read by codegen/kernels/aderdg.py:37-41); this file regenerates them from
// their definitions so no reference data has to travel with the library:
//
//   kDivM_d [k][l]  = (1/M_k) * int_T  d_d(phi_k) * phi_l          (B x B)
//   kDivMT_d[k][l]  = (1/M_k) * int_T  phi_k * d_d(phi_l)          (B x B)
//   rDivM_f [k][m]  = (1/(M_k * M2_m)) * int_F_f phi_k * psi_m     (B x B2)
//   fMrT_f  [m][l]  =                    int_F_f psi_m * phi_l     (B2 x B)
//   rT_f    [m][l]  = (1/M2_m)         * int_F_f psi_m * phi_l     (B2 x B)
//   fP_h    [m][n]  = int_tri psi_m(chi,tau) * psi_n(P_h(chi,tau)) (B2 x B2)
//
// with M_k = int_T phi_k^2 (the 3D mass matrix is diagonal) and M2_m the
// diagonal 2D mass matrix.  Face parameterisation and orientation maps follow
// src/Numerical/Transformation.cpp:300-349 (chiTau2XiEtaZeta); basis ordering
// follows src/Numerical/BasisFunction.h:133-143 (3D) and the matching 2D order.
// Gradients come from forward-mode dual numbers, integrals from a collapsed
// (Duffy) Gauss-Legendre rule that is exact for the polynomial degrees used.
// The definitions were checked entry by entry against the reference XML files
// (tests/test_global_matrices.py).
#include "global_matrices.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace sdg {
namespace {

// Value plus gradient in (xi, eta, zeta).
struct Dual3 {
  double v, d[3];
  Dual3(double x = 0.0) : v(x), d{0.0, 0.0, 0.0} {}
};
inline Dual3 operator+(const Dual3& a, const Dual3& b) {
  Dual3 r(a.v + b.v);
  for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
inline Dual3 operator-(const Dual3& a, const Dual3& b) {
  Dual3 r(a.v - b.v);
  for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
inline Dual3 operator*(const Dual3& a, const Dual3& b) {
  Dual3 r(a.v * b.v);
  for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
inline Dual3 operator*(double s, const Dual3& a) {
  Dual3 r(s * a.v);
  for (int i = 0; i < 3; ++i) r.d[i] = s * a.d[i];
  return r;
}

// Homogenised Jacobi polynomial y^n * P_n^{(a,b)}(x / y), evaluated by the
// three-term recurrence multiplied through by powers of y; it stays regular at
// y = 0, which is where the collapsed coordinates of the simplex degenerate.
template <typename T>
T jacobiHomogeneous(int n, int a, int b, const T& x, const T& y) {
  if (n == 0) return T(1.0);
  T prev2(0.0);
  T prev1(1.0);
  T cur = (0.5 * (a - b)) * y + (1.0 + 0.5 * (a + b)) * x;
  for (int m = 2; m <= n; ++m) {
    const double s = 2.0 * m + a + b;
    const double c1 = (s - 1.0);
    const double c2 = static_cast<double>(a * a - b * b);
    const double c3 = s * (s - 2.0);
    const double c4 = 2.0 * (m + a - 1.0) * (m + b - 1.0) * s;
    const double c5 = 2.0 * m * (m + a + b) * (s - 2.0);
    prev2 = prev1;
    prev1 = cur;
    cur = (1.0 / c5) * (c1 * ((c2 * y + c3 * x) * prev1) - c4 * ((y * y) * prev2));
  }
  return cur;
}

// Orthogonal (Dubiner) basis function with multi-index (i, j, k) on the
// tetrahedron {xi, eta, zeta >= 0, xi + eta + zeta <= 1}.
template <typename T>
T tetBasis(int i, int j, int k, const T& xi, const T& eta, const T& zeta) {
  const T one(1.0);
  const T r = 2.0 * xi - one + eta + zeta;
  const T s = 2.0 * eta - one + zeta;
  const T t = 2.0 * zeta - one;
  const T sigma = one - eta - zeta;
  const T theta = one - zeta;
  return jacobiHomogeneous(i, 0, 0, r, sigma) * jacobiHomogeneous(j, 2 * i + 1, 0, s, theta) *
         jacobiHomogeneous(k, 2 * i + 2 * j + 2, 0, t, one);
}

// Orthogonal basis on the triangle {chi, tau >= 0, chi + tau <= 1}.
double triBasis(int i, int j, double chi, double tau) {
  const double r = 2.0 * chi - 1.0 + tau;
  const double s = 2.0 * tau - 1.0;
  return jacobiHomogeneous(i, 0, 0, r, 1.0 - tau) * jacobiHomogeneous(j, 2 * i + 1, 0, s, 1.0);
}

struct Idx3 {
  int i, j, k;
};
struct Idx2 {
  int i, j;
};

std::vector<Idx3> tetIndices(int order) {
  std::vector<Idx3> out;
  for (int deg = 0; deg < order; ++deg)
    for (int k = 0; k <= deg; ++k)
      for (int j = 0; j <= deg - k; ++j) out.push_back({deg - j - k, j, k});
  return out;
}

std::vector<Idx2> triIndices(int order) {
  std::vector<Idx2> out;
  for (int deg = 0; deg < order; ++deg)
    for (int j = 0; j <= deg; ++j) out.push_back({deg - j, j});
  return out;
}

// Gauss-Legendre nodes/weights on [0, 1] (Newton on P_n, Tricomi initial guess).
void gaussLegendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int m = 2; m <= n; ++m) {
        const double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      const double dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int m = 2; m <= n; ++m) {
      const double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
      p0 = p1;
      p1 = p2;
    }
    const double dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);  // = 0.5 * 2/((1-z^2) dp^2)
  }
}

struct TetRule {
  std::vector<double> xi, eta, zeta, w;
};
struct TriRule {
  std::vector<double> chi, tau, w;
};

// Collapsed product rule: xi = a(1-b)(1-c), eta = b(1-c), zeta = c.
TetRule tetRule(int n) {
  std::vector<double> g, gw;
  gaussLegendre01(n, g, gw);
  TetRule r;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int c = 0; c < n; ++c) {
        r.xi.push_back(g[a] * (1.0 - g[b]) * (1.0 - g[c]));
        r.eta.push_back(g[b] * (1.0 - g[c]));
        r.zeta.push_back(g[c]);
        r.w.push_back(gw[a] * gw[b] * gw[c] * (1.0 - g[b]) * (1.0 - g[c]) * (1.0 - g[c]));
      }
  return r;
}

TriRule triRule(int n) {
  std::vector<double> g, gw;
  gaussLegendre01(n, g, gw);
  TriRule r;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      r.chi.push_back(g[a] * (1.0 - g[b]));
      r.tau.push_back(g[b]);
      r.w.push_back(gw[a] * gw[b] * (1.0 - g[b]));
    }
  return r;
}

}  // namespace

void faceToReference(int face, double chi, double tau, int orientation, double out[3]) {
  double a = chi, b = tau;
  switch (orientation) {
    case 0: a = tau; b = chi; break;
    case 1: a = 1.0 - chi - tau; b = tau; break;
    case 2: a = chi; b = 1.0 - chi - tau; break;
    default: break;
  }
  switch (face) {
    case 0: out[0] = b; out[1] = a; out[2] = 0.0; break;
    case 1: out[0] = a; out[1] = 0.0; out[2] = b; break;
    case 2: out[0] = 0.0; out[1] = b; out[2] = a; break;
    default: out[0] = 1.0 - a - b; out[1] = a; out[2] = b; break;
  }
}

void orientationMap(int orientation, double chi, double tau, double* a, double* b) {
  switch (orientation) {
    case 0: *a = tau; *b = chi; break;
    case 1: *a = 1.0 - chi - tau; *b = tau; break;
    case 2: *a = chi; *b = 1.0 - chi - tau; break;
    default: *a = chi; *b = tau; break;
  }
}

double evalTetBasis(int order, int index, double xi, double eta, double zeta) {
  const auto idx = tetIndices(order);
  const auto& t = idx[index];
  return tetBasis<double>(t.i, t.j, t.k, xi, eta, zeta);
}

void evalTetBasisGradient(int order, int index, double xi, double eta, double zeta, double out[3]) {
  const auto idx = tetIndices(order);
  const auto& t = idx[index];
  Dual3 x(xi), y(eta), z(zeta);
  x.d[0] = 1.0;
  y.d[1] = 1.0;
  z.d[2] = 1.0;
  const Dual3 v = tetBasis<Dual3>(t.i, t.j, t.k, x, y, z);
  for (int d = 0; d < 3; ++d) out[d] = v.d[d];
}

double evalTriBasis(int order, int index, double chi, double tau) {
  const auto idx = triIndices(order);
  return triBasis(idx[index].i, idx[index].j, chi, tau);
}

GlobalMatrixLayout globalMatrixLayout(int order) {
  GlobalMatrixLayout l{};
  const int b3 = numBasis3(order);
  const int b2 = numBasis2(order);
  std::size_t off = 0;
  l.kDivM = off;  off += 3 * b3 * b3;
  l.kDivMT = off; off += 3 * b3 * b3;
  l.rDivM = off;  off += 4 * b3 * b2;
  l.fMrT = off;   off += 4 * b2 * b3;
  l.rT = off;     off += 4 * b2 * b3;
  l.fP = off;     off += 3 * b2 * b2;
  l.mass = off;   off += b3;
  l.mass2 = off;  off += b2;
  l.total = off;
  return l;
}

int generateGlobalMatrices(int order, double* out) {
  if (order < 1 || order > 8) return -1;
  const int b3 = numBasis3(order);
  const int b2 = numBasis2(order);
  const auto layout = globalMatrixLayout(order);
  std::memset(out, 0, layout.total * sizeof(double));

  const auto idx3 = tetIndices(order);
  const auto idx2 = triIndices(order);
  // Degree of every integrand is <= 2(order-1) in each collapsed direction
  // plus at most 2 from the Duffy Jacobian: order+2 points are exact.
  const TetRule tet = tetRule(order + 2);
  const TriRule tri = triRule(order + 2);
  const std::size_t nq = tet.w.size();
  const std::size_t nf = tri.w.size();

  std::vector<double> phi(b3 * nq), dphi(3 * b3 * nq);
  for (int k = 0; k < b3; ++k)
    for (std::size_t q = 0; q < nq; ++q) {
      Dual3 x(tet.xi[q]), y(tet.eta[q]), z(tet.zeta[q]);
      x.d[0] = 1.0;
      y.d[1] = 1.0;
      z.d[2] = 1.0;
      const Dual3 v = tetBasis<Dual3>(idx3[k].i, idx3[k].j, idx3[k].k, x, y, z);
      phi[k * nq + q] = v.v;
      for (int d = 0; d < 3; ++d) dphi[(d * b3 + k) * nq + q] = v.d[d];
    }

  double* mass = out + layout.mass;
  for (int k = 0; k < b3; ++k) {
    double s = 0.0;
    for (std::size_t q = 0; q < nq; ++q) s += tet.w[q] * phi[k * nq + q] * phi[k * nq + q];
    mass[k] = s;
  }
  std::vector<double> psi(b2 * nf);
  for (int m = 0; m < b2; ++m)
    for (std::size_t q = 0; q < nf; ++q) psi[m * nf + q] = triBasis(idx2[m].i, idx2[m].j, tri.chi[q], tri.tau[q]);
  double* mass2 = out + layout.mass2;
  for (int m = 0; m < b2; ++m) {
    double s = 0.0;
    for (std::size_t q = 0; q < nf; ++q) s += tri.w[q] * psi[m * nf + q] * psi[m * nf + q];
    mass2[m] = s;
  }

  for (int d = 0; d < 3; ++d) {
    double* km = out + layout.kDivM + d * b3 * b3;
    double* kmt = out + layout.kDivMT + d * b3 * b3;
    for (int k = 0; k < b3; ++k)
      for (int l = 0; l < b3; ++l) {
        double s1 = 0.0, s2 = 0.0;
        for (std::size_t q = 0; q < nq; ++q) {
          s1 += tet.w[q] * dphi[(d * b3 + k) * nq + q] * phi[l * nq + q];
          s2 += tet.w[q] * phi[k * nq + q] * dphi[(d * b3 + l) * nq + q];
        }
        km[k * b3 + l] = s1 / mass[k];
        kmt[k * b3 + l] = s2 / mass[k];
      }
  }

  std::vector<double> trace(b3 * nf);
  for (int f = 0; f < 4; ++f) {
    for (int k = 0; k < b3; ++k)
      for (std::size_t q = 0; q < nf; ++q) {
        double p[3];
        faceToReference(f, tri.chi[q], tri.tau[q], -1, p);
        trace[k * nf + q] = tetBasis<double>(idx3[k].i, idx3[k].j, idx3[k].k, p[0], p[1], p[2]);
      }
    double* rdm = out + layout.rDivM + f * b3 * b2;
    double* fmrt = out + layout.fMrT + f * b2 * b3;
    double* rt = out + layout.rT + f * b2 * b3;
    for (int k = 0; k < b3; ++k)
      for (int m = 0; m < b2; ++m) {
        double s = 0.0;
        for (std::size_t q = 0; q < nf; ++q) s += tri.w[q] * trace[k * nf + q] * psi[m * nf + q];
        rdm[k * b2 + m] = s / (mass[k] * mass2[m]);
        fmrt[m * b3 + k] = s;
        rt[m * b3 + k] = s / mass2[m];
      }
  }

  for (int h = 0; h < 3; ++h) {
    double* fp = out + layout.fP + h * b2 * b2;
    for (int m = 0; m < b2; ++m)
      for (int n = 0; n < b2; ++n) {
        double s = 0.0;
        for (std::size_t q = 0; q < nf; ++q) {
          double a, b;
          orientationMap(h, tri.chi[q], tri.tau[q], &a, &b);
          s += tri.w[q] * psi[m * nf + q] * triBasis(idx2[n].i, idx2[n].j, a, b);
        }
        fp[m * b2 + n] = s;
      }
  }
  return 0;
}

}  // namespace sdg
