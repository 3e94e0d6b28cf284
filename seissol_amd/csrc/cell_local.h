// Per-element FP64 precompute of the elastic ADER-DG operators: the three
// star matrices and the local/neighbour flux solvers of the four faces.
//
// Follows src/Initializer/Model/CellLocalMatrices.cpp:66-296 (driver),
// src/Equations/elastic/Model/ElasticSetup.h:27-198 (coefficient matrices,
// Godunov state, face rotation), src/Numerical/Transformation.cpp:66-290
// (Jacobian, rotation matrices) and src/Geometry/MeshTools.cpp:20-133
// (face nodes, normals, tangents, areas, volume).  Flux-solver contraction as
// specified in codegen/kernels/aderdg.py:221-246.
//
// All 9x9 matrices are stored column-major (entry (row, col) at row + 9*col),
// which is the tensor layout the reference kernels consume.
#pragma once

#ifndef SDG_HD
#if defined(__HIPCC__)
#define SDG_HD __host__ __device__
#else
#define SDG_HD
#endif
#endif

namespace sdg {

struct ElasticMaterial {
  double rho, lambda, mu;
};

constexpr int FaceRegular = 0;
constexpr int FaceDynamicRupture = 1;

// Vertices of the reference tetrahedron spanning each face
// (src/Geometry/MeshTools.cpp:20).
SDG_HD inline int faceNode(int face, int i) {
  const int table[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
  return table[face][i];
}

SDG_HD inline void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Gradients of the reference coordinates (xi, eta, zeta) with respect to
// (x, y, z): the rows of the inverse of J = [v1-v0 | v2-v0 | v3-v0].
SDG_HD inline double referenceGradients(const double v[4][3], double grad[3][3]) {
  double e[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) e[c][r] = v[c + 1][r] - v[0][r];
  // inverse of J (columns e[0], e[1], e[2]): rows are (e1 x e2, e2 x e0, e0 x e1) / det
  double c0[3], c1[3], c2[3];
  cross3(e[1], e[2], c0);
  cross3(e[2], e[0], c1);
  cross3(e[0], e[1], c2);
  const double det = e[0][0] * c0[0] + e[0][1] * c0[1] + e[0][2] * c0[2];
  const double inv = 1.0 / det;
  for (int k = 0; k < 3; ++k) {
    grad[0][k] = c0[k] * inv;
    grad[1][k] = c1[k] * inv;
    grad[2][k] = c2[k] * inv;
  }
  return det;
}

// Transposed elastic coefficient matrix of direction `dim`, column-major
// (ElasticSetup.h:27-77: entries (q, p) such that dQ/dt(:,p) gets
// sum_q dQ/dx_dim(:,q) * M(q,p)).
SDG_HD inline void transposedCoefficientMatrix(const ElasticMaterial& m, int dim, double out[81]) {
  for (int i = 0; i < 81; ++i) out[i] = 0.0;
  const double l2m = m.lambda + 2.0 * m.mu;
  const double ri = 1.0 / m.rho;
  auto at = [&](int r, int c) -> double& { return out[r + 9 * c]; };
  if (dim == 0) {
    at(6, 0) = -l2m; at(6, 1) = -m.lambda; at(6, 2) = -m.lambda;
    at(7, 3) = -m.mu; at(8, 5) = -m.mu;
    at(0, 6) = -ri; at(3, 7) = -ri; at(5, 8) = -ri;
  } else if (dim == 1) {
    at(7, 0) = -m.lambda; at(7, 1) = -l2m; at(7, 2) = -m.lambda;
    at(6, 3) = -m.mu; at(8, 4) = -m.mu;
    at(1, 7) = -ri; at(3, 6) = -ri; at(4, 8) = -ri;
  } else {
    at(8, 0) = -m.lambda; at(8, 1) = -m.lambda; at(8, 2) = -l2m;
    at(7, 4) = -m.mu; at(6, 5) = -m.mu;
    at(2, 8) = -ri; at(5, 6) = -ri; at(4, 7) = -ri;
  }
}

// Solve A X = B for 9x9 column-major A, B (B overwritten by X); Gaussian
// elimination with partial pivoting on a private copy of A.
SDG_HD inline void solve9(const double A[81], double B[81]) {
  double a[81];
  for (int i = 0; i < 81; ++i) a[i] = A[i];
  for (int c = 0; c < 9; ++c) {
    int piv = c;
    double best = a[c + 9 * c] < 0 ? -a[c + 9 * c] : a[c + 9 * c];
    for (int r = c + 1; r < 9; ++r) {
      const double v = a[r + 9 * c] < 0 ? -a[r + 9 * c] : a[r + 9 * c];
      if (v > best) { best = v; piv = r; }
    }
    if (piv != c) {
      for (int k = 0; k < 9; ++k) {
        double t = a[c + 9 * k]; a[c + 9 * k] = a[piv + 9 * k]; a[piv + 9 * k] = t;
        t = B[c + 9 * k]; B[c + 9 * k] = B[piv + 9 * k]; B[piv + 9 * k] = t;
      }
    }
    const double d = 1.0 / a[c + 9 * c];
    for (int r = c + 1; r < 9; ++r) {
      const double f = a[r + 9 * c] * d;
      if (f == 0.0) continue;
      for (int k = c; k < 9; ++k) a[r + 9 * k] -= f * a[c + 9 * k];
      for (int k = 0; k < 9; ++k) B[r + 9 * k] -= f * B[c + 9 * k];
    }
  }
  for (int c = 8; c >= 0; --c) {
    const double d = 1.0 / a[c + 9 * c];
    for (int k = 0; k < 9; ++k) {
      double s = B[c + 9 * k];
      for (int j = c + 1; j < 9; ++j) s -= a[c + 9 * j] * B[j + 9 * k];
      B[c + 9 * k] = s * d;
    }
  }
}

// Transposed Godunov states at a regular interface between the local and the
// neighbouring elastic material (ElasticSetup.h:79-180): with the eigenvector
// basis R of the face-normal Riemann problem and the selector chi of the waves
// leaving the local element, QgodNeighbor = R^-T chi R^T, QgodLocal = I - that.
SDG_HD inline void godunovStates(const ElasticMaterial& loc, const ElasticMaterial& nb,
                                 double qLocal[81], double qNeighbor[81]) {
  double R[81];
  for (int i = 0; i < 81; ++i) R[i] = 0.0;
  auto r = [&](int i, int j) -> double& { return R[i + 9 * j]; };
  const double l2mL = loc.lambda + 2.0 * loc.mu;
  r(0, 0) = l2mL; r(1, 0) = loc.lambda; r(2, 0) = loc.lambda;
  r(6, 0) = __builtin_sqrt(l2mL / loc.rho);
  r(3, 1) = loc.mu; r(7, 1) = __builtin_sqrt(loc.mu / loc.rho);
  r(5, 2) = loc.mu; r(8, 2) = __builtin_sqrt(loc.mu / loc.rho);
  r(4, 3) = l2mL; r(1, 4) = l2mL; r(2, 5) = l2mL;
  const double l2mN = nb.lambda + 2.0 * nb.mu;
  r(5, 6) = nb.mu; r(8, 6) = -__builtin_sqrt(nb.mu / nb.rho);
  r(3, 7) = nb.mu; r(7, 7) = -__builtin_sqrt(nb.mu / nb.rho);
  r(0, 8) = l2mN; r(1, 8) = nb.lambda; r(2, 8) = nb.lambda;
  r(6, 8) = -__builtin_sqrt(l2mN / nb.rho);

  double RT[81];
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) RT[i + 9 * j] = R[j + 9 * i];
  // right-hand side chi * R^T: keep rows 0..2 of R^T (outgoing P, S1, S2 waves)
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) qNeighbor[i + 9 * j] = (i < 3) ? RT[i + 9 * j] : 0.0;
  solve9(RT, qNeighbor);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) qLocal[i + 9 * j] = (i == j ? 1.0 : 0.0) - qNeighbor[i + 9 * j];
}

// Rotation from the face-aligned (n, s, t) frame to (x, y, z) for the 6 stress
// and 3 velocity components, and its inverse (Transformation.cpp:104-290).
SDG_HD inline void faceRotation(const double n[3], const double s[3], const double t[3],
                                double T[81], double Tinv[81]) {
  for (int i = 0; i < 81; ++i) { T[i] = 0.0; Tinv[i] = 0.0; }
  auto tt = [&](int i, int j) -> double& { return T[i + 9 * j]; };
  auto ti = [&](int i, int j) -> double& { return Tinv[i + 9 * j]; };
  const double* b[3] = {n, s, t};
  // symmetric tensor (Voigt order xx, yy, zz, xy, yz, xz)
  const int pa[6] = {0, 1, 2, 0, 1, 0};
  const int pb[6] = {0, 1, 2, 1, 2, 2};
  // T: local-frame Voigt component `col` to global component `row`.
  // Diagonal local components: sigma_aa contributes a_x a_y to global (x, y).
  for (int row = 0; row < 6; ++row)
    for (int col = 0; col < 3; ++col) {
      const double* u = b[col];
      tt(row, col) = u[pa[row]] * u[pb[row]];
    }
  for (int row = 0; row < 6; ++row)
    for (int col = 3; col < 6; ++col) {
      const double* u = b[pa[col]];
      const double* w = b[pb[col]];
      tt(row, col) = u[pa[row]] * w[pb[row]] + u[pb[row]] * w[pa[row]];
    }
  // inverse: global component `col` to local component `row`
  for (int row = 0; row < 6; ++row) {
    const double* u = b[pa[row]];
    const double* w = b[pb[row]];
    for (int col = 0; col < 6; ++col) {
      const int x = pa[col], y = pb[col];
      ti(row, col) = (col < 3) ? u[x] * w[y] : u[x] * w[y] + u[y] * w[x];
    }
  }
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      tt(6 + i, 6 + k) = b[k][i];
      ti(6 + k, 6 + i) = b[k][i];
    }
}

// Star matrices and flux solvers of one element.
//   v:        vertex coordinates
//   loc:      material of the element; nb[f]: material behind face f
//   faceType: FaceRegular / FaceDynamicRupture per face (DR: no local flux)
//   star:     [3][81]; aplus, aminus: [4][81]
SDG_HD inline void cellLocalMatrices(const double v[4][3], const ElasticMaterial& loc,
                                     const ElasticMaterial nb[4], const int faceType[4],
                                     double* star, double* aplus, double* aminus) {
  double grad[3][3];
  const double det = referenceGradients(v, grad);
  double A[3][81];
  for (int d = 0; d < 3; ++d) transposedCoefficientMatrix(loc, d, A[d]);
  for (int r = 0; r < 3; ++r)
    for (int i = 0; i < 81; ++i)
      star[81 * r + i] = grad[r][0] * A[0][i] + grad[r][1] * A[1][i] + grad[r][2] * A[2][i];
  const double volume = (det < 0 ? -det : det) / 6.0;

  for (int f = 0; f < 4; ++f) {
    const double* p0 = v[faceNode(f, 0)];
    const double* p1 = v[faceNode(f, 1)];
    const double* p2 = v[faceNode(f, 2)];
    double ab[3], ac[3], n[3], s[3], t[3];
    for (int k = 0; k < 3; ++k) { ab[k] = p1[k] - p0[k]; ac[k] = p2[k] - p0[k]; }
    cross3(ab, ac, n);
    for (int k = 0; k < 3; ++k) s[k] = ab[k];
    cross3(n, s, t);
    const double nn = __builtin_sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double ns = __builtin_sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double nt = __builtin_sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const double area = 0.5 * nn;
    for (int k = 0; k < 3; ++k) { n[k] /= nn; s[k] /= ns; t[k] /= nt; }

    double qL[81], qN[81], T[81], Ti[81];
    godunovStates(loc, nb[f], qL, qN);
    faceRotation(n, s, t, T, Ti);
    const double scale = -2.0 * area / (6.0 * volume);
    const double scalePlus = faceType[f] == FaceDynamicRupture ? 0.0 : scale;

    // G = Qgod * Atilde (Atilde = A[0] for the isotropic material), then
    // flux = scale * Tinv^T * G * T^T.
    for (int side = 0; side < 2; ++side) {
      const double* q = side == 0 ? qL : qN;
      double G[81], H[81];
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          double acc = 0.0;
          for (int k = 0; k < 9; ++k) acc += q[i + 9 * k] * A[0][k + 9 * j];
          G[i + 9 * j] = acc;
        }
      // H = G * T^T : H(k, j) = sum_l G(k, l) T(j, l)
      for (int k = 0; k < 9; ++k)
        for (int j = 0; j < 9; ++j) {
          double acc = 0.0;
          for (int l = 0; l < 9; ++l) acc += G[k + 9 * l] * T[j + 9 * l];
          H[k + 9 * j] = acc;
        }
      double* out = (side == 0 ? aplus : aminus) + 81 * f;
      const double sc = side == 0 ? scalePlus : scale;
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          double acc = 0.0;
          for (int k = 0; k < 9; ++k) acc += Ti[k + 9 * i] * H[k + 9 * j];
          out[i + 9 * j] = sc * acc;
        }
    }
  }
}

// Largest stable time step of the element (GlobalTimestep.cpp:27-47):
// cfl * 2 * r_in / (vp * (2*order - 1)), r_in = 3 V / (sum of face areas).
SDG_HD inline double cflTimestep(const double v[4][3], const ElasticMaterial& m, int order, double cfl) {
  double e[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) e[c][r] = v[c + 1][r] - v[0][r];
  double c0[3];
  cross3(e[1], e[2], c0);
  double det = e[0][0] * c0[0] + e[0][1] * c0[1] + e[0][2] * c0[2];
  if (det < 0) det = -det;
  double areaSum = 0.0;
  for (int f = 0; f < 4; ++f) {
    double ab[3], ac[3], n[3];
    for (int k = 0; k < 3; ++k) {
      ab[k] = v[faceNode(f, 1)][k] - v[faceNode(f, 0)][k];
      ac[k] = v[faceNode(f, 2)][k] - v[faceNode(f, 0)][k];
    }
    cross3(ab, ac, n);
    areaSum += 0.5 * __builtin_sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  }
  const double rIn = 3.0 * (det / 6.0) / areaSum;
  const double vp = __builtin_sqrt((m.lambda + 2.0 * m.mu) / m.rho);
  return cfl * 2.0 * rIn / (vp * (2.0 * order - 1.0));
}

}  // namespace sdg
