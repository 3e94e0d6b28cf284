// ORACLE -- test infrastructure only.
//
// CPU FP64 restatement of the SeisSol elastic ADER-DG element update and of
// the per-element precompute, used exclusively as the checker of the HIP
// library (tests, the smoke check and the CPU-baseline leg of the benchmark).
// Nothing in the library links, loads or calls this code.
//
// Parity status of each function (what pins it) is documented next to it.
//
// Layouts: DOFs and time-integrated DOFs [n][9][B] (basis index fastest),
// derivatives [n][sum_i B(O-i)][9] stored per derivative block as [9][B(O-i)],
// 9x9 matrices column-major, global matrices row-major as produced by the
// generator (kDivMT not negated).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline int nb3(int o) { return o * (o + 1) * (o + 2) / 6; }
inline int nb2(int o) { return o * (o + 1) / 2; }

struct Globals {
  int order, B, B2;
  const double *kDivM, *kDivMT, *rDivM, *fMrT, *rT, *fP;
  Globals(int o, const double* gm) : order(o), B(nb3(o)), B2(nb2(o)) {
    std::size_t off = 0;
    kDivM = gm + off; off += 3 * B * B;
    kDivMT = gm + off; off += 3 * B * B;
    rDivM = gm + off; off += 4 * B * B2;
    fMrT = gm + off; off += 4 * B2 * B;
    rT = gm + off; off += 4 * B2 * B;
    fP = gm + off;
  }
};

// ---------------------------------------------------------------------------
// Per-element precompute.
// src/Initializer/Model/CellLocalMatrices.cpp:46-62 (star), :139-286 (fluxes);
// src/Numerical/Transformation.cpp:66-102 (Jacobian), :104-290 (rotations);
// src/Equations/elastic/Model/ElasticSetup.h:27-198;
// src/Geometry/MeshTools.cpp:20, 43-69, 111-133.
// Pinned by: Godunov golden matrices of tests/Model/Values.h (elastic),
// invariants QgodLocal + QgodNeighbor = I and P^2 = P, and the constant- and
// polynomial-state invariants of the full update on a conforming mesh.
// ---------------------------------------------------------------------------

void coefficientMatrixT(double rho, double lambda, double mu, int dim, double M[9][9]) {
  // M[row][col] of the transposed Jacobian (ElasticSetup.h:27-77)
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) M[i][j] = 0.0;
  const double l2m = lambda + 2 * mu;
  if (dim == 0) {
    M[6][0] = -l2m; M[6][1] = -lambda; M[6][2] = -lambda; M[7][3] = -mu; M[8][5] = -mu;
    M[0][6] = -1 / rho; M[3][7] = -1 / rho; M[5][8] = -1 / rho;
  } else if (dim == 1) {
    M[7][0] = -lambda; M[7][1] = -l2m; M[7][2] = -lambda; M[6][3] = -mu; M[8][4] = -mu;
    M[1][7] = -1 / rho; M[3][6] = -1 / rho; M[4][8] = -1 / rho;
  } else {
    M[8][0] = -lambda; M[8][1] = -lambda; M[8][2] = -l2m; M[7][4] = -mu; M[6][5] = -mu;
    M[2][8] = -1 / rho; M[5][6] = -1 / rho; M[4][7] = -1 / rho;
  }
}

void jacobianGradients(const double* x, const double* y, const double* z, double gXi[3],
                       double gEta[3], double gZeta[3]) {
  // explicit cofactor expansion (Transformation.cpp:66-102)
  const double det = x[0] * (y[1] * (z[3] - z[2]) + y[2] * (z[1] - z[3]) + y[3] * (z[2] - z[1])) +
                     x[1] * (y[0] * (z[2] - z[3]) + y[2] * (z[3] - z[0]) + y[3] * (z[0] - z[2])) +
                     x[2] * (y[0] * (z[3] - z[1]) + y[1] * (z[0] - z[3]) + y[3] * (z[1] - z[0])) +
                     x[3] * (y[0] * (z[1] - z[2]) + y[1] * (z[2] - z[0]) + y[2] * (z[0] - z[1]));
  const double id = 1.0 / det;
  gXi[0] = id * (y[0] * (z[2] - z[3]) + y[2] * (z[3] - z[0]) + y[3] * (z[0] - z[2]));
  gXi[1] = id * (x[0] * (z[3] - z[2]) + x[2] * (z[0] - z[3]) + x[3] * (z[2] - z[0]));
  gXi[2] = id * (x[0] * (y[2] - y[3]) + x[2] * (y[3] - y[0]) + x[3] * (y[0] - y[2]));
  gEta[0] = id * (y[0] * (z[3] - z[1]) + y[1] * (z[0] - z[3]) + y[3] * (z[1] - z[0]));
  gEta[1] = id * (x[0] * (z[1] - z[3]) + x[1] * (z[3] - z[0]) + x[3] * (z[0] - z[1]));
  gEta[2] = id * (x[0] * (y[3] - y[1]) + x[1] * (y[0] - y[3]) + x[3] * (y[1] - y[0]));
  gZeta[0] = id * (y[0] * (z[1] - z[2]) + y[1] * (z[2] - z[0]) + y[2] * (z[0] - z[1]));
  gZeta[1] = id * (x[0] * (z[2] - z[1]) + x[1] * (z[0] - z[2]) + x[2] * (z[1] - z[0]));
  gZeta[2] = id * (x[0] * (y[1] - y[2]) + x[1] * (y[2] - y[0]) + x[2] * (y[0] - y[1]));
}

// LU with partial pivoting, solve A X = B (row-major 9x9, in place on copies)
void luSolve9(double A[9][9], double Bm[9][9]) {
  int perm[9];
  for (int i = 0; i < 9; ++i) perm[i] = i;
  for (int k = 0; k < 9; ++k) {
    int p = k;
    for (int i = k + 1; i < 9; ++i)
      if (std::fabs(A[i][k]) > std::fabs(A[p][k])) p = i;
    if (p != k) {
      for (int j = 0; j < 9; ++j) std::swap(A[k][j], A[p][j]);
      std::swap(perm[k], perm[p]);
      for (int j = 0; j < 9; ++j) std::swap(Bm[k][j], Bm[p][j]);
    }
    for (int i = k + 1; i < 9; ++i) {
      A[i][k] /= A[k][k];
      for (int j = k + 1; j < 9; ++j) A[i][j] -= A[i][k] * A[k][j];
    }
  }
  for (int c = 0; c < 9; ++c) {
    for (int i = 1; i < 9; ++i)
      for (int k = 0; k < i; ++k) Bm[i][c] -= A[i][k] * Bm[k][c];
    for (int i = 8; i >= 0; --i) {
      for (int k = i + 1; k < 9; ++k) Bm[i][c] -= A[i][k] * Bm[k][c];
      Bm[i][c] /= A[i][i];
    }
  }
}

void godunov(const double* loc, const double* nb, double qL[9][9], double qN[9][9]) {
  // loc/nb = (rho, lambda, mu); eigenvectors as in ElasticSetup.h:79-128
  double R[9][9] = {};
  const double rl = loc[0], ll = loc[1], ml = loc[2];
  const double rn = nb[0], ln = nb[1], mn = nb[2];
  R[0][0] = ll + 2 * ml; R[1][0] = ll; R[2][0] = ll; R[6][0] = std::sqrt((ll + 2 * ml) / rl);
  R[3][1] = ml; R[7][1] = std::sqrt(ml / rl);
  R[5][2] = ml; R[8][2] = std::sqrt(ml / rl);
  R[4][3] = ll + 2 * ml; R[1][4] = ll + 2 * ml; R[2][5] = ll + 2 * ml;
  R[5][6] = mn; R[8][6] = -std::sqrt(mn / rn);
  R[3][7] = mn; R[7][7] = -std::sqrt(mn / rn);
  R[0][8] = ln + 2 * mn; R[1][8] = ln; R[2][8] = ln; R[6][8] = -std::sqrt((ln + 2 * mn) / rn);
  double RT[9][9], rhs[9][9];
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) RT[i][j] = R[j][i];
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) rhs[i][j] = (i < 3) ? RT[i][j] : 0.0;
  luSolve9(RT, rhs);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      qN[i][j] = rhs[i][j];
      qL[i][j] = (i == j ? 1.0 : 0.0) - rhs[i][j];
    }
}

void rotationT(const double n[3], const double s[3], const double t[3], double T[9][9], double Ti[9][9]) {
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) T[i][j] = Ti[i][j] = 0.0;
  const double nx = n[0], ny = n[1], nz = n[2], sx = s[0], sy = s[1], sz = s[2], tx = t[0], ty = t[1], tz = t[2];
  // symmetricTensor2RotationMatrix (Transformation.cpp:157-195)
  T[0][0] = nx * nx; T[1][0] = ny * ny; T[2][0] = nz * nz; T[3][0] = ny * nx; T[4][0] = nz * ny; T[5][0] = nz * nx;
  T[0][1] = sx * sx; T[1][1] = sy * sy; T[2][1] = sz * sz; T[3][1] = sy * sx; T[4][1] = sz * sy; T[5][1] = sz * sx;
  T[0][2] = tx * tx; T[1][2] = ty * ty; T[2][2] = tz * tz; T[3][2] = ty * tx; T[4][2] = tz * ty; T[5][2] = tz * tx;
  T[0][3] = 2 * nx * sx; T[1][3] = 2 * ny * sy; T[2][3] = 2 * nz * sz;
  T[3][3] = ny * sx + nx * sy; T[4][3] = nz * sy + ny * sz; T[5][3] = nz * sx + nx * sz;
  T[0][4] = 2 * sx * tx; T[1][4] = 2 * sy * ty; T[2][4] = 2 * sz * tz;
  T[3][4] = sy * tx + sx * ty; T[4][4] = sz * ty + sy * tz; T[5][4] = sz * tx + sx * tz;
  T[0][5] = 2 * nx * tx; T[1][5] = 2 * ny * ty; T[2][5] = 2 * nz * tz;
  T[3][5] = ny * tx + nx * ty; T[4][5] = nz * ty + ny * tz; T[5][5] = nz * tx + nx * tz;
  // inverseSymmetricTensor2RotationMatrix (:220-260)
  Ti[0][0] = nx * nx; Ti[1][0] = sx * sx; Ti[2][0] = tx * tx; Ti[3][0] = nx * sx; Ti[4][0] = sx * tx; Ti[5][0] = nx * tx;
  Ti[0][1] = ny * ny; Ti[1][1] = sy * sy; Ti[2][1] = ty * ty; Ti[3][1] = ny * sy; Ti[4][1] = sy * ty; Ti[5][1] = ny * ty;
  Ti[0][2] = nz * nz; Ti[1][2] = sz * sz; Ti[2][2] = tz * tz; Ti[3][2] = nz * sz; Ti[4][2] = sz * tz; Ti[5][2] = nz * tz;
  Ti[0][3] = 2 * ny * nx; Ti[1][3] = 2 * sy * sx; Ti[2][3] = 2 * ty * tx;
  Ti[3][3] = ny * sx + nx * sy; Ti[4][3] = sy * tx + sx * ty; Ti[5][3] = ny * tx + nx * ty;
  Ti[0][4] = 2 * nz * ny; Ti[1][4] = 2 * sz * sy; Ti[2][4] = 2 * tz * ty;
  Ti[3][4] = nz * sy + ny * sz; Ti[4][4] = sz * ty + sy * tz; Ti[5][4] = nz * ty + ny * tz;
  Ti[0][5] = 2 * nz * nx; Ti[1][5] = 2 * sz * sx; Ti[2][5] = 2 * tz * tx;
  Ti[3][5] = nz * sx + nx * sz; Ti[4][5] = sz * tx + sx * tz; Ti[5][5] = nz * tx + nx * tz;
  // tensor1RotationMatrix / inverse (:104-155)
  for (int i = 0; i < 3; ++i) {
    T[6 + i][6] = n[i]; T[6 + i][7] = s[i]; T[6 + i][8] = t[i];
    Ti[6][6 + i] = n[i]; Ti[7][6 + i] = s[i]; Ti[8][6 + i] = t[i];
  }
}

const int kFaceNodes[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};

void cellLocal(const double* vtx, const double* mat, const double* const nbMat[4], const std::uint8_t* faceType,
               double* star, double* aplus, double* aminus) {
  double x[4], y[4], z[4];
  for (int v = 0; v < 4; ++v) { x[v] = vtx[3 * v]; y[v] = vtx[3 * v + 1]; z[v] = vtx[3 * v + 2]; }
  double g[3][3];
  jacobianGradients(x, y, z, g[0], g[1], g[2]);
  double AT[3][9][9];
  for (int d = 0; d < 3; ++d) coefficientMatrixT(mat[0], mat[1], mat[2], d, AT[d]);
  for (int r = 0; r < 3; ++r)
    for (int i = 0; i < 9; ++i)
      for (int j = 0; j < 9; ++j)
        star[81 * r + i + 9 * j] = g[r][0] * AT[0][i][j] + g[r][1] * AT[1][i][j] + g[r][2] * AT[2][i][j];
  double ab[3], ac[3], ad[3], cr[3];
  for (int k = 0; k < 3; ++k) {
    ab[k] = vtx[3 + k] - vtx[k]; ac[k] = vtx[6 + k] - vtx[k]; ad[k] = vtx[9 + k] - vtx[k];
  }
  cr[0] = ab[1] * ac[2] - ab[2] * ac[1];
  cr[1] = ab[2] * ac[0] - ab[0] * ac[2];
  cr[2] = ab[0] * ac[1] - ab[1] * ac[0];
  const double volume = std::fabs(ad[0] * cr[0] + ad[1] * cr[1] + ad[2] * cr[2]) / 6.0;
  for (int f = 0; f < 4; ++f) {
    const double* a = vtx + 3 * kFaceNodes[f][0];
    const double* b = vtx + 3 * kFaceNodes[f][1];
    const double* c = vtx + 3 * kFaceNodes[f][2];
    double u[3], w[3], n[3], s[3], t[3];
    for (int k = 0; k < 3; ++k) { u[k] = b[k] - a[k]; w[k] = c[k] - a[k]; }
    n[0] = u[1] * w[2] - u[2] * w[1]; n[1] = u[2] * w[0] - u[0] * w[2]; n[2] = u[0] * w[1] - u[1] * w[0];
    for (int k = 0; k < 3; ++k) s[k] = u[k];
    t[0] = n[1] * s[2] - n[2] * s[1]; t[1] = n[2] * s[0] - n[0] * s[2]; t[2] = n[0] * s[1] - n[1] * s[0];
    const double area = 0.5 * std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double* vec : {n, s, t}) {
      const double l = std::sqrt(vec[0] * vec[0] + vec[1] * vec[1] + vec[2] * vec[2]);
      for (int k = 0; k < 3; ++k) vec[k] /= l;
    }
    double qL[9][9], qN[9][9], T[9][9], Ti[9][9], Atil[9][9];
    godunov(mat, nbMat[f], qL, qN);
    rotationT(n, s, t, T, Ti);
    coefficientMatrixT(mat[0], mat[1], mat[2], 0, Atil);
    const double fluxScale = -2.0 * area / (6.0 * volume);
    // A(i,j) = s * sum_{k,l} Tinv(k,i) * (Qgod*Atil)(k,l) * T(j,l)   (aderdg.py:221-246)
    for (int side = 0; side < 2; ++side) {
      double (*q)[9] = side == 0 ? qL : qN;
      const double sc = (side == 0 && faceType && faceType[f] == 1) ? 0.0 : fluxScale;
      double* out = (side == 0 ? aplus : aminus) + 81 * f;
      for (int i = 0; i < 9; ++i)
        for (int j = 0; j < 9; ++j) {
          double acc = 0.0;
          for (int k = 0; k < 9; ++k)
            for (int l = 0; l < 9; ++l) {
              double gkl = 0.0;
              for (int m = 0; m < 9; ++m) gkl += q[k][m] * Atil[m][l];
              acc += Ti[k][i] * gkl * T[j][l];
            }
          out[i + 9 * j] = sc * acc;
        }
    }
  }
}

}  // namespace

extern "C" {

int oracle_godunov(const double* local, const double* neighbor, double* qLocal, double* qNeighbor) {
  double qL[9][9], qN[9][9];
  godunov(local, neighbor, qL, qN);
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      qLocal[i + 9 * j] = qL[i][j];
      qNeighbor[i + 9 * j] = qN[i][j];
    }
  return 0;
}

// vertices [n][4][3]; material [n][3] (rho, lambda, mu); neighbors [n][4]
// (material behind a face = material of the neighbour); faceTypes [n][4] or NULL
int oracle_cell_local_matrices(std::int64_t n, const double* vertices, const double* material,
                               const std::int32_t* neighbors, const std::uint8_t* faceTypes, double* star,
                               double* aplus, double* aminus) {
#pragma omp parallel for schedule(static)
  for (std::int64_t e = 0; e < n; ++e) {
    const double* nbm[4];
    for (int f = 0; f < 4; ++f) nbm[f] = material + 3 * neighbors[4 * e + f];
    cellLocal(vertices + 12 * e, material + 3 * e, nbm, faceTypes ? faceTypes + 4 * e : nullptr, star + 243 * e,
              aplus + 324 * e, aminus + 324 * e);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Cauchy-Kowalevski derivatives + time integral.
// src/Kernels/LinearCK/Time.cpp:65-140; spec codegen/kernels/aderdg.py:463-537;
// kDivMT negated once (src/Memory/GlobalData.cpp:51-60, 279).
// Pinned by: exactness of one ADER step for polynomial plane states on a
// conforming mesh (tests), constant states have zero derivatives, and
// TimeBasis monomial coefficients (tests/Numerical/TimeBasis.t.h).
// ---------------------------------------------------------------------------
std::int64_t oracle_derivatives_size(int order) {
  std::int64_t s = 0;
  for (int i = 0; i < order; ++i) s += 9 * nb3(order - i);
  return s;
}

int oracle_ader(int order, const double* gm, std::int64_t n, const double* coeffs, const double* dofs,
                const double* star, double* timeIntegrated, double* derivatives) {
  const Globals G(order, gm);
  const int B = G.B;
  const std::int64_t dsize = oracle_derivatives_size(order);
#pragma omp parallel
  {
    std::vector<double> dq(dsize), tmp(3 * B * 9);
#pragma omp for schedule(static)
    for (std::int64_t e = 0; e < n; ++e) {
      const double* Q = dofs + e * 9 * B;
      const double* S = star + e * 243;
      double* I = timeIntegrated + e * 9 * B;
      // block 0 = Q
      for (int q = 0; q < 9; ++q)
        for (int k = 0; k < B; ++k) dq[q * B + k] = Q[q * B + k];
      std::int64_t prevOff = 0;
      std::int64_t off = 9 * B;
      for (int q = 0; q < 9; ++q)
        for (int k = 0; k < B; ++k) I[q * B + k] = coeffs[0] * Q[q * B + k];
      for (int d = 1; d < order; ++d) {
        const int m = nb3(order - d), nIn = nb3(order - d + 1);
        const double* prev = dq.data() + prevOff;  // [9][nIn]
        double* cur = dq.data() + off;             // [9][m]
        // D_j[l][p] = sum_q prev[q][l] * star_j(q,p)
        for (int j = 0; j < 3; ++j)
          for (int p = 0; p < 9; ++p)
            for (int l = 0; l < nIn; ++l) {
              double acc = 0.0;
              for (int q = 0; q < 9; ++q) acc += prev[q * nIn + l] * S[81 * j + q + 9 * p];
              tmp[(j * 9 + p) * B + l] = acc;
            }
        for (int p = 0; p < 9; ++p)
          for (int k = 0; k < m; ++k) {
            double acc = 0.0;
            for (int j = 0; j < 3; ++j) {
              const double* km = G.kDivMT + j * B * B + k * B;
              for (int l = 0; l < nIn; ++l) acc -= km[l] * tmp[(j * 9 + p) * B + l];
            }
            cur[p * m + k] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int k = 0; k < m; ++k) I[p * B + k] += coeffs[d] * cur[p * m + k];
        prevOff = off;
        off += 9 * m;
      }
      if (derivatives) std::memcpy(derivatives + e * dsize, dq.data(), dsize * sizeof(double));
    }
  }
  return 0;
}

// Time::evaluate (Time.cpp:237-255): out = sum_i coeffs[i] * dQ_i
int oracle_taylor(int order, std::int64_t n, const double* coeffs, const double* derivatives, double* out) {
  const int B = nb3(order);
  const std::int64_t dsize = oracle_derivatives_size(order);
#pragma omp parallel for schedule(static)
  for (std::int64_t e = 0; e < n; ++e) {
    double* o = out + e * 9 * B;
    for (int i = 0; i < 9 * B; ++i) o[i] = 0.0;
    std::int64_t off = 0;
    for (int d = 0; d < order; ++d) {
      const int m = nb3(order - d);
      const double* blk = derivatives + e * dsize + off;
      for (int p = 0; p < 9; ++p)
        for (int k = 0; k < m; ++k) o[p * B + k] += coeffs[d] * blk[p * m + k];
      off += 9 * m;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Volume + local flux.  src/Kernels/LinearCK/Local.cpp:124-160; spec
// aderdg.py:321-386 (rDivM_f * fMrT_f * I * AplusT_f, CPU form); DR faces
// contribute no local flux (Local.cpp:150-155).
// Pinned by: constant state on a conforming periodic mesh keeps zero update,
// polynomial-state exactness of the whole step (tests).
// ---------------------------------------------------------------------------
int oracle_local(int order, const double* gm, std::int64_t n, const double* timeIntegrated, const double* star,
                 const double* aplus, const std::uint8_t* faceTypes, double* dofs) {
  const Globals G(order, gm);
  const int B = G.B, B2 = G.B2;
#pragma omp parallel
  {
    std::vector<double> IS(9 * B), tr(9 * B2), trA(9 * B2);
#pragma omp for schedule(static)
    for (std::int64_t e = 0; e < n; ++e) {
      const double* I = timeIntegrated + e * 9 * B;
      const double* S = star + e * 243;
      double* Q = dofs + e * 9 * B;
      for (int j = 0; j < 3; ++j) {
        for (int p = 0; p < 9; ++p)
          for (int l = 0; l < B; ++l) {
            double acc = 0.0;
            for (int q = 0; q < 9; ++q) acc += I[q * B + l] * S[81 * j + q + 9 * p];
            IS[p * B + l] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int k = 0; k < B; ++k) {
            double acc = 0.0;
            const double* km = G.kDivM + j * B * B + k * B;
            for (int l = 0; l < B; ++l) acc += km[l] * IS[p * B + l];
            Q[p * B + k] += acc;
          }
      }
      for (int f = 0; f < 4; ++f) {
        if (faceTypes && faceTypes[4 * e + f] == 1) continue;
        const double* A = aplus + e * 324 + 81 * f;
        for (int q = 0; q < 9; ++q)
          for (int m = 0; m < B2; ++m) {
            double acc = 0.0;
            const double* fm = G.fMrT + f * B2 * B + m * B;
            for (int l = 0; l < B; ++l) acc += fm[l] * I[q * B + l];
            tr[q * B2 + m] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int m = 0; m < B2; ++m) {
            double acc = 0.0;
            for (int q = 0; q < 9; ++q) acc += tr[q * B2 + m] * A[q + 9 * p];
            trA[p * B2 + m] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int k = 0; k < B; ++k) {
            double acc = 0.0;
            const double* rd = G.rDivM + f * B * B2 + k * B2;
            for (int m = 0; m < B2; ++m) acc += rd[m] * trA[p * B2 + m];
            Q[p * B + k] += acc;
          }
      }
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Neighbour flux.  src/Kernels/LinearCK/Neighbor.cpp:66-92; spec
// aderdg.py:414-431: Q += rDivM_f * fP_h * rT_j * I_nb * AminusT_f with
// j = faceRelations[f][0], h = faceRelations[f][1]; only regular faces.
// Pinned by the same mesh invariants as the local flux (they fail for a wrong
// side or orientation convention).
// ---------------------------------------------------------------------------
int oracle_neighbor(int order, const double* gm, std::int64_t n, const double* timeIntegrated,
                    const std::int32_t* neighbors, const std::uint8_t* faceRelations, const double* aminus,
                    const std::uint8_t* faceTypes, double* dofs) {
  const Globals G(order, gm);
  const int B = G.B, B2 = G.B2;
#pragma omp parallel
  {
    std::vector<double> t1(9 * B2), t2(9 * B2), t3(9 * B2);
#pragma omp for schedule(static)
    for (std::int64_t e = 0; e < n; ++e) {
      double* Q = dofs + e * 9 * B;
      for (int f = 0; f < 4; ++f) {
        if (faceTypes && faceTypes[4 * e + f] != 0) continue;
        const double* In = timeIntegrated + static_cast<std::int64_t>(neighbors[4 * e + f]) * 9 * B;
        const int j = faceRelations[(4 * e + f) * 2 + 0];
        const int h = faceRelations[(4 * e + f) * 2 + 1];
        const double* A = aminus + e * 324 + 81 * f;
        for (int q = 0; q < 9; ++q)
          for (int m = 0; m < B2; ++m) {
            double acc = 0.0;
            const double* r = G.rT + j * B2 * B + m * B;
            for (int l = 0; l < B; ++l) acc += r[l] * In[q * B + l];
            t1[q * B2 + m] = acc;
          }
        for (int q = 0; q < 9; ++q)
          for (int m = 0; m < B2; ++m) {
            double acc = 0.0;
            const double* fp = G.fP + h * B2 * B2 + m * B2;
            for (int l = 0; l < B2; ++l) acc += fp[l] * t1[q * B2 + l];
            t2[q * B2 + m] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int m = 0; m < B2; ++m) {
            double acc = 0.0;
            for (int q = 0; q < 9; ++q) acc += t2[q * B2 + m] * A[q + 9 * p];
            t3[p * B2 + m] = acc;
          }
        for (int p = 0; p < 9; ++p)
          for (int k = 0; k < B; ++k) {
            double acc = 0.0;
            const double* rd = G.rDivM + f * B * B2 + k * B2;
            for (int m = 0; m < B2; ++m) acc += rd[m] * t3[p * B2 + m];
            Q[p * B + k] += acc;
          }
      }
    }
  }
  return 0;
}

// Proxy kernel `all` (src/Proxy/KernelHost.h:54, KernelHost.cpp:108-158):
// per step, ader + local integral on every cell, then the neighbour integral
// on every cell from the time-integrated DOFs of this step.
int oracle_proxy_all(int order, const double* gm, std::int64_t n, int steps, const double* coeffs,
                     const double* star, const double* aplus, const double* aminus,
                     const std::int32_t* neighbors, const std::uint8_t* faceRelations,
                     const std::uint8_t* faceTypes, double* dofs, double* timeIntegrated) {
  for (int s = 0; s < steps; ++s) {
    oracle_ader(order, gm, n, coeffs, dofs, star, timeIntegrated, nullptr);
    oracle_local(order, gm, n, timeIntegrated, star, aplus, faceTypes, dofs);
    oracle_neighbor(order, gm, n, timeIntegrated, neighbors, faceRelations, aminus, faceTypes, dofs);
  }
  return 0;
}

// Dynamic rupture: space-time interpolation of both sides of every DR face
// (DynamicRupture::spaceTimeInterpolation, src/Kernels/DynamicRupture.cpp:57-109;
// spec dynamic_rupture.py:108-170): for time point o the Taylor sum of the
// side's derivatives, evaluated at the face points with V3mTo2n(side, h)
// (h = 0 plus side, faceRelation minus side) and rotated with TinvT.
// v3m [4][4][np][B]; coeffs [O][O]; out [f][O][9][points].
int oracle_dr_interpolate(int order, std::int64_t nf, int points, const double* v3m, const double* derivatives,
                          const std::int32_t* plus, const std::int32_t* minus, const std::uint8_t* info,
                          const double* tinvT, const double* coeffs, double* qPlus, double* qMinus) {
  const int B = nb3(order), np = (order + 1) * (order + 1);
  const std::int64_t dsize = oracle_derivatives_size(order);
#pragma omp parallel for schedule(static)
  for (std::int64_t f = 0; f < nf; ++f) {
    std::vector<double> Qt(9 * B);
    for (int side = 0; side < 2; ++side) {
      const std::int64_t e = side == 0 ? plus[f] : minus[f];
      const int fs = info[3 * f + side], h = side == 0 ? 0 : info[3 * f + 2];
      const double* V = v3m + ((fs * 4 + h) * np) * static_cast<std::int64_t>(B);
      double* out = side == 0 ? qPlus : qMinus;
      for (int o = 0; o < order; ++o) {
        oracle_taylor(order, 1, coeffs + o * order, derivatives + e * dsize, Qt.data());
        for (int k = 0; k < points; ++k)
          for (int p = 0; p < 9; ++p) {
            double acc = 0.0;
            if (k < np)
              for (int q = 0; q < 9; ++q) {
                double at = 0.0;
                for (int l = 0; l < B; ++l) at += V[k * B + l] * Qt[q * B + l];
                acc += at * tinvT[81 * f + q + 9 * p];
              }
            out[((f * order + o) * 9 + p) * points + k] = acc;
          }
      }
    }
  }
  return 0;
}

// Nodal flux of the DR faces into their cells (Neighbor.cpp:93-101; spec
// dynamic_rupture.py:174-190): for every cell face mapped to DR face m,
// Q[p][k] += sum_l V3mTo2nTWDivM(face, h)[k][l] sum_q G[q][l] F(q, p) with the
// imposed state G and flux solver F of the cell's side.  v3mtw [4][4][B][np].
int oracle_dr_nodal_flux(int order, std::int64_t n, int points, const double* v3mtw, const std::int32_t* drMap,
                         const std::int32_t* plus, const std::uint8_t* info, const double* fluxPlus,
                         const double* fluxMinus, const double* impPlus, const double* impMinus, double* dofs) {
  const int B = nb3(order), np = (order + 1) * (order + 1);
#pragma omp parallel for schedule(static)
  for (std::int64_t e = 0; e < n; ++e) {
    std::vector<double> gf(np * 9);
    for (int face = 0; face < 4; ++face) {
      const std::int64_t m = drMap[4 * e + face];
      if (m < 0) continue;
      const bool isPlus = plus[m] == e && info[3 * m] == face;
      const int h = isPlus ? 0 : info[3 * m + 2];
      const double* G = (isPlus ? impPlus : impMinus) + m * 9 * points;
      const double* F = (isPlus ? fluxPlus : fluxMinus) + m * 81;
      for (int l = 0; l < np; ++l)
        for (int p = 0; p < 9; ++p) {
          double acc = 0.0;
          for (int q = 0; q < 9; ++q) acc += G[q * points + l] * F[q + 9 * p];
          gf[l * 9 + p] = acc;
        }
      const double* W = v3mtw + ((face * 4 + h) * static_cast<std::int64_t>(B)) * np;
      for (int p = 0; p < 9; ++p)
        for (int k = 0; k < B; ++k) {
          double acc = 0.0;
          for (int l = 0; l < np; ++l) acc += W[k * np + l] * gf[l * 9 + p];
          dofs[(e * 9 + p) * B + k] += acc;
        }
    }
  }
  return 0;
}

// Local time stepping with rate-2 clusters (cluster c advances with
// dt0 * 2^c), restated from TimeCluster::computeLocalIntegration /
// computeNeighboringIntegration (src/Solver/TimeStepping/TimeCluster.cpp:358-628),
// TimeCommon::computeIntegrals (src/Kernels/TimeCommon.cpp:40-87) and the
// LtsSetup bits (src/Initializer/LtsSetup.h:18-110): bits 0-3 neighbour
// provides derivatives, 4-7 same-cluster relation, 8 buffers, 9 derivatives,
// 10 accumulated buffers.  Time-driven order: at every multiple k of dt0 the
// cells whose step begins at k predict (ader + local integral; buffer stored,
// or added for accumulating cells on the second of two steps), and the cells
// whose step ends at k + 1 correct (neighbour integral from the neighbour's
// buffer, or its derivatives over [k_s - T, k_s - T + dt_c], T the start of
// the neighbour's step).  Cells need not be sorted by cluster.
int oracle_lts(int order, const double* gm, std::int64_t n, int macroSteps, int numClusters,
               const std::int32_t* clusterIds, double dt0, const std::uint16_t* setups, const double* star,
               const double* aplus, const double* aminus, const std::int32_t* neighbors,
               const std::uint8_t* faceRelations, const std::uint8_t* faceTypes, double* dofs) {
  const int B = nb3(order);
  const std::int64_t q = 9LL * B, dsize = oracle_derivatives_size(order);
  std::vector<double> buffers(n * q, 0.0), derivs(n * dsize, 0.0);
  const std::int64_t fine = std::int64_t(1) << (numClusters - 1);
  auto integrate = [&](double t0, double t1, std::vector<double>& c) {
    c.assign(order, 0.0);
    double fact = 1.0;
    for (int i = 0; i < order; ++i) {
      fact *= (i + 1);
      c[i] = (std::pow(t1, i + 1) - std::pow(t0, i + 1)) / fact;
    }
  };
  for (int m = 0; m < macroSteps; ++m)
    for (std::int64_t k = 0; k < fine; ++k) {
#pragma omp parallel for schedule(dynamic, 64)
      for (std::int64_t e = 0; e < n; ++e) {
        const int c = clusterIds[e];
        const std::int64_t span = std::int64_t(1) << c;
        if (k % span != 0) continue;
        std::vector<double> coeffs, I(q), d(dsize);
        integrate(0.0, dt0 * span, coeffs);
        oracle_ader(order, gm, 1, coeffs.data(), dofs + e * q, star + e * 243, I.data(), d.data());
        oracle_local(order, gm, 1, I.data(), star + e * 243, aplus + e * 324, faceTypes ? faceTypes + 4 * e : nullptr,
                     dofs + e * q);
        const std::uint16_t s = setups[e];
        const bool reset = ((k / span) % 2) == 0;
        if (s & (1u << 8)) {
          double* b = buffers.data() + e * q;
          const bool add = (s & (1u << 10)) && !reset;
          for (std::int64_t i = 0; i < q; ++i) b[i] = add ? b[i] + I[i] : I[i];
        }
        if (s & (1u << 9)) std::memcpy(derivs.data() + e * dsize, d.data(), dsize * sizeof(double));
      }
#pragma omp parallel for schedule(dynamic, 64)
      for (std::int64_t e = 0; e < n; ++e) {
        const int c = clusterIds[e];
        const std::int64_t span = std::int64_t(1) << c;
        if ((k + 1) % span != 0) continue;
        const std::int64_t start = k + 1 - span;
        const std::uint16_t s = setups[e];
        std::vector<double> faceI(4 * q, 0.0), coeffs, tmp(q);
        for (int f = 0; f < 4; ++f) {
          if (faceTypes && faceTypes[4 * e + f] != 0) continue;
          const std::int64_t nb = neighbors[4 * e + f];
          if (s & (1u << f)) {
            const std::int64_t nspan = std::int64_t(1) << clusterIds[nb];
            const std::int64_t nStart = (start / nspan) * nspan;
            const double sub = dt0 * static_cast<double>(start - nStart);
            integrate(sub, sub + dt0 * span, coeffs);
            oracle_taylor(order, 1, coeffs.data(), derivs.data() + nb * dsize, tmp.data());
            std::memcpy(faceI.data() + f * q, tmp.data(), q * sizeof(double));
          } else {
            std::memcpy(faceI.data() + f * q, buffers.data() + nb * q, q * sizeof(double));
          }
        }
        const std::int32_t local[4] = {0, 1, 2, 3};
        oracle_neighbor(order, gm, 1, faceI.data(), local, faceRelations + 8 * e, aminus + 324 * e,
                        faceTypes ? faceTypes + 4 * e : nullptr, dofs + e * q);
      }
    }
  return 0;
}

}  // extern "C"
