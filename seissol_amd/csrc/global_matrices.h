// Global ADER-DG matrices of the modal Dubiner basis (host side, FP64).
#pragma once
#include <cstddef>

namespace sdg {

constexpr int numBasis3(int order) { return order * (order + 1) * (order + 2) / 6; }
constexpr int numBasis2(int order) { return order * (order + 1) / 2; }

// Offsets (in doubles) of each matrix family in the flat buffer written by
// generateGlobalMatrices.  All matrices are row-major [rows][cols].
struct GlobalMatrixLayout {
  std::size_t kDivM;   // [3][B][B]
  std::size_t kDivMT;  // [3][B][B]   (as in the reference data, i.e. not negated)
  std::size_t rDivM;   // [4][B][B2]
  std::size_t fMrT;    // [4][B2][B]
  std::size_t rT;      // [4][B2][B]
  std::size_t fP;      // [3][B2][B2]
  std::size_t mass;    // [B]   diagonal of the 3D mass matrix
  std::size_t mass2;   // [B2]  diagonal of the 2D (face) mass matrix
  std::size_t total;
};

GlobalMatrixLayout globalMatrixLayout(int order);
int generateGlobalMatrices(int order, double* out);

// Point on face `face` of the reference tetrahedron for face coordinates
// (chi, tau) seen with orientation `orientation` (-1: none).
void faceToReference(int face, double chi, double tau, int orientation, double out[3]);
void orientationMap(int orientation, double chi, double tau, double* a, double* b);
double evalTetBasis(int order, int index, double xi, double eta, double zeta);
void evalTetBasisGradient(int order, int index, double xi, double eta, double zeta, double out[3]);
double evalTriBasis(int order, int index, double chi, double tau);

}  // namespace sdg
