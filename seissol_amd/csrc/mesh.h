// Synthetic meshes, materials and DOFs for the proxy harness (host side).
#pragma once
#include <cstdint>

namespace sdg {

struct MaterialRange {
  double rhoMin, rhoMax, vp, vs;
};

// material: [n][3] = (rho, lambda, mu)
void randomMaterials(std::int64_t n, std::uint64_t seed, const MaterialRange& range, double* material);
// vertices [n][4][3], neighbors [n][4], faceRelations [n][4][2] = (neighbour side, orientation)
void proxyMesh(std::int64_t n, std::uint64_t seed, double* vertices, std::int32_t* neighbors,
               std::uint8_t* faceRelations);
std::int64_t periodicMeshSize(int nx, int ny, int nz);
int periodicMesh(int nx, int ny, int nz, double h, double jitter, std::uint64_t seed, double* vertices,
                 std::int32_t* neighbors, std::uint8_t* faceRelations);
double minCflTimestep(std::int64_t n, int order, double cfl, const double* vertices, const double* material);
void uniformDofs(std::int64_t count, std::uint64_t seed, double* out);

}  // namespace sdg
