// Synthetic tetrahedral meshes for the proxy harness.
//
// Two generators:
//  * proxyMesh: independent, seeded random affine images of the reference
//    tetrahedron (det J in [0.5, 2]) with random face neighbours and random
//    face relations, the data the reference proxy fabricates
//    (src/Proxy/Allocator.cpp:49-89), but with physical geometry and material
//    so the per-element operators are genuine.
//  * periodicMesh: a conforming periodic box of cubes, each split into five
//    tetrahedra (mirrored split on alternating cubes), vertices jittered with a
//    seeded periodic displacement.  Neighbours, neighbour sides and relative
//    orientations are derived from shared vertices the way a mesh reader does
//    (src/Geometry/PUMLReader.cpp:404-427, orientation = position of this
//    face's first vertex in the neighbour's face vertex list).
#include "mesh.h"

#include "cell_local.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <random>
#include <unordered_map>
#include <vector>

namespace sdg {

namespace {
double det3(const double e[3][3]) {
  return e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
         e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
         e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
}
}  // namespace

void randomMaterials(std::int64_t n, std::uint64_t seed, const MaterialRange& range, double* material) {
  std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ULL);
  std::uniform_real_distribution<double> rhoDist(range.rhoMin, range.rhoMax);
  for (std::int64_t e = 0; e < n; ++e) {
    const double rho = range.rhoMin == range.rhoMax ? range.rhoMin : rhoDist(rng);
    const double mu = rho * range.vs * range.vs;
    const double lambda = rho * range.vp * range.vp - 2.0 * mu;
    material[3 * e + 0] = rho;
    material[3 * e + 1] = lambda;
    material[3 * e + 2] = mu;
  }
}

void proxyMesh(std::int64_t n, std::uint64_t seed, double* vertices, std::int32_t* neighbors,
               std::uint8_t* faceRelations) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::uniform_real_distribution<double> detDist(0.5, 2.0);
  std::uniform_int_distribution<int> side(0, 3);
  std::uniform_int_distribution<int> orient(0, 2);
  std::uniform_int_distribution<std::int64_t> cell(0, n - 1);
  const double ref[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (std::int64_t e = 0; e < n; ++e) {
    double J[3][3];
    double d = 0.0;
    // a well-shaped random affine map: identity plus a bounded perturbation,
    // rescaled to the drawn determinant
    do {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) J[r][c] = (r == c ? 1.0 : 0.0) + 0.35 * u(rng);
      d = det3(J);
    } while (d < 0.2);
    const double target = detDist(rng);
    const double s = std::cbrt(target / d);
    double x0[3] = {10.0 * u(rng), 10.0 * u(rng), 10.0 * u(rng)};
    for (int v = 0; v < 4; ++v)
      for (int r = 0; r < 3; ++r) {
        double acc = x0[r];
        for (int c = 0; c < 3; ++c) acc += s * J[r][c] * ref[v][c];
        vertices[(e * 4 + v) * 3 + r] = acc;
      }
    for (int f = 0; f < 4; ++f) {
      faceRelations[(e * 4 + f) * 2 + 0] = static_cast<std::uint8_t>(side(rng));
      faceRelations[(e * 4 + f) * 2 + 1] = static_cast<std::uint8_t>(orient(rng));
      neighbors[e * 4 + f] = static_cast<std::int32_t>(cell(rng));
    }
  }
}

std::int64_t periodicMeshSize(int nx, int ny, int nz) {
  return 5LL * nx * ny * nz;
}

int periodicMesh(int nx, int ny, int nz, double h, double jitter, std::uint64_t seed, double* vertices,
                 std::int32_t* neighbors, std::uint8_t* faceRelations) {
  if (nx < 4 || ny < 4 || nz < 4 || (nx | ny | nz) & 1) return -1;
  const std::int64_t nv = static_cast<std::int64_t>(nx) * ny * nz;
  std::vector<double> disp(3 * nv);
  {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto& x : disp) x = jitter * h * u(rng);
  }
  auto vid = [&](int i, int j, int k) -> std::int64_t {
    i = ((i % nx) + nx) % nx;
    j = ((j % ny) + ny) % ny;
    k = ((k % nz) + nz) % nz;
    return (static_cast<std::int64_t>(k) * ny + j) * nx + i;
  };
  // corners c = a + 2b + 4c (a, b, c in {0,1})
  const int evenTets[5][4] = {{0, 3, 5, 6}, {1, 0, 3, 5}, {2, 0, 3, 6}, {4, 0, 5, 6}, {7, 3, 5, 6}};
  const int oddTets[5][4] = {{1, 2, 4, 7}, {0, 1, 2, 4}, {3, 1, 2, 7}, {5, 1, 4, 7}, {6, 2, 4, 7}};
  std::vector<std::array<std::int64_t, 4>> ids(periodicMeshSize(nx, ny, nz));
  std::int64_t e = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const auto& tets = ((i + j + k) & 1) ? oddTets : evenTets;
        for (int t = 0; t < 5; ++t, ++e) {
          double v[4][3];
          std::int64_t gid[4];
          for (int q = 0; q < 4; ++q) {
            const int c = tets[t][q];
            const int a = c & 1, b = (c >> 1) & 1, cc = (c >> 2) & 1;
            gid[q] = vid(i + a, j + b, k + cc);
            v[q][0] = (i + a) * h + disp[3 * gid[q] + 0];
            v[q][1] = (j + b) * h + disp[3 * gid[q] + 1];
            v[q][2] = (k + cc) * h + disp[3 * gid[q] + 2];
          }
          double ed[3][3];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) ed[r][c] = v[r + 1][c] - v[0][c];
          if (det3(ed) < 0) {  // positive orientation gives outward face normals
            std::swap(gid[2], gid[3]);
            for (int c = 0; c < 3; ++c) std::swap(v[2][c], v[3][c]);
          }
          for (int q = 0; q < 4; ++q)
            for (int c = 0; c < 3; ++c) vertices[(e * 4 + q) * 3 + c] = v[q][c];
          ids[e] = {gid[0], gid[1], gid[2], gid[3]};
        }
      }
  // match faces through their sorted vertex ids
  struct Key {
    std::int64_t a, b, c;
    bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c; }
  };
  struct KeyHash {
    std::size_t operator()(const Key& k) const {
      return std::hash<std::int64_t>()(k.a * 73856093LL ^ k.b * 19349663LL ^ k.c * 83492791LL);
    }
  };
  std::unordered_map<Key, std::vector<std::pair<std::int64_t, int>>, KeyHash> faces;
  const std::int64_t n = static_cast<std::int64_t>(ids.size());
  faces.reserve(2 * n);
  auto keyOf = [&](std::int64_t el, int f) {
    std::int64_t v[3] = {ids[el][faceNode(f, 0)], ids[el][faceNode(f, 1)], ids[el][faceNode(f, 2)]};
    std::sort(v, v + 3);
    return Key{v[0], v[1], v[2]};
  };
  for (std::int64_t el = 0; el < n; ++el)
    for (int f = 0; f < 4; ++f) faces[keyOf(el, f)].push_back({el, f});
  for (std::int64_t el = 0; el < n; ++el)
    for (int f = 0; f < 4; ++f) {
      const auto& list = faces[keyOf(el, f)];
      if (list.size() != 2) return -2;
      const auto& other = list[0].first == el && list[0].second == f ? list[1] : list[0];
      const std::int64_t nb = other.first;
      const int side = other.second;
      const std::int64_t first = ids[el][faceNode(f, 0)];
      int orientation = -1;
      for (int q = 0; q < 3; ++q)
        if (ids[nb][faceNode(side, q)] == first) orientation = q;
      if (orientation < 0) return -3;
      neighbors[el * 4 + f] = static_cast<std::int32_t>(nb);
      faceRelations[(el * 4 + f) * 2 + 0] = static_cast<std::uint8_t>(side);
      faceRelations[(el * 4 + f) * 2 + 1] = static_cast<std::uint8_t>(orientation);
    }
  return 0;
}

double minCflTimestep(std::int64_t n, int order, double cfl, const double* vertices, const double* material) {
  double dt = 1e300;
  for (std::int64_t e = 0; e < n; ++e) {
    double v[4][3];
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 3; ++c) v[q][c] = vertices[(e * 4 + q) * 3 + c];
    const ElasticMaterial m{material[3 * e], material[3 * e + 1], material[3 * e + 2]};
    dt = std::min(dt, cflTimestep(v, m, order, cfl));
  }
  return dt;
}

void uniformDofs(std::int64_t count, std::uint64_t seed, double* out) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  for (std::int64_t i = 0; i < count; ++i) out[i] = u(rng);
}

}  // namespace sdg
