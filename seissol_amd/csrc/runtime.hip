// C ABI of the engine: errors, global matrices, synthetic inputs, precompute,
// engine lifetime, batched kernels and the proxy kernel selection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/seissol_amd.h"
#include "cell_local.h"
#include "engine.h"
#include "global_matrices.h"
#include "mesh.h"
#include "mfma_common.h"

namespace sdg {

namespace {
thread_local std::string gLastError;
}

void setError(const std::string& msg) { gLastError = msg; }

int hipFail(hipError_t err, const char* what) {
  gLastError = std::string(what) + ": " + hipGetErrorString(err);
  return SDG_ERR_HIP;
}

namespace {

// Reorder the generated matrices into the transposed device layout.
std::vector<double> deviceMatrices(int order, const double* gm) {
  const int B = numBasis3(order), B2 = numBasis2(order);
  const auto g = globalMatrixLayout(order);
  const auto d = deviceMatrixLayout(order);
  std::vector<double> out(d.total, 0.0);
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < B; ++k)
      for (int l = 0; l < B; ++l) {
        out[d.ck + (j * B + l) * B + k] = -gm[g.kDivMT + (j * B + k) * B + l];
        out[d.vol + (j * B + l) * B + k] = gm[g.kDivM + (j * B + k) * B + l];
      }
  for (int f = 0; f < 4; ++f)
    for (int m = 0; m < B2; ++m)
      for (int l = 0; l < B; ++l) {
        out[d.fmrt + (f * B + l) * B2 + m] = gm[g.fMrT + (f * B2 + m) * B + l];
        out[d.rt + (f * B + l) * B2 + m] = gm[g.rT + (f * B2 + m) * B + l];
        out[d.rdm + (f * B2 + m) * B + l] = gm[g.rDivM + (f * B + l) * B2 + m];
      }
  for (int h = 0; h < 3; ++h)
    for (int m = 0; m < B2; ++m)
      for (int nn = 0; nn < B2; ++nn) out[d.fp + (h * B2 + nn) * B2 + m] = gm[g.fP + (h * B2 + m) * B2 + nn];
  return out;
}

int uploadMatrices(EngineState& s, const double* gm) {
  const auto host = deviceMatrices(s.order, gm);
  s.layout = deviceMatrixLayout(s.order);
  const size_t bytes = host.size() * (s.arithmetic == SDG_FP64 ? sizeof(double) : sizeof(float));
  if (s.matrices == nullptr) SDG_HIP_CHECK(hipMalloc(&s.matrices, bytes));
  if (s.arithmetic == SDG_FP64) {
    SDG_HIP_CHECK(hipMemcpy(s.matrices, host.data(), bytes, hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(host.begin(), host.end());
    SDG_HIP_CHECK(hipMemcpy(s.matrices, f.data(), bytes, hipMemcpyHostToDevice));
  }
  return SDG_OK;
}

uint16_t bf16Bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                                      // round to nearest even
  return static_cast<uint16_t>(u >> 16);
}
float bf16Value(uint16_t h) {
  const uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

uint16_t f16Bits(float f) {
  const _Float16 h = static_cast<_Float16>(f);  // round to nearest even
  uint16_t b;
  std::memcpy(&b, &h, 2);
  return b;
}
float f16Value(uint16_t b) {
  _Float16 h;
  std::memcpy(&h, &b, 2);
  return static_cast<float>(h);
}

// Split a row-major FP64 matrix (times `scale`) into three planes of bf16 or
// fp16 (x = hi + mid + lo, each plane the round-to-nearest of the remainder)
// laid out in the lane order of v_mfma_f32_16x16x32_{bf16,f16}'s A operand:
// lane l of fragment (mt, ks) holds A[16 mt + l % 16][32 ks + 8 (l / 16) + i], i < 8.
void appendFragments(const std::function<double(int, int)>& A, int rows, int cols, int mtiles, int ksteps,
                     uint16_t* out, bool f16, double scale) {
  for (int mt = 0; mt < mtiles; ++mt)
    for (int ks = 0; ks < ksteps; ++ks) {
      uint16_t* base = out + (static_cast<int64_t>(mt) * ksteps + ks) * mfma::kFragPair;
      for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 8; ++i) {
          const int r = 16 * mt + l % 16, c = 32 * ks + 8 * (l / 16) + i;
          double rest = (r < rows && c < cols) ? A(r, c) * scale : 0.0;
          for (int plane = 0; plane < 3; ++plane) {
            const uint16_t h = f16 ? f16Bits(static_cast<float>(rest)) : bf16Bits(static_cast<float>(rest));
            base[plane * mfma::kFragPlane + l * 8 + i] = h;
            rest -= static_cast<double>(f16 ? f16Value(h) : bf16Value(h));
          }
        }
    }
}

// power of two that brings the largest |entry| into [2^13, 2^14) (fp16 form), else 1
double fragmentScale(const std::function<double(int, int)>& A, int rows, int cols, bool f16) {
  if (!f16) return 1.0;
  double mx = 0.0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) mx = std::max(mx, std::fabs(A(r, c)));
  if (mx == 0.0) return 1.0;
  return std::ldexp(1.0, 13 - std::ilogb(mx));
}

template <int O>
int uploadMfmaData(EngineState& s, const double* gm) {
  using S = mfma::MShape<O>;
  const auto g = globalMatrixLayout(O);
  const int B = S::B, B2 = S::B2;
  const bool f16 = s.arithmetic == SDG_F16X3;
  std::vector<uint16_t> frags(S::total, 0);
  std::vector<float> tables(S::tTotal);
  auto add = [&](const std::function<double(int, int)>& A, int rows, int cols, int mtiles, int ksteps,
                 int64_t offset, int unscaleSlot) {
    const double sc = fragmentScale(A, rows, cols, f16);
    appendFragments(A, rows, cols, mtiles, ksteps, frags.data() + offset, f16, sc);
    tables[S::tAUn + unscaleSlot] = static_cast<float>(1.0 / sc);
  };
  for (int j = 0; j < 3; ++j) {
    add([&](int r, int c) { return -gm[g.kDivMT + (j * B + r) * B + c]; }, B, B, S::MT, S::KSB,
        S::ck + static_cast<int64_t>(j) * S::MT * S::KSB * mfma::kFragPair, j);
    add([&](int r, int c) { return gm[g.kDivM + (j * B + r) * B + c]; }, B, B, S::MT, S::KSB,
        S::vol + static_cast<int64_t>(j) * S::MT * S::KSB * mfma::kFragPair, 3 + j);
  }
  // the four face operands stacked with B2P rows each (rows B2..B2P-1 zero).
  // The local pass forms the traces rT_f I = (fMrT_f / M2) I directly, so its
  // flux matrices carry the mass factor instead: rDivM_f diag(M2) (the
  // neighbour pass uses rDivM_f itself, per face, below).
  constexpr int B2P = S::B2P;
  add([&](int r, int c) {
        const int m = r % B2P;
        return m < B2 ? gm[g.fMrT + ((r / B2P) * B2 + m) * B + c] / gm[g.mass2 + m] : 0.0;
      },
      S::FR, B, S::MTF, S::KSB, S::fmrt, 6);
  add([&](int r, int c) {
        const int m = c % B2P;
        return m < B2 ? gm[g.rDivM + ((c / B2P) * B + r) * B2 + m] * gm[g.mass2 + m] : 0.0;
      },
      B, S::FR, S::MT, S::KSF, S::rdm, 7);
  // rDivM_f per face (k-step f: face f's B2 columns), for the neighbour kernel's face-by-face product
  static_assert(B2P <= 32, "one face's rows fit one k-step");
  add([&](int r, int c) { return c % 32 < B2 ? gm[g.rDivM + ((c / 32) * B + r) * B2 + c % 32] : 0.0; }, B, 128,
      S::MT, 4, S::rdmF, 8);
  for (int m = 0; m < B2; ++m) tables[S::tInvM2 + m] = static_cast<float>(1.0 / gm[g.mass2 + m]);
  // fP_h rows as (value, column) lists of O entries, padded with zeros
  for (int h = 0; h < 3; ++h) {
    double mx = 0.0;
    for (int i = 0; i < B2 * B2; ++i) mx = std::max(mx, std::fabs(gm[g.fP + h * B2 * B2 + i]));
    for (int r = 0; r < B2; ++r) {
      int cnt = 0;
      for (int c = 0; c < B2; ++c) {
        const double v = gm[g.fP + (h * B2 + r) * B2 + c];
        if (std::fabs(v) <= 1e-13 * mx) continue;
        if (cnt == O) {
          setError("face permutation matrix with more than `order` non-zeros in a row");
          return SDG_ERR_INVALID;
        }
        tables[S::tFP + (h * B2 + r) * O + cnt] = static_cast<float>(v);
        tables[S::tFPCol + (h * B2 + r) * O + cnt] = static_cast<float>(c);
        ++cnt;
      }
    }
  }
  if (s.mfmaFrags == nullptr) SDG_HIP_CHECK(hipMalloc(&s.mfmaFrags, frags.size() * sizeof(uint16_t)));
  if (s.mfmaTables == nullptr) SDG_HIP_CHECK(hipMalloc(&s.mfmaTables, tables.size() * sizeof(float)));
  SDG_HIP_CHECK(hipMemcpy(s.mfmaFrags, frags.data(), frags.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  SDG_HIP_CHECK(hipMemcpy(s.mfmaTables, tables.data(), tables.size() * sizeof(float), hipMemcpyHostToDevice));
  return SDG_OK;
}

int uploadMfma(EngineState& s, const double* gm) {
  switch (s.order) {
    case 3: return uploadMfmaData<3>(s, gm);
    case 4: return uploadMfmaData<4>(s, gm);
    case 5: return uploadMfmaData<5>(s, gm);
    case 6: return uploadMfmaData<6>(s, gm);
    case 7: return uploadMfmaData<7>(s, gm);
    default:
      setError("matrix-core arithmetic supports orders 3..7");
      return SDG_ERR_UNSUPPORTED;
  }
}

}  // namespace

CellLayout engineLayout(const EngineState& e) {
  CellLayout c;
  const int O = e.order;
  const int rs = e.arithmetic == SDG_FP64 ? 8 : 4;
  c.align = e.alignment > rs ? e.alignment / rs : 1;
  c.ldq = c.rows(numBasis3(O));
  c.qStride = 9LL * c.ldq;
  c.starCompact = e.starFormat == SDG_STAR_COMPACT ? 1 : 0;
  c.starStride = c.starCompact ? 72 : 243;
  for (int i = 0; i < O; ++i) c.dsize += 9LL * c.rows(numBasis3(O - i));
  return c;
}

CellLayout cellLayout(const EngineState& e, const sdg_layer* L) {
  CellLayout c = engineLayout(e);
  if (L != nullptr) {
    if (L->star_stride > 0) c.starStride = L->star_stride;
    if (L->aplus_stride > 0) c.aplusStride = L->aplus_stride;
    if (L->aminus_stride > 0) c.aminusStride = L->aminus_stride;
  }
  return c;
}

bool layerStridesOk(const EngineState& e, const sdg_layer* L, const char* who) {
  if (L == nullptr) return true;
  const int64_t starMin = e.starFormat == SDG_STAR_COMPACT ? 72 : 243;
  if (L->star_stride < 0 || L->aplus_stride < 0 || L->aminus_stride < 0 ||
      (L->star_stride > 0 && L->star_stride < starMin) || (L->aplus_stride > 0 && L->aplus_stride < 324) ||
      (L->aminus_stride > 0 && L->aminus_stride < 324)) {
    setError(std::string(who) + ": per-cell strides must be 0 (dense) or at least the block size (star " +
             std::to_string(starMin) + ", aplus/aminus 324 reals)");
    return false;
  }
  return true;
}

namespace {

bool validLayer(const EngineState& st, const sdg_layer* L, bool needNeighbors) {
  if (L == nullptr) { setError("layer is NULL"); return false; }
  if (!layerStridesOk(st, L, "layer")) return false;
  if (L->num_elements < 0) { setError("negative element count"); return false; }
  if (L->num_elements > (int64_t(1) << 31) - 1) { setError("too many elements for 32-bit neighbour ids"); return false; }
  if (L->num_elements == 0) return true;
  // the matrix-core forms pass face traces between the passes: there the
  // time integral is an optional output of the local pass
  const bool tracesOnly = matrixCore(st.arithmetic) && L->face_traces != nullptr;
  if (L->dofs == nullptr || (L->time_integrated == nullptr && !tracesOnly)) {
    setError("dofs/time_integrated missing");
    return false;
  }
  if (needNeighbors && (L->face_neighbors == nullptr || L->face_relations == nullptr || L->aminus == nullptr)) {
    setError("neighbour data missing");
    return false;
  }
  return true;
}

// cell-local precompute on the device, one thread per element
__global__ void kCellLocal(int64_t n, const double* __restrict__ vertices, const double* __restrict__ material,
                           const int32_t* __restrict__ neighbors, const uint8_t* __restrict__ faceTypes,
                           double* __restrict__ star, double* __restrict__ aplus, double* __restrict__ aminus) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double v[4][3];
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < 3; ++c) v[q][c] = vertices[(e * 4 + q) * 3 + c];
  const ElasticMaterial loc{material[3 * e], material[3 * e + 1], material[3 * e + 2]};
  ElasticMaterial nb[4];
  int ft[4];
  for (int f = 0; f < 4; ++f) {
    const int64_t o = neighbors[e * 4 + f];
    nb[f] = ElasticMaterial{material[3 * o], material[3 * o + 1], material[3 * o + 2]};
    ft[f] = faceTypes ? faceTypes[e * 4 + f] : FaceRegular;
  }
  cellLocalMatrices(v, loc, nb, ft, star + e * 243, aplus + e * 324, aminus + e * 324);
}

template <typename Dst>
__global__ void kConvert(int64_t n, const double* __restrict__ src, Dst* __restrict__ dst) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = static_cast<Dst>(src[i]);
}

}  // namespace

// Face traces of sdg_proxy_run_static: one array per GPU, held until the
// process ends.  External linkage keeps it in the code object; host code uses
// only the per-device address from hipGetSymbolAddress, never the name.
constexpr int64_t kScratchFloats = int64_t(16) << 20;  // 64 MiB
__device__ __align__(256) float scratch[kScratchFloats];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex gScratchMutex;
void* gScratchAddress[kMaxScratchDevices] = {};
}  // namespace

// device address of the current GPU's scratch array, looked up once per device
int scratchAddress(void** out) {
  int dev = 0;
  SDG_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) {
    setError("scratch array: device index beyond " + std::to_string(kMaxScratchDevices));
    return SDG_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (gScratchAddress[dev] == nullptr) {
    void* p = nullptr;
    SDG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    gScratchAddress[dev] = p;
  }
  *out = gScratchAddress[dev];
  return SDG_OK;
}

// the kernels of proxy kernels 0-4 (validated layer); no allocation
int proxyRun(const EngineState& st, int kernel_id, int steps, double dt, const sdg_layer* layer, hipStream_t s) {
  double coeffs[kMaxOrder];
  sdg_monomial_integrate(st.order, 0.0, dt, coeffs);
  const bool mc = matrixCore(st.arithmetic);
  if (mc && layer->num_elements > 0 && layer->face_traces == nullptr && (kernel_id == 0 || kernel_id == 1 || kernel_id == 2)) {
    setError("sdg_proxy_run: matrix-core arithmetic needs layer->face_traces");
    return SDG_ERR_INVALID;
  }
  if (layer->num_elements > 0 && layer->time_integrated == nullptr && (!mc || kernel_id >= 3)) {
    setError("sdg_proxy_run: this kernel needs layer->time_integrated");
    return SDG_ERR_INVALID;
  }
  // the FP64 matrix-core passes hand FP64 face traces from the local pass to
  // the neighbour pass (the local pass still writes I); as for the
  // matrix-core forms, `neigh` alone reads the traces of the last local pass
  const bool mc64 = hasMfma64(st) && layer->face_traces != nullptr;
  for (int step = 0; step < steps; ++step) {
    int rc = SDG_OK;
    if (mc64 && kernel_id <= 2) {
      if (kernel_id == 0 || kernel_id == 1) rc = launchAderLocal(st, coeffs, layer, s);
      if (rc == SDG_OK && (kernel_id == 0 || kernel_id == 2)) rc = launchNeighborMfma64(st, layer, s);
      if (rc != SDG_OK) return rc;
      continue;
    }
    if (mc && kernel_id <= 2) {
      if (kernel_id == 0 || kernel_id == 1) rc = launchLocalMfma(st, coeffs, layer, false, s);
      if (rc == SDG_OK && (kernel_id == 0 || kernel_id == 2)) rc = launchNeighborMfma(st, layer, s);
      if (rc != SDG_OK) return rc;
      continue;
    }
    switch (kernel_id) {
      case 0:
        rc = launchAderLocal(st, coeffs, layer, s);
        if (rc == SDG_OK) rc = launchNeighbor(st, layer, s);
        break;
      case 1: rc = launchAderLocal(st, coeffs, layer, s); break;
      case 2: rc = launchNeighbor(st, layer, s); break;
      case 3: rc = launchAder(st, coeffs, layer, s); break;
      case 4: rc = launchLocal(st, layer, s); break;
    }
    if (rc != SDG_OK) return rc;
  }
  return SDG_OK;
}

}  // namespace sdg

using namespace sdg;

extern "C" {

const char* sdg_last_error(void) { return gLastError.c_str(); }
int sdg_version(void) { return 1; }
int sdg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

size_t sdg_global_matrices_size(int order) {
  if (order < 1 || order > 8) return 0;
  return globalMatrixLayout(order).total;
}

int sdg_global_matrices(int order, double* out) {
  if (out == nullptr || order < 1 || order > 8) {
    setError("sdg_global_matrices: order must be in [1, 8] and out non-NULL");
    return SDG_ERR_INVALID;
  }
  return generateGlobalMatrices(order, out) == 0 ? SDG_OK : SDG_ERR_INVALID;
}

int sdg_basis_values(int order, double xi, double eta, double zeta, double* out) {
  if (out == nullptr || order < 1 || order > 8) { setError("invalid arguments"); return SDG_ERR_INVALID; }
  for (int k = 0; k < numBasis3(order); ++k) out[k] = evalTetBasis(order, k, xi, eta, zeta);
  return SDG_OK;
}

int sdg_basis_gradients(int order, double xi, double eta, double zeta, double* out) {
  if (out == nullptr || order < 1 || order > 8) { setError("invalid arguments"); return SDG_ERR_INVALID; }
  const int B = numBasis3(order);
  for (int k = 0; k < B; ++k) {
    double g[3];
    evalTetBasisGradient(order, k, xi, eta, zeta, g);
    for (int d = 0; d < 3; ++d) out[d * B + k] = g[d];
  }
  return SDG_OK;
}

int sdg_mesh_proxy(int64_t n, uint64_t seed, double* vertices, int32_t* neighbors, uint8_t* face_relations) {
  if (n <= 0 || vertices == nullptr || neighbors == nullptr || face_relations == nullptr) {
    setError("sdg_mesh_proxy: invalid arguments");
    return SDG_ERR_INVALID;
  }
  proxyMesh(n, seed, vertices, neighbors, face_relations);
  return SDG_OK;
}

int64_t sdg_mesh_periodic_size(int nx, int ny, int nz) { return periodicMeshSize(nx, ny, nz); }

int sdg_mesh_periodic(int nx, int ny, int nz, double h, double jitter, uint64_t seed, double* vertices,
                      int32_t* neighbors, uint8_t* face_relations) {
  if (vertices == nullptr || neighbors == nullptr || face_relations == nullptr || jitter < 0 || jitter > 0.2) {
    setError("sdg_mesh_periodic: invalid arguments (jitter must be in [0, 0.2])");
    return SDG_ERR_INVALID;
  }
  const int rc = periodicMesh(nx, ny, nz, h, jitter, seed, vertices, neighbors, face_relations);
  if (rc != 0) {
    setError("sdg_mesh_periodic: dimensions must be even and >= 4 (code " + std::to_string(rc) + ")");
    return SDG_ERR_INVALID;
  }
  return SDG_OK;
}

int sdg_random_materials(int64_t n, uint64_t seed, double rho_min, double rho_max, double vp, double vs,
                         double* material) {
  if (n < 0 || material == nullptr || rho_min <= 0 || rho_max < rho_min || vs <= 0 || vp <= vs * std::sqrt(2.0)) {
    setError("sdg_random_materials: invalid arguments (need vp > sqrt(2) vs > 0)");
    return SDG_ERR_INVALID;
  }
  randomMaterials(n, seed, MaterialRange{rho_min, rho_max, vp, vs}, material);
  return SDG_OK;
}

double sdg_cfl_timestep(int64_t n, int order, double cfl, const double* vertices, const double* material) {
  if (n <= 0 || vertices == nullptr || material == nullptr) return -1.0;
  return minCflTimestep(n, order, cfl, vertices, material);
}

int sdg_uniform_dofs(int64_t count, uint64_t seed, double* out) {
  if (count < 0 || out == nullptr) { setError("invalid arguments"); return SDG_ERR_INVALID; }
  uniformDofs(count, seed, out);
  return SDG_OK;
}

int sdg_monomial_integrate(int order, double t0, double t1, double* coeffs) {
  if (order < 1 || coeffs == nullptr) { setError("invalid arguments"); return SDG_ERR_INVALID; }
  double p0 = t0, p1 = t1, fact = 1.0;
  for (int i = 0; i < order; ++i) {
    fact *= (i + 1);
    coeffs[i] = (p1 - p0) / fact;
    p0 *= t0;
    p1 *= t1;
  }
  return SDG_OK;
}

int sdg_cell_local_matrices(int64_t n, const double* vertices, const double* material, const int32_t* neighbors,
                            const uint8_t* face_types, double* star, double* aplus, double* aminus, void* stream) {
  if (n < 0 || (n > 0 && (!vertices || !material || !neighbors || !star || !aplus || !aminus))) {
    setError("sdg_cell_local_matrices: invalid arguments");
    return SDG_ERR_INVALID;
  }
  if (n == 0) return SDG_OK;
  const unsigned grid = static_cast<unsigned>((n + 63) / 64);
  kCellLocal<<<grid, 64, 0, static_cast<hipStream_t>(stream)>>>(n, vertices, material, neighbors, face_types,
                                                                  star, aplus, aminus);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

int64_t sdg_derivatives_size(int order) {
  int64_t s = 0;
  for (int i = 0; i < order; ++i) s += 9LL * numBasis3(order - i);
  return s;
}

int sdg_engine_create(sdg_engine** engine, int order, int arithmetic) {
  if (engine == nullptr) { setError("engine is NULL"); return SDG_ERR_INVALID; }
  *engine = nullptr;
  if (order < kMinOrder || order > kMaxOrder) {
    setError("sdg_engine_create: order must be in [2, 8]");
    return SDG_ERR_UNSUPPORTED;
  }
  if (arithmetic < SDG_FP64 || arithmetic > SDG_F16X3) {
    setError("sdg_engine_create: unknown arithmetic");
    return SDG_ERR_INVALID;
  }
  // the current GPU's scratch address (sdg_proxy_run_static) is resolved at set-up
  void* scratchPtr = nullptr;
  if (const int src = scratchAddress(&scratchPtr); src != SDG_OK) return src;
  auto* e = new sdg_engine();
  e->s.order = order;
  e->s.arithmetic = arithmetic;
  std::vector<double> gm(globalMatrixLayout(order).total);
  generateGlobalMatrices(order, gm.data());
  int rc = uploadMatrices(e->s, gm.data());
  if (rc == SDG_OK && matrixCore(arithmetic)) rc = uploadMfma(e->s, gm.data());
  if (rc == SDG_OK && arithmetic == SDG_FP64) rc = uploadMfma64(e->s, gm.data());
  if (rc != SDG_OK) {
    if (e->s.matrices) (void)hipFree(e->s.matrices);
    if (e->s.mfmaFrags) (void)hipFree(e->s.mfmaFrags);
    if (e->s.mfmaTables) (void)hipFree(e->s.mfmaTables);
    if (e->s.mfma64Frags) (void)hipFree(e->s.mfma64Frags);
    delete e;
    return rc;
  }
  *engine = e;
  return SDG_OK;
}

int sdg_engine_set_global_matrices(sdg_engine* engine, const double* host_matrices) {
  if (engine == nullptr || host_matrices == nullptr) { setError("invalid arguments"); return SDG_ERR_INVALID; }
  const int rc = uploadMatrices(engine->s, host_matrices);
  if (rc != SDG_OK) return rc;
  if (engine->s.arithmetic == SDG_FP64) return uploadMfma64(engine->s, host_matrices);
  if (!matrixCore(engine->s.arithmetic)) return rc;
  return uploadMfma(engine->s, host_matrices);
}

int sdg_engine_set_kernels(sdg_engine* engine, int family) {
  if (engine == nullptr || (family != SDG_KERNELS_AUTO && family != SDG_KERNELS_VALU)) {
    setError("sdg_engine_set_kernels: engine is NULL or the family is unknown");
    return SDG_ERR_INVALID;
  }
  if (family == SDG_KERNELS_VALU && matrixCore(engine->s.arithmetic)) {
    setError("sdg_engine_set_kernels: the split-16-bit forms run on the matrix cores only");
    return SDG_ERR_UNSUPPORTED;
  }
  engine->s.kernels = family;
  return SDG_OK;
}

int sdg_engine_set_layout(sdg_engine* engine, int alignment, int star_format) {
  if (engine == nullptr) { setError("sdg_engine_set_layout: engine is NULL"); return SDG_ERR_INVALID; }
  const int rs = engine->s.arithmetic == SDG_FP64 ? 8 : 4;
  if (alignment < 0 || alignment > 1024 || (alignment > 0 && ((alignment & (alignment - 1)) != 0 || alignment % rs != 0))) {
    setError("sdg_engine_set_layout: alignment must be 0 or a power of two, a multiple of sizeof(real), <= 1024");
    return SDG_ERR_INVALID;
  }
  if (star_format != SDG_STAR_DENSE && star_format != SDG_STAR_COMPACT) {
    setError("sdg_engine_set_layout: star_format must be SDG_STAR_DENSE or SDG_STAR_COMPACT");
    return SDG_ERR_INVALID;
  }
  engine->s.alignment = alignment;
  engine->s.starFormat = star_format;
  return SDG_OK;
}

int sdg_layout_sizes(const sdg_engine* engine, int32_t* dofs_ld, int64_t* dofs_reals, int64_t* derivatives_reals) {
  if (engine == nullptr) { setError("sdg_layout_sizes: engine is NULL"); return SDG_ERR_INVALID; }
  const CellLayout c = engineLayout(engine->s);
  if (dofs_ld) *dofs_ld = c.ldq;
  if (dofs_reals) *dofs_reals = c.qStride;
  if (derivatives_reals) *derivatives_reals = c.dsize;
  return SDG_OK;
}

int sdg_engine_destroy(sdg_engine* engine) {
  if (engine == nullptr) return SDG_OK;
  if (engine->s.matrices) SDG_HIP_CHECK(hipFree(engine->s.matrices));
  if (engine->s.mfmaFrags) SDG_HIP_CHECK(hipFree(engine->s.mfmaFrags));
  if (engine->s.mfmaTables) SDG_HIP_CHECK(hipFree(engine->s.mfmaTables));
  if (engine->s.drMatrices) SDG_HIP_CHECK(hipFree(engine->s.drMatrices));
  if (engine->s.mfma64Frags) SDG_HIP_CHECK(hipFree(engine->s.mfma64Frags));
  delete engine;
  return SDG_OK;
}

int sdg_compute_batched_ader(sdg_engine* engine, const double* coeffs, double /*time_step_width*/,
                             const sdg_layer* layer, void* stream) {
  if (engine == nullptr || coeffs == nullptr || !validLayer(engine->s, layer, false)) return SDG_ERR_INVALID;
  if (layer->num_elements > 0 && (layer->star == nullptr || layer->time_integrated == nullptr)) {
    setError("star matrices or time_integrated missing");
    return SDG_ERR_INVALID;
  }
  return launchAder(engine->s, coeffs, layer, static_cast<hipStream_t>(stream));
}

int sdg_compute_batched_ader_local(sdg_engine* engine, const double* coeffs, double /*time_step_width*/,
                                   const sdg_layer* layer, void* stream) {
  if (engine == nullptr || coeffs == nullptr || !validLayer(engine->s, layer, false)) return SDG_ERR_INVALID;
  if (layer->num_elements > 0 && (layer->star == nullptr || layer->aplus == nullptr)) {
    setError("star/aplus matrices missing");
    return SDG_ERR_INVALID;
  }
  if (matrixCore(engine->s.arithmetic) && layer->face_traces != nullptr)
    return launchLocalMfma(engine->s, coeffs, layer, true, static_cast<hipStream_t>(stream));
  return launchAderLocal(engine->s, coeffs, layer, static_cast<hipStream_t>(stream));
}

int sdg_evaluate_batched(sdg_engine* engine, const double* coeffs, const void* derivatives, void* time_evaluated,
                         int64_t num_elements, void* stream) {
  if (engine == nullptr || coeffs == nullptr || num_elements < 0 ||
      (num_elements > 0 && (derivatives == nullptr || time_evaluated == nullptr))) {
    setError("sdg_evaluate_batched: invalid arguments");
    return SDG_ERR_INVALID;
  }
  return launchTaylor(engine->s, coeffs, derivatives, time_evaluated, num_elements,
                      static_cast<hipStream_t>(stream));
}

int sdg_compute_batched_integral(sdg_engine* engine, const sdg_layer* layer, void* stream) {
  if (engine == nullptr || !validLayer(engine->s, layer, false)) return SDG_ERR_INVALID;
  if (layer->num_elements > 0 &&
      (layer->star == nullptr || layer->aplus == nullptr || layer->time_integrated == nullptr)) {
    setError("star/aplus matrices or time_integrated missing");
    return SDG_ERR_INVALID;
  }
  return launchLocal(engine->s, layer, static_cast<hipStream_t>(stream));
}

int sdg_compute_batched_neighbors_integral(sdg_engine* engine, const sdg_layer* layer, void* stream) {
  if (engine == nullptr || !validLayer(engine->s, layer, true)) return SDG_ERR_INVALID;
  if (matrixCore(engine->s.arithmetic) && layer->face_traces != nullptr)
    return launchNeighborMfma(engine->s, layer, static_cast<hipStream_t>(stream));
  // the FP64 engine reads the time-integrated DOFs here, as the reference's
  // neighbour integral does; its matrix-core neighbour pass runs only inside
  // the fused proxy step, whose local pass writes the traces it reads
  return launchNeighbor(engine->s, layer, static_cast<hipStream_t>(stream));
}

int sdg_proxy_kernel_id(const char* name) {
  if (name == nullptr) return -1;
  const char* names[] = {"all", "local", "neigh", "ader", "localwoader", "neigh_dr", "godunov_dr", "all_dr"};
  for (int i = 0; i < 8; ++i)
    if (std::strcmp(name, names[i]) == 0) return i;
  return -1;
}

int sdg_proxy_run(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer, void* stream) {
  if (kernel_id >= 5 && kernel_id <= 7) {
    setError("sdg_proxy_run: the DR kernels need DR faces; call sdg_proxy_run_dr");
    return SDG_ERR_INVALID;
  }
  if (engine == nullptr || steps < 0 || kernel_id < 0 || kernel_id > 4 || !validLayer(engine->s, layer, kernel_id == 0 || kernel_id == 2)) {
    if (engine == nullptr || steps < 0 || kernel_id < 0 || kernel_id > 4) setError("sdg_proxy_run: invalid arguments");
    return SDG_ERR_INVALID;
  }
  return proxyRun(engine->s, kernel_id, steps, dt, layer, static_cast<hipStream_t>(stream));
}

int sdg_proxy_run_static(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer,
                         void* stream) {
  if (engine == nullptr || layer == nullptr || steps < 0 || kernel_id < 0 || kernel_id > 4) {
    setError("sdg_proxy_run_static: invalid arguments (kernels 0-4)");
    return SDG_ERR_INVALID;
  }
  if (!matrixCore(engine->s.arithmetic)) return sdg_proxy_run(engine, kernel_id, steps, dt, layer, stream);
  if (kernel_id == 2) {
    setError("sdg_proxy_run_static: the neighbour kernel alone would read the face traces of an earlier call");
    return SDG_ERR_UNSUPPORTED;
  }
  int b2p = 0;
  switch (engine->s.order) {
    case 3: b2p = mfma::MShape<3>::B2P; break;
    case 4: b2p = mfma::MShape<4>::B2P; break;
    case 5: b2p = mfma::MShape<5>::B2P; break;
    case 6: b2p = mfma::MShape<6>::B2P; break;
    case 7: b2p = mfma::MShape<7>::B2P; break;
    default: setError("sdg_proxy_run_static: matrix-core path supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
  const int64_t perElement = 36LL * b2p;
  if (layer->num_elements < 0 || layer->num_elements > kScratchFloats / perElement) {
    setError("sdg_proxy_run_static: at most " + std::to_string(kScratchFloats / perElement) +
             " elements at this order (64 MiB of face traces)");
    return SDG_ERR_INVALID;
  }
  // the local pass of every step writes each trace the neighbour pass reads
  // (padding rows included), so the array needs no clearing between calls
  void* traces = nullptr;
  if (const int rc = scratchAddress(&traces); rc != SDG_OK) return rc;
  sdg_layer L = *layer;
  L.face_traces = traces;
  if (!validLayer(engine->s, &L, kernel_id == 0)) return SDG_ERR_INVALID;
  return proxyRun(engine->s, kernel_id, steps, dt, &L, static_cast<hipStream_t>(stream));
}

int sdg_upload_reals(sdg_engine* engine, void* device_dst, const double* host_src, int64_t count, void* stream) {
  if (engine == nullptr || count < 0 || (count > 0 && (!device_dst || !host_src))) {
    setError("sdg_upload_reals: invalid arguments");
    return SDG_ERR_INVALID;
  }
  auto st = static_cast<hipStream_t>(stream);
  if (engine->s.arithmetic == SDG_FP64) {
    SDG_HIP_CHECK(hipMemcpyAsync(device_dst, host_src, count * sizeof(double), hipMemcpyHostToDevice, st));
    SDG_HIP_CHECK(hipStreamSynchronize(st));
  } else {
    std::vector<float> f(host_src, host_src + count);
    SDG_HIP_CHECK(hipMemcpyAsync(device_dst, f.data(), count * sizeof(float), hipMemcpyHostToDevice, st));
    SDG_HIP_CHECK(hipStreamSynchronize(st));
  }
  return SDG_OK;
}

int sdg_download_reals(sdg_engine* engine, double* host_dst, const void* device_src, int64_t count, void* stream) {
  if (engine == nullptr || count < 0 || (count > 0 && (!device_src || !host_dst))) {
    setError("sdg_download_reals: invalid arguments");
    return SDG_ERR_INVALID;
  }
  auto st = static_cast<hipStream_t>(stream);
  if (engine->s.arithmetic == SDG_FP64) {
    SDG_HIP_CHECK(hipMemcpyAsync(host_dst, device_src, count * sizeof(double), hipMemcpyDeviceToHost, st));
    SDG_HIP_CHECK(hipStreamSynchronize(st));
  } else {
    std::vector<float> f(count);
    SDG_HIP_CHECK(hipMemcpyAsync(f.data(), device_src, count * sizeof(float), hipMemcpyDeviceToHost, st));
    SDG_HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < count; ++i) host_dst[i] = f[i];
  }
  return SDG_OK;
}

int sdg_convert_fp64_to_real(sdg_engine* engine, void* device_dst, const double* device_src, int64_t count,
                             void* stream) {
  if (engine == nullptr || count < 0 || (count > 0 && (!device_dst || !device_src))) {
    setError("sdg_convert_fp64_to_real: invalid arguments");
    return SDG_ERR_INVALID;
  }
  if (count == 0) return SDG_OK;
  auto st = static_cast<hipStream_t>(stream);
  if (engine->s.arithmetic == SDG_FP64) {
    SDG_HIP_CHECK(hipMemcpyAsync(device_dst, device_src, count * sizeof(double), hipMemcpyDeviceToDevice, st));
  } else {
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    kConvert<float><<<grid, 256, 0, st>>>(count, device_src, static_cast<float*>(device_dst));
    SDG_HIP_CHECK(hipGetLastError());
  }
  return SDG_OK;
}

}  // extern "C"
