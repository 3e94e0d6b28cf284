// Internal engine state and kernel entry points shared by the HIP sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/seissol_amd.h"

namespace sdg {

void setError(const std::string& msg);
int hipFail(hipError_t err, const char* what);
#define SDG_HIP_CHECK(call)                                  \
  do {                                                       \
    hipError_t err_ = (call);                                \
    if (err_ != hipSuccess) return ::sdg::hipFail(err_, #call); \
  } while (0)

constexpr int kMaxOrder = 8;
constexpr int kMinOrder = 2;

// Offsets (in reals) of the device copies of the global matrices.  All
// matrices are stored transposed relative to the generator output so that
// consecutive output rows are consecutive in memory:
//   ck  [3][B][B]    ck[j][l][k]   = -kDivMT_j[k][l]
//   vol [3][B][B]    vol[j][l][k]  =  kDivM_j[k][l]
//   fmrt[4][B][B2]   fmrt[f][l][m] =  fMrT_f[m][l]
//   rt  [4][B][B2]   rt[j][l][m]   =  rT_j[m][l]
//   fp  [3][B2][B2]  fp[h][n][m]   =  fP_h[m][n]
//   rdm [4*B2][B]    rdm[f*B2+m][k] = rDivM_f[k][m]
struct DeviceMatrixLayout {
  int64_t ck, vol, fmrt, rt, fp, rdm, total;
};
DeviceMatrixLayout deviceMatrixLayout(int order);

inline bool matrixCore(int arithmetic) {
  return arithmetic == SDG_BF16X3 || arithmetic == SDG_BF16X6 || arithmetic == SDG_F16X3;
}

// Storage layout of a layer's per-cell arrays as the kernels see them.
// DOF-like blocks (dofs, buffers, Taylor outputs) are [9][ldq] per cell, a
// derivative block of m rows is [9][rows(m)], the blocks one after another;
// the per-cell operators sit starStride / aplusStride / aminusStride reals
// apart, the star matrices as 3 x 81 dense or 3 x 24 non-zeros.
struct CellLayout {
  int32_t ldq = 0;
  int32_t align = 1;      // rows of a block are rounded up to a multiple of this (a power of two)
  int64_t qStride = 0;    // 9 * ldq
  int64_t dsize = 0;      // reals of one cell's derivatives
  int64_t starStride = 0, aplusStride = 324, aminusStride = 324;
  int32_t starCompact = 0;
  int32_t pad_ = 0;
  __host__ __device__ int rows(int m) const { return (m + align - 1) & ~(align - 1); }
};

struct EngineState {
  int order = 0;
  int arithmetic = SDG_FP64;
  // layout of the cells' arrays (sdg_engine_set_layout): alignment in bytes
  // (0 = packed) and the star format
  int alignment = 0;
  int starFormat = SDG_STAR_DENSE;
  void* matrices = nullptr;  // device, real type of the engine
  DeviceMatrixLayout layout{};
  // matrix-core arithmetics: pre-split 16-bit fragments and FP32 tables
  void* mfmaFrags = nullptr;
  void* mfmaTables = nullptr;
  // dynamic rupture face matrices (V3mTo2n, V3mTo2nTWDivM), on first use
  void* drMatrices = nullptr;
  // FP64 engines of orders 3-7: FP64 fragments of the matrix-core local pass
  void* mfma64Frags = nullptr;
  // SDG_KERNELS_AUTO: the matrix-core passes where the engine has them;
  // SDG_KERNELS_VALU: the VALU kernels only (sdg_engine_set_kernels)
  int kernels = SDG_KERNELS_AUTO;
};

// the layout of `layer` (its strides, 0 = the dense defaults) under the engine's settings
CellLayout cellLayout(const EngineState& e, const sdg_layer* layer);
// proxy kernels 0-4 on a validated layer (sdg_proxy_run); allocates nothing
int proxyRun(const EngineState& st, int kernelId, int steps, double dt, const sdg_layer* layer, hipStream_t s);
// the same for arrays handed over without a layer (derivatives, Taylor outputs)
CellLayout engineLayout(const EngineState& e);
// checks the per-cell strides of `layer` against the minimum sizes; sets the error text
bool layerStridesOk(const EngineState& e, const sdg_layer* layer, const char* who);

}  // namespace sdg

struct sdg_engine {
  sdg::EngineState s;
};

namespace sdg {
// kernels_valu.hip
int launchAder(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream);
int launchAderLocal(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream);
int launchLocal(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
int launchNeighbor(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
int launchTaylor(const EngineState& e, const double* coeffs, const void* derivatives, void* out, int64_t n,
                 hipStream_t stream);
// local time stepping kernels; `variant` selects a tile shape (0..3, 0 = default)
constexpr int kTileVariants = 4;
// neighbour pass: the kTileVariants tiles of kNeighbor, then kNeighborRows
// (one element per 16-lane row)
constexpr int kNeighborRowsVariant = kTileVariants;
constexpr int kNeighborVariants = kTileVariants + 1;
int launchLocalIntegration(const EngineState& e, const double* coeffs, const sdg_layer* layer, int resetBuffers,
                           hipStream_t stream, int variant = 0);
int launchNeighboringIntegration(const EngineState& e, const double* timeCoeffs, const double* subtimeCoeffs,
                                 const sdg_layer* layer, hipStream_t stream, int variant = 0);
// kernels_mfma.hip
int64_t mfmaFragmentCount(int order);
int launchLocalMfma(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                    hipStream_t stream);
int launchNeighborMfma(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
// kernels_mfma64.hip: the FP64 engines' fused ader + local pass on the matrix
// cores (time integral written when the layer has time_integrated, face
// traces when it has face_traces (FP64 [n][4][9][B2P]); with
// lts_setup the LTS buffers and derivatives, else derivatives of every cell
// when the layer has them)
bool hasMfma64(const EngineState& e);
int launchAderLocalMfma64(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream);
// the FP64 neighbour pass on the matrix cores from the FP64 face traces the
// local pass wrote (layer->face_traces)
int launchNeighborMfma64(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
int launchLocalIntegrationMfma64(const EngineState& e, const double* coeffs, const sdg_layer* layer,
                                 int resetBuffers, hipStream_t stream);
int uploadMfma64(EngineState& s, const double* globalMatrices);
// the LTS local pass on the matrix cores: buffers and derivatives per the
// layer's lts_setup (TimeCluster::computeLocalIntegration), no face traces
int launchLocalIntegrationMfma(const EngineState& e, const double* coeffs, const sdg_layer* layer, int resetBuffers,
                               hipStream_t stream);
}  // namespace sdg
