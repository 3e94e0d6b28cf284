// Tile geometry shared by the matrix-core kernels and the host code that
// prepares their pre-split, pre-swizzled matrix fragments.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.h"

namespace sdg {
namespace mfma {

constexpr int basis3(int o) { return o * (o + 1) * (o + 2) / 6; }
constexpr int basis2(int o) { return o * (o + 1) / 2; }
constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// 16-bit elements (bf16 or fp16) of one (m-tile, k-step) fragment set: 64
// lanes x 8 values for each of three split planes, one plane after the other.
constexpr int kFragPlane = 512;
constexpr int kFragPair = 3 * kFragPlane;
constexpr int kTileElems = 16;        // elements per workgroup tile
constexpr int kCols = 9 * kTileElems; // GEMM columns (quantity, element)

template <int O>
struct MShape {
  static constexpr int B = basis3(O);
  static constexpr int B2 = basis2(O);
  static constexpr int MT = cdiv(B, 16);            // row tiles of the basis index
  static constexpr int KPB = cdiv(B, 32) * 32;      // padded K over the basis index
  static constexpr int KSB = KPB / 32;
  static constexpr int B2P = (B2 + 3) / 4 * 4;      // face rows padded to a multiple of 4
  static constexpr int FR = 4 * B2P;                // rows of the four stacked (padded) face operands
  static constexpr int MTF = cdiv(FR, 16);
  static constexpr int KPF = cdiv(FR, 32) * 32;
  static constexpr int KSF = KPF / 32;
  static constexpr int KMAX = KPB > KPF ? KPB : KPF;
  static constexpr int LDX = KMAX + 8;              // bf16 stride of one X column (bank spread)
  static constexpr int THREADS = 64 * MT;
  // fragment offsets (bf16 elements)
  static constexpr int64_t ck = 0;                                   // [3][MT][KSB]
  static constexpr int64_t vol = ck + 3LL * MT * KSB * kFragPair;    // [3][MT][KSB]
  static constexpr int64_t fmrt = vol + 3LL * MT * KSB * kFragPair;  // [MTF][KSB]
  static constexpr int64_t rdm = fmrt + 1LL * MTF * KSB * kFragPair; // [MT][KSF]
  // rDivM_f one face at a time: k-step f holds face f's B2 columns, zero-padded to 32
  static constexpr int64_t rdmF = rdm + 1LL * MT * KSF * kFragPair;  // [MT][4]
  static constexpr int64_t total = rdmF + 4LL * MT * kFragPair;
  // float tables: invM2[B2], fP[3][B2][B2] as at most O non-zeros per row
  // (every fP_h row has at most O), unscale factors of the stationary
  // matrices [ck0, ck1, ck2, vol0, vol1, vol2, fmrt, rdm, rdmF] (1 unless
  // the fragments were scaled by a power of two)
  static constexpr int tInvM2 = 0;
  static constexpr int tFP = B2;                  // fP_h rows as O (value) entries: [3][B2][O]
  static constexpr int tFPCol = tFP + 3 * B2 * O; // their columns (exact small integers): [3][B2][O]
  static constexpr int tAUn = tFPCol + 3 * B2 * O;
  static constexpr int tTotal = tAUn + 9;
};

// local-time-stepping outputs of the local pass (TimeCluster.cpp:380-434):
// per cell its LtsSetup bits say whether it stores its buffer (accumulated
// over the sub-steps a slower neighbour sees unless `reset`) and its
// derivatives; setup == nullptr: buffer (if any) stored, derivatives (if any)
// stored for every cell
struct LtsOut {
  const uint16_t* setup = nullptr;
  float* deriv = nullptr;
  int reset = 1;
  int pad_ = 0;
};

// host entry points of one order's kernels
template <int O>
struct MfmaOrder {
  // local pass; lts == nullptr: the GTS pass (face traces, no derivatives)
  static int local(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                   hipStream_t stream, const LtsOut* lts);
  static int neighbor(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
};

}  // namespace mfma

namespace mfma64 {
using mfma::basis2;
using mfma::basis3;
using mfma::cdiv;
using mfma::kCols;
using mfma::kTileElems;

constexpr int cceil4(int a) { return (a + 3) / 4 * 4; }

// geometry of the FP64 local kernel per order
template <int O>
struct Shape64 {
  static constexpr int B = basis3(O);
  static constexpr int B2 = basis2(O);
  static constexpr int MT = cdiv(B, 16);        // row tiles of the basis index
  static constexpr int KB = cceil4(B);          // K over the basis index
  static constexpr int KSB = KB / 4;
  static constexpr int B2P = cceil4(B2);        // face rows padded to a multiple of 4
  static constexpr int FR = 4 * B2P;            // rows of the four stacked face operands
  static constexpr int MTF = cdiv(FR, 16);
  static constexpr int KSF = B2P / 4;           // k-steps of one face's flux product
  // one wave per basis row tile and per trace row tile (wave w owns both tiles w)
  static constexpr int WAVES = MT > MTF ? MT : MTF;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int LDX = KB + 2;            // doubles per X column (bank spread)
  static constexpr int LDW = B2P + 2;           // doubles per W column of one face
  static_assert(kCols * LDW <= kCols * LDX, "a face's W operand fits the X area");
  // fragment offsets (doubles): one fragment = 64 lanes x 1 double
  static constexpr int64_t ck = 0;                                  // [3][MT][KSB]
  static constexpr int64_t vol = ck + 3LL * MT * KSB * 64;          // [3][MT][KSB]
  static constexpr int64_t fmrt = vol + 3LL * MT * KSB * 64;        // [MTF][KSB]
  static constexpr int64_t rdm = fmrt + 1LL * MTF * KSB * 64;       // [MT][4][KSF], rDivM_f diag(M2)
  // the neighbour pass: rDivM_f itself, and fP_h rows as O (value, column) entries
  static constexpr int64_t rdmN = rdm + 4LL * MT * KSF * 64;        // [MT][4][KSF]
  static constexpr int64_t tFP = rdmN + 4LL * MT * KSF * 64;        // [3][B2][O] values
  static constexpr int64_t tFPCol = tFP + 3LL * B2 * O;             // [3][B2][O] columns (exact integers)
  static constexpr int64_t total = tFPCol + 3LL * B2 * O;
  static_assert((kCols * LDX + kTileElems * 324 + kTileElems * 72) * 8 <= 163840, "LDS of one workgroup per CU");
  // neighbour kernel: one wave per basis row tile; the face operand, the
  // gathered traces, AminusT and the fP tables in LDS
  static constexpr int NWAVES = MT;
  static constexpr int NTHREADS = 64 * NWAVES;
};

// LTS outputs of the FP64 matrix-core local pass (TimeCluster.cpp:380-434);
// setup == nullptr: the buffer and (if any) the derivatives of every cell
struct LtsOut64 {
  const uint16_t* setup = nullptr;
  double* deriv = nullptr;
  int reset = 1;
  int pad_ = 0;
};
// host entry of one order's FP64 matrix-core local pass
template <int O>
struct MfmaOrder64 {
  static int local(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream,
                   const LtsOut64* lts);
  static int neighbor(const EngineState& e, const sdg_layer* layer, hipStream_t stream);
};
// fragment count (doubles) of the FP64 matrix-core kernels of `order` (0: none)
int64_t fragmentCount(int order);
}  // namespace mfma64
}  // namespace sdg
