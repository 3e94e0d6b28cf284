// Plain FP64 element update on the matrix cores (v_mfma_f64_16x16x4_f64):
// the fused ader + local pass of the FP64 drop-in (arithmetic SDG_FP64), FP64
// throughout, one product per contraction.  Included by the per-order sources
// kernels_mfma64_o<O>.hip.
//
// Every contraction with a global matrix A (kDivMT, kDivM, fMrT, rDivM) is a
// GEMM  Y(M x 144) += A(M x K) X(K x 144)  over a tile of 16 elements, the 144
// columns being (quantity p, element e) -> 16 p + e, in k-steps of 4.  One
// wave owns one 16-row tile of the basis index for all 9 quantities of the
// 16 elements: lane l holds rows g, g + 4, g + 8, g + 12 (g = l / 16) of the
// tile for element e = l % 16 (the accumulator layout of v_mfma_f64_16x16x4),
// so the 9x9 star and AplusT products that mix quantities are register
// arithmetic; wave w also owns the w-th 16-row tile of the stacked face traces.
// X lives in LDS column-major ([column][k], k contiguous); A is
// pre-swizzled in global memory so one k-step's fragment is one coalesced
// 8-byte load per lane (L2-resident, shared by all workgroups).
//
// Phases of a tile (Time.cpp:65-140, Local.cpp:124-160; spec
// codegen/kernels/aderdg.py:321-537): CK derivatives and time integral I,
// face traces (fMrT_f / M2) I, volume sum_j kDivM_j I star_j, W_f = traces_f
// AplusT_f, local flux rDivM_f diag(M2) W_f face by face; Q += volume + flux.
// The time integral (and, for LTS, the buffer and the derivatives) are
// written for the neighbour pass.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "engine.h"
#include "mfma_common.h"

namespace sdg {
namespace mfma64 {

using mfma::basis3;
using mfma::cdiv;
using mfma::kCols;
using mfma::kTileElems;

typedef double doublex4 __attribute__((ext_vector_type(4)));

__device__ constexpr int kStarQ[9][3] = {{6, 7, 8}, {6, 7, 8}, {6, 7, 8}, {6, 7, -1}, {7, 8, -1},
                                         {6, 8, -1}, {0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
__device__ constexpr int kStarOff[9] = {0, 3, 6, 9, 11, 13, 15, 18, 21};
__device__ constexpr int kStarCnt[9] = {3, 3, 3, 2, 2, 2, 3, 3, 3};
__device__ constexpr int kStarPos[24] = {6 + 0,  7 + 0,  8 + 0,  6 + 9,  7 + 9,  8 + 9,  6 + 18, 7 + 18,
                                         8 + 18, 6 + 27, 7 + 27, 7 + 36, 8 + 36, 6 + 45, 8 + 45, 0 + 54,
                                         3 + 54, 5 + 54, 1 + 63, 3 + 63, 4 + 63, 2 + 72, 4 + 72, 5 + 72};

template <int O>
struct Coeffs64 {
  double c[O];
};

__device__ inline doublex4 mma(double a, double b, doublex4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// the lane's rows base + 4 r (r < 4, those below `rows`) of one column into X
template <int LD>
__device__ inline void storeX(double* xs, int col, int base, int rows, const double (&v)[4]) {
  double* d = xs + col * LD + base;
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (base + 4 * r < rows) d[4 * r] = v[r];
}

// q[p * ld + row0 + 4 j] += acc[p][j] for the lane's rows below B: all 36 old
// values are loaded first, unconditionally from in-bounds (clamped) rows, then
// stored where real (a read-modify-write under a branch waits on every load)
template <int B>
__device__ inline void addRows(double* q, int ld, int row0, bool live, const doublex4 (&acc)[9]) {
  double old[9][4];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) old[p][j] = q[p * ld + min(row0 + 4 * j, B - 1)];
  if (live)
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (row0 + 4 * j < B) q[p * ld + row0 + 4 * j] = old[p][j] + acc[p][j];
}

// acc[q] += A_tile X(:, quantity q) over `ksteps` k-steps, q < 9
template <int LD>
__device__ inline void gemm9(doublex4 (&acc)[9], const double* __restrict__ a, int ksteps, const double* xs,
                             int lane) {
  const int g = lane >> 4, e = lane & 15;
  double an = a[lane];
  for (int kk = 0; kk < ksteps; ++kk) {
    const double ac = an;
    if (kk + 1 < ksteps) an = a[(kk + 1) * 64 + lane];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = mma(ac, xs[(q * 16 + e) * LD + 4 * kk + g], acc[q]);
  }
}

// One pass of y += sum_j (A_j X) star_j for the lane's element: the input
// quantities of group PASS ({6,7,8} -> stresses 0-5, {0,1,2} and {3,4,5} ->
// velocities 6-8; the elastic star matrix couples velocities and stresses
// only), one matrix at a time with 3 accumulator tiles
template <int LD, int PASS>
__device__ inline void gemmPass(const double* const (&a)[3], int ks, const double* xs, int lane, const double* sS,
                                double (&y)[9][4]) {
  const int g = lane >> 4, e = lane & 15;
  constexpr int q0 = PASS == 0 ? 6 : (PASS == 1 ? 0 : 3);
#pragma unroll
  for (int jj = 0; jj < 3; ++jj) {
    doublex4 acc[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[i] = doublex4{0, 0, 0, 0};
    const double* aj = a[jj];
    double an = aj[lane];
    for (int kk = 0; kk < ks; ++kk) {
      const double ac = an;
      if (kk + 1 < ks) an = aj[(kk + 1) * 64 + lane];
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[i] = mma(ac, xs[((q0 + i) * 16 + e) * LD + 4 * kk + g], acc[i]);
    }
    // the star coefficients are read where they are used (else the compiler
    // keeps all 72 live across the recursion)
    const double* sSp = sS;
    asm volatile("" : "+v"(sSp));
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (k < kStarCnt[p]) {
          const int qi = kStarQ[p][k];
          if (qi >= q0 && qi < q0 + 3) {
            const double sv = sSp[jj * 24 + kStarOff[p] + k];
#pragma unroll
            for (int j = 0; j < 4; ++j) y[p][j] += acc[qi - q0][j] * sv;
          }
        }
    // finish this contraction before the next product starts: otherwise the
    // compiler defers the contractions past the k-loops and keeps all their
    // accumulator tiles live at once
#pragma unroll
    for (int p = (PASS == 0 ? 0 : 6); p < (PASS == 0 ? 6 : 9); ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) asm volatile("" : "+v"(y[p][j]));
  }
}

template <int LD>
__device__ inline void gemmPassAt(int pass, const double* const (&a)[3], int ks, const double* xs, int lane,
                                  const double* sS, double (&y)[9][4]) {
  if (pass == 0) gemmPass<LD, 0>(a, ks, xs, lane, sS, y);
  else if (pass == 1) gemmPass<LD, 1>(a, ks, xs, lane, sS, y);
  else gemmPass<LD, 2>(a, ks, xs, lane, sS, y);
}

// the wave of the pass-1/2 unit k (tile k / 2, pass 1 + k % 2) of a star GEMM
// over mt m-tiles: after the owners' pass-0 units, round-robin from wave mt on
constexpr int unitWave(int mt, int k, int waves) { return (mt + k) % waves; }
// staging slots a star GEMM over mt m-tiles needs (units not run by their owner)
constexpr int stageSlots(int mt, int waves) {
  int n = 0;
  for (int k = 0; k < 2 * mt; ++k) n += unitWave(mt, k, waves) != k / 2 ? 1 : 0;
  return n;
}
template <int O, int WAVES>
constexpr int maxStageSlots() {
  int n = stageSlots(cdiv(basis3(O), 16), WAVES);  // the volume
  for (int d = 1; d < O; ++d) n = n > stageSlots(cdiv(basis3(O - d), 16), WAVES) ? n : stageSlots(cdiv(basis3(O - d), 16), WAVES);
  return n;
}
constexpr int kSlotDoubles = 3 * 4 * 64;  // a lane's 4 rows of the 3 velocities, 64 lanes

// y += sum_j (A_j X) star_j for m-tile w (w < mt) with the 3 mt (tile, pass)
// units spread over the waves: wave t runs pass 0 of tile t, the pass-1/2
// units go round-robin from wave mt on; a unit run by a wave other than its
// tile's owner hands its velocity partial sums over `stage`.  Contains the
// barrier after which every wave has finished reading X; on return the
// owner's y holds the whole sum.  a(j, t): the fragments of matrix j, tile t.
template <int LD, int WAVES, bool BALANCE, typename Frag>
__device__ inline void starsBalanced(const Frag& fragOf, int mt, int ks, const double* xs, double* stage, int w,
                                     int lane, const double* sS, double (&y)[9][4]) {
  if constexpr (!BALANCE) {
    // every pass of tile w by its owner (where the staging area does not fit)
    if (w < mt) {
      const double* a[3] = {fragOf(0, w), fragOf(1, w), fragOf(2, w)};
      gemmPass<LD, 0>(a, ks, xs, lane, sS, y);
      gemmPass<LD, 1>(a, ks, xs, lane, sS, y);
      gemmPass<LD, 2>(a, ks, xs, lane, sS, y);
    }
    __syncthreads();
    return;
  }
  if (w < mt) {
    const double* a[3] = {fragOf(0, w), fragOf(1, w), fragOf(2, w)};
    gemmPass<LD, 0>(a, ks, xs, lane, sS, y);
  }
  int slot = 0;
  for (int k = 0; k < 2 * mt; ++k) {
    const int t = k / 2, pass = 1 + k % 2, unitW = unitWave(mt, k, WAVES);
    if (unitW == w) {
      const double* a[3] = {fragOf(0, t), fragOf(1, t), fragOf(2, t)};
      if (unitW == t) {
        gemmPassAt<LD>(pass, a, ks, xs, lane, sS, y);
      } else {
        double tmp[9][4];
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j) tmp[p][j] = 0.0;
        gemmPassAt<LD>(pass, a, ks, xs, lane, sS, tmp);
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int j = 0; j < 4; ++j) stage[((slot * 3 + q) * 4 + j) * 64 + lane] = tmp[6 + q][j];
      }
    }
    if (unitW != t) ++slot;
  }
  __syncthreads();
  slot = 0;
  for (int k = 0; k < 2 * mt; ++k) {
    const int t = k / 2, unitW = unitWave(mt, k, WAVES);
    if (unitW != t) {
      if (w == t)
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int j = 0; j < 4; ++j) y[6 + q][j] += stage[((slot * 3 + q) * 4 + j) * 64 + lane];
      ++slot;
    }
  }
}

template <int O, bool LTS, bool DEF>
__global__ __launch_bounds__(Shape64<O>::THREADS, 1) void kLocalMfma64(
    int64_t n, const double* __restrict__ frag, Coeffs64<O> cf, double* __restrict__ dofs,
    double* __restrict__ integ, double* __restrict__ traces, const double* __restrict__ star,
    const double* __restrict__ aplus, const uint8_t* __restrict__ faceTypes, CellLayout lay, LtsOut64 lts) {
  using S = Shape64<O>;
  constexpr int B = S::B, MT = S::MT, LDX = S::LDX, LDW = S::LDW, B2P = S::B2P;
  // LDS: X, AplusT of the four faces ([element][face][81]), the compact star
  // matrices ([element][3][24]) and the helper waves' velocity partial sums;
  // where all four do not fit (ORDER 7) the partial sums share AplusT's area
  // and AplusT is loaded after the volume product
  constexpr int kA = kTileElems * 324, kS = kTileElems * 72;
  constexpr int kBalancedStage = maxStageSlots<O, S::WAVES>() * kSlotDoubles;
  // the helper waves' partial sums need their own area or AplusT's; without
  // room for either (ORDER 7) each owner runs all passes of its tile
  constexpr bool kBalance = (kCols * LDX + kS + (kA > kBalancedStage ? kA : kBalancedStage)) * 8 <= 163840;
  constexpr int kStage = kBalance ? kBalancedStage : 0;
  constexpr bool kShareA = (kCols * LDX + kA + kS + kStage) * 8 > 163840;
  static_assert((kCols * LDX + kS + (kShareA ? (kA > kStage ? kA : kStage) : kA + kStage)) * 8 <= 163840,
                "LDS of one workgroup per CU");
  __shared__ __attribute__((aligned(16))) double xs[kCols * LDX];
  __shared__ double sA[kShareA ? (kA > kStage ? kA : kStage) : kA];
  __shared__ double sStar[kS];
  __shared__ double sStageOwn[kShareA || kStage == 0 ? 1 : kStage];
  double* const sStage = kShareA ? sA : sStageOwn;

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const int64_t elem = E0 + e;
  const bool valid = elem < n;
  const int ldq = DEF ? B : lay.ldq;
  const int64_t qStride = DEF ? 9 * B : lay.qStride;
  const int row0 = 16 * w + g;  // the lane's rows row0 + 4 r
  const bool ownsRows = w < MT;  // waves beyond MT own only a trace tile
  double* const laneDofs = dofs + (valid ? elem : 0) * qStride + (ownsRows ? row0 : 0);

  // the tile's compact star matrices and AplusT, staged through registers: every
  // load of a thread is issued before the first store, from a clamped address
  // (in bounds: a launched tile has E0 < n), since a loop that stores each
  // value as it arrives waits on every load in turn
  {
    constexpr int SPT = cdiv(kTileElems * 72, S::THREADS);
    const int64_t starStride = DEF ? (lay.starCompact ? 72 : 243) : lay.starStride;
    double sv[SPT];
    if (lay.starCompact) {  // wave-uniform: the compact format needs no position table
#pragma unroll
      for (int u = 0; u < SPT; ++u) {
        const int i = min(static_cast<int>(threadIdx.x) + u * S::THREADS, kTileElems * 72 - 1);
        const int64_t el = E0 + i / 72 < n ? E0 + i / 72 : n - 1;
        sv[u] = star[el * starStride + i % 72];
      }
    } else {
#pragma unroll
      for (int u = 0; u < SPT; ++u) {
        const int i = min(static_cast<int>(threadIdx.x) + u * S::THREADS, kTileElems * 72 - 1);
        const int64_t el = E0 + i / 72 < n ? E0 + i / 72 : n - 1;
        const int r = i % 72;
        sv[u] = star[el * starStride + 81 * (r / 24) + kStarPos[r % 24]];
      }
    }
#pragma unroll
    for (int u = 0; u < SPT; ++u) {
      const int i = threadIdx.x + u * S::THREADS;
      if (i < kTileElems * 72) sStar[i] = E0 + i / 72 < n ? sv[u] : 0.0;
    }
  }
  auto loadAplus = [&]() {
    constexpr int APT = cdiv(kTileElems * 324, S::THREADS);
    double av[APT];
#pragma unroll
    for (int u = 0; u < APT; ++u) {
      const int i = min(static_cast<int>(threadIdx.x) + u * S::THREADS, kTileElems * 324 - 1);
      const int64_t el = E0 + i / 324 < n ? E0 + i / 324 : n - 1;
      av[u] = aplus[DEF ? el * 324 + i % 324 : el * lay.aplusStride + i % 324];
    }
#pragma unroll
    for (int u = 0; u < APT; ++u) {
      const int i = threadIdx.x + u * S::THREADS;
      if (i < kTileElems * 324) sA[i] = E0 + i / 324 < n ? av[u] : 0.0;
    }
  };
  if constexpr (!kShareA) loadAplus();

  // Q: the lane's rows (zero beyond B and for padding elements); read again
  // for the final update, so only the time integral stays live
  double q[9][4], iacc[9][4];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      q[p][j] = (valid && ownsRows && row0 + 4 * j < B) ? laneDofs[p * ldq + 4 * j] : 0.0;
      iacc[p][j] = cf.c[0] * q[p][j];
    }
  [[maybe_unused]] unsigned setup = 0;
  [[maybe_unused]] bool storeDeriv = false;
  [[maybe_unused]] double* cellDeriv = nullptr;
  if constexpr (LTS) {
    setup = (lts.setup != nullptr && valid) ? lts.setup[elem] : 0u;
    storeDeriv = lts.deriv != nullptr && valid && (lts.setup == nullptr || (setup & SDG_LTS_HAS_DERIVATIVES));
    cellDeriv = lts.deriv + (valid ? elem : 0) * lay.dsize;
    if (storeDeriv && ownsRows)
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (row0 + 4 * j < B) cellDeriv[p * ldq + row0 + 4 * j] = q[p][j];
  }
  [[maybe_unused]] int64_t doff = 9LL * ldq;
  // X holds rows [0, KB) (rows beyond are never read)
  if (ownsRows)
#pragma unroll
    for (int p = 0; p < 9; ++p) storeX<LDX>(xs, p * 16 + e, row0, S::KB, q[p]);
  __syncthreads();
  const double* sS = sStar + e * 72;

  // Cauchy-Kowalevski recursion: wave w owns m-tile w of dQ_d
#pragma unroll 1
  for (int d = 1; d < O; ++d) {
    const int m = basis3(O - d), nIn = basis3(O - d + 1);
    const int mtiles = cdiv(m, 16), ks = cdiv(nIn, 4);
    double dq[9][4];
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) dq[p][j] = 0.0;
    starsBalanced<LDX, S::WAVES, kBalance>(
        [&](int j, int t) { return frag + S::ck + (static_cast<int64_t>(j) * MT + t) * S::KSB * 64; }, mtiles, ks, xs,
        sStage, w, lane, sS, dq);  // ends with the barrier after which dQ_{d-1} is no longer read
    if (w < mtiles) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (row0 + 4 * j >= m)
#pragma unroll
          for (int p = 0; p < 9; ++p) dq[p][j] = 0.0;
    }
    if constexpr (LTS) {
      const int ldm = lay.rows(m);
      if (storeDeriv && w < mtiles)
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (row0 + 4 * j < m) cellDeriv[doff + p * ldm + row0 + 4 * j] = dq[p][j];
      doff += 9LL * ldm;
    }
    if (w < mtiles) {
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        storeX<LDX>(xs, p * 16 + e, row0, S::KB, dq[p]);
#pragma unroll
        for (int j = 0; j < 4; ++j) iacc[p][j] += cf.c[d] * dq[p][j];
      }
    }
    __syncthreads();
  }

  // time integral: output (LTS: the buffer, stored or accumulated) and X
  bool storeBuffer = integ != nullptr && valid;
  [[maybe_unused]] bool accumulate = false;
  if constexpr (LTS) {
    storeBuffer = storeBuffer && (lts.setup == nullptr || (setup & SDG_LTS_HAS_BUFFERS));
    accumulate = lts.setup != nullptr && !lts.reset && (setup & SDG_LTS_ACCUMULATE_BUFFERS);
  }
  if (storeBuffer && ownsRows) {
    double* const elemInteg = integ + elem * qStride;
    // accumulated buffers: all old values loaded first (clamped rows, in
    // bounds), then stored, so that the loads are in flight together
    double old[9][4] = {};
    if constexpr (LTS) {
      if (accumulate)
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j) old[p][j] = elemInteg[p * ldq + min(row0 + 4 * j, B - 1)];
    }
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (row0 + 4 * j < B) elemInteg[p * ldq + row0 + 4 * j] = old[p][j] + iacc[p][j];
  }
  if (ownsRows)
#pragma unroll
    for (int p = 0; p < 9; ++p) storeX<LDX>(xs, p * 16 + e, row0, S::KB, iacc[p]);
  __syncthreads();

  // face traces (fMrT_f / M2) I: trace tile w, rows f * B2P + m
  // the lane's trace rows r0 + 4 r: face r / B2P of the stacked operand
  const int r0 = 16 * w + g;
  const bool ownsTrace = w < S::MTF;
  double tr[9][4];
  if (w < S::MTF) {
    doublex4 acc[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[p] = doublex4{0, 0, 0, 0};
    gemm9<LDX>(acc, frag + S::fmrt + static_cast<int64_t>(w) * S::KSB * 64, S::KSB, xs, lane);
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) tr[p][j] = acc[p][j];
    // the traces for the neighbour pass: [element][face][quantity][B2P]
    if (traces != nullptr && valid)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = r0 + 4 * j;
        if (r < S::FR) {
          double* t = traces + ((elem * 4 + r / B2P) * 9) * B2P + r % B2P;
#pragma unroll
          for (int p = 0; p < 9; ++p) t[p * B2P] = tr[p][j];
        }
      }
  }
  // volume: sum_j kDivM_j I star_j (the initial value of the flux accumulators)
  doublex4 acc[9];
  {
    double vol[9][4];
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) vol[p][j] = 0.0;
    starsBalanced<LDX, S::WAVES, kBalance>(
        [&](int j, int t) { return frag + S::vol + (static_cast<int64_t>(j) * MT + t) * S::KSB * 64; }, MT, S::KSB,
        xs, sStage, w, lane, sS, vol);
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[p] = doublex4{vol[p][0], vol[p][1], vol[p][2], vol[p][3]};
  }
  if constexpr (kShareA) {
    __syncthreads();  // the owners have read the staged partial sums
    loadAplus();
    __syncthreads();
  }
  // W = traces AplusT_f, each row with its own face's matrix; DR faces give none
  if (ownsTrace) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int f = min((r0 + 4 * j) / B2P, 3);  // rows beyond FR: zero fragment rows, any face
      const bool dr = valid && faceTypes != nullptr && faceTypes[elem * 4 + f] == SDG_FACE_DYNAMIC_RUPTURE;
      double wt[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        // one column of AplusT_f at a time (else all 81 entries are loaded at once)
        const double* A = sA + e * 324 + f * 81 + 9 * p;
        asm volatile("" : "+v"(A));
        double s = 0.0;
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) s += tr[qq][j] * A[qq];
        wt[p] = dr ? 0.0 : s;
        asm volatile("" : "+v"(wt[p]));
      }
#pragma unroll
      for (int p = 0; p < 9; ++p) tr[p][j] = wt[p];
    }
  }

  // local flux face by face: acc += rDivM_f diag(M2) W_f (X area reused for W_f)
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    __syncthreads();  // the X area is free (f = 0: I no longer read; else: W_{f-1} read)
    if (ownsTrace)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = r0 + 4 * j;
        if (r >= f * B2P && r < (f + 1) * B2P)
#pragma unroll
          for (int p = 0; p < 9; ++p) xs[(p * 16 + e) * LDW + r - f * B2P] = tr[p][j];
      }
    __syncthreads();
    if (ownsRows)
      gemm9<LDW>(acc, frag + S::rdm + (static_cast<int64_t>(w) * 4 + f) * S::KSF * 64, S::KSF, xs, lane);
  }
  addRows<B>(dofs + (valid ? elem : 0) * qStride, ldq, row0, valid && ownsRows, acc);
}

// Neighbour flux from the neighbours' face traces (Neighbor.cpp:66-109; spec
// aderdg.py:414-431): face by face, the 16 neighbours' traces rT_j I of the
// shared side are gathered, W_f = (fP_h Tr) AminusT_f is formed by 16-lane
// groups (one element, two rows per lane), and Q += rDivM_f W_f on the
// matrix cores; the next face's traces and AminusT are loaded into registers
// while the current face is computed.
template <int O, bool DEF>
__global__ __launch_bounds__(Shape64<O>::NTHREADS, 2) void kNeighborMfma64(
    int64_t n, const double* __restrict__ frag, double* __restrict__ dofs, const double* __restrict__ traces,
    const int32_t* __restrict__ neighbors, const uint8_t* __restrict__ faceRel, const double* __restrict__ aminus,
    const uint8_t* __restrict__ faceTypes, CellLayout lay) {
  using S = Shape64<O>;
  constexpr int B = S::B, B2 = S::B2, B2P = S::B2P, LDW = S::LDW, THREADS = S::NTHREADS;
  constexpr int HALF = B2P / 2;                 // row pairs of one (padded) face
  static_assert(HALF <= 16, "a face's row pairs fit a 16-lane group");
  constexpr int GROUPS = THREADS / 16;          // 16-lane groups per pass, one element each
  constexpr int PASSES = cdiv(kTileElems, GROUPS);
  constexpr int TR = 9 * B2P;                   // doubles of one face trace
  constexpr int TR2 = TR / 2;                   // 16-byte pieces of one face trace
  constexpr int GPT = cdiv(kTileElems * TR2, THREADS);
  constexpr int APT = cdiv(kTileElems * 81, THREADS);
  __shared__ __attribute__((aligned(16))) double xs[kCols * LDW];  // face operand, rows [0, B2P)
  __shared__ __attribute__((aligned(16))) double sTr[kTileElems * TR];
  __shared__ double sAm[kTileElems * 81];
  __shared__ double sFP[3 * B2 * O];
  __shared__ int sFPCol[3 * B2 * O];
  __shared__ int sLink[kTileElems * 4];  // trace index nb * 4 + side of (element, face), -1: none
  __shared__ int sRel[kTileElems * 4];   // orientation h of (element, face)

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const bool valid = E0 + e < n;
  const int ldq = DEF ? B : lay.ldq;
  const int64_t qStride = DEF ? 9 * B : lay.qStride;
  const int row0 = 16 * w + g;  // the lane's rows row0 + 4 r

  // the fP tables and the neighbour links, every load issued unconditionally
  // from an in-bounds (clamped) address before the first store
  {
    constexpr int FPN = 3 * B2 * O, FPI = cdiv(FPN, THREADS);
    double fv[FPI], fc[FPI];
#pragma unroll
    for (int u = 0; u < FPI; ++u) {
      const int i = min(static_cast<int>(threadIdx.x) + u * THREADS, FPN - 1);
      fv[u] = frag[S::tFP + i];
      fc[u] = frag[S::tFPCol + i];
    }
    static_assert(kTileElems * 4 <= THREADS, "one thread per (element, face) link");
    const int s = threadIdx.x >> 2 & (kTileElems - 1), f = threadIdx.x & 3;
    const int64_t el = E0 + s < n ? E0 + s : n - 1;
    const int type = faceTypes != nullptr ? faceTypes[el * 4 + f] : SDG_FACE_REGULAR;
    const int nb = neighbors[el * 4 + f];
    const int side = faceRel[(el * 4 + f) * 2 + 0], rel = faceRel[(el * 4 + f) * 2 + 1];
#pragma unroll
    for (int u = 0; u < FPI; ++u) {
      const int i = threadIdx.x + u * THREADS;
      if (i < FPN) {
        sFP[i] = fv[u];
        sFPCol[i] = static_cast<int>(fc[u]);
      }
    }
    if (threadIdx.x < kTileElems * 4) {
      const bool use = E0 + s < n && type == SDG_FACE_REGULAR;
      sLink[threadIdx.x] = use ? nb * 4 + min(side, 3) : -1;
      sRel[threadIdx.x] = use ? min(rel, 2) : 0;
    }
  }
  __syncthreads();
  double2 gv[GPT];
  double av[APT];
  auto gather = [&](int f) {
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      double2 v{0.0, 0.0};
      if (idx < kTileElems * TR2) {
        const int link = sLink[(idx / TR2) * 4 + f];
        if (link >= 0) v = *reinterpret_cast<const double2*>(traces + static_cast<int64_t>(link) * TR + 2 * (idx % TR2));
      }
      gv[i] = v;
    }
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      const int64_t at = (E0 + idx / 81) * (DEF ? 324 : lay.aminusStride) + f * 81 + idx % 81;
      av[i] = (idx < kTileElems * 81 && E0 + idx / 81 < n) ? aminus[at] : 0.0;
    }
  };
  gather(0);
  doublex4 acc[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) acc[p] = doublex4{0, 0, 0, 0};
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    // the previous face's items have read sTr / sAm and its product xs
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      if (idx < kTileElems * TR2) *reinterpret_cast<double2*>(sTr + 2 * idx) = gv[i];
    }
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = threadIdx.x + i * THREADS;
      if (idx < kTileElems * 81) sAm[idx] = av[i];
    }
    if (f < 3) gather(f + 1);
    __syncthreads();
#pragma unroll
    for (int ps = 0; ps < PASSES; ++ps) {
      // rows 2 rp and 2 rp + 1 of face f of element s (rows >= B2 are padding: zero)
      const int s = static_cast<int>(threadIdx.x >> 4) + ps * GROUPS;
      const int rp = static_cast<int>(threadIdx.x & 15);
      const int r0 = 2 * rp;
      const bool act = s < kTileElems && rp < HALF;
      const int sl = s < kTileElems ? s : 0;
      const bool regular = act && sLink[sl * 4 + f] >= 0;
      const int h = sRel[sl * 4 + f];
      const double* tr = sTr + sl * TR;
      double u[2][9];
#pragma unroll
      for (int ri = 0; ri < 2; ++ri) {
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) u[ri][qq] = 0.0;
        const int row = r0 + ri;
        if (regular && row < B2) {
          const double* fp = sFP + (h * B2 + row) * O;
          const int* fc = sFPCol + (h * B2 + row) * O;
#pragma unroll
          for (int kk = 0; kk < O; ++kk) {
            const double c = fp[kk];
            const int mm = fc[kk];
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) u[ri][qq] += c * tr[qq * B2P + mm];
          }
        }
      }
      const double* A = sAm + sl * 81;
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        double w0 = 0.0, w1 = 0.0;
        if (regular)
#pragma unroll
          for (int qq = 0; qq < 9; ++qq) {
            const double a = A[qq + 9 * p];
            w0 += u[0][qq] * a;
            w1 += u[1][qq] * a;
          }
        if (act) *reinterpret_cast<double2*>(xs + (p * 16 + s) * LDW + r0) = double2{w0, w1};
      }
    }
    __syncthreads();
    gemm9<LDW>(acc, frag + S::rdmN + (static_cast<int64_t>(w) * 4 + f) * S::KSF * 64, S::KSF, xs, lane);
  }
  addRows<B>(dofs + (valid ? E0 + e : 0) * qStride, ldq, row0, valid, acc);
}

template <int B>
bool defaultLayout64(const CellLayout& lay) {
  return lay.ldq == B && lay.qStride == 9 * B && lay.aplusStride == 324 && lay.starStride == (lay.starCompact ? 72 : 243);
}

template <int O>
int MfmaOrder64<O>::local(const EngineState& st, const double* coeffs, const sdg_layer* L, hipStream_t stream,
                          const LtsOut64* lts) {
  using S = Shape64<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  Coeffs64<O> cf{};
  for (int i = 0; i < O; ++i) cf.c[i] = coeffs[i];
  const CellLayout lay = cellLayout(st, L);
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  const bool def = defaultLayout64<S::B>(lay);
  auto launch = [&](auto ltsFlag, auto defFlag) {
    kLocalMfma64<O, decltype(ltsFlag)::value, decltype(defFlag)::value><<<grid, S::THREADS, 0, stream>>>(
        n, static_cast<const double*>(st.mfma64Frags), cf, static_cast<double*>(L->dofs),
        static_cast<double*>(L->time_integrated), static_cast<double*>(L->face_traces),
        static_cast<const double*>(L->star), static_cast<const double*>(L->aplus), L->face_types, lay,
        lts != nullptr ? *lts : LtsOut64{});
  };
  if (lts != nullptr) {
    if (def) launch(std::true_type{}, std::true_type{});
    else launch(std::true_type{}, std::false_type{});
  } else {
    if (def) launch(std::false_type{}, std::true_type{});
    else launch(std::false_type{}, std::false_type{});
  }
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <int O>
int MfmaOrder64<O>::neighbor(const EngineState& st, const sdg_layer* L, hipStream_t stream) {
  using S = Shape64<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const CellLayout lay = cellLayout(st, L);
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  auto launch = [&](auto defFlag) {
    kNeighborMfma64<O, decltype(defFlag)::value><<<grid, S::NTHREADS, 0, stream>>>(
        n, static_cast<const double*>(st.mfma64Frags), static_cast<double*>(L->dofs),
        static_cast<const double*>(L->face_traces), L->face_neighbors, L->face_relations,
        static_cast<const double*>(L->aminus), L->face_types, lay);
  };
  if (defaultLayout64<S::B>(lay) && lay.aminusStride == 324) launch(std::true_type{});
  else launch(std::false_type{});
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

}  // namespace mfma64
}  // namespace sdg
