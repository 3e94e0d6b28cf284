// Element-update kernels on the vector ALUs (FP64 drop-in and FP32).
//
// Work decomposition: a wave owns EPW elements; each element is handled by
// TPE consecutive lanes and lane t of an element owns the basis rows
// t, t + TPE, ..., so every small GEMM Y(M x 9) += A(M x K) X(K x 9) is a
// register-blocked R x 9 outer-product accumulation per lane.  The global
// matrix A is read transposed (consecutive rows = consecutive addresses, all
// elements of a wave read the same lines); the per-element operand X lives in
// LDS in [row][quantity] order so one row of 9 quantities is a broadcast read.
//
// The CK recursion and the volume integral are evaluated as
//   sum_j (A_j X) star_j
// i.e. three accumulators sets followed by the sparse 9x9 star product
// (24 non-zeros, pattern of codegen/matrices/star.xml), which keeps the star
// matrices out of the inner loop.  Face terms use the factored forms of the
// reference CPU kernels (codegen/kernels/aderdg.py:354-431):
//   local:     rDivM_f ((fMrT_f I) AplusT_f)
//   neighbour: rDivM_f ((fP_h (rT_j I_nb)) AminusT_f)
// with the four faces' rDivM products fused into one K = 4*B2 contraction.
#include <hip/hip_runtime.h>

#include "dpp_rows.h"
#include "engine.h"

namespace sdg {
namespace {

constexpr int basis3(int o) { return o * (o + 1) * (o + 2) / 6; }
constexpr int basis2(int o) { return o * (o + 1) / 2; }

// Tile shape: R basis rows per lane (TPE lanes per element, EPW elements per
// wave) and WAVES waves per block.  RX = 1 takes one more row per lane than
// the default; WV is the number of waves.  Local time stepping picks the
// shape per cluster and kernel by timing (sdg_lts_autotune).
template <int O, int RX = 0, int WV = 4>
struct Shape {
  static constexpr int B = basis3(O);
  static constexpr int B2 = basis2(O);
  static constexpr int R = (B <= 10 ? 1 : (B <= 56 ? 2 : 3)) + RX;
  static constexpr int TPE = (B + R - 1) / R;
  static constexpr int EPW = 64 / TPE;
  static constexpr int WAVES = WV;
  static constexpr int EPB = EPW * WAVES;
  static constexpr int THREADS = 64 * WAVES;
  static constexpr int QS = 9 * B;       // reals per DOF block
  static constexpr int FS = 4 * B2 * 9;  // reals of the four face operands
  // device matrix offsets (must match deviceMatrixLayout)
  static constexpr int64_t ck = 0;
  static constexpr int64_t vol = ck + 3LL * B * B;
  static constexpr int64_t fmrt = vol + 3LL * B * B;
  static constexpr int64_t rt = fmrt + 4LL * B * B2;
  static constexpr int64_t fp = rt + 4LL * B * B2;
  static constexpr int64_t rdm = fp + 3LL * B2 * B2;
  static constexpr int64_t total = rdm + 4LL * B2 * B;
};

// Sparsity of the elastic star matrix: for output quantity p the non-zero
// input quantities q (codegen/matrices/star.xml, 0-based).
__device__ constexpr int kStarQ[9][3] = {{6, 7, 8}, {6, 7, 8}, {6, 7, 8}, {6, 7, -1}, {7, 8, -1},
                                         {6, 8, -1}, {0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
constexpr int kStarNnz = 24;  // per direction

template <int O, typename real>
struct Coeffs {
  real c[O];
};

// local time stepping of the predictor (ader + local integral)
struct LtsArgs {
  const uint16_t* setup;  // [n] LTS bits (SDG_LTS_*) or nullptr: global time stepping
  int reset;              // accumulating cells store their buffer instead of adding to it
};

// local time stepping of the neighbour integral
template <int O, typename real>
struct NeighborLts {
  const uint16_t* setup;  // [n] LTS bits or nullptr: every neighbour provides a buffer
  const real* deriv;      // derivatives of all cells (the layer's layout), indexed like the buffers
  Coeffs<O, real> gts, sub;  // time-basis coefficients: GTS relation, sub-interval of a larger step
};

// Gather the 24 non-zeros of the three star matrices of the elements of the
// block into LDS (compact order: output quantity p major, then the input
// quantities of its row of the pattern, the reference's sparse order), from
// dense [3][81] or compact [3][24] storage, cells starStride reals apart.
template <typename real>
__device__ inline void loadStars(real* sStar, const real* __restrict__ star, int64_t e0, int64_t n, int epb,
                                 const CellLayout& lay) {
  for (int idx = threadIdx.x; idx < epb * 3 * kStarNnz; idx += blockDim.x) {
    const int s = idx / (3 * kStarNnz);
    const int w = idx % (3 * kStarNnz);
    const int j = w / kStarNnz;
    int z = w % kStarNnz;
    const int64_t e = e0 + s;
    if (lay.starCompact) {
      sStar[idx] = (e < n) ? star[e * lay.starStride + w] : real(0);
      continue;
    }
    int p = 0, k = 0;
    // position z -> (p, k)
    for (p = 0; p < 9; ++p) {
      const int cnt = (p >= 3 && p <= 5) ? 2 : 3;
      if (z < cnt) { k = z; break; }
      z -= cnt;
    }
    const int q = kStarQ[p][k];
    sStar[idx] = (e < n) ? star[e * lay.starStride + 81 * j + q + 9 * p] : real(0);
  }
}

// y[r][p] = sum_j sum_q acc[j][r][q] * star_j(q, p) with the compact stars of one element
template <int R, typename real>
__device__ inline void applyStars(const real (&acc)[3][R][9], const real* sS, real (&y)[R][9]) {
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int p = 0; p < 9; ++p) y[r][p] = real(0);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int z = 0;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      const int cnt = (p >= 3 && p <= 5) ? 2 : 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (k < cnt) {
          const real sv = sS[j * kStarNnz + z + k];
          const int q = kStarQ[p][k];
#pragma unroll
          for (int r = 0; r < R; ++r) y[r][p] += acc[j][r][q] * sv;
        }
      }
      z += cnt;
    }
  }
}

// acc[j][r][p] += sum_{l<K} AT_j[l][row_r] * X[l][p], j < J, rows < M.
// AheadX: both operands of a step are loaded ahead of it, the matrix entries
// (L1/L2) and the step's row of X (LDS); else X is read at its step, and in
// modes 0 and 1 the matrix entries too.  Mode 2: two steps per iteration with
// two register sets, the operands of steps l + 2 and l + 3 loaded while steps
// l and l + 1 are computed; else one step at a time, unrolled as the compiler
// chooses (mode 0) or not unrolled (mode 1).
template <int J, int R, int Mode, typename real, bool AheadX>
__device__ inline void gemmRows(real (&acc)[J][R][9], const real* __restrict__ AT, int64_t strideJ, int ldAT,
                                const real* X, int K, int t, int tpe, int M) {
  if constexpr (!AheadX) {
    if (K <= 0) return;
    if constexpr (Mode == 1) {
#pragma unroll 1
      for (int l = 0; l < K; ++l) {
        real x[9];
#pragma unroll
        for (int p = 0; p < 9; ++p) x[p] = X[l * 9 + p];
#pragma unroll
        for (int j = 0; j < J; ++j) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int row = t + r * tpe;
            const int rr = row < M ? row : M - 1;
            const real av = AT[j * strideJ + static_cast<int64_t>(l) * ldAT + rr];
            const real a = row < M ? av : real(0);
#pragma unroll
            for (int p = 0; p < 9; ++p) acc[j][r][p] += a * x[p];
          }
        }
      }
      return;
    }
    if constexpr (Mode == 0) {
      for (int l = 0; l < K; ++l) {
        real x[9];
#pragma unroll
        for (int p = 0; p < 9; ++p) x[p] = X[l * 9 + p];
#pragma unroll
        for (int j = 0; j < J; ++j) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int row = t + r * tpe;
            // clamped address: lanes past the last row read a valid entry and drop it
            const int rr = row < M ? row : M - 1;
            const real av = AT[j * strideJ + static_cast<int64_t>(l) * ldAT + rr];
            const real a = row < M ? av : real(0);
#pragma unroll
            for (int p = 0; p < 9; ++p) acc[j][r][p] += a * x[p];
          }
        }
      }
      return;
    }
    int rr[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * tpe;
      // clamped address: lanes past the last row read a valid entry and drop it
      ok[r] = row < M;
      rr[r] = ok[r] ? row : M - 1;
    }
    auto load = [&](int l, real (&v)[J][R]) {
      const int lc = l < K ? l : K - 1;  // steps past the end reload the last one (in bounds, unused)
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r) v[j][r] = AT[j * strideJ + static_cast<int64_t>(lc) * ldAT + rr[r]];
    };
    auto step = [&](int l, const real (&v)[J][R]) {
      real x[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) x[p] = X[l * 9 + p];
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const real a = ok[r] ? v[j][r] : real(0);
#pragma unroll
          for (int p = 0; p < 9; ++p) acc[j][r][p] += a * x[p];
        }
    };
    real a0[J][R], a1[J][R];
    load(0, a0);
    load(1, a1);
    int l = 0;
    for (; l + 1 < K; l += 2) {
      real b0[J][R], b1[J][R];
      load(l + 2, b0);
      load(l + 3, b1);
      __builtin_amdgcn_sched_barrier(0);
      step(l, a0);
      step(l + 1, a1);
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          a0[j][r] = b0[j][r];
          a1[j][r] = b1[j][r];
        }
    }
    if (l < K) step(l, a0);
  } else {
    if (K <= 0) return;
    int rr[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * tpe;
      // clamped address: lanes past the last row read a valid entry and drop it
      ok[r] = row < M;
      rr[r] = ok[r] ? row : M - 1;
    }
    // steps past the end reload the last one (in bounds, unused)
    auto load = [&](int l, real (&v)[J][R]) {
      const int lc = l < K ? l : K - 1;
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r) v[j][r] = AT[j * strideJ + static_cast<int64_t>(lc) * ldAT + rr[r]];
    };
    auto loadX = [&](int l, real (&x)[9]) {
      const int lc = l < K ? l : K - 1;
#pragma unroll
      for (int p = 0; p < 9; ++p) x[p] = X[lc * 9 + p];
    };
    auto step = [&](const real (&x)[9], const real (&v)[J][R]) {
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const real a = ok[r] ? v[j][r] : real(0);
#pragma unroll
          for (int p = 0; p < 9; ++p) acc[j][r][p] += a * x[p];
        }
    };
    auto copy = [](auto& dst, const auto& src) {
      constexpr int N = sizeof(src) / sizeof(real);
      const real* s = reinterpret_cast<const real*>(&src);
      real* d = reinterpret_cast<real*>(&dst);
#pragma unroll
      for (int i = 0; i < N; ++i) d[i] = s[i];
    };
    if constexpr (Mode == 2) {
      real a0[J][R], a1[J][R], x0[9], x1[9];
      load(0, a0);
      load(1, a1);
      loadX(0, x0);
      loadX(1, x1);
      int l = 0;
      for (; l + 1 < K; l += 2) {
        real b0[J][R], b1[J][R], y0[9], y1[9];
        load(l + 2, b0);
        load(l + 3, b1);
        loadX(l + 2, y0);
        loadX(l + 3, y1);
        __builtin_amdgcn_sched_barrier(0);
        step(x0, a0);
        step(x1, a1);
        copy(a0, b0);
        copy(a1, b1);
        copy(x0, y0);
        copy(x1, y1);
      }
      if (l < K) step(x0, a0);
    } else {
      real a0[J][R], x0[9];
      load(0, a0);
      loadX(0, x0);
      auto body = [&](int l) {
        real b0[J][R], y0[9];
        load(l + 1, b0);
        loadX(l + 1, y0);
        step(x0, a0);
        copy(a0, b0);
        copy(x0, y0);
      };
      if constexpr (Mode == 1) {
#pragma unroll 1
        for (int l = 0; l < K; ++l) body(l);
      } else {
        for (int l = 0; l < K; ++l) body(l);
      }
    }
  }
}

// product loop of gemmRows per kernel, each choice measured faster than the
// others where it is used: two-step loads (2) at ORDER 6 (FP64: the neighbour
// kernel with its default tile only, the other tile spills) and for FP64
// ORDER 5 with the tile of more rows per lane; FP32 at ORDER <= 5 not unrolled
// (1); else one step at a time as the compiler unrolls it (0).  The FP64
// ORDER 4 kernels then spill about 0.9 KB per lane and are still faster than
// their spill-free rolled form.
template <typename real, int O, int RX, bool Neighbor>
constexpr int productMode() {
  if constexpr (sizeof(real) == 8) {
    if (O == 6) return (RX == 0 || !Neighbor) ? 2 : 0;
    return (O == 5 && RX == 1) ? 2 : 0;
  } else {
    return O == 6 ? 2 : (O <= 5 ? 1 : 0);
  }
}

// a tile's DOF blocks of src (elements e0, e0 + 1, ..., each [9][ldq], the
// B rows of each quantity read) into LDS as [element][row][quantity], zeros
// past element n - 1; the loads go in chunks of six per thread, each chunk
// issued before its first value is stored (clamped addresses, in bounds: a
// launched tile has e0 < n)
template <int EPB, int QS, int B, int THREADS, typename real>
__device__ inline void stageBlock(real* sA, const real* __restrict__ src, int64_t e0, int64_t n, int ldq,
                                  int64_t qStride) {
  constexpr int TOT = EPB * QS;
  constexpr int NI = (TOT + THREADS - 1) / THREADS;
#pragma unroll
  for (int c = 0; c < NI; c += 6) {
    real v[6];
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int idx = min(static_cast<int>(threadIdx.x) + (c + u) * THREADS, TOT - 1);
      const int s = idx / QS, w = idx % QS;
      const int64_t el = e0 + s < n ? e0 + s : n - 1;
      v[u] = src[el * qStride + (w / B) * ldq + w % B];
    }
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int idx = static_cast<int>(threadIdx.x) + (c + u) * THREADS;
      if (c + u < NI && idx < TOT) {
        const int s = idx / QS, w = idx % QS;
        sA[s * QS + (w % B) * 9 + w / B] = e0 + s < n ? v[u] : real(0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// ader (+ optional volume and local flux): Time.cpp:65-140 and Local.cpp:124-160
// ---------------------------------------------------------------------------
template <typename real, int O, bool WithLocal, int RX = 0, int WV = 4>
__global__ __launch_bounds__(64 * WV) void kAder(int64_t n, const real* __restrict__ mats, Coeffs<O, real> cf,
                                             real* __restrict__ dofs, real* __restrict__ integ,
                                             real* __restrict__ deriv, CellLayout lay,
                                             const real* __restrict__ star, const real* __restrict__ aplus,
                                             const uint8_t* __restrict__ faceTypes, LtsArgs lts) {
  using S = Shape<O, RX, WV>;
  constexpr int B = S::B, B2 = S::B2, R = S::R, TPE = S::TPE;
  constexpr int kProd = productMode<real, O, RX, false>();
  constexpr bool kAheadX = false;  // X loaded at the step (FP64: ahead it spills; FP32: no faster)
  __shared__ real sA[S::EPB * S::QS];
  __shared__ real sB[S::EPB * (S::FS > S::QS ? S::FS : S::QS)];
  __shared__ real sStar[S::EPB * 3 * kStarNnz];
  constexpr int ldB = (S::FS > S::QS ? S::FS : S::QS);

  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * S::EPB;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int slot = wave * S::EPW + lane / TPE;
  const int t = lane % TPE;
  const bool laneOk = lane < S::EPW * TPE;
  const int64_t e = e0 + slot;
  const bool active = laneOk && e < n;
  // LTS (TimeCluster.cpp:380-434): which of buffer / derivatives this cell
  // provides, and whether its buffer accumulates over the sub-steps
  const unsigned setup = (lts.setup != nullptr && active) ? lts.setup[e] : 0u;
  const bool storeDeriv = deriv != nullptr && active && (lts.setup == nullptr || (setup & SDG_LTS_HAS_DERIVATIVES));
  const bool storeBuffer = active && (lts.setup == nullptr || (setup & SDG_LTS_HAS_BUFFERS));
  const bool accumulate = lts.setup != nullptr && !lts.reset && (setup & SDG_LTS_ACCUMULATE_BUFFERS);

  // stage Q as [row][quantity]
  stageBlock<S::EPB, S::QS, B, S::THREADS>(sA, dofs, e0, n, lay.ldq, lay.qStride);
  loadStars(sStar, star, e0, n, S::EPB, lay);
  __syncthreads();

  const int sl = laneOk ? slot : 0;
  real* xa = sA + sl * S::QS;
  real* xb = sB + sl * ldB;
  const real* sS = sStar + sl * 3 * kStarNnz;

  real qreg[R][9], iacc[R][9];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = t + r * TPE;
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      qreg[r][p] = row < B ? xa[row * 9 + p] : real(0);
      iacc[r][p] = cf.c[0] * qreg[r][p];
    }
  }
  if (storeDeriv) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (row < B)
        for (int p = 0; p < 9; ++p) deriv[e * lay.dsize + p * lay.ldq + row] = qreg[r][p];
    }
  }

  int64_t doff = 9LL * lay.ldq;
  int nIn = B;
#pragma unroll 1
  for (int d = 1; d < O; ++d) {
    const int m = basis3(O - d);
    const real* src = (d & 1) ? xa : xb;
    real* dst = (d & 1) ? xb : xa;
    real acc[3][R][9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) acc[j][r][p] = real(0);
    gemmRows<3, R, kProd, real, kAheadX>(acc, mats + S::ck, static_cast<int64_t>(B) * B, B, src, nIn, t, TPE, m);
    real dq[R][9];
    applyStars<R>(acc, sS, dq);
    __syncthreads();  // all reads of dst (previous src) done
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (laneOk && row < m) {
#pragma unroll
        for (int p = 0; p < 9; ++p) {
          dst[row * 9 + p] = dq[r][p];
          iacc[r][p] += cf.c[d] * dq[r][p];
        }
        if (storeDeriv)
          for (int p = 0; p < 9; ++p) deriv[e * lay.dsize + doff + p * lay.rows(m) + row] = dq[r][p];
      }
    }
    __syncthreads();
    doff += 9LL * lay.rows(m);
    nIn = m;
  }

  if (storeBuffer) {
    real* const elemInteg = integ + e * lay.qStride;
    // accumulated buffers: all old values loaded first (clamped rows, in
    // bounds), then stored; a read-modify-write per value through runtime
    // strides waits on every load in turn
    real old[R][9] = {};
    if (accumulate)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) old[r][p] = elemInteg[p * lay.ldq + min(t + r * TPE, B - 1)];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (row < B)
#pragma unroll
        for (int p = 0; p < 9; ++p) elemInteg[p * lay.ldq + row] = old[r][p] + iacc[r][p];
    }
  }
  if constexpr (WithLocal) {
    // I as [row][quantity] in sA
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (laneOk && row < B)
#pragma unroll
        for (int p = 0; p < 9; ++p) xa[row * 9 + p] = iacc[r][p];
    }
    __syncthreads();
    // volume
    {
      real acc[3][R][9];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int p = 0; p < 9; ++p) acc[j][r][p] = real(0);
      gemmRows<3, R, kProd, real, kAheadX>(acc, mats + S::vol, static_cast<int64_t>(B) * B, B, xa, B, t, TPE, B);
      real y[R][9];
      applyStars<R>(acc, sS, y);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) qreg[r][p] += y[r][p];
    }
    // face traces times AplusT
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
      real tr[1][R][9];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) tr[0][r][p] = real(0);
      gemmRows<1, R, kProd, real, kAheadX>(tr, mats + S::fmrt + static_cast<int64_t>(f) * B * B2, 0, B2, xa, B, t, TPE, B2);
      const bool skip = !active || (faceTypes != nullptr && faceTypes[e * 4 + f] == SDG_FACE_DYNAMIC_RUPTURE);
      const real* A = aplus + (active ? e : 0) * lay.aplusStride + 81 * f;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int row = t + r * TPE;
        if (laneOk && row < B2) {
#pragma unroll
          for (int p = 0; p < 9; ++p) {
            real w = real(0);
            if (!skip)
#pragma unroll
              for (int q = 0; q < 9; ++q) w += tr[0][r][q] * A[q + 9 * p];
            xb[(f * B2 + row) * 9 + p] = w;
          }
        }
      }
    }
    __syncthreads();
    {
      real acc[1][R][9];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) acc[0][r][p] = qreg[r][p];
      gemmRows<1, R, kProd, real, kAheadX>(acc, mats + S::rdm, 0, B, xb, 4 * B2, t, TPE, B);
      if (active) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int row = t + r * TPE;
          if (row < B)
#pragma unroll
            for (int p = 0; p < 9; ++p) dofs[e * lay.qStride + p * lay.ldq + row] = acc[0][r][p];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// volume + local flux from a given I (Local.cpp:124-160)
// ---------------------------------------------------------------------------
template <typename real, int O>
__global__ __launch_bounds__(256) void kLocal(int64_t n, const real* __restrict__ mats, real* __restrict__ dofs,
                                              const real* __restrict__ integ, const real* __restrict__ star,
                                              const real* __restrict__ aplus,
                                              const uint8_t* __restrict__ faceTypes, CellLayout lay) {
  using S = Shape<O>;
  constexpr int B = S::B, B2 = S::B2, R = S::R, TPE = S::TPE;
  constexpr int kProd = productMode<real, O, 0, false>();
  constexpr bool kAheadX = false;  // X loaded at the step (FP64: ahead it spills; FP32: no faster)
  __shared__ real sA[S::EPB * S::QS];
  __shared__ real sB[S::EPB * S::FS];
  __shared__ real sStar[S::EPB * 3 * kStarNnz];
  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * S::EPB;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int slot = wave * S::EPW + lane / TPE;
  const int t = lane % TPE;
  const bool laneOk = lane < S::EPW * TPE;
  const int64_t e = e0 + slot;
  const bool active = laneOk && e < n;
  stageBlock<S::EPB, S::QS, B, S::THREADS>(sA, integ, e0, n, lay.ldq, lay.qStride);
  loadStars(sStar, star, e0, n, S::EPB, lay);
  __syncthreads();
  const int sl = laneOk ? slot : 0;
  real* xa = sA + sl * S::QS;
  real* xb = sB + sl * S::FS;
  const real* sS = sStar + sl * 3 * kStarNnz;

  real qreg[R][9];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = t + r * TPE;
#pragma unroll
    for (int p = 0; p < 9; ++p) qreg[r][p] = (active && row < B) ? dofs[e * lay.qStride + p * lay.ldq + row] : real(0);
  }
  {
    real acc[3][R][9];
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) acc[j][r][p] = real(0);
    gemmRows<3, R, kProd, real, kAheadX>(acc, mats + S::vol, static_cast<int64_t>(B) * B, B, xa, B, t, TPE, B);
    real y[R][9];
    applyStars<R>(acc, sS, y);
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) qreg[r][p] += y[r][p];
  }
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    real tr[1][R][9];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) tr[0][r][p] = real(0);
    gemmRows<1, R, kProd, real, kAheadX>(tr, mats + S::fmrt + static_cast<int64_t>(f) * B * B2, 0, B2, xa, B, t, TPE, B2);
    const bool skip = !active || (faceTypes != nullptr && faceTypes[e * 4 + f] == SDG_FACE_DYNAMIC_RUPTURE);
    const real* A = aplus + (active ? e : 0) * lay.aplusStride + 81 * f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (laneOk && row < B2) {
#pragma unroll
        for (int p = 0; p < 9; ++p) {
          real w = real(0);
          if (!skip)
#pragma unroll
            for (int q = 0; q < 9; ++q) w += tr[0][r][q] * A[q + 9 * p];
          xb[(f * B2 + row) * 9 + p] = w;
        }
      }
    }
  }
  __syncthreads();
  real acc[1][R][9];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[0][r][p] = qreg[r][p];
  gemmRows<1, R, kProd, real, kAheadX>(acc, mats + S::rdm, 0, B, xb, 4 * B2, t, TPE, B);
  if (active) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (row < B)
#pragma unroll
        for (int p = 0; p < 9; ++p) dofs[e * lay.qStride + p * lay.ldq + row] = acc[0][r][p];
    }
  }
}

// ---------------------------------------------------------------------------
// neighbour flux (Neighbor.cpp:66-92)
// ---------------------------------------------------------------------------
template <typename real, int O, int RX = 0, int WV = 4>
__global__ __launch_bounds__(64 * WV) void kNeighbor(int64_t n, const real* __restrict__ mats, real* __restrict__ dofs,
                                                 const real* __restrict__ integ,
                                                 const int32_t* __restrict__ neighbors,
                                                 const uint8_t* __restrict__ faceRel,
                                                 const real* __restrict__ aminus,
                                                 const uint8_t* __restrict__ faceTypes, NeighborLts<O, real> lts,
                                                 CellLayout lay) {
  using S = Shape<O, RX, WV>;
  constexpr int B = S::B, B2 = S::B2, R = S::R, TPE = S::TPE;
  constexpr int kProd = productMode<real, O, RX, true>();
  // FP64: X loaded ahead (the neighbour kernels need 130 fewer registers at
  // ORDER 5); FP32: at the step (ahead measured slower at ORDER 6)
  constexpr bool kAheadX = sizeof(real) == 8;
  __shared__ real sA[S::EPB * S::QS];
  __shared__ real sU[S::EPB * B2 * 9];
  __shared__ real sB[S::EPB * S::FS];
  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * S::EPB;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int slot = wave * S::EPW + lane / TPE;
  const int t = lane % TPE;
  const bool laneOk = lane < S::EPW * TPE;
  const int64_t e = e0 + slot;
  const bool active = laneOk && e < n;
  const int sl = laneOk ? slot : 0;
  real* xa = sA + sl * S::QS;
  real* xu = sU + sl * B2 * 9;
  real* xb = sB + sl * S::FS;

  // the four faces' links, read once (clamped: off regular faces nb = 0)
  int64_t nbs[4];
  bool regs[4];
  uint8_t rels[4][2];
  const int64_t ec = active ? e : 0;
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    regs[f] = active && (faceTypes == nullptr || faceTypes[ec * 4 + f] == SDG_FACE_REGULAR);
    const int32_t nbRaw = neighbors[ec * 4 + f];
    nbs[f] = regs[f] ? nbRaw : 0;
    rels[f][0] = faceRel[(ec * 4 + f) * 2 + 0];
    rels[f][1] = faceRel[(ec * 4 + f) * 2 + 1];
  }
  const unsigned setupAll = (lts.setup != nullptr && active) ? lts.setup[e] : 0u;
  // the neighbour buffer of a face, gathered into registers (unconditional,
  // clamped addresses; a face that takes derivatives or is not regular
  // discards the values); FP32: one face ahead, so that the loads are in
  // flight during the previous face's products (FP64: at the face, the
  // registers of a prefetch make the kernel spill)
  constexpr bool kAhead = sizeof(real) == 4;
  constexpr int GI = (S::QS + TPE - 1) / TPE;
  real gv[GI];
  // entry f of a per-face register array (a select chain, not an indexed access)
  auto pick = [](const auto (&v)[4], int f) { return f == 0 ? v[0] : (f == 1 ? v[1] : (f == 2 ? v[2] : v[3])); };
  auto gather = [&](int64_t nb) {
#pragma unroll
    for (int c = 0; c < GI; ++c) {
      const int idx = min(t + c * TPE, S::QS - 1);
      gv[c] = integ[nb * lay.qStride + (idx / B) * lay.ldq + idx % B];
    }
  };
  if constexpr (kAhead) gather(nbs[0]);

#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    const bool regular = pick(regs, f);
    const int64_t nb = pick(nbs, f);
    const uint8_t side = f == 0 ? rels[0][0] : (f == 1 ? rels[1][0] : (f == 2 ? rels[2][0] : rels[3][0]));
    const uint8_t orient = f == 0 ? rels[0][1] : (f == 1 ? rels[1][1] : (f == 2 ? rels[2][1] : rels[3][1]));
    // clamp so malformed relations can never index outside the matrix table
    const int j = regular ? min(static_cast<int>(side), 3) : 0;
    const int h = regular ? min(static_cast<int>(orient), 2) : 0;
    // the neighbour's time integral: its buffer, or its derivatives integrated
    // over this step (TimeCommon.cpp:40-87): GTS relation -> time coefficients,
    // else the sub-interval coefficients
    const unsigned setup = regular ? setupAll : 0u;
    const bool fromDerivatives = (setup >> f) & 1u;
    const Coeffs<O, real>& cf = ((setup >> (4 + f)) & 1u) ? lts.gts : lts.sub;
    if constexpr (!kAhead) gather(nb);
    if (laneOk && !fromDerivatives) {
#pragma unroll
      for (int c = 0; c < GI; ++c) {
        const int idx = t + c * TPE;
        if (idx < S::QS) xa[(idx % B) * 9 + idx / B] = regular ? gv[c] : real(0);
      }
    } else if (laneOk) {
      // the neighbour's derivatives integrated over the interval: the O
      // loads of an entry are issued together, branch-free (an entry past a
      // block's rows reads the block's first entry, in bounds, with weight 0)
      for (int idx = t; idx < S::QS; idx += TPE) {
        const int p = idx / B, l = idx % B;
        real dv[O];
        int64_t off = nb * lay.dsize;
#pragma unroll
        for (int i = 0; i < O; ++i) {
          const int m = basis3(O - i), ld = lay.rows(m);
          dv[i] = lts.deriv[off + (l < m ? p * ld + l : 0)];
          off += 9LL * ld;
        }
        real v = real(0);
#pragma unroll
        for (int i = 0; i < O; ++i) v += (l < basis3(O - i) ? static_cast<real>(cf.c[i]) : real(0)) * dv[i];
        xa[l * 9 + p] = v;
      }
    }
    if constexpr (kAhead)
      if (f + 1 < 4) gather(pick(nbs, f + 1));  // in flight during this face's products
    __syncthreads();
    {
      real u[1][R][9];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) u[0][r][p] = real(0);
      gemmRows<1, R, kProd, real, kAheadX>(u, mats + S::rt + static_cast<int64_t>(j) * B * B2, 0, B2, xa, B, t, TPE, B2);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int row = t + r * TPE;
        if (laneOk && row < B2)
#pragma unroll
          for (int p = 0; p < 9; ++p) xu[row * 9 + p] = u[0][r][p];
      }
    }
    __syncthreads();
    {
      real v[1][R][9];
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int p = 0; p < 9; ++p) v[0][r][p] = real(0);
      gemmRows<1, R, kProd, real, kAheadX>(v, mats + S::fp + static_cast<int64_t>(h) * B2 * B2, 0, B2, xu, B2, t, TPE, B2);
      const real* A = aminus + (active ? e : 0) * lay.aminusStride + 81 * f;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int row = t + r * TPE;
        if (laneOk && row < B2) {
#pragma unroll
          for (int p = 0; p < 9; ++p) {
            real w = real(0);
            if (regular)
#pragma unroll
              for (int q = 0; q < 9; ++q) w += v[0][r][q] * A[q + 9 * p];
            xb[(f * B2 + row) * 9 + p] = w;
          }
        }
      }
    }
    __syncthreads();
  }
  real acc[1][R][9];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[0][r][p] = real(0);
  gemmRows<1, R, kProd, real, kAheadX>(acc, mats + S::rdm, 0, B, xb, 4 * B2, t, TPE, B);
  if (active) {
    // all old values loaded first (clamped rows, in bounds), then stored
    real* const q = dofs + e * lay.qStride;
    real old[R][9];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) old[r][p] = q[p * lay.ldq + min(t + r * TPE, B - 1)];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = t + r * TPE;
      if (row < B)
#pragma unroll
        for (int p = 0; p < 9; ++p) q[p * lay.ldq + row] = old[r][p] + acc[0][r][p];
    }
  }
}

// Neighbour flux with one element per 16-lane row, four per wave, and no
// block barriers (Neighbor.cpp:66-92, the same products as kNeighbor, in the
// same order).  Lane c of an element's row owns output rows c, c + 16, ...;
// each product step's per-element operand X[l][p] is read once per run of 16
// steps into nine row registers (lane c holds X[l0 + c][p]) and broadcast
// across the row (DPP row_newbcast), and so is the face's AminusT.  An
// element's operands stay in LDS areas of its own wave.
template <typename real, int O>
__global__ __launch_bounds__(64) void kNeighborRows(int64_t n, const real* __restrict__ mats, real* __restrict__ dofs,
                                                    const real* __restrict__ integ,
                                                    const int32_t* __restrict__ neighbors,
                                                    const uint8_t* __restrict__ faceRel,
                                                    const real* __restrict__ aminus,
                                                    const uint8_t* __restrict__ faceTypes, NeighborLts<O, real> lts,
                                                    CellLayout lay) {
  using S = Shape<O, 0, 1>;
  constexpr int B = S::B, B2 = S::B2;
  constexpr int RB = (B + 15) / 16, RF = (B2 + 15) / 16;  // rows per lane: volume, face basis
  constexpr int GI = (9 * B + 15) / 16;                   // operand entries staged per lane
  // per element: X [16 RB][9], then W in its place, and U [16 RF][9] (rows
  // past B / B2 only read by lanes whose values are never broadcast)
  constexpr int XS = 16 * RB * 9, FS = 16 * RF * 9;
  __shared__ real sX[4][XS];
  __shared__ real sU[4][FS];
  const int lane = threadIdx.x;
  const int slot = lane >> 4, c = lane & 15;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * 4 + slot;
  const bool active = e < n;
  const int64_t ec = active ? e : 0;
  real* const xs = sX[slot];
  real* const us = sU[slot];
  real* const ws = xs;  // W after X: the first product has read X

  // the four faces' links, read once (clamped: off regular faces nb = 0)
  int64_t nbs[4];
  bool regs[4];
  uint8_t rels[4][2];
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    regs[f] = active && (faceTypes == nullptr || faceTypes[ec * 4 + f] == SDG_FACE_REGULAR);
    const int32_t nbRaw = neighbors[ec * 4 + f];
    nbs[f] = regs[f] ? nbRaw : 0;
    rels[f][0] = faceRel[(ec * 4 + f) * 2 + 0];
    rels[f][1] = faceRel[(ec * 4 + f) * 2 + 1];
  }
  const unsigned setupAll = (lts.setup != nullptr && active) ? lts.setup[e] : 0u;
  auto pick = [](const auto (&v)[4], int f) { return f == 0 ? v[0] : (f == 1 ? v[1] : (f == 2 ? v[2] : v[3])); };
  // the matrix entries of a product are loaded in groups of kGroup steps.
  // FP32 at ORDER 5: groups of 8 and the neighbour buffer at its face, which
  // takes the kernel to 160 registers and three waves per SIMD (measured 9 %
  // faster there; slower at orders 4 and 6, which keep groups of 16 and the
  // buffer one face ahead).  FP64: groups of 4 (16 cost the second wave per
  // SIMD; 4 measured 11 % faster than loading each entry at its step)
  constexpr bool kLean = sizeof(real) == 4 && O == 5;
  constexpr int kGroup = sizeof(real) == 8 ? 4 : (kLean ? 8 : 16);
  // the neighbour buffer of a face gathered into registers (clamped
  // addresses; FP32: one face ahead, in flight during the previous face)
  constexpr bool kAhead = sizeof(real) == 4 && !kLean;
  real gv[GI];
  auto gather = [&](int64_t nb) {
#pragma unroll
    for (int i = 0; i < GI; ++i) {
      const int idx = min(c + 16 * i, 9 * B - 1);
      gv[i] = integ[nb * lay.qStride + (idx / B) * lay.ldq + idx % B];
    }
  };
  if constexpr (kAhead) gather(nbs[0]);

  // acc[r][p] += sum_{l < K} AT[l][row_r] * Xs[l][p] over runs of 16 steps
  auto product = [&](auto rowsI, auto kI, real(&acc)[decltype(rowsI)::value][9], const real* __restrict__ AT,
                     int ldAT, int M, const real* Xs) {
    constexpr int R = decltype(rowsI)::value, K = decltype(kI)::value;
    int rr[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = c + 16 * r;
      ok[r] = row < M;
      rr[r] = ok[r] ? row : M - 1;  // clamped: read, dropped
    }
    staticFor<(K + 15) / 16>([&](auto runI) {
      constexpr int l0 = 16 * decltype(runI)::value;
      constexpr int nl = K - l0 < 16 ? K - l0 : 16;
      // the matrix entries of a group of kGroup steps all loaded before its
      // first FMA (the compiler does not move loads across the FMA
      // statements: loaded at their step, each step waits on its load)
      real av[kGroup][R];
      real row[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) row[p] = Xs[(l0 + c) * 9 + p];
      staticFor<nl>([&](auto jI) {
        constexpr int jj = decltype(jI)::value;
        if constexpr (jj % kGroup == 0) {
#pragma unroll
          for (int u = 0; u < kGroup; ++u)
            if (jj + u < nl)
#pragma unroll
              for (int r = 0; r < R; ++r) av[u][r] = AT[static_cast<int64_t>(l0 + jj + u) * ldAT + rr[r]];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) fmacRowBcast9<jj>(acc[r], row, ok[r] ? av[jj % kGroup][r] : real(0));
      });
    });
  };

  real acc[RB][9];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[r][p] = real(0);

#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    const bool regular = pick(regs, f);
    const int64_t nb = pick(nbs, f);
    const uint8_t side = f == 0 ? rels[0][0] : (f == 1 ? rels[1][0] : (f == 2 ? rels[2][0] : rels[3][0]));
    const uint8_t orient = f == 0 ? rels[0][1] : (f == 1 ? rels[1][1] : (f == 2 ? rels[2][1] : rels[3][1]));
    // clamp so malformed relations can never index outside the matrix table
    const int j = regular ? min(static_cast<int>(side), 3) : 0;
    const int h = regular ? min(static_cast<int>(orient), 2) : 0;
    // the neighbour's time integral: its buffer, or its derivatives integrated
    // over this step (TimeCommon.cpp:40-87): GTS relation -> time coefficients,
    // else the sub-interval coefficients
    const unsigned setup = regular ? setupAll : 0u;
    const bool fromDerivatives = (setup >> f) & 1u;
    const Coeffs<O, real>& cf = ((setup >> (4 + f)) & 1u) ? lts.gts : lts.sub;
    if constexpr (!kAhead) gather(nb);
    waveSync();  // the previous face's reads of the element's areas are done
    if (!fromDerivatives) {
#pragma unroll
      for (int i = 0; i < GI; ++i) {
        const int idx = c + 16 * i;
        if (idx < 9 * B) xs[(idx % B) * 9 + idx / B] = regular ? gv[i] : real(0);
      }
    } else {
      // the neighbour's derivatives integrated over the interval: the O loads
      // of an entry are issued together (an entry past a block's rows reads
      // the block's first entry, in bounds, with weight 0)
      for (int idx = c; idx < 9 * B; idx += 16) {
        const int p = idx / B, l = idx % B;
        real dv[O];
        int64_t off = nb * lay.dsize;
#pragma unroll
        for (int i = 0; i < O; ++i) {
          const int m = basis3(O - i), ld = lay.rows(m);
          dv[i] = lts.deriv[off + (l < m ? p * ld + l : 0)];
          off += 9LL * ld;
        }
        real v = real(0);
#pragma unroll
        for (int i = 0; i < O; ++i) v += (l < basis3(O - i) ? static_cast<real>(cf.c[i]) : real(0)) * dv[i];
        xs[l * 9 + p] = v;
      }
    }
    if constexpr (kAhead)
      if (f + 1 < 4) gather(pick(nbs, f + 1));  // in flight during this face's products
    // AminusT of the face: rowA[p] holds A(min(c, 8), p)
    const real* A = aminus + ec * lay.aminusStride + 81 * f;
    real rowA[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) rowA[p] = A[min(c, 8) + 9 * p];
    waveSync();
    // U = rT_j X (rows < B2)
    real u[RF][9];
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) u[r][p] = real(0);
    product(std::integral_constant<int, RF>{}, std::integral_constant<int, B>{}, u,
            mats + S::rt + static_cast<int64_t>(j) * B * B2, B2, B2, xs);
#pragma unroll
    for (int r = 0; r < RF; ++r)
      if (c + 16 * r < B2)
#pragma unroll
        for (int p = 0; p < 9; ++p) us[(c + 16 * r) * 9 + p] = u[r][p];
    waveSync();
    // V = fP_h U, then W = V AminusT (zero off regular faces)
    real v[RF][9];
#pragma unroll
    for (int r = 0; r < RF; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) v[r][p] = real(0);
    product(std::integral_constant<int, RF>{}, std::integral_constant<int, B2>{}, v,
            mats + S::fp + static_cast<int64_t>(h) * B2 * B2, B2, B2, us);
#pragma unroll
    for (int r = 0; r < RF; ++r) {
      real w[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) w[p] = real(0);
      staticFor<9>([&](auto qI) {
        constexpr int q = decltype(qI)::value;
        fmacRowBcast9<q>(w, rowA, v[r][q]);
      });
      if (c + 16 * r < B2)
#pragma unroll
        for (int p = 0; p < 9; ++p) ws[(c + 16 * r) * 9 + p] = regular ? w[p] : real(0);
    }
    waveSync();
    // Q += rDivM_f W, accumulated face by face (rows f B2 .. f B2 + B2 - 1 of
    // the stacked [rDivM_0 .. rDivM_3])
    product(std::integral_constant<int, RB>{}, std::integral_constant<int, B2>{}, acc,
            mats + S::rdm + static_cast<int64_t>(f) * B2 * B, B, B, ws);
  }
  if (active) {
    // all old values loaded first (clamped rows, in bounds), then stored
    real* const q = dofs + e * lay.qStride;
    real old[RB][9];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
      for (int p = 0; p < 9; ++p) old[r][p] = q[p * lay.ldq + min(c + 16 * r, B - 1)];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int row = c + 16 * r;
      if (row < B)
#pragma unroll
        for (int p = 0; p < 9; ++p) q[p * lay.ldq + row] = old[r][p] + acc[r][p];
    }
  }
}

// Taylor sum out = sum_i coeffs[i] dQ_i (KernelsAux.cpp:117-184 semantics)
template <typename real, int O>
__global__ __launch_bounds__(256) void kTaylor(int64_t n, Coeffs<O, real> cf, const real* __restrict__ deriv,
                                               CellLayout lay, real* __restrict__ out) {
  constexpr int B = basis3(O);
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * 9 * B) return;
  const int64_t e = idx / (9 * B);
  const int w = static_cast<int>(idx % (9 * B));
  const int p = w / B, k = w % B;
  real acc = real(0);
  int64_t off = 0;
#pragma unroll
  for (int d = 0; d < O; ++d) {
    const int m = basis3(O - d), ld = lay.rows(m);
    if (k < m) acc += cf.c[d] * deriv[e * lay.dsize + off + p * ld + k];
    off += 9LL * ld;
  }
  out[e * lay.qStride + p * lay.ldq + k] = acc;
}

template <typename real, int O>
Coeffs<O, real> makeCoeffs(const double* c) {
  Coeffs<O, real> r{};
  for (int i = 0; i < O; ++i) r.c[i] = static_cast<real>(c[i]);
  return r;
}

template <typename real, int O, int RX = 0, int WV = 4>
int aderImpl(const EngineState& st, const double* coeffs, const sdg_layer* L, hipStream_t stream, bool withLocal,
             int reset) {
  using S = Shape<O, RX, WV>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const dim3 grid(static_cast<unsigned>((n + S::EPB - 1) / S::EPB));
  const auto cf = makeCoeffs<real, O>(coeffs);
  const auto* mats = static_cast<const real*>(st.matrices);
  const LtsArgs lts{L->lts_setup, reset};
  const CellLayout lay = cellLayout(st, L);
  if (withLocal)
    kAder<real, O, true, RX, WV><<<grid, S::THREADS, 0, stream>>>(
        n, mats, cf, static_cast<real*>(L->dofs), static_cast<real*>(L->time_integrated),
        static_cast<real*>(L->derivatives), lay, static_cast<const real*>(L->star),
        static_cast<const real*>(L->aplus), L->face_types, lts);
  else
    kAder<real, O, false, RX, WV><<<grid, S::THREADS, 0, stream>>>(
        n, mats, cf, static_cast<real*>(L->dofs), static_cast<real*>(L->time_integrated),
        static_cast<real*>(L->derivatives), lay, static_cast<const real*>(L->star),
        static_cast<const real*>(L->aplus), L->face_types, lts);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <typename real, int O>
int localImpl(const EngineState& st, const sdg_layer* L, hipStream_t stream) {
  using S = Shape<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const dim3 grid(static_cast<unsigned>((n + S::EPB - 1) / S::EPB));
  kLocal<real, O><<<grid, S::THREADS, 0, stream>>>(
      n, static_cast<const real*>(st.matrices), static_cast<real*>(L->dofs),
      static_cast<const real*>(L->time_integrated), static_cast<const real*>(L->star),
      static_cast<const real*>(L->aplus), L->face_types, cellLayout(st, L));
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <typename real, int O, int RX = 0, int WV = 4>
int neighborImpl(const EngineState& st, const sdg_layer* L, hipStream_t stream, const double* gtsCoeffs,
                 const double* subCoeffs) {
  using S = Shape<O, RX, WV>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const dim3 grid(static_cast<unsigned>((n + S::EPB - 1) / S::EPB));
  NeighborLts<O, real> lts{};
  if (gtsCoeffs != nullptr) {
    lts.setup = L->lts_setup;
    lts.deriv = static_cast<const real*>(L->derivatives);
    lts.gts = makeCoeffs<real, O>(gtsCoeffs);
    lts.sub = makeCoeffs<real, O>(subCoeffs);
  }
  kNeighbor<real, O, RX, WV><<<grid, S::THREADS, 0, stream>>>(
      n, static_cast<const real*>(st.matrices), static_cast<real*>(L->dofs),
      static_cast<const real*>(L->time_integrated), L->face_neighbors, L->face_relations,
      static_cast<const real*>(L->aminus), L->face_types, lts, cellLayout(st, L));
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <typename real, int O>
int neighborRowsImpl(const EngineState& st, const sdg_layer* L, hipStream_t stream, const double* gtsCoeffs,
                     const double* subCoeffs) {
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const dim3 grid(static_cast<unsigned>((n + 3) / 4));
  NeighborLts<O, real> lts{};
  if (gtsCoeffs != nullptr) {
    lts.setup = L->lts_setup;
    lts.deriv = static_cast<const real*>(L->derivatives);
    lts.gts = makeCoeffs<real, O>(gtsCoeffs);
    lts.sub = makeCoeffs<real, O>(subCoeffs);
  }
  kNeighborRows<real, O><<<grid, 64, 0, stream>>>(
      n, static_cast<const real*>(st.matrices), static_cast<real*>(L->dofs),
      static_cast<const real*>(L->time_integrated), L->face_neighbors, L->face_relations,
      static_cast<const real*>(L->aminus), L->face_types, lts, cellLayout(st, L));
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <typename real, int O>
int taylorImpl(const EngineState& st, const double* coeffs, const void* deriv, void* out, int64_t n,
               hipStream_t stream) {
  if (n == 0) return SDG_OK;
  const int64_t total = n * 9 * basis3(O);
  const dim3 grid(static_cast<unsigned>((total + 255) / 256));
  kTaylor<real, O><<<grid, 256, 0, stream>>>(n, makeCoeffs<real, O>(coeffs), static_cast<const real*>(deriv),
                                              engineLayout(st), static_cast<real*>(out));
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <template <typename, int> class F, typename... Args>
int dispatch(const EngineState& st, Args&&... args) {
  const bool dbl = st.arithmetic == SDG_FP64;
  switch (st.order) {
#define SDG_CASE(o)                                                            \
  case o:                                                                      \
    return dbl ? F<double, o>::run(st, args...) : F<float, o>::run(st, args...);
    SDG_CASE(2) SDG_CASE(3) SDG_CASE(4) SDG_CASE(5) SDG_CASE(6) SDG_CASE(7) SDG_CASE(8)
#undef SDG_CASE
    default:
      setError("unsupported order");
      return SDG_ERR_UNSUPPORTED;
  }
}

template <typename real, int O>
struct AderOp {
  static int run(const EngineState& st, const double* c, const sdg_layer* L, hipStream_t s, bool wl, int reset) {
    return aderImpl<real, O>(st, c, L, s, wl, reset);
  }
};
// static LDS of kAder / kNeighbor for a tile shape
template <typename real, int O, int RX, int WV>
constexpr int64_t tileLds() {
  using S = Shape<O, RX, WV>;
  const int64_t ader = S::EPB * (S::QS + (S::FS > S::QS ? S::FS : S::QS) + 3 * kStarNnz);
  const int64_t nb = S::EPB * (S::QS + S::B2 * 9 + S::FS);
  return static_cast<int64_t>(sizeof(real)) * (ader > nb ? ader : nb);
}
// waves per block of a variant: 4 unless its LDS does not fit, then 2
template <typename real, int O, int RX, int WV>
constexpr int fitWaves() {
  return tileLds<real, O, RX, WV>() <= 163840 ? WV : 2;
}

// tile variants of the local time stepping kernels: bit 0 -> one more row per
// lane, bit 1 -> two waves per block instead of four (orders 4..6)
template <typename real, int O>
struct LtsLocalOp {
  static int run(const EngineState& st, const double* c, const sdg_layer* L, hipStream_t s, int reset, int v) {
    if constexpr (O >= 4 && O <= 6) {
      switch (v) {
        case 1: return aderImpl<real, O, 1, fitWaves<real, O, 1, 4>()>(st, c, L, s, true, reset);
        case 2: return aderImpl<real, O, 0, 2>(st, c, L, s, true, reset);
        case 3: return aderImpl<real, O, 1, 2>(st, c, L, s, true, reset);
        default: break;
      }
    }
    return aderImpl<real, O>(st, c, L, s, true, reset);
  }
};
template <typename real, int O>
struct LtsNeighborOp {
  static int run(const EngineState& st, const sdg_layer* L, hipStream_t s, const double* g, const double* sub, int v) {
    if (v == kNeighborRowsVariant) return neighborRowsImpl<real, O>(st, L, s, g, sub);
    if constexpr (O >= 4 && O <= 6) {
      switch (v) {
        case 1: return neighborImpl<real, O, 1, fitWaves<real, O, 1, 4>()>(st, L, s, g, sub);
        case 2: return neighborImpl<real, O, 0, 2>(st, L, s, g, sub);
        case 3: return neighborImpl<real, O, 1, 2>(st, L, s, g, sub);
        default: break;
      }
    }
    return neighborImpl<real, O>(st, L, s, g, sub);
  }
};
template <typename real, int O>
struct LocalOp {
  static int run(const EngineState& st, const sdg_layer* L, hipStream_t s) { return localImpl<real, O>(st, L, s); }
};
template <typename real, int O>
struct NeighborOp {
  static int run(const EngineState& st, const sdg_layer* L, hipStream_t s, const double* g, const double* sub) {
    // FP32 at orders 3-6: one element per 16-lane row (measured 1.05-1.5x
    // faster there; slower at order 8 and in FP64, where its accumulators and
    // operand rows need more registers than kNeighbor's tiles)
    if constexpr (sizeof(real) == 4 && O >= 3 && O <= 6) return neighborRowsImpl<real, O>(st, L, s, g, sub);
    return neighborImpl<real, O>(st, L, s, g, sub);
  }
};
template <typename real, int O>
struct TaylorOp {
  static int run(const EngineState& st, const double* c, const void* d, void* o, int64_t n, hipStream_t s) {
    return taylorImpl<real, O>(st, c, d, o, n, s);
  }
};

}  // namespace

DeviceMatrixLayout deviceMatrixLayout(int order) {
  const int B = basis3(order), B2 = basis2(order);
  DeviceMatrixLayout l{};
  l.ck = 0;
  l.vol = l.ck + 3LL * B * B;
  l.fmrt = l.vol + 3LL * B * B;
  l.rt = l.fmrt + 4LL * B * B2;
  l.fp = l.rt + 4LL * B * B2;
  l.rdm = l.fp + 3LL * B2 * B2;
  l.total = l.rdm + 4LL * B2 * B;
  return l;
}

int launchAder(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream) {
  return dispatch<AderOp>(e, coeffs, layer, stream, false, 1);
}
int launchAderLocal(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream) {
  if (hasMfma64(e)) return launchAderLocalMfma64(e, coeffs, layer, stream);
  return dispatch<AderOp>(e, coeffs, layer, stream, true, 1);
}
int launchLocalIntegration(const EngineState& e, const double* coeffs, const sdg_layer* layer, int resetBuffers,
                           hipStream_t stream, int variant) {
  return dispatch<LtsLocalOp>(e, coeffs, layer, stream, resetBuffers, variant);
}
int launchNeighboringIntegration(const EngineState& e, const double* timeCoeffs, const double* subtimeCoeffs,
                                 const sdg_layer* layer, hipStream_t stream, int variant) {
  return dispatch<LtsNeighborOp>(e, layer, stream, timeCoeffs, subtimeCoeffs, variant);
}
int launchLocal(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  return dispatch<LocalOp>(e, layer, stream);
}
int launchNeighbor(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  return dispatch<NeighborOp>(e, layer, stream, static_cast<const double*>(nullptr),
                              static_cast<const double*>(nullptr));
}
int launchTaylor(const EngineState& e, const double* coeffs, const void* derivatives, void* out, int64_t n,
                 hipStream_t stream) {
  return dispatch<TaylorOp>(e, coeffs, derivatives, out, n, stream);
}

}  // namespace sdg
