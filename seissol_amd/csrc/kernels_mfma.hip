// Element update on the matrix cores: FP32 data, split-BF16 products, FP32
// accumulation (arithmetics SDG_BF16X3 and SDG_BF16X6).
//
// Every contraction with a global matrix A (kDivMT, kDivM, fMrT, rDivM) is a
// GEMM  Y(M x 144) += A(M x K) X(K x 144)  over a tile of 16 elements, the 144
// columns being (quantity p, element e) -> 16 p + e.  It runs as
// v_mfma_f32_16x16x32_bf16 with both operands split into NS bf16 planes
// (x = x0 + x1 [+ x2], each plane the rounding of the remainder):
//   NS = 2: A0X0 + A0X1 + A1X0                      (3 products, ~2^-17)
//   NS = 3: A0X0 + A0X1 + A1X0 + A0X2 + A2X0 + A1X1  (6 products, ~2^-24)
//
// One wave owns one 16-row tile of the basis index for all 9 quantities of the
// 16 elements, so a lane holds rows 4g..4g+3 (g = lane / 16) of element
// e = lane % 16 for every quantity: the 9x9 star, AplusT and AminusT products
// that mix quantities are plain register arithmetic.  X lives in LDS
// column-major ([column][k], k contiguous) as bf16 hi/lo; A is pre-split and
// pre-swizzled in global memory so one fragment is one coalesced 16-byte load
// per lane (L2-resident, shared by all workgroups).
//
// Two kernels form one time step of the proxy `all` kernel:
//  kLocalMfma:    CK derivatives + time integral, volume, local flux of the
//                 four faces; writes Q and the face traces rT_f I.
//  kNeighborMfma: neighbour flux rDivM_f ((fP_h Tr) AminusT_f) from the traces
//                 the neighbours wrote; updates Q.
// Operator sequence: codegen/kernels/aderdg.py:321-537 (the neighbour's rT_j I
// is produced by the neighbour itself, rT_j = diag(1/M2) fMrT_j).
#include <hip/hip_runtime.h>

#include "engine.h"
#include "mfma_common.h"

namespace sdg {
namespace mfma {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ constexpr int kStarQ[9][3] = {{6, 7, 8}, {6, 7, 8}, {6, 7, 8}, {6, 7, -1}, {7, 8, -1},
                                         {6, 8, -1}, {0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
__device__ constexpr int kStarOff[9] = {0, 3, 6, 9, 11, 13, 15, 18, 21};
__device__ constexpr int kStarCnt[9] = {3, 3, 3, 2, 2, 2, 3, 3, 3};

template <int O>
struct CoeffsF {
  float c[O];
};

template <int NS>
__device__ inline void split(float x, __bf16 (&parts)[NS]) {
  float rest = x;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    parts[i] = static_cast<__bf16>(rest);
    rest -= static_cast<float>(parts[i]);
  }
}

// X planes: NS consecutive arrays of kCols * LDX bf16
template <int NS, int LDX>
__device__ inline void storeX4(__bf16* xs, int col, int k0, const float (&v)[4]) {
  bf16x4 out[NS];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    __bf16 parts[NS];
    split<NS>(v[j], parts);
#pragma unroll
    for (int i = 0; i < NS; ++i) out[i][j] = parts[i];
  }
#pragma unroll
  for (int i = 0; i < NS; ++i) *reinterpret_cast<bf16x4*>(xs + i * kCols * LDX + col * LDX + k0) = out[i];
}

template <int NS, int LDX>
__device__ inline void storeX1(__bf16* xs, int col, int k, float v) {
  __bf16 parts[NS];
  split<NS>(v, parts);
#pragma unroll
  for (int i = 0; i < NS; ++i) xs[i * kCols * LDX + col * LDX + k] = parts[i];
}

template <int NS, int LDX>
__device__ inline void storeZero4(__bf16* xs, int col, int k0) {
  bf16x4 z;
#pragma unroll
  for (int j = 0; j < 4; ++j) z[j] = static_cast<__bf16>(0.0f);
#pragma unroll
  for (int i = 0; i < NS; ++i) *reinterpret_cast<bf16x4*>(xs + i * kCols * LDX + col * LDX + k0) = z;
}

__device__ inline bf16x8 loadFrag(const __bf16* __restrict__ base, int lane) {
  return *reinterpret_cast<const bf16x8*>(base + lane * 8);
}

// split product, smallest terms first
template <int NS>
__device__ inline floatx4 mfmaSplit(const bf16x8 (&a)[NS], const bf16x8 (&b)[NS], floatx4 c) {
  if constexpr (NS == 3) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
  }
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
  return c;
}

// acc[jj][i] += A_jj(tile) X(:, quantity Q0 + i)  for JJ matrices with
// fragment bases a[jj] and NQ consecutive quantities
template <int NS, int JJ, int Q0, int NQ, int LDX>
__device__ inline void gemmTile(floatx4 (&acc)[JJ][NQ], const __bf16* const (&a)[JJ], int ksteps,
                                const __bf16* xs, int lane) {
  const int g = lane >> 4, e = lane & 15;
  for (int kk = 0; kk < ksteps; ++kk) {
    bf16x8 af[JJ][NS];
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj)
#pragma unroll
      for (int i = 0; i < NS; ++i) af[jj][i] = loadFrag(a[jj] + kk * kFragPair + i * kFragPlane, lane);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int off = ((Q0 + q) * 16 + e) * LDX + kk * 32 + 8 * g;
      bf16x8 bf[NS];
#pragma unroll
      for (int i = 0; i < NS; ++i) bf[i] = *reinterpret_cast<const bf16x8*>(xs + i * kCols * LDX + off);
#pragma unroll
      for (int jj = 0; jj < JJ; ++jj) acc[jj][q] = mfmaSplit<NS>(af[jj], bf, acc[jj][q]);
    }
  }
}

// Y += sum_j (A_j X) star_j for the lane's element.  The elastic star matrix
// maps velocities (6..8) to stresses (0..5) and stresses to velocities, so the
// contraction runs in two passes of 3 and 6 input quantities, which bounds the
// live accumulators at 3 x 6 tiles.
template <int NS, int LDX>
__device__ inline void gemmStars(const __bf16* const (&a)[3], int ksteps, const __bf16* xs, int lane,
                                 const float* sS, float (&y)[9][4]) {
  {
    floatx4 acc[3][3];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[jj][i] = floatx4{0, 0, 0, 0};
    gemmTile<NS, 3, 6, 3, LDX>(acc, a, ksteps, xs, lane);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int p = 0; p < 6; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (k < kStarCnt[p]) {
            const float sv = sS[jj * 24 + kStarOff[p] + k];
            const int qi = kStarQ[p][k] - 6;
#pragma unroll
            for (int j = 0; j < 4; ++j) y[p][j] += acc[jj][qi][j] * sv;
          }
  }
  {
    floatx4 acc[3][6];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int i = 0; i < 6; ++i) acc[jj][i] = floatx4{0, 0, 0, 0};
    gemmTile<NS, 3, 0, 6, LDX>(acc, a, ksteps, xs, lane);
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int p = 6; p < 9; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float sv = sS[jj * 24 + kStarOff[p] + k];
          const int qi = kStarQ[p][k];
#pragma unroll
          for (int j = 0; j < 4; ++j) y[p][j] += acc[jj][qi][j] * sv;
        }
  }
}

template <int O, int NS>
__global__ __launch_bounds__(MShape<O>::THREADS, 1) void kLocalMfma(
    int64_t n, const __bf16* __restrict__ frag, const float* __restrict__ tables, CoeffsF<O> cf,
    float* __restrict__ dofs, float* __restrict__ integ, float* __restrict__ traces,
    const float* __restrict__ star, const float* __restrict__ aplus, const uint8_t* __restrict__ faceTypes) {
  using S = MShape<O>;
  constexpr int B = S::B, B2 = S::B2, MT = S::MT, LDX = S::LDX, FR = S::FR;
  constexpr int TPW = cdiv(S::MTF, MT);
  __shared__ __bf16 xs[NS * kCols * LDX];
  __shared__ float sStar[kTileElems * 72];
  __shared__ float sAp[kTileElems * 81];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const int64_t elem = E0 + e;
  const bool valid = elem < n;

  for (int i = threadIdx.x; i < kTileElems * 72; i += blockDim.x) {
    const int s = i / 72, r = i % 72, jj = r / 24;
    int z = r % 24, p = 0;
    while (z >= kStarCnt[p]) { z -= kStarCnt[p]; ++p; }
    const int q = kStarQ[p][z];
    sStar[i] = (E0 + s < n) ? star[(E0 + s) * 243 + 81 * jj + q + 9 * p] : 0.0f;
  }

  float q[9][4], iacc[9][4];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = 16 * w + 4 * g + j;
      q[p][j] = (valid && row < B) ? dofs[elem * 9 * B + p * B + row] : 0.0f;
      iacc[p][j] = cf.c[0] * q[p][j];
    }
#pragma unroll
  for (int p = 0; p < 9; ++p) storeX4<NS, LDX>(xs, p * 16 + e, 16 * w + 4 * g, q[p]);
  for (int t = MT + w; t < S::KPB / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<NS, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();
  const float* sS = sStar + e * 72;

  // Cauchy-Kowalevski recursion
#pragma unroll 1
  for (int d = 1; d < O; ++d) {
    const int m = basis3(O - d), nIn = basis3(O - d + 1);
    const int mtiles = cdiv(m, 16), ks = cdiv(nIn, 32), kpNext = cdiv(m, 32) * 32;
    float dq[9][4];
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) dq[p][j] = 0.0f;
    if (w < mtiles) {
      const __bf16* a[3] = {frag + S::ck + (0LL * MT + w) * S::KSB * kFragPair,
                            frag + S::ck + (1LL * MT + w) * S::KSB * kFragPair,
                            frag + S::ck + (2LL * MT + w) * S::KSB * kFragPair};
      gemmStars<NS, LDX>(a, ks, xs, lane, sS, dq);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (16 * w + 4 * g + j >= m)
#pragma unroll
          for (int p = 0; p < 9; ++p) dq[p][j] = 0.0f;
    }
    __syncthreads();
    if (w < mtiles) {
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        storeX4<NS, LDX>(xs, p * 16 + e, 16 * w + 4 * g, dq[p]);
#pragma unroll
        for (int j = 0; j < 4; ++j) iacc[p][j] += cf.c[d] * dq[p][j];
      }
    }
    for (int t = w; t < kpNext / 16; t += MT)
      if (t >= mtiles)
        for (int p = 0; p < 9; ++p) storeZero4<NS, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
    __syncthreads();
  }

  // time-integrated DOFs: optional output, and the operand of the next GEMMs
  if (integ != nullptr && valid) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 16 * w + 4 * g + j;
        if (row < B) integ[elem * 9 * B + p * B + row] = iacc[p][j];
      }
  }
#pragma unroll
  for (int p = 0; p < 9; ++p) storeX4<NS, LDX>(xs, p * 16 + e, 16 * w + 4 * g, iacc[p]);
  for (int t = MT + w; t < S::KPB / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<NS, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();

  // volume: Q += sum_j kDivM_j I star_j
  {
    const __bf16* a[3] = {frag + S::vol + (0LL * MT + w) * S::KSB * kFragPair,
                          frag + S::vol + (1LL * MT + w) * S::KSB * kFragPair,
                          frag + S::vol + (2LL * MT + w) * S::KSB * kFragPair};
    gemmStars<NS, LDX>(a, S::KSB, xs, lane, sS, q);
  }

  // face traces T = fMrT_f I (rows f * B2 + m)
  float tr[TPW][9][4];
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti) {
    const int t = w + ti * MT;
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) tr[ti][p][j] = 0.0f;
    if (t < S::MTF) {
      floatx4 acc[1][9];
#pragma unroll
      for (int p = 0; p < 9; ++p) acc[0][p] = floatx4{0, 0, 0, 0};
      const __bf16* a[1] = {frag + S::fmrt + static_cast<int64_t>(t) * S::KSB * kFragPair};
      gemmTile<NS, 1, 0, 9, LDX>(acc, a, S::KSB, xs, lane);
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) tr[ti][p][j] = acc[0][p][j];
      // neighbour-side traces rT_f I = T / M2
      if (valid) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int r = 16 * t + 4 * g + j;
          if (r < FR) {
            const int f = r / B2, mm = r % B2;
            const float s = tables[S::tInvM2 + mm];
#pragma unroll
            for (int p = 0; p < 9; ++p) traces[((elem * 4 + f) * 9 + p) * B2 + mm] = tr[ti][p][j] * s;
          }
        }
      }
    }
  }

  // W = T AplusT_f, face by face (AplusT staged in LDS)
  float wv[TPW][9][4];
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti)
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) wv[ti][p][j] = 0.0f;
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    __syncthreads();
    for (int i = threadIdx.x; i < kTileElems * 81; i += blockDim.x) {
      const int s = i / 81, r = i % 81;
      sAp[i] = (E0 + s < n) ? aplus[(E0 + s) * 324 + f * 81 + r] : 0.0f;
    }
    __syncthreads();
    const bool dr = valid && faceTypes != nullptr && faceTypes[elem * 4 + f] == SDG_FACE_DYNAMIC_RUPTURE;
    const float* A = sAp + e * 81;
#pragma unroll
    for (int ti = 0; ti < TPW; ++ti) {
      const int t = w + ti * MT;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = 16 * t + 4 * g + j;
        if (t < S::MTF && r < FR && r / B2 == f && !dr) {
#pragma unroll
          for (int p = 0; p < 9; ++p) {
            float acc = 0.0f;
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) acc += tr[ti][qq][j] * A[qq + 9 * p];
            wv[ti][p][j] = acc;
          }
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti) {
    const int t = w + ti * MT;
    if (t < S::MTF)
#pragma unroll
      for (int p = 0; p < 9; ++p) storeX4<NS, LDX>(xs, p * 16 + e, 16 * t + 4 * g, wv[ti][p]);
  }
  for (int t = S::MTF + w; t < S::KPF / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<NS, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();

  // local flux: Q += [rDivM_0 .. rDivM_3] W
  {
    floatx4 acc[1][9];
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[0][p] = floatx4{0, 0, 0, 0};
    const __bf16* a[1] = {frag + S::rdm + static_cast<int64_t>(w) * S::KSF * kFragPair};
    gemmTile<NS, 1, 0, 9, LDX>(acc, a, S::KSF, xs, lane);
    if (valid) {
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int row = 16 * w + 4 * g + j;
          if (row < B) dofs[elem * 9 * B + p * B + row] = q[p][j] + acc[0][p][j];
        }
    }
  }
}

template <int O, int NS>
__global__ __launch_bounds__(MShape<O>::THREADS, 1) void kNeighborMfma(
    int64_t n, const __bf16* __restrict__ frag, const float* __restrict__ tables, float* __restrict__ dofs,
    const float* __restrict__ traces, const int32_t* __restrict__ neighbors, const uint8_t* __restrict__ faceRel,
    const float* __restrict__ aminus, const uint8_t* __restrict__ faceTypes) {
  using S = MShape<O>;
  constexpr int B = S::B, B2 = S::B2, MT = S::MT, LDX = S::LDX, FR = S::FR;
  __shared__ __bf16 xs[NS * kCols * LDX];
  __shared__ float sTr[kTileElems * 9 * B2];
  __shared__ float sAm[kTileElems * 81];
  __shared__ float sFP[3 * B2 * B2];

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const int64_t elem = E0 + e;
  const bool valid = elem < n;

  for (int i = threadIdx.x; i < 3 * B2 * B2; i += blockDim.x) sFP[i] = tables[S::tFP + i];
  // zero padding rows of the face operand
  for (int i = threadIdx.x; i < kCols * (S::KPF - FR); i += blockDim.x) {
    const int col = i / (S::KPF - FR), k = FR + i % (S::KPF - FR);
#pragma unroll
    for (int pl = 0; pl < NS; ++pl) xs[pl * kCols * LDX + col * LDX + k] = static_cast<__bf16>(0.0f);
  }
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    __syncthreads();
    for (int i = threadIdx.x; i < kTileElems * 9 * B2; i += blockDim.x) {
      const int s = i / (9 * B2), r = i % (9 * B2);
      const int64_t el = E0 + s;
      float v = 0.0f;
      if (el < n && (faceTypes == nullptr || faceTypes[el * 4 + f] == SDG_FACE_REGULAR)) {
        const int64_t nb = neighbors[el * 4 + f];
        const int side = min(static_cast<int>(faceRel[(el * 4 + f) * 2 + 0]), 3);
        v = traces[(nb * 4 + side) * 9 * B2 + r];
      }
      sTr[i] = v;
    }
    for (int i = threadIdx.x; i < kTileElems * 81; i += blockDim.x) {
      const int s = i / 81, r = i % 81;
      sAm[i] = (E0 + s < n) ? aminus[(E0 + s) * 324 + f * 81 + r] : 0.0f;
    }
    __syncthreads();
    for (int item = threadIdx.x; item < kTileElems * B2; item += blockDim.x) {
      const int s = item / B2, row = item % B2;
      const int64_t el = E0 + s;
      const bool regular = el < n && (faceTypes == nullptr || faceTypes[el * 4 + f] == SDG_FACE_REGULAR);
      const int h = regular ? min(static_cast<int>(faceRel[(el * 4 + f) * 2 + 1]), 2) : 0;
      float u[9];
#pragma unroll
      for (int qq = 0; qq < 9; ++qq) u[qq] = 0.0f;
      const float* fp = sFP + (h * B2 + row) * B2;
      const float* tr = sTr + s * 9 * B2;
      for (int mm = 0; mm < B2; ++mm) {
        const float c = fp[mm];
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) u[qq] += c * tr[qq * B2 + mm];
      }
      const float* A = sAm + s * 81;
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        float acc = 0.0f;
        if (regular)
#pragma unroll
          for (int qq = 0; qq < 9; ++qq) acc += u[qq] * A[qq + 9 * p];
        storeX1<NS, LDX>(xs, p * 16 + s, f * B2 + row, acc);
      }
    }
  }
  __syncthreads();
  floatx4 acc[1][9];
#pragma unroll
  for (int p = 0; p < 9; ++p) acc[0][p] = floatx4{0, 0, 0, 0};
  const __bf16* a[1] = {frag + S::rdm + static_cast<int64_t>(w) * S::KSF * kFragPair};
  gemmTile<NS, 1, 0, 9, LDX>(acc, a, S::KSF, xs, lane);
  if (valid) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = 16 * w + 4 * g + j;
        if (row < B) dofs[elem * 9 * B + p * B + row] += acc[0][p][j];
      }
  }
}

template <int O, int NS>
int localImpl(const EngineState& st, const double* coeffs, const sdg_layer* L, bool writeIntegrated,
              hipStream_t stream) {
  using S = MShape<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  CoeffsF<O> cf{};
  for (int i = 0; i < O; ++i) cf.c[i] = static_cast<float>(coeffs[i]);
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  kLocalMfma<O, NS><<<grid, S::THREADS, 0, stream>>>(
      n, static_cast<const __bf16*>(st.mfmaFrags), static_cast<const float*>(st.mfmaTables), cf,
      static_cast<float*>(L->dofs), writeIntegrated ? static_cast<float*>(L->time_integrated) : nullptr, static_cast<float*>(L->face_traces),
      static_cast<const float*>(L->star), static_cast<const float*>(L->aplus), L->face_types);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <int O, int NS>
int neighborImpl(const EngineState& st, const sdg_layer* L, hipStream_t stream) {
  using S = MShape<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  kNeighborMfma<O, NS><<<grid, S::THREADS, 0, stream>>>(
      n, static_cast<const __bf16*>(st.mfmaFrags), static_cast<const float*>(st.mfmaTables),
      static_cast<float*>(L->dofs), static_cast<const float*>(L->face_traces), L->face_neighbors, L->face_relations,
      static_cast<const float*>(L->aminus), L->face_types);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

}  // namespace
}  // namespace mfma

int64_t mfmaFragmentCount(int order) {
  switch (order) {
    case 3: return mfma::MShape<3>::total;
    case 4: return mfma::MShape<4>::total;
    case 5: return mfma::MShape<5>::total;
    case 6: return mfma::MShape<6>::total;
    case 7: return mfma::MShape<7>::total;
    default: return 0;
  }
}

int launchLocalMfma(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                    hipStream_t stream) {
  switch (e.order) {
    case 3: return e.arithmetic == SDG_BF16X6 ? mfma::localImpl<3, 3>(e, coeffs, layer, writeIntegrated, stream)
                                          : mfma::localImpl<3, 2>(e, coeffs, layer, writeIntegrated, stream);
    case 4: return e.arithmetic == SDG_BF16X6 ? mfma::localImpl<4, 3>(e, coeffs, layer, writeIntegrated, stream)
                                          : mfma::localImpl<4, 2>(e, coeffs, layer, writeIntegrated, stream);
    case 5: return e.arithmetic == SDG_BF16X6 ? mfma::localImpl<5, 3>(e, coeffs, layer, writeIntegrated, stream)
                                          : mfma::localImpl<5, 2>(e, coeffs, layer, writeIntegrated, stream);
    case 6: return e.arithmetic == SDG_BF16X6 ? mfma::localImpl<6, 3>(e, coeffs, layer, writeIntegrated, stream)
                                          : mfma::localImpl<6, 2>(e, coeffs, layer, writeIntegrated, stream);
    case 7: return e.arithmetic == SDG_BF16X6 ? mfma::localImpl<7, 3>(e, coeffs, layer, writeIntegrated, stream)
                                          : mfma::localImpl<7, 2>(e, coeffs, layer, writeIntegrated, stream);
    default: setError("matrix-core path supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}

int launchNeighborMfma(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  switch (e.order) {
    case 3: return e.arithmetic == SDG_BF16X6 ? mfma::neighborImpl<3, 3>(e, layer, stream)
                                          : mfma::neighborImpl<3, 2>(e, layer, stream);
    case 4: return e.arithmetic == SDG_BF16X6 ? mfma::neighborImpl<4, 3>(e, layer, stream)
                                          : mfma::neighborImpl<4, 2>(e, layer, stream);
    case 5: return e.arithmetic == SDG_BF16X6 ? mfma::neighborImpl<5, 3>(e, layer, stream)
                                          : mfma::neighborImpl<5, 2>(e, layer, stream);
    case 6: return e.arithmetic == SDG_BF16X6 ? mfma::neighborImpl<6, 3>(e, layer, stream)
                                          : mfma::neighborImpl<6, 2>(e, layer, stream);
    case 7: return e.arithmetic == SDG_BF16X6 ? mfma::neighborImpl<7, 3>(e, layer, stream)
                                          : mfma::neighborImpl<7, 2>(e, layer, stream);
    default: setError("matrix-core path supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}

}  // namespace sdg
