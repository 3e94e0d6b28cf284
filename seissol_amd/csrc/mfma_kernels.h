// Kernel templates of the element update on the matrix cores (included by
// the per-order sources kernels_mfma_o<O>.hip): FP32 data, split 16-bit products,
// FP32 accumulation (arithmetics SDG_BF16X3, SDG_BF16X6 and SDG_F16X3).
//
// Every contraction with a global matrix A (kDivMT, kDivM, fMrT, rDivM) is a
// GEMM  Y(M x 144) += A(M x K) X(K x 144)  over a tile of 16 elements, the 144
// columns being (quantity p, element e) -> 16 p + e.  Both operands are split
// into planes of a 16-bit type (x = x0 + x1 [+ x2], each plane the rounding of
// the remainder) and the plane products run as v_mfma_f32_16x16x32_{bf16,f16}:
//   V = 2  (bf16, 2 planes): A0X0 + A0X1 + A1X0                      ~2^-17
//   V = 3  (bf16, 3 planes): + A0X2 + A2X0 + A1X1                     ~2^-24
//   V = 16 (fp16, 2 planes): A0X0 + A0X1 + A1X0                      ~2^-21
// The fp16 form scales every X column (and every stationary matrix) by a
// power of two that brings its largest magnitude into [2^13, 2^14), so the
// 5-bit exponent of fp16 never limits; the scales are exact and are divided
// out of the FP32 results.
//
// One wave owns one 16-row tile of the basis index for all 9 quantities of the
// 16 elements, so a lane holds rows 4g..4g+3 (g = lane / 16) of element
// e = lane % 16 for every quantity: the 9x9 star, AplusT and AminusT products
// that mix quantities are plain register arithmetic.  X lives in LDS
// column-major ([column][k], k contiguous) as 16-bit planes; A is pre-split and
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
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <string>
#include <type_traits>

#include "dpp_rows.h"
#include "engine.h"
#include "mfma_common.h"

namespace sdg {
namespace mfma {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

template <int V>
struct Prec;
template <>
struct Prec<2> {
  using T = __bf16;
  using V8 = bf16x8;
  using V4 = bf16x4;
  static constexpr int NX = 2;
  static constexpr bool scaled = false;
};
template <>
struct Prec<3> {
  using T = __bf16;
  using V8 = bf16x8;
  using V4 = bf16x4;
  static constexpr int NX = 3;
  static constexpr bool scaled = false;
};
template <>
struct Prec<16> {
  using T = _Float16;
  using V8 = halfx8;
  using V4 = halfx4;
  static constexpr int NX = 2;
  static constexpr bool scaled = true;
};

// largest |x| of a scaled column: [2^13, 2^14)
constexpr int kScaleExp = 13;

__device__ constexpr int kStarQ[9][3] = {{6, 7, 8}, {6, 7, 8}, {6, 7, 8}, {6, 7, -1}, {7, 8, -1},
                                         {6, 8, -1}, {0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
__device__ constexpr int kStarOff[9] = {0, 3, 6, 9, 11, 13, 15, 18, 21};
__device__ constexpr int kStarCnt[9] = {3, 3, 3, 2, 2, 2, 3, 3, 3};
// entry z of a compact star matrix: its column-major position q + 9 p
__device__ constexpr int kStarPos[24] = {6 + 0,  7 + 0,  8 + 0,  6 + 9,  7 + 9,  8 + 9,  6 + 18, 7 + 18,
                                         8 + 18, 6 + 27, 7 + 27, 7 + 36, 8 + 36, 6 + 45, 8 + 45, 0 + 54,
                                         3 + 54, 5 + 54, 1 + 63, 3 + 63, 4 + 63, 2 + 72, 4 + 72, 5 + 72};

template <int O>
struct CoeffsF {
  float c[O];
};


// geometry of the two kernels per order and variant
template <int O, int V>
struct Cfg {
  using S = MShape<O>;
  static constexpr int NX = Prec<V>::NX;
  static constexpr bool scaled = Prec<V>::scaled;
  static constexpr int LDX_L = S::KMAX + 8;  // local kernel: padded column stride
  static constexpr int LDS_L = NX * kCols * LDX_L * 2 + kTileElems * 72 * 4 + 9 * 16 * kTileElems * 4 +
                               (scaled ? S::MT * kCols * 4 : 0);
  static constexpr int WAVES_L = (V != 3 && LDS_L <= 81920) ? 2 : 1;
  // neighbour kernel: one face's operand at a time (32 rows, 8 of padding in the stride)
  static constexpr int LDX_N = 40;
  static constexpr int LDS_N = NX * kCols * LDX_N * 2 + kTileElems * 9 * S::B2P * 4 + kTileElems * 81 * 4 +
                               3 * S::B2 * O * 8 + (scaled ? kCols * 4 : 0) + 2 * kTileElems * 4 * 4;
  static constexpr int WGS_N = 163840 / LDS_N;  // workgroups per CU that the LDS allows
  // minimum waves per SIMD asked of the register allocator: three when the LDS
  // lets three or more waves per SIMD run
  static constexpr int WAVES_N = V == 3 ? 1 : (WGS_N * S::MT >= 12 ? 3 : 2);
};

__device__ inline floatx4 mma(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline floatx4 mma(halfx8 a, halfx8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// split product, smallest terms first
template <int V>
__device__ inline floatx4 mfmaSplit(const typename Prec<V>::V8 (&a)[Prec<V>::NX],
                                    const typename Prec<V>::V8 (&b)[Prec<V>::NX], floatx4 c) {
  if constexpr (V == 3) {
    c = mma(a[1], b[1], c);
    c = mma(a[2], b[0], c);
    c = mma(a[0], b[2], c);
  }
  c = mma(a[1], b[0], c);
  c = mma(a[0], b[1], c);
  c = mma(a[0], b[0], c);
  return c;
}

template <int V>
__device__ inline void split(float x, typename Prec<V>::T (&parts)[Prec<V>::NX]) {
  using T = typename Prec<V>::T;
  float rest = x;
#pragma unroll
  for (int i = 0; i < Prec<V>::NX; ++i) {
    parts[i] = static_cast<T>(rest);
    rest -= static_cast<float>(parts[i]);
  }
}

// X planes: NX consecutive arrays of kCols * LDX 16-bit values
template <int V, int LDX>
__device__ inline void storeX4(typename Prec<V>::T* xs, int col, int k0, const float (&v)[4], float scale) {
  using P = Prec<V>;
  typename P::V4 out[P::NX];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    typename P::T parts[P::NX];
    split<V>(v[j] * scale, parts);
#pragma unroll
    for (int i = 0; i < P::NX; ++i) out[i][j] = parts[i];
  }
#pragma unroll
  for (int i = 0; i < P::NX; ++i)
    *reinterpret_cast<typename P::V4*>(xs + i * kCols * LDX + col * LDX + k0) = out[i];
}

template <int V, int LDX>
__device__ inline void storeX1(typename Prec<V>::T* xs, int col, int k, float v) {
  typename Prec<V>::T parts[Prec<V>::NX];
  split<V>(v, parts);
#pragma unroll
  for (int i = 0; i < Prec<V>::NX; ++i) xs[i * kCols * LDX + col * LDX + k] = parts[i];
}

// two consecutive rows k0, k0 + 1 (k0 even) of one column: one 32-bit store per plane
template <int V, int LDX>
__device__ inline void storeX2(typename Prec<V>::T* xs, int col, int k0, float v0, float v1) {
  using P = Prec<V>;
  typedef typename P::T T2 __attribute__((ext_vector_type(2)));
  typename P::T a[P::NX], b[P::NX];
  split<V>(v0, a);
  split<V>(v1, b);
#pragma unroll
  for (int i = 0; i < P::NX; ++i) *reinterpret_cast<T2*>(xs + i * kCols * LDX + col * LDX + k0) = T2{a[i], b[i]};
}

template <int V, int LDX>
__device__ inline void storeZero4(typename Prec<V>::T* xs, int col, int k0) {
  using P = Prec<V>;
  typename P::V4 z;
#pragma unroll
  for (int j = 0; j < 4; ++j) z[j] = static_cast<typename P::T>(0.0f);
#pragma unroll
  for (int i = 0; i < P::NX; ++i) *reinterpret_cast<typename P::V4*>(xs + i * kCols * LDX + col * LDX + k0) = z;
}

// power-of-two scale of a column with largest magnitude cm, and its inverse,
// built from the exponent bits (exponent clamped to [-100, 127])
__device__ inline void columnScale(float cm, float& scale, float& unscale) {
  const int biased = static_cast<int>((__float_as_uint(cm) >> 23) & 0xffu);
  if (cm > 0.0f && biased != 0xff) {
    const int k = max(biased - 127, -100);
    scale = __uint_as_float(static_cast<unsigned>(127 + kScaleExp - k) << 23);
    unscale = __uint_as_float(static_cast<unsigned>(127 - kScaleExp + k) << 23);
  } else {
    scale = 1.0f;
    unscale = 1.0f;
  }
}

// Lane reductions in registers (no LDS round trip); every lane of the wave
// must be active.
// max over the wave's four 16-lane rows (lanes e, e + 16, e + 32, e + 48):
// gfx950's row swaps exchange rows 0-1 with 2-3, then rows 0 with 1 and 2 with 3
__device__ inline float rowsMax4(float m) {
  const unsigned u = __float_as_uint(m);
  const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
  m = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
  const unsigned u2 = __float_as_uint(m);
  const auto b = __builtin_amdgcn_permlane16_swap(u2, u2, false, false);
  return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}
// max over each 16-lane row through DPP: partners lane ^ 1 and lane ^ 2
// (quad_perm), then the mirrored lane within 8 and within 16 lanes
template <int CTRL>
__device__ inline float dppMove(float m) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(m), __float_as_int(m), CTRL, 0xF, 0xF, false));
}
__device__ inline float rowMax16(float m) {
  m = fmaxf(m, dppMove<0xB1>(m));   // quad_perm [1, 0, 3, 2]
  m = fmaxf(m, dppMove<0x4E>(m));   // quad_perm [2, 3, 0, 1]
  m = fmaxf(m, dppMove<0x141>(m));  // row_half_mirror
  return fmaxf(m, dppMove<0x140>(m));  // row_mirror
}

// largest magnitude of each quantity column over the rows this wave holds
// (lanes of one element differ in lane / 16), written to sMax[w][16 q + e]
template <int R>
__device__ inline void waveColumnMax(const float (&v)[R][9][4], bool has, float* sMax, int w, int lane) {
  const int g = lane >> 4, e = lane & 15;
  float m[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    m[q] = 0.0f;
    if (has)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int j = 0; j < 4; ++j) m[q] = fmaxf(m[q], fabsf(v[r][q][j]));
    m[q] = rowsMax4(m[q]);
  }
  if (g == 0)
#pragma unroll
    for (int q = 0; q < 9; ++q) sMax[w * kCols + q * 16 + e] = m[q];
}

// column scales of this lane's element from the maxima of the WAVES waves
template <int WAVES>
__device__ inline void columnScales(const float* sMax, int e, float (&scale)[9], float (&unscale)[9]) {
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    float cm = 0.0f;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) cm = fmaxf(cm, sMax[w * kCols + q * 16 + e]);
    columnScale(cm, scale[q], unscale[q]);
  }
}

template <int V>
__device__ inline typename Prec<V>::V8 loadFrag(const uint16_t* __restrict__ base, int lane) {
  return *reinterpret_cast<const typename Prec<V>::V8*>(base + lane * 8);
}

// acc[jj][i] += A_jj(tile) X(:, quantity Q0 + i)  for JJ matrices with
// fragment bases a[jj] and NQ consecutive quantities
template <int V, int JJ, int Q0, int NQ, int LDX>
__device__ inline void gemmTile(floatx4 (&acc)[JJ][NQ], const uint16_t* const (&a)[JJ], int ksteps,
                                const typename Prec<V>::T* xs, int lane) {
  using P = Prec<V>;
  constexpr int NX = P::NX;
  const int g = lane >> 4, e = lane & 15;
  for (int kk = 0; kk < ksteps; ++kk) {
    typename P::V8 af[JJ][NX];
#pragma unroll
    for (int jj = 0; jj < JJ; ++jj)
#pragma unroll
      for (int i = 0; i < NX; ++i) af[jj][i] = loadFrag<V>(a[jj] + kk * kFragPair + i * kFragPlane, lane);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int off = ((Q0 + q) * 16 + e) * LDX + kk * 32 + 8 * g;
      typename P::V8 bf[NX];
#pragma unroll
      for (int i = 0; i < NX; ++i) bf[i] = *reinterpret_cast<const typename P::V8*>(xs + i * kCols * LDX + off);
#pragma unroll
      for (int jj = 0; jj < JJ; ++jj) acc[jj][q] = mfmaSplit<V>(af[jj], bf, acc[jj][q]);
    }
  }
}

// The lane's four consecutive basis rows row0..row0+3 of the 9 quantities of
// its element at q[p * ld + j] (row0 = 16 w + 4 g): one 16-byte access per
// quantity when V4 (ld a multiple of 4, the block 16-byte aligned), else per
// value; rows from B on (the padding of an aligned layout) are never used.
// V4: qSafe is an in-bounds quad of the same quantity block (the lane's own
// when it is real): every load is issued unconditionally and the value
// selected after, since a load under a branch makes the compiler wait on it
// before the next one.
template <int B, bool V4>
__device__ inline void loadRows(const float* q, const float* qSafe, int ld, int row0, bool valid,
                                float (&v)[9][4]) {
  if constexpr (V4) {
    const bool live = valid && row0 < B;
    float4 raw[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) raw[p] = *reinterpret_cast<const float4*>(qSafe + p * ld);
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      float4 x = live ? raw[p] : float4{0, 0, 0, 0};
      if constexpr (B % 4 != 0) {
        if (row0 + 1 >= B) x.y = 0.0f;
        if (row0 + 2 >= B) x.z = 0.0f;
        if (row0 + 3 >= B) x.w = 0.0f;
      }
      v[p][0] = x.x;
      v[p][1] = x.y;
      v[p][2] = x.z;
      v[p][3] = x.w;
    }
  } else {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) v[p][j] = (valid && row0 + j < B) ? q[p * ld + j] : 0.0f;
  }
}

// q[p * ld + j] += (add ? add[p][j] : 0) + acc[p][j] * s * un[p] for the
// lane's rows below B (16-byte accesses when V4 and the four rows are all real)
template <int B, bool V4>
__device__ inline void updateRows(float* q, int ld, int row0, const floatx4 (&acc)[9], float s, const float (&un)[9],
                                  const float (*add)[4]) {
  if (V4 && (B % 4 == 0 || row0 + 3 < B)) {
    if (row0 < B)
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        float4* d = reinterpret_cast<float4*>(q + p * ld);
        float4 x = *d;
        const float sc = s * un[p];
        x.x += (add ? add[p][0] : 0.0f) + acc[p][0] * sc;
        x.y += (add ? add[p][1] : 0.0f) + acc[p][1] * sc;
        x.z += (add ? add[p][2] : 0.0f) + acc[p][2] * sc;
        x.w += (add ? add[p][3] : 0.0f) + acc[p][3] * sc;
        *d = x;
      }
  } else {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (row0 + j < B) q[p * ld + j] += (add ? add[p][j] : 0.0f) + acc[p][j] * (s * un[p]);
  }
}

// Y += sum_j (A_j X) star_j for the lane's element, with the matrix and
// column scales divided out (the star values in LDS carry the CK matrices'
// scales; kFactor: times aUn[j] as well).  The products run in three passes over input
// quantity groups {6,7,8}, {0,1,2}, {3,4,5}, each holding 3 x 3 accumulator
// tiles (the elastic star matrix maps velocities (6..8) to stresses (0..5)
// and stresses to velocities); the three matrices' fragments of a k-step are
// loaded inside the pass (L2-resident), which keeps them out of the live set.
template <int V, int KSMAX, int LDX, bool kFactor>
__device__ inline void gemmStars(const uint16_t* const (&a)[3], int ks, const typename Prec<V>::T* xs, int lane,
                                 const float* sS, const float (&aUn)[3], const float (&colUn)[9], float (&y)[9][4],
                                 int passMask = 7) {
  using P = Prec<V>;
  constexpr int NX = P::NX;
  const int g = lane >> 4, e = lane & 15;
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
    if (((passMask >> pass) & 1) == 0) continue;  // wave-uniform
    const int q0 = pass == 0 ? 6 : (pass == 1 ? 0 : 3);
    floatx4 acc[3][3];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int i = 0; i < 3; ++i) acc[jj][i] = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int kk = 0; kk < KSMAX; ++kk)
      if (kk < ks) {
        typename P::V8 af[1][3][NX];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj)
#pragma unroll
          for (int i = 0; i < NX; ++i) af[0][jj][i] = loadFrag<V>(a[jj] + kk * kFragPair + i * kFragPlane, lane);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          const int off = ((q0 + i) * 16 + e) * LDX + kk * 32 + 8 * g;
          typename P::V8 bf[NX];
#pragma unroll
          for (int pl = 0; pl < NX; ++pl) bf[pl] = *reinterpret_cast<const typename P::V8*>(xs + pl * kCols * LDX + off);
#pragma unroll
          for (int jj = 0; jj < 3; ++jj) acc[jj][i] = mfmaSplit<V>(af[0][jj], bf, acc[jj][i]);
        }
      }
    // contraction with the star matrices for the outputs this group feeds
#pragma unroll
    for (int jj = 0; jj < 3; ++jj)
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (k < kStarCnt[p]) {
            const int qi = kStarQ[p][k];
            if (qi >= q0 && qi < q0 + 3) {
              const float sv = kFactor ? sS[jj * 24 + kStarOff[p] + k] * aUn[jj] * colUn[qi]
                                       : sS[jj * 24 + kStarOff[p] + k] * colUn[qi];
#pragma unroll
              for (int j = 0; j < 4; ++j) y[p][j] += acc[jj][qi - q0][j] * sv;
            }
          }
  }
}

// DEF: the arrays have the default layout (DOF blocks [9][B], operators 243 /
// 72, 324 reals apart), so every stride is a compile-time constant
template <int O, int V, bool V4, bool LTS, bool DEF>
__global__ __launch_bounds__(MShape<O>::THREADS, (Cfg<O, V>::WAVES_L)) void kLocalMfma(
    int64_t n, const uint16_t* __restrict__ frag, const float* __restrict__ tables, CoeffsF<O> cf,
    float* __restrict__ dofs, float* __restrict__ integ, float* __restrict__ traces,
    const float* __restrict__ star, const float* __restrict__ aplus, const uint8_t* __restrict__ faceTypes,
    CellLayout lay, LtsOut lts) {
  using S = MShape<O>;
  using C = Cfg<O, V>;
  using T = typename Prec<V>::T;
  constexpr int B = S::B, B2 = S::B2, MT = S::MT, LDX = C::LDX_L, FR = S::FR, NX = C::NX;
  constexpr bool scaled = C::scaled;
  constexpr int TPW = cdiv(S::MTF, MT);
  __shared__ __attribute__((aligned(16))) T xs[NX * kCols * LDX];
  __shared__ float sStar[kTileElems * 72];
  __shared__ float sStage[3 * 3 * 16 * kTileElems];  // CK partial sums handed between waves (3 slots)
  __shared__ float sMax[scaled ? MT * kCols : 1];

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));  // wave-uniform
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const int64_t elem = E0 + e;
  const bool valid = elem < n;
  // per-lane addresses as 32-bit offsets from uniform tile bases (one SGPR
  // pair each) rather than 36 live 64-bit addresses
  const int ldq = DEF ? B : lay.ldq;
  const int64_t qStride = DEF ? 9 * B : lay.qStride;
  float* const tileDofs = dofs + E0 * qStride;
  const int laneQ = e * static_cast<int>(qStride) + 16 * w + 4 * g;
  float* const tileTraces = traces + E0 * 36 * S::B2P;
  const float* aUn = tables + S::tAUn;
  const float aUnCk[3] = {aUn[0], aUn[1], aUn[2]};
  // the volume products' factor relative to the staged star values (powers
  // of two; 1 when kDivM's fragments carry the scales of kDivMT's)
  const float aUnVol[3] = {aUn[3] / aUn[0], aUn[4] / aUn[1], aUn[5] / aUn[2]};
  const float aUnF = aUn[6], aUnR = aUn[7];

  {
    // the tile's compact star matrices, every load of a thread issued before
    // the first store (clamped addresses, in bounds: a launched tile has E0 < n)
    constexpr int SI = cdiv(kTileElems * 72, S::THREADS);
    float sv[SI];
    const int64_t starStride = DEF ? (lay.starCompact ? 72 : 243) : lay.starStride;
    if (lay.starCompact) {  // wave-uniform: the compact format needs no position table
#pragma unroll
      for (int u = 0; u < SI; ++u) {
        const int i = threadIdx.x + u * S::THREADS;
        const int s = min(i / 72, kTileElems - 1), r = i % 72;
        const int64_t el = E0 + s < n ? E0 + s : n - 1;
        sv[u] = star[el * starStride + r];
      }
    } else {
#pragma unroll
      for (int u = 0; u < SI; ++u) {
        const int i = threadIdx.x + u * S::THREADS;
        const int s = min(i / 72, kTileElems - 1), r = i % 72;
        const int64_t el = E0 + s < n ? E0 + s : n - 1;
        sv[u] = star[el * starStride + 81 * (r / 24) + kStarPos[r % 24]];
      }
    }
    // staged times the CK matrices' unscale factor of their direction (a
    // power of two: exact), which the CK products then need not apply
#pragma unroll
    for (int u = 0; u < SI; ++u) {
      const int i = threadIdx.x + u * S::THREADS;
      const int jd = (i % 72) / 24;
      const float f = jd == 0 ? aUnCk[0] : (jd == 1 ? aUnCk[1] : aUnCk[2]);
      if (i < kTileElems * 72) sStar[i] = E0 + i / 72 < n ? sv[u] * f : 0.0f;
    }
  }

  // Q itself is read again at the end, so only the time integral stays live
  float q[1][9][4], iacc[1][9][4];
  {
    // in-bounds quad for lanes without real rows: the last quad of a real
    // element's block (V4: ldq a multiple of 4)
    const int64_t safeElem = valid ? elem : n - 1;
    const int safeRow = V4 ? min(16 * w + 4 * g, (B - 1) / 4 * 4) : 0;
    loadRows<B, V4>(tileDofs + laneQ, dofs + safeElem * qStride + safeRow, ldq, 16 * w + 4 * g, valid, q[0]);
  }
  // LTS: which of buffer and derivatives this cell provides
  [[maybe_unused]] unsigned setup = 0;
  [[maybe_unused]] bool storeDeriv = false;
  [[maybe_unused]] float* cellDeriv = nullptr;
  if constexpr (LTS) {
    setup = (lts.setup != nullptr && valid) ? lts.setup[elem] : 0u;
    storeDeriv = lts.deriv != nullptr && valid && (lts.setup == nullptr || (setup & SDG_LTS_HAS_DERIVATIVES));
    cellDeriv = lts.deriv + (valid ? elem : 0) * lay.dsize;
    if (storeDeriv)
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (16 * w + 4 * g + j < B) cellDeriv[p * ldq + 16 * w + 4 * g + j] = q[0][p][j];
  }
  [[maybe_unused]] int64_t doff = 9LL * ldq;
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) iacc[0][p][j] = cf.c[0] * q[0][p][j];
  // X <- Q
  float cs[9], un[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) cs[p] = un[p] = 1.0f;
  if constexpr (scaled) {
    waveColumnMax<1>(q, true, sMax, w, lane);
    __syncthreads();
    columnScales<MT>(sMax, e, cs, un);
  }
#pragma unroll
  for (int p = 0; p < 9; ++p) storeX4<V, LDX>(xs, p * 16 + e, 16 * w + 4 * g, q[0][p], cs[p]);
  for (int t = MT + w; t < S::KPB / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<V, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();
  const float* sS = sStar + e * 72;

  // Cauchy-Kowalevski recursion.  Wave w owns m-tile w of dQ_d.  When dQ_d
  // has fewer m-tiles than there are waves, idle waves take over velocity
  // passes (input groups {0,1,2}, {3,4,5}) of the owned tiles and hand their
  // partial sums back through LDS: mode 2 = passes 1 and 2 of a single tile, mode 1 = pass 2 of
  // every tile, 0 = none.
  constexpr int kSlotFloats = 3 * 16 * kTileElems;  // 3 velocities x 16 rows x 16 elements
  constexpr int kSlots = 3;
  float* const stage = sStage;
#pragma unroll 1
  for (int d = 1; d < O; ++d) {
    const int m = basis3(O - d), nIn = basis3(O - d + 1);
    const int mtiles = cdiv(m, 16), ks = cdiv(nIn, 32), kpNext = cdiv(m, 32) * 32;
    const int helpers = MT - mtiles;
    const int mode = (mtiles == 1 && helpers >= 2 && kSlots >= 2) ? 2 : ((helpers >= mtiles && mtiles <= kSlots) ? 1 : 0);
    int tile = -1, mask = 0, slot = -1;
    if (w < mtiles) {
      tile = w;
      mask = mode == 2 ? 1 : (mode == 1 ? 3 : 7);
    } else if (mode == 2 && w - mtiles < 2) {
      tile = 0;
      mask = 2 << (w - mtiles);
      slot = w - mtiles;
    } else if (mode == 1 && w - mtiles < mtiles) {
      tile = w - mtiles;
      mask = 4;
      slot = tile;
    }
    float dq[1][9][4];
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j) dq[0][p][j] = 0.0f;
    if (tile >= 0) {
      const uint16_t* a[3] = {frag + S::ck + (0LL * MT + tile) * S::KSB * kFragPair,
                              frag + S::ck + (1LL * MT + tile) * S::KSB * kFragPair,
                              frag + S::ck + (2LL * MT + tile) * S::KSB * kFragPair};
      gemmStars<V, S::KSB, LDX, false>(a, ks, xs, lane, sS, aUnCk, un, dq[0], mask);
    }
    if (mode != 0) {
      if (slot >= 0)
#pragma unroll
        for (int p = 6; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j) stage[slot * kSlotFloats + (p - 6) * 256 + (4 * g + j) * 16 + e] = dq[0][p][j];
      __syncthreads();
      if (w < mtiles) {
        const int s0 = mode == 2 ? 0 : w, ns = mode == 2 ? 2 : 1;
        for (int sl = s0; sl < s0 + ns; ++sl)
#pragma unroll
          for (int p = 6; p < 9; ++p)
#pragma unroll
            for (int j = 0; j < 4; ++j) dq[0][p][j] += stage[sl * kSlotFloats + (p - 6) * 256 + (4 * g + j) * 16 + e];
      }
    }
    if (w < mtiles) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (16 * w + 4 * g + j >= m)
#pragma unroll
          for (int p = 0; p < 9; ++p) dq[0][p][j] = 0.0f;
    }
    if constexpr (LTS) {
      // derivative block d: [9][rows(m)] after the blocks before it
      const int ldm = lay.rows(m);
      if (storeDeriv && w < mtiles)
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (16 * w + 4 * g + j < m) cellDeriv[doff + p * ldm + 16 * w + 4 * g + j] = dq[0][p][j];
      doff += 9LL * ldm;
    }
    if constexpr (scaled) waveColumnMax<1>(dq, w < mtiles, sMax, w, lane);
    __syncthreads();
    if constexpr (scaled) columnScales<MT>(sMax, e, cs, un);
    if (w < mtiles) {
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        storeX4<V, LDX>(xs, p * 16 + e, 16 * w + 4 * g, dq[0][p], cs[p]);
#pragma unroll
        for (int j = 0; j < 4; ++j) iacc[0][p][j] += cf.c[d] * dq[0][p][j];
      }
    }
    for (int t = w; t < kpNext / 16; t += MT)
      if (t >= mtiles)
        for (int p = 0; p < 9; ++p) storeZero4<V, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
    __syncthreads();
  }

  // time-integrated DOFs: optional output (LTS: the buffer, stored or
  // accumulated), and the operand of the next GEMMs
  bool storeBuffer = integ != nullptr && valid;
  [[maybe_unused]] bool accumulate = false;
  if constexpr (LTS) {
    storeBuffer = storeBuffer && (lts.setup == nullptr || (setup & SDG_LTS_HAS_BUFFERS));
    accumulate = lts.setup != nullptr && !lts.reset && (setup & SDG_LTS_ACCUMULATE_BUFFERS);
  }
  if (storeBuffer) {
    float* const elemInteg = integ + elem * qStride;
    const int row0 = 16 * w + 4 * g;
    // accumulated buffers: all old values loaded first (clamped rows, in
    // bounds), then stored, so that the loads are in flight together
    float old[9][4] = {};
    if constexpr (LTS) {
      if (accumulate)
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
          for (int j = 0; j < 4; ++j) old[p][j] = elemInteg[p * ldq + min(row0 + j, B - 1)];
    }
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (row0 + j < B) elemInteg[p * ldq + row0 + j] = old[p][j] + iacc[0][p][j];
  }
  if constexpr (scaled) {
    waveColumnMax<1>(iacc, true, sMax, w, lane);
    __syncthreads();
    columnScales<MT>(sMax, e, cs, un);
  }
#pragma unroll
  for (int p = 0; p < 9; ++p) storeX4<V, LDX>(xs, p * 16 + e, 16 * w + 4 * g, iacc[0][p], cs[p]);
  for (int t = MT + w; t < S::KPB / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<V, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();

  // face traces rT_f I = (fMrT_f / M2) I (rows f * B2P + m): the fragments
  // carry 1/M2, and the local flux's rDivM fragments carry M2
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
      const uint16_t* a[1] = {frag + S::fmrt + static_cast<int64_t>(t) * S::KSB * kFragPair};
      if (E0 < n) gemmTile<V, 1, 0, 9, LDX>(acc, a, S::KSB, xs, lane);
#pragma unroll
      for (int p = 0; p < 9; ++p)
#pragma unroll
        for (int j = 0; j < 4; ++j) tr[ti][p][j] = acc[0][p][j] * (aUnF * un[p]);
      // the traces: the lane's four rows lie in one
      // face (faces padded to B2P rows, a multiple of 4), one 16-byte store
      // per quantity into the padded layout [element][face][quantity][B2P]
      const int r0 = 16 * t + 4 * g;
      if (valid && r0 < FR && (!LTS || traces != nullptr)) {
        const int f = r0 / S::B2P, mm = r0 % S::B2P;
#pragma unroll
        for (int p = 0; p < 9; ++p)
          *reinterpret_cast<float4*>(tileTraces + (e * 36 + f * 9 + p) * S::B2P + mm) =
              float4{tr[ti][p][0], tr[ti][p][1], tr[ti][p][2], tr[ti][p][3]};
      }
    }
  }

  // volume: Q += sum_j kDivM_j I star_j (X holds I)
  float vol[9][4];
#pragma unroll
  for (int p = 0; p < 9; ++p)
#pragma unroll
    for (int j = 0; j < 4; ++j) vol[p][j] = 0.0f;
  {
    const uint16_t* a[3] = {frag + S::vol + (0LL * MT + w) * S::KSB * kFragPair,
                            frag + S::vol + (1LL * MT + w) * S::KSB * kFragPair,
                            frag + S::vol + (2LL * MT + w) * S::KSB * kFragPair};
    // `E0 < n` holds for every launched tile; the guard gives this GEMM (and
    // the trace GEMM above) a block of its own, with which the register
    // allocator keeps the kernel free of spills
    if (E0 < n) gemmStars<V, S::KSB, LDX, true>(a, S::KSB, xs, lane, sS, aUnVol, un, vol);
  }

  // Once every wave has finished the volume GEMM, the X area is free until W
  // is stored into it: it takes the four faces' AplusT and (when it fits) the
  // volume term, so that W = (rT_f I) AplusT_f is formed by every lane at once with
  // its own face's matrix (a lane's four rows of a tile lie in one face).
  constexpr int kXBytes = NX * kCols * LDX * 2;
  constexpr bool kParkVol = (kTileElems * 324 + kTileElems * 9 * B) * 4 <= kXBytes;
  static_assert(kTileElems * 324 * 4 <= kXBytes, "the X area holds the tile's AplusT");
  float* const sA4 = reinterpret_cast<float*>(xs);
  float* const sVol = sA4 + kTileElems * 324;  // [element][quantity][B]
  constexpr int APT = cdiv(kTileElems * 324, S::THREADS);  // AplusT of the four faces, [element][face][81]
  float apv[APT];
#pragma unroll
  for (int i = 0; i < APT; ++i) {
    const int idx = threadIdx.x + i * S::THREADS;
    const int64_t at = DEF ? E0 * 324 + idx : (E0 + idx / 324) * lay.aplusStride + idx % 324;
    apv[i] = (idx < kTileElems * 324 && E0 + idx / 324 < n) ? aplus[at] : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < APT; ++i) {
    const int idx = threadIdx.x + i * S::THREADS;
    if (idx < kTileElems * 324) sA4[idx] = apv[i];
  }
  if constexpr (kParkVol) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (16 * w + 4 * g + j < B) sVol[(e * 9 + p) * B + 16 * w + 4 * g + j] = vol[p][j];
  }
  __syncthreads();
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti) {
    const int t = w + ti * MT;
    const int r0 = 16 * t + 4 * g;
    if (t < S::MTF && r0 < FR) {
      const int f = r0 / S::B2P;
      const bool dr = valid && faceTypes != nullptr && faceTypes[elem * 4 + f] == SDG_FACE_DYNAMIC_RUPTURE;
      const float* A = sA4 + e * 324 + f * 81;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float wrow[9];
#pragma unroll
        for (int p = 0; p < 9; ++p) {
          float acc = 0.0f;
#pragma unroll
          for (int qq = 0; qq < 9; ++qq) acc += tr[ti][qq][j] * A[qq + 9 * p];
          wrow[p] = dr ? 0.0f : acc;
        }
#pragma unroll
        for (int p = 0; p < 9; ++p) tr[ti][p][j] = wrow[p];
      }
    }
  }
  if constexpr (kParkVol) {
#pragma unroll
    for (int p = 0; p < 9; ++p)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        vol[p][j] = 16 * w + 4 * g + j < B ? sVol[(e * 9 + p) * B + 16 * w + 4 * g + j] : 0.0f;
  }

  if constexpr (scaled) waveColumnMax<TPW>(tr, true, sMax, w, lane);
  __syncthreads();
  if constexpr (scaled) columnScales<MT>(sMax, e, cs, un);
#pragma unroll
  for (int ti = 0; ti < TPW; ++ti) {
    const int t = w + ti * MT;
    if (t < S::MTF)
#pragma unroll
      for (int p = 0; p < 9; ++p) storeX4<V, LDX>(xs, p * 16 + e, 16 * t + 4 * g, tr[ti][p], cs[p]);
  }
  for (int t = S::MTF + w; t < S::KPF / 16; t += MT)
    for (int p = 0; p < 9; ++p) storeZero4<V, LDX>(xs, p * 16 + e, 16 * t + 4 * g);
  __syncthreads();

  // local flux: Q += [rDivM_0 .. rDivM_3] W, the accumulators starting from the
  // volume term divided by the (power-of-two) unscale factors, so that the
  // update needs no separately live volume registers
  {
    floatx4 acc[1][9];
#pragma unroll
    for (int p = 0; p < 9; ++p) {
      const float inv = 1.0f / (aUnR * un[p]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[0][p][j] = vol[p][j] * inv;
    }
    const uint16_t* a[1] = {frag + S::rdm + static_cast<int64_t>(w) * S::KSF * kFragPair};
    gemmTile<V, 1, 0, 9, LDX>(acc, a, S::KSF, xs, lane);
    if (valid) updateRows<B, V4>(tileDofs + laneQ, ldq, 16 * w + 4 * g, acc[0], aUnR, un, nullptr);
  }
}

template <int O, int V, bool V4, bool DEF>
__global__ __launch_bounds__(MShape<O>::THREADS, (Cfg<O, V>::WAVES_N)) void kNeighborMfma(
    int64_t n, const uint16_t* __restrict__ frag, const float* __restrict__ tables, float* __restrict__ dofs,
    const float* __restrict__ traces, const int32_t* __restrict__ neighbors, const uint8_t* __restrict__ faceRel,
    const float* __restrict__ aminus, const uint8_t* __restrict__ faceTypes, CellLayout lay) {
  using S = MShape<O>;
  using C = Cfg<O, V>;
  using T = typename Prec<V>::T;
  constexpr int B = S::B, B2 = S::B2, B2P = S::B2P, LDX = C::LDX_N, NX = C::NX;
  constexpr bool scaled = C::scaled;
  constexpr int HALF = B2P / 2;                  // row pairs of one (padded) face
  static_assert(HALF <= 16 && B2P <= 32 && LDX >= 32, "a face's row pairs fit a 16-lane group, its rows a k-step");
  constexpr int GROUPS = S::THREADS / 16;        // 16-lane groups per pass, one element each
  constexpr int PASSES = cdiv(kTileElems, GROUPS);
  __shared__ __attribute__((aligned(16))) T xs[NX * kCols * LDX];  // face operand: rows [0, B2P) of one face, the rest zero
  __shared__ __attribute__((aligned(16))) float sTr[kTileElems * 9 * B2P];
  __shared__ float sAm[kTileElems * 81];
  __shared__ float sFP[3 * B2 * O];   // fP_h rows: O (value, column) entries each
  __shared__ int sFPCol[3 * B2 * O];
  __shared__ float sUn[scaled ? kCols : 1];  // unscale factor of each operand column of the current face
  __shared__ int sLink[kTileElems * 4];  // trace index nb * 4 + side of (element, face), -1: none
  __shared__ int sRel[kTileElems * 4];   // orientation h of (element, face)
  constexpr int TR = 9 * B2P;                                  // floats of one (padded) face trace
  constexpr int TR4 = TR / 4;                                  // 16-byte pieces of one face trace
  constexpr int GPT = cdiv(kTileElems * TR4, S::THREADS);      // gathered pieces per thread
  constexpr int APT = cdiv(kTileElems * 81, S::THREADS);       // AminusT floats per thread

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));  // wave-uniform
  const int g = lane >> 4, e = lane & 15;
  const int64_t E0 = static_cast<int64_t>(blockIdx.x) * kTileElems;
  const bool valid = E0 + e < n;
  const int ldq = DEF ? B : lay.ldq;
  const int64_t qStride = DEF ? 9 * B : lay.qStride;
  float* const tileDofs = dofs + E0 * qStride;
  const int laneQ = e * static_cast<int>(qStride) + 16 * w + 4 * g;

  // the fP tables and the neighbour links of the four faces, read once; every
  // load is issued unconditionally from an in-bounds (clamped) address and its
  // value selected after, so that all of them are in flight together (a load
  // under a branch makes the compiler wait on it before the next)
  {
    constexpr int FPN = 3 * B2 * O, FPI = cdiv(FPN, S::THREADS);
    float fv[FPI], fc[FPI];
#pragma unroll
    for (int u = 0; u < FPI; ++u) {
      const int i = min(static_cast<int>(threadIdx.x) + u * S::THREADS, FPN - 1);
      fv[u] = tables[S::tFP + i];
      fc[u] = tables[S::tFPCol + i];
    }
    static_assert(kTileElems * 4 <= S::THREADS, "one thread per (element, face) link");
    const int s = threadIdx.x >> 2 & (kTileElems - 1), f = threadIdx.x & 3;
    const int64_t el = E0 + s < n ? E0 + s : n - 1;
    const int type = faceTypes != nullptr ? faceTypes[el * 4 + f] : SDG_FACE_REGULAR;
    const int nb = neighbors[el * 4 + f];
    const int side = faceRel[(el * 4 + f) * 2 + 0], rel = faceRel[(el * 4 + f) * 2 + 1];
#pragma unroll
    for (int u = 0; u < FPI; ++u) {
      const int i = threadIdx.x + u * S::THREADS;
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
  // rows [B2P, 32) of the operand stay zero (rows [B2, B2P) are written as zeros by the items)
  if constexpr (B2P < 32) {
    for (int i = threadIdx.x; i < kCols * (32 - B2P); i += blockDim.x) {
      const int col = i / (32 - B2P), k = B2P + i % (32 - B2P);
#pragma unroll
      for (int pl = 0; pl < NX; ++pl) xs[pl * kCols * LDX + col * LDX + k] = static_cast<T>(0.0f);
    }
  }
  __syncthreads();
  // the next face's traces and AminusT are loaded into registers while the
  // current face is computed
  float4 gv[GPT];
  float av[APT];
  auto gather = [&](int f) {
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int idx = threadIdx.x + i * S::THREADS;
      float4 v{0.0f, 0.0f, 0.0f, 0.0f};
      if (idx < kTileElems * TR4) {
        const int link = sLink[(idx / TR4) * 4 + f];
        if (link >= 0) v = *reinterpret_cast<const float4*>(traces + static_cast<int64_t>(link) * TR + 4 * (idx % TR4));
      }
      gv[i] = v;
    }
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = threadIdx.x + i * S::THREADS;
      const int64_t at = (E0 + idx / 81) * (DEF ? 324 : lay.aminusStride) + f * 81 + idx % 81;
      av[i] = (idx < kTileElems * 81 && E0 + idx / 81 < n) ? aminus[at] : 0.0f;
    }
  };
  gather(0);
  floatx4 tot[1][9];
#pragma unroll
  for (int p = 0; p < 9; ++p) tot[0][p] = floatx4{0, 0, 0, 0};
  // face by face: W_f = (fP_h Tr_f) AminusT_f into the operand (the fp16 form
  // scales each column by a power of two), then Q += rDivM_f W_f as one k-step
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    // the previous face's items have read sTr / sAm (barrier before its product)
#pragma unroll
    for (int i = 0; i < GPT; ++i) {
      const int idx = threadIdx.x + i * S::THREADS;
      if (idx < kTileElems * TR4) *reinterpret_cast<float4*>(sTr + 4 * idx) = gv[i];
    }
#pragma unroll
    for (int i = 0; i < APT; ++i) {
      const int idx = threadIdx.x + i * S::THREADS;
      if (idx < kTileElems * 81) sAm[idx] = av[i];
    }
    if (f < 3) gather(f + 1);
    __syncthreads();  // also: every wave's product of the previous face has read xs and sUn
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
      const float* tr = sTr + sl * TR;
      float u[2][9];
#pragma unroll
      for (int ri = 0; ri < 2; ++ri) {
#pragma unroll
        for (int qq = 0; qq < 9; ++qq) u[ri][qq] = 0.0f;
        const int row = r0 + ri;
        if (regular && row < B2) {
          const float* fp = sFP + (h * B2 + row) * O;
          const int* fc = sFPCol + (h * B2 + row) * O;
#pragma unroll
          for (int kk = 0; kk < O; ++kk) {
            const float c = fp[kk];
            const int mm = fc[kk];
#pragma unroll
            for (int qq = 0; qq < 9; ++qq) u[ri][qq] += c * tr[qq * B2P + mm];
          }
        }
      }
      // W = U AminusT: rowA[p] holds A(min(rp, 8), p), broadcast across the
      // element's 16 lanes (DPP row_newbcast) instead of read per (q, p)
      const float* A = sAm + sl * 81;
      float rowA[9], w0[9], w1[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        rowA[p] = A[min(rp, 8) + 9 * p];
        w0[p] = 0.0f;
        w1[p] = 0.0f;
      }
      staticFor<9>([&](auto qI) {
        constexpr int qq = decltype(qI)::value;
        fmacRowBcast9<qq>(w0, rowA, u[0][qq]);
        fmacRowBcast9<qq>(w1, rowA, u[1][qq]);
      });
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        w0[p] = regular ? w0[p] : 0.0f;
        w1[p] = regular ? w1[p] : 0.0f;
      }
      if constexpr (scaled) {
        // column maxima over the 16 lanes of the element (idle lanes hold zeros)
#pragma unroll
        for (int p = 0; p < 9; ++p) {
          const float m = rowMax16(fmaxf(fabsf(w0[p]), fabsf(w1[p])));
          float sc, inv;
          columnScale(m, sc, inv);
          if (act) storeX2<V, LDX>(xs, p * 16 + s, r0, w0[p] * sc, w1[p] * sc);
          if (s < kTileElems && rp == 0) sUn[p * 16 + s] = inv;
        }
      } else {
#pragma unroll
        for (int p = 0; p < 9; ++p)
          if (act) storeX2<V, LDX>(xs, p * 16 + s, r0, w0[p], w1[p]);
      }
    }
    __syncthreads();
    const uint16_t* a[1] = {frag + S::rdmF + (static_cast<int64_t>(w) * 4 + f) * kFragPair};
    if constexpr (scaled) {
      floatx4 acc[1][9];
#pragma unroll
      for (int p = 0; p < 9; ++p) acc[0][p] = floatx4{0, 0, 0, 0};
      gemmTile<V, 1, 0, 9, LDX>(acc, a, 1, xs, lane);
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        const float un = sUn[p * 16 + e];
#pragma unroll
        for (int j = 0; j < 4; ++j) tot[0][p][j] += acc[0][p][j] * un;
      }
    } else {
      gemmTile<V, 1, 0, 9, LDX>(tot, a, 1, xs, lane);
    }
  }
  float one[9];
#pragma unroll
  for (int p = 0; p < 9; ++p) one[p] = 1.0f;
  const float aUnR = tables[S::tAUn + 8];
  if (valid) updateRows<B, V4>(tileDofs + laneQ, ldq, 16 * w + 4 * g, tot[0], aUnR, one, nullptr);
}

// the arrays have the default layout: DOF blocks [9][B] back to back, the
// operators 243 (72 compact) and 324 reals apart
template <int B>
bool defaultLayout(const CellLayout& lay) {
  return lay.ldq == B && lay.qStride == 9 * B && lay.aplusStride == 324 && lay.aminusStride == 324 &&
         lay.starStride == (lay.starCompact ? 72 : 243);
}

// launch(vec, def) with compile-time flags: 16-byte DOF accesses (always for B
// a multiple of 4) and the default layout
template <int B, typename F>
void launchVariants(bool v4, bool def, F&& launch) {
  if constexpr (B % 4 != 0) {
    if (!v4) {
      if (def) launch(std::false_type{}, std::true_type{});
      else launch(std::false_type{}, std::false_type{});
      return;
    }
  }
  if (def) launch(std::true_type{}, std::true_type{});
  else launch(std::true_type{}, std::false_type{});
}

// 16-byte DOF accesses: every block starts on a multiple of 4 reals and the
// arrays are 16-byte aligned (always so for B a multiple of 4 on aligned arrays)
inline bool vectorAccess(const CellLayout& lay, const void* dofs, const void* integ, const char* who, bool& v4) {
  const bool aligned = (reinterpret_cast<uintptr_t>(dofs) % 16 == 0) &&
                       (integ == nullptr || reinterpret_cast<uintptr_t>(integ) % 16 == 0);
  v4 = lay.ldq % 4 == 0 && aligned;
  if (lay.ldq % 4 == 0 && !aligned) {
    setError(std::string(who) + ": dofs and time_integrated must be 16-byte aligned");
    return false;
  }
  return true;
}

// lts == nullptr: the proxy/GTS local pass (traces written, derivatives never);
// otherwise the LTS local pass: buffers and derivatives per the cells' setup,
// traces written only when the layer has them
template <int O, int V>
int localImpl(const EngineState& st, const double* coeffs, const sdg_layer* L, bool writeIntegrated,
              hipStream_t stream, const LtsOut* lts = nullptr) {
  using S = MShape<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  CoeffsF<O> cf{};
  for (int i = 0; i < O; ++i) cf.c[i] = static_cast<float>(coeffs[i]);
  const CellLayout lay = cellLayout(st, L);
  float* integ = writeIntegrated ? static_cast<float*>(L->time_integrated) : nullptr;
  bool v4 = false;
  if (!vectorAccess(lay, L->dofs, integ, "matrix-core local pass", v4)) return SDG_ERR_INVALID;
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  auto launch = [&](auto vec, auto def) {
    constexpr bool kV4 = decltype(vec)::value, kDef = decltype(def)::value;
    if (lts != nullptr)
      kLocalMfma<O, V, kV4, true, kDef><<<grid, S::THREADS, 0, stream>>>(
          n, static_cast<const uint16_t*>(st.mfmaFrags), static_cast<const float*>(st.mfmaTables), cf,
          static_cast<float*>(L->dofs), integ, static_cast<float*>(L->face_traces),
          static_cast<const float*>(L->star), static_cast<const float*>(L->aplus), L->face_types, lay, *lts);
    else
      kLocalMfma<O, V, kV4, false, kDef><<<grid, S::THREADS, 0, stream>>>(
          n, static_cast<const uint16_t*>(st.mfmaFrags), static_cast<const float*>(st.mfmaTables), cf,
          static_cast<float*>(L->dofs), integ, static_cast<float*>(L->face_traces),
          static_cast<const float*>(L->star), static_cast<const float*>(L->aplus), L->face_types, lay, LtsOut{});
  };
  launchVariants<S::B>(v4, defaultLayout<S::B>(lay), launch);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <int O, int V>
int neighborImpl(const EngineState& st, const sdg_layer* L, hipStream_t stream) {
  using S = MShape<O>;
  const int64_t n = L->num_elements;
  if (n == 0) return SDG_OK;
  const CellLayout lay = cellLayout(st, L);
  bool v4 = false;
  if (!vectorAccess(lay, L->dofs, nullptr, "matrix-core neighbour pass", v4)) return SDG_ERR_INVALID;
  const dim3 grid(static_cast<unsigned>((n + kTileElems - 1) / kTileElems));
  auto launch = [&](auto vec, auto def) {
    kNeighborMfma<O, V, decltype(vec)::value, decltype(def)::value><<<grid, S::THREADS, 0, stream>>>(
        n, static_cast<const uint16_t*>(st.mfmaFrags), static_cast<const float*>(st.mfmaTables),
        static_cast<float*>(L->dofs), static_cast<const float*>(L->face_traces), L->face_neighbors,
        L->face_relations, static_cast<const float*>(L->aminus), L->face_types, lay);
  };
  launchVariants<S::B>(v4, defaultLayout<S::B>(lay), launch);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <int O>
int localDispatch(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                  hipStream_t stream, const LtsOut* lts) {
  switch (e.arithmetic) {
    case SDG_BF16X6: return localImpl<O, 3>(e, coeffs, layer, writeIntegrated, stream, lts);
    case SDG_F16X3: return localImpl<O, 16>(e, coeffs, layer, writeIntegrated, stream, lts);
    default: return localImpl<O, 2>(e, coeffs, layer, writeIntegrated, stream, lts);
  }
}

template <int O>
int neighborDispatch(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  switch (e.arithmetic) {
    case SDG_BF16X6: return neighborImpl<O, 3>(e, layer, stream);
    case SDG_F16X3: return neighborImpl<O, 16>(e, layer, stream);
    default: return neighborImpl<O, 2>(e, layer, stream);
  }
}

}  // namespace

// the host entry of one order (explicitly instantiated by its own source)
template <int O>
int MfmaOrder<O>::local(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                        hipStream_t stream, const LtsOut* lts) {
  return localDispatch<O>(e, coeffs, layer, writeIntegrated, stream, lts);
}
template <int O>
int MfmaOrder<O>::neighbor(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  return neighborDispatch<O>(e, layer, stream);
}

}  // namespace mfma
}  // namespace sdg
