// Dynamic rupture faces around the friction solve:
//  * sdg_dr_face_data: per DR face the transposed inverse face rotation, the
//    two flux solvers and the impedances of the friction law
//    (src/Initializer/Model/DynamicRuptureMatrices.cpp:150-470);
//  * kDrInterpolate: space-time interpolation of both sides' derivatives at
//    the Gauss-Legendre time points and the face quadrature points, rotated to
//    the face frame, batched by face-matrix variant (DynamicRupture::spaceTimeInterpolation,
//    src/Kernels/DynamicRupture.cpp:57-168; spec dynamic_rupture.py:108-170);
//  * kDrNodalFlux: the flux of the imposed states into the adjacent cells
//    (Neighbor.cpp:93-101 with the nodalFlux spec, dynamic_rupture.py:174-190).
#include <hip/hip_runtime.h>

#include <cmath>
#include <random>
#include <vector>

#include "cell_local.h"
#include "dpp_rows.h"
#include "engine.h"
#include "global_matrices.h"

namespace sdg {
namespace {

constexpr int kDrThreads = 64;

// time-basis coefficients of the O time points: c[o][d] = t_o^d / d!
template <int O>
struct StepCoeffs {
  double c[O][O];
};

int stroudPoints(int order) { return (order + 1) * (order + 1); }

bool drLayerOk(const sdg_dr_layer* L, int order, const char* who) {
  if (L == nullptr || L->num_faces < 0) {
    setError(std::string(who) + ": invalid DR layer");
    return false;
  }
  if (stroudPoints(order) > 64) {
    setError(std::string(who) + ": dynamic rupture supports orders up to 7 (the (order+1)^2 points of a face "
             "are the lanes of one wave)");
    return false;
  }
  if (L->num_faces == 0) return true;
  if (L->points < stroudPoints(order) || L->points > 64 || L->plus_element == nullptr ||
      L->minus_element == nullptr || L->face_info == nullptr) {
    setError(std::string(who) + ": DR layer needs points in [(order+1)^2, 64], elements and face info");
    return false;
  }
  return true;
}

// face matrices of the engine's order on the device (real type of the engine):
// V3mTo2n [4][4][n][B], then V3mTo2nTWDivM transposed, [4][4][n][B] (the
// nodal flux reads it over the lanes' basis functions)
template <typename real>
int uploadDrMatrices(EngineState& s) {
  const int O = s.order, B = numBasis3(O), n = stroudPoints(O);
  std::vector<double> v(16LL * n * B), vt(16LL * n * B);
  if (sdg_dr_face_matrices(O, v.data(), vt.data()) != SDG_OK) return SDG_ERR_INVALID;
  std::vector<real> both(v.size() + vt.size());
  for (size_t i = 0; i < v.size(); ++i) both[i] = static_cast<real>(v[i]);
  for (int fh = 0; fh < 16; ++fh)
    for (int k = 0; k < B; ++k)
      for (int l = 0; l < n; ++l)
        both[v.size() + (static_cast<size_t>(fh) * n + l) * B + k] =
            static_cast<real>(vt[(static_cast<size_t>(fh) * B + k) * n + l]);
  SDG_HIP_CHECK(hipMalloc(&s.drMatrices, both.size() * sizeof(real)));
  SDG_HIP_CHECK(hipMemcpy(s.drMatrices, both.data(), both.size() * sizeof(real), hipMemcpyHostToDevice));
  return SDG_OK;
}

int ensureDrMatrices(EngineState& s) {
  if (s.drMatrices != nullptr) return SDG_OK;
  return s.arithmetic == SDG_FP64 ? uploadDrMatrices<double>(s) : uploadDrMatrices<float>(s);
}

constexpr int kDrWaves = 4;   // sides in flight per interpolation workgroup (one per wave)
constexpr int kDrScan = 1;    // faces of the chunk per thread
constexpr int kDrChunk = 64 * kDrWaves * kDrScan;  // faces scanned by one interpolation workgroup

// position of entry i of the packed derivatives (blocks [9][B(O-d)] one
// after another) in the layout's derivatives (blocks [9][rows(B(O-d))])
template <int O>
__device__ inline int derivativeOffset(int i, const CellLayout& lay) {
  if (lay.align == 1) return i;
  int rem = i, off = 0, at = 0;
#pragma unroll
  for (int d = 0; d < O; ++d) {
    const int m = numBasis3(O - d);
    if (rem >= 0 && rem < 9 * m) at = off + (rem / m) * lay.rows(m) + rem % m;
    rem -= 9 * m;
    off += 9 * lay.rows(m);
  }
  return at;
}

// first column (d, l = 0) of derivative order d among the columns of all
// orders, sum_{e < d} B(O - e); the packed block of order d starts at 9 times it
template <int O>
__host__ __device__ constexpr int columnOffset(int d) {
  int t = 0;
  for (int e = 0; e < d; ++e) t += numBasis3(O - e);
  return t;
}

// Space-time interpolation, batched by the 16 variants of V3mTo2n(i, h)
// (h = 0: plus side, h = 1..3: minus side with that orientation), as the
// reference batches per (side, faceRelation): workgroup (chunk, variant)
// stages its matrix in LDS once and serves every side of its variant among
// the chunk's kDrChunk faces, listed in face order and dealt round-robin to
// its waves (one side per wave at a time).  Per side the derivatives are
// rotated to the face frame first, dQ'_d[p][l] = sum_q dQ_d[q][l] TinvT(q, p)
// (the rotation commutes with the interpolation; columns (d, l) spread over
// the lanes), then interpolated, w_d[k][p] = sum_l V[k][l] dQ'_d[p][l] (lane
// k = Stroud point), and combined at the time points,
// out[f][o][p][k] = sum_d c[o][d] w_d[k][p] (FP64, the friction law's
// q_interpolated layout; points k in [(O+1)^2, P) are written as zero).
// The side's factors shared by all points (the derivatives, the rotation)
// are read once per 16-lane row and broadcast with DPP row_newbcast.  The
// gather of a wave's next side is issued before the arithmetic of its
// current one, so that the loads are in flight while it computes.
template <typename real, int O>
__global__ __launch_bounds__(64 * kDrWaves, 2) void kDrInterpolate(int64_t nf, const real* __restrict__ V,
                                                                const real* __restrict__ deriv, CellLayout lay,
                                                                sdg_dr_layer L, StepCoeffs<O> cf,
                                                                double* __restrict__ outPlus,
                                                                double* __restrict__ outMinus) {
  constexpr int B = numBasis3(O), NP = (O + 1) * (O + 1), LDV = B + 1;
  constexpr int DS = [] {
    int t = 0;
    for (int d = 0; d < O; ++d) t += 9 * numBasis3(O - d);
    return t;
  }();
  constexpr int DI = (DS + 63) / 64;  // derivative values per lane
  constexpr int NCOL = DS / 9;        // columns (d, l) of the derivatives
  constexpr int NT = 64 * kDrWaves;   // threads
  constexpr int VI = (NP * B + NT - 1) / NT;  // matrix entries staged per thread
  __shared__ real sV[NP * LDV];
  // padded by 16: a row read starting at the last run of a block reaches past it
  __shared__ real sD[kDrWaves][DS + 16];
  __shared__ real sT[kDrWaves][96];
  __shared__ int sFace[kDrChunk];  // the chunk's sides of this variant: chunk-local face index
  __shared__ int sElem[kDrChunk];  // and the side's element
  __shared__ int sCount[kDrScan][kDrWaves];
  const int variant = blockIdx.y, vi = variant >> 2, vh = variant & 3;
  const int side = vh == 0 ? 0 : 1;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const real* Vv = V + static_cast<int64_t>(vi * 4 + vh) * NP * B;
  const int64_t f0 = static_cast<int64_t>(blockIdx.x) * kDrChunk;
  // the start-up loads (the variant's matrix, the chunk's face data), all
  // issued before the first is used
  real vv[VI];
#pragma unroll
  for (int u = 0; u < VI; ++u) vv[u] = Vv[min(static_cast<int>(threadIdx.x) + NT * u, NP * B - 1)];
  bool match[kDrScan];
  int elem[kDrScan];
#pragma unroll
  for (int s = 0; s < kDrScan; ++s) {
    const int64_t f = min(f0 + threadIdx.x + NT * s, nf - 1);  // nf >= 1 here
    const int faceSide = min(static_cast<int>(L.face_info[f * 3 + side]), 3);
    // faceRelation = orientation + 1 lies in [1, 3]; clamped so that a malformed
    // entry still names a minus-side variant (0 would match no workgroup)
    const int h = side == 0 ? 0 : min(max(static_cast<int>(L.face_info[f * 3 + 2]), 1), 3);
    const int el = side == 0 ? L.plus_element[f] : L.minus_element[f];
    match[s] = f0 + threadIdx.x + NT * s < nf && faceSide == vi && h == vh;
    elem[s] = el;
  }
  unsigned long long mask[kDrScan];
#pragma unroll
  for (int s = 0; s < kDrScan; ++s) {
    mask[s] = __ballot(match[s]);
    if (lane == 0) sCount[s][wave] = __popcll(mask[s]);
  }
#pragma unroll
  for (int u = 0; u < VI; ++u) {
    const int i = threadIdx.x + NT * u;
    if (u + 1 < VI || i < NP * B) sV[(i / B) * LDV + i % B] = vv[u];
  }
  __syncthreads();
  // the chunk's sides of this variant in face order: scan s, then wave
  int total = 0;
#pragma unroll
  for (int s = 0; s < kDrScan; ++s)
#pragma unroll
    for (int w = 0; w < kDrWaves; ++w) total += sCount[s][w];
  total = __builtin_amdgcn_readfirstlane(total);
#pragma unroll
  for (int s = 0; s < kDrScan; ++s) {
    int base = 0;
#pragma unroll
    for (int s1 = 0; s1 < kDrScan; ++s1)
#pragma unroll
      for (int w = 0; w < kDrWaves; ++w)
        base += (s1 < s || (s1 == s && w < wave)) ? sCount[s1][w] : 0;
    if (match[s]) {
      const int at = base + static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                                static_cast<unsigned>(mask[s] >> 32),
                                __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask[s]), 0u)));
      sFace[at] = threadIdx.x + NT * s;
      sElem[at] = elem[s];
    }
  }
  __syncthreads();
  const int P = L.points;
  double* out = side == 0 ? outPlus : outMinus;
  real pre[DI];  // the next side's derivatives, in flight during the current side
  double preT[2];
  auto fetch = [&](int j) {
    const int64_t e = __builtin_amdgcn_readfirstlane(sElem[j]);
    const int64_t f = f0 + __builtin_amdgcn_readfirstlane(sFace[j]);
    const real* src = deriv + e * lay.dsize;
    // an opaque copy of the lane index: the gather offsets are recomputed per
    // side (hoisted out of the side loop they are spilled, and every load then
    // waits on the reload of its address)
    int ln = lane;
    asm volatile("" : "+v"(ln));
    if (lay.align == 1) {
#pragma unroll
      for (int c = 0; c < DI; ++c) pre[c] = src[min(ln + 64 * c, DS - 1)];  // clamped: in bounds
    } else {
#pragma unroll
      for (int c = 0; c < DI; ++c) pre[c] = src[derivativeOffset<O>(min(ln + 64 * c, DS - 1), lay)];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) preT[c] = L.tinv_t[f * 81 + min(ln + 64 * c, 80)];
  };
  if (wave < total) fetch(wave);
  for (int j = wave; j < total; j += kDrWaves) {
    const int64_t f = f0 + __builtin_amdgcn_readfirstlane(sFace[j]);
    waveSync();  // the previous side's reads of sD/sT are done
#pragma unroll
    for (int c = 0; c < DI; ++c) {
      const int i = lane + 64 * c;
      if (c + 1 < DI || i < DS) sD[wave][i] = pre[c];
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int i = lane + 64 * c;
      if (i < 81) sT[wave][i] = static_cast<real>(preT[c]);
    }
    waveSync();
    if (j + kDrWaves < total) fetch(j + kDrWaves);
    const int c16 = lane & 15;
    // rotation to the face frame first (it commutes with the interpolation):
    // dQ_d[p][l] <- sum_q dQ_d[q][l] TinvT(q, p) in place, one column (d, l)
    // per lane and pass; rowT[p] holds TinvT(c16, p)
    {
      real rowT[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) rowT[p] = sT[wave][9 * p + c16];
      // one column at a time (two in flight would hold 36 more registers)
#pragma unroll 1
      for (int cc = 0; cc < (NCOL + 63) / 64; ++cc) {
        const int col = min(lane + 64 * cc, NCOL - 1);  // clamped: read, not written
        int at = 0, stride = 1;
        staticFor<O>([&](auto dI) {
          constexpr int d = decltype(dI)::value;
          constexpr int m = numBasis3(O - d);
          constexpr int c0 = columnOffset<O>(d);
          if (col >= c0 && col < c0 + m) {
            at = 9 * c0 + (col - c0);  // the packed block of order d starts at 9 * c0
            stride = m;
          }
        });
        real x[9], y[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
          x[q] = sD[wave][at + q * stride];
          y[q] = real(0);
        }
        staticFor<9>([&](auto qI) {
          constexpr int q = decltype(qI)::value;
          fmacRowBcast9<q>(y, rowT, x[q]);
        });
        if (lane + 64 * cc < NCOL)
#pragma unroll
          for (int p = 0; p < 9; ++p) sD[wave][at + p * stride] = y[p];
      }
    }
    waveSync();
    // lanes k >= NP compute with the last point's row and store zero or nothing.
    // The row offset is an opaque copy per side: the basis values are loop
    // invariant, and hoisted out of the side loop they are spilled
    int vk = min(lane, NP - 1) * LDV;
    asm volatile("" : "+v"(vk));
    real w[O][9];
#pragma unroll
    for (int d = 0; d < O; ++d)
#pragma unroll
      for (int q = 0; q < 9; ++q) w[d][q] = real(0);
    // w_d[q] += V[k][l] dQ_d[q][l] over runs of 16 basis functions: row q of a
    // run holds dQ_d[q][l0 + c16], and step j broadcasts its lane j
    staticFor<O>([&](auto dI) {
      constexpr int d = decltype(dI)::value;
      constexpr int m = numBasis3(O - d);
      constexpr int off = 9 * columnOffset<O>(d);
      staticFor<(m + 15) / 16>([&](auto rI) {
        constexpr int l0 = 16 * decltype(rI)::value;
        constexpr int nl = m - l0 < 16 ? m - l0 : 16;
        // the run's rows and basis values are all read before its first FMA
        real row[9], xv[nl];
#pragma unroll
        for (int q = 0; q < 9; ++q) row[q] = sD[wave][off + q * m + l0 + c16];
#pragma unroll
        for (int jj = 0; jj < nl; ++jj) xv[jj] = sV[vk + l0 + jj];
        staticFor<nl>([&](auto jI) {
          constexpr int jj = decltype(jI)::value;
          fmacRowBcast9<jj>(w[d], row, xv[jj]);
        });
      });
    });
    // per time point: out[o][p] = sum_d c[o][d] w_d[p] (the rotated w).  The
    // lane index of the store offsets is an opaque copy per side: hoisted out
    // of the side loop, the 54 offsets are spilled
    double* const dst = out + f * (O * 9) * P;
    int sl = lane;
    asm volatile("" : "+v"(sl));
#pragma unroll
    for (int o = 0; o < O; ++o) {
      real acc[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        real a = real(0);
#pragma unroll
        for (int d = 0; d < O; ++d) a += static_cast<real>(cf.c[o][d]) * w[d][p];
        acc[p] = a;
      }
#pragma unroll
      for (int p = 0; p < 9; ++p)
        if (lane < P) dst[(o * 9 + p) * P + sl] = lane < NP ? static_cast<double>(acc[p]) : 0.0;
      // one time point at a time: interleaved, the time points' sums and
      // store addresses outgrow the registers
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// One workgroup per cell: for each DR face of the cell,
// Q[p][k] += sum_l VTW[k][l] sum_q G[q][l] F(q, p) with the imposed state G
// and flux solver F of the cell's side (plus: h = 0; minus: h = faceRelation).
// Both contractions take their face-uniform factor (F, then GF) from 16-lane
// rows broadcast with DPP row_newbcast; VTW is read as VTWt[l][k].
template <typename real, int O>
__global__ __launch_bounds__(kDrThreads) void kDrNodalFlux(int64_t n, const real* __restrict__ VTWt,
                                                           const int32_t* __restrict__ drMap, sdg_dr_layer L,
                                                           const double* __restrict__ impPlus,
                                                           const double* __restrict__ impMinus,
                                                           real* __restrict__ dofs, CellLayout lay) {
  static_assert(kDrThreads == 64, "one wave per cell");
  constexpr int B = numBasis3(O), NP = (O + 1) * (O + 1);
  constexpr int R = (B + kDrThreads - 1) / kDrThreads;  // basis functions per lane
  // GF as [p][l]; both arrays padded by 16 for the row reads of a last run
  __shared__ real sGF[9 * 64 + 16];
  __shared__ real sF[81 + 16];
  const int64_t e = blockIdx.x;
  if (e >= n) return;
  const int tid = threadIdx.x, c16 = tid & 15;
  const int P = L.points;
  real acc[R][9];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int p = 0; p < 9; ++p) acc[r][p] = real(0);
  bool any = false;
  for (int face = 0; face < 4; ++face) {
    const int32_t m = drMap[e * 4 + face];
    if (m < 0 || m >= L.num_faces) continue;
    any = true;
    // the face's side data as wave-uniform values: the three selects below
    // then stay scalar.  (With the byte loaded per lane, the compiler lowered
    // `plus` as a divergent branch and kept the plus flux solver on the path
    // where plus_element matches and the side does not.)
    const int plusElem = __builtin_amdgcn_readfirstlane(L.plus_element[m]);
    const int plusSide = __builtin_amdgcn_readfirstlane(static_cast<int>(L.face_info[m * 3 + 0]));
    const int relation = __builtin_amdgcn_readfirstlane(static_cast<int>(L.face_info[m * 3 + 2]));
    const bool plus = plusElem == e && plusSide == face;
    const int h = plus ? 0 : min(max(relation, 1), 3);  // minus side: faceRelation in [1, 3]
    const double* G = (plus ? impPlus : impMinus) + static_cast<int64_t>(m) * 9 * P;
    const double* F = (plus ? L.flux_solver_plus : L.flux_solver_minus) + static_cast<int64_t>(m) * 81;
    // the face's loads, all issued before the first wait: G at the lane's
    // point (clamped: lanes l >= NP compute a copy of the last point), F
    const int l = min(tid, NP - 1);
    real g[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) g[q] = static_cast<real>(G[q * P + l]);
    real fv[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) fv[c] = static_cast<real>(F[min(tid + 64 * c, 80)]);
    __syncthreads();  // the previous face's reads of sF and sGF are done
#pragma unroll
    for (int c = 0; c < 2; ++c)
      if (tid + 64 * c < 81) sF[tid + 64 * c] = fv[c];
    __syncthreads();
    // GF[l][p] = sum_q G[q][l] F(q, p); rowF[p] holds F(c16, p)
    {
      real rowF[9], gf[9];
#pragma unroll
      for (int p = 0; p < 9; ++p) {
        rowF[p] = sF[9 * p + c16];
        gf[p] = real(0);
      }
      staticFor<9>([&](auto qI) {
        constexpr int q = decltype(qI)::value;
        fmacRowBcast9<q>(gf, rowF, g[q]);
      });
#pragma unroll
      for (int p = 0; p < 9; ++p)
        if (tid < NP) sGF[p * 64 + tid] = gf[p];
    }
    __syncthreads();
    // Q[p][k] += sum_l VTW[k][l] GF[l][p] over runs of 16 points: row p of a
    // run holds GF[l0 + c16][p], and step j broadcasts its lane j
    const real* W = VTWt + static_cast<int64_t>(face * 4 + h) * NP * B;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int k = min(tid + kDrThreads * r, B - 1);
      staticFor<(NP + 15) / 16>([&](auto rI) {
        constexpr int l0 = 16 * decltype(rI)::value;
        constexpr int nl = NP - l0 < 16 ? NP - l0 : 16;
        real row[9], xv[nl];
#pragma unroll
        for (int jj = 0; jj < nl; ++jj) xv[jj] = W[(l0 + jj) * B + k];
#pragma unroll
        for (int p = 0; p < 9; ++p) row[p] = sGF[p * 64 + l0 + c16];
        staticFor<nl>([&](auto jI) {
          constexpr int jj = decltype(jI)::value;
          fmacRowBcast9<jj>(acc[r], row, xv[jj]);
        });
      });
    }
  }
  if (!any) return;
  // all old values loaded first (clamped rows, in bounds), then stored
  real* const q = dofs + e * lay.qStride;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int k = tid + r * kDrThreads;
    real old[9];
#pragma unroll
    for (int p = 0; p < 9; ++p) old[p] = q[p * lay.ldq + min(k, B - 1)];
    if (k < B)
#pragma unroll
      for (int p = 0; p < 9; ++p) q[p * lay.ldq + k] = old[p] + acc[r][p];
  }
}

template <typename real, int O>
int interpolateImpl(EngineState& st, const double* coeffs, const void* deriv, const sdg_dr_layer* L, double* qp,
                    double* qm, hipStream_t stream) {
  if (L->num_faces == 0) return SDG_OK;
  StepCoeffs<O> cf{};
  for (int o = 0; o < O; ++o)
    for (int d = 0; d < O; ++d) cf.c[o][d] = coeffs[o * O + d];
  const dim3 grid(static_cast<unsigned>((L->num_faces + kDrChunk - 1) / kDrChunk), 16);
  kDrInterpolate<real, O><<<grid, 64 * kDrWaves, 0, stream>>>(L->num_faces, static_cast<const real*>(st.drMatrices),
                                                              static_cast<const real*>(deriv), engineLayout(st), *L,
                                                              cf, qp, qm);
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

template <typename real, int O>
int nodalFluxImpl(EngineState& st, const sdg_layer* layer, const int32_t* drMap, const sdg_dr_layer* L,
                  const double* ip, const double* im, hipStream_t stream) {
  const int64_t n = layer->num_elements;
  if (n == 0 || L->num_faces == 0) return SDG_OK;
  const int64_t v3mSize = 16LL * stroudPoints(O) * numBasis3(O);
  kDrNodalFlux<real, O><<<static_cast<unsigned>(n), kDrThreads, 0, stream>>>(
      n, static_cast<const real*>(st.drMatrices) + v3mSize, drMap, *L, ip, im, static_cast<real*>(layer->dofs),
      cellLayout(st, layer));
  SDG_HIP_CHECK(hipGetLastError());
  return SDG_OK;
}

}  // namespace
}  // namespace sdg

using namespace sdg;

extern "C" int sdg_dr_face_data(int64_t num_faces, const double* vertices, const double* material,
                                const int32_t* plus_element, const int32_t* minus_element, const uint8_t* face_info,
                                double* tinv_t, double* flux_solver_plus, double* flux_solver_minus,
                                double* imp_and_eta) {
  if (num_faces < 0 || (num_faces > 0 && (vertices == nullptr || material == nullptr || plus_element == nullptr ||
                                          minus_element == nullptr || face_info == nullptr || tinv_t == nullptr ||
                                          flux_solver_plus == nullptr || flux_solver_minus == nullptr))) {
    setError("sdg_dr_face_data: invalid arguments");
    return SDG_ERR_INVALID;
  }
  for (int64_t m = 0; m < num_faces; ++m)
    if (face_info[m * 3] > 3 || face_info[m * 3 + 1] > 3 || face_info[m * 3 + 2] < 1 || face_info[m * 3 + 2] > 3) {
      setError("sdg_dr_face_data: face " + std::to_string(m) +
               ": sides must lie in [0, 3] and faceRelation (orientation + 1) in [1, 3]");
      return SDG_ERR_INVALID;
    }
  for (int64_t m = 0; m < num_faces; ++m) {
    const int64_t ep = plus_element[m], em = minus_element[m];
    const int sp = face_info[m * 3] & 3;
    double v[4][3], vm[4][3];
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 3; ++c) {
        v[q][c] = vertices[(ep * 4 + q) * 3 + c];
        vm[q][c] = vertices[(em * 4 + q) * 3 + c];
      }
    // fault normal and tangents from the plus side's face (MeshTools::normalAndTangents)
    const double* p0 = v[faceNode(sp, 0)];
    const double* p1 = v[faceNode(sp, 1)];
    const double* p2 = v[faceNode(sp, 2)];
    double ab[3], ac[3], nrm[3], s[3], t[3];
    for (int k = 0; k < 3; ++k) {
      ab[k] = p1[k] - p0[k];
      ac[k] = p2[k] - p0[k];
    }
    cross3(ab, ac, nrm);
    for (int k = 0; k < 3; ++k) s[k] = ab[k];
    cross3(nrm, s, t);
    const double nn = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const double ns = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const double nt = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const double area = 0.5 * nn;
    for (int k = 0; k < 3; ++k) {
      nrm[k] /= nn;
      s[k] /= ns;
      t[k] /= nt;
    }
    double T[81], Ti[81], grad[3][3];
    faceRotation(nrm, s, t, T, Ti);
    for (int q = 0; q < 9; ++q)
      for (int p = 0; p < 9; ++p) tinv_t[m * 81 + q + 9 * p] = Ti[p + 9 * q];
    const double volPlus = std::fabs(referenceGradients(v, grad)) / 6.0;
    const double volMinus = std::fabs(referenceGradients(vm, grad)) / 6.0;
    const ElasticMaterial matP{material[3 * ep], material[3 * ep + 1], material[3 * ep + 2]};
    const ElasticMaterial matM{material[3 * em], material[3 * em + 1], material[3 * em + 2]};
    // fluxSolver(q, p) = scale * sum_k A0(q, k) T(p, k)  (rotateFluxMatrix)
    for (int side = 0; side < 2; ++side) {
      double A0[81];
      transposedCoefficientMatrix(side == 0 ? matP : matM, 0, A0);
      const double scale = side == 0 ? -2.0 * area / (6.0 * volPlus) : 2.0 * area / (6.0 * volMinus);
      double* F = (side == 0 ? flux_solver_plus : flux_solver_minus) + m * 81;
      for (int q = 0; q < 9; ++q)
        for (int p = 0; p < 9; ++p) {
          double acc = 0.0;
          for (int k = 0; k < 9; ++k) acc += A0[q + 9 * k] * T[p + 9 * k];
          F[q + 9 * p] = scale * acc;
        }
    }
    if (imp_and_eta != nullptr) {
      auto speeds = [](const ElasticMaterial& mt, double& vp, double& vs) {
        vp = std::sqrt((mt.lambda + 2.0 * mt.mu) / mt.rho);
        vs = std::sqrt(mt.mu / mt.rho);
      };
      double vpP, vsP, vpM, vsM;
      speeds(matP, vpP, vsP);
      speeds(matM, vpM, vsM);
      const double zp = matP.rho * vpP, zs = matP.rho * vsP, zpN = matM.rho * vpM, zsN = matM.rho * vsM;
      double* ie = imp_and_eta + m * 11;
      ie[0] = zp;
      ie[1] = zs;
      ie[2] = zpN;
      ie[3] = zsN;
      ie[4] = 1.0 / (1.0 / zp + 1.0 / zpN);
      ie[5] = 1.0 / (1.0 / zs + 1.0 / zsN);
      ie[6] = 1.0 / zs + 1.0 / zsN;
      ie[7] = 1.0 / zp;
      ie[8] = 1.0 / zs;
      ie[9] = 1.0 / zpN;
      ie[10] = 1.0 / zsN;
    }
  }
  return SDG_OK;
}

extern "C" int sdg_dr_space_time_interpolation(sdg_engine* engine, const double* coeffs, const void* derivatives,
                                               const sdg_dr_layer* layer, double* q_plus, double* q_minus,
                                               void* stream) {
  if (engine == nullptr || coeffs == nullptr) {
    setError("sdg_dr_space_time_interpolation: NULL argument");
    return SDG_ERR_INVALID;
  }
  auto& st = engine->s;
  if (!drLayerOk(layer, st.order, "sdg_dr_space_time_interpolation")) return SDG_ERR_INVALID;
  if (layer->num_faces > 0 && (derivatives == nullptr || q_plus == nullptr || q_minus == nullptr ||
                               layer->tinv_t == nullptr)) {
    setError("sdg_dr_space_time_interpolation: derivatives, outputs and TinvT are required");
    return SDG_ERR_INVALID;
  }
  if (const int rc = ensureDrMatrices(st); rc != SDG_OK) return rc;
  const auto s = static_cast<hipStream_t>(stream);
  const bool dbl = st.arithmetic == SDG_FP64;
  switch (st.order) {
#define SDG_DR_CASE(o)                                                                              \
  case o:                                                                                           \
    return dbl ? interpolateImpl<double, o>(st, coeffs, derivatives, layer, q_plus, q_minus, s)   \
               : interpolateImpl<float, o>(st, coeffs, derivatives, layer, q_plus, q_minus, s);
    SDG_DR_CASE(2) SDG_DR_CASE(3) SDG_DR_CASE(4) SDG_DR_CASE(5) SDG_DR_CASE(6) SDG_DR_CASE(7)
#undef SDG_DR_CASE
    default:
      setError("sdg_dr_space_time_interpolation: unsupported order");
      return SDG_ERR_UNSUPPORTED;
  }
}

extern "C" int sdg_dr_nodal_flux(sdg_engine* engine, const sdg_layer* layer, const int32_t* dr_mapping,
                                 const sdg_dr_layer* dr_layer, const double* imposed_plus,
                                 const double* imposed_minus, void* stream) {
  if (engine == nullptr || layer == nullptr || layer->num_elements < 0) {
    setError("sdg_dr_nodal_flux: invalid arguments");
    return SDG_ERR_INVALID;
  }
  auto& st = engine->s;
  if (!drLayerOk(dr_layer, st.order, "sdg_dr_nodal_flux")) return SDG_ERR_INVALID;
  if (layer->num_elements > 0 && dr_layer->num_faces > 0 &&
      (layer->dofs == nullptr || dr_mapping == nullptr || imposed_plus == nullptr || imposed_minus == nullptr ||
       dr_layer->flux_solver_plus == nullptr || dr_layer->flux_solver_minus == nullptr)) {
    setError("sdg_dr_nodal_flux: dofs, dr_mapping, imposed states and flux solvers are required");
    return SDG_ERR_INVALID;
  }
  if (const int rc = ensureDrMatrices(st); rc != SDG_OK) return rc;
  const auto s = static_cast<hipStream_t>(stream);
  const bool dbl = st.arithmetic == SDG_FP64;
  switch (st.order) {
#define SDG_DR_CASE(o)                                                                                            \
  case o:                                                                                                         \
    return dbl ? nodalFluxImpl<double, o>(st, layer, dr_mapping, dr_layer, imposed_plus, imposed_minus, s)      \
               : nodalFluxImpl<float, o>(st, layer, dr_mapping, dr_layer, imposed_plus, imposed_minus, s);
    SDG_DR_CASE(2) SDG_DR_CASE(3) SDG_DR_CASE(4) SDG_DR_CASE(5) SDG_DR_CASE(6) SDG_DR_CASE(7)
#undef SDG_DR_CASE
    default:
      setError("sdg_dr_nodal_flux: unsupported order");
      return SDG_ERR_UNSUPPORTED;
  }
}

// The proxy's dynamic-rupture data (src/Proxy/Allocator.cpp:186-294): 4 n DR
// faces between random cells with random sides and orientations, and every
// cell face mapped to a random DR face.
extern "C" int sdg_proxy_dr_faces(int64_t n, uint64_t seed, int32_t* dr_mapping, int32_t* plus_element,
                                  int32_t* minus_element, uint8_t* face_info) {
  if (n <= 0 || 4 * n > INT32_MAX || dr_mapping == nullptr || plus_element == nullptr || minus_element == nullptr ||
      face_info == nullptr) {
    setError("sdg_proxy_dr_faces: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const int64_t nf = 4 * n;
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> side(0, 3), orientation(0, 2);
  std::uniform_int_distribution<int64_t> cell(0, n - 1), face(0, nf - 1);
  for (int64_t i = 0; i < 4 * n; ++i) dr_mapping[i] = static_cast<int32_t>(face(rng));
  for (int64_t f = 0; f < nf; ++f) {
    plus_element[f] = static_cast<int32_t>(cell(rng));
    minus_element[f] = static_cast<int32_t>(cell(rng));
    face_info[3 * f] = static_cast<uint8_t>(side(rng));
    face_info[3 * f + 1] = static_cast<uint8_t>(side(rng));
    face_info[3 * f + 2] = static_cast<uint8_t>(orientation(rng) + 1);
  }
  return SDG_OK;
}

// Proxy kernels with dynamic rupture (src/Proxy/KernelDevice.cpp:109-123,
// KernelHost.cpp:219-244): neigh_dr = neighbour integral of the regular faces
// + nodal flux of the DR faces; godunov_dr = space-time interpolation at the
// Gauss-Legendre points of [0, dt]; all_dr = ader + local integral, then
// godunov_dr, then neigh_dr.
extern "C" int sdg_proxy_run_dr(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer,
                                const sdg_proxy_dr* dr, void* stream) {
  if (engine == nullptr || steps < 0 || kernel_id < 0 || kernel_id > 7) {
    setError("sdg_proxy_run_dr: invalid arguments");
    return SDG_ERR_INVALID;
  }
  if (kernel_id < 5) return sdg_proxy_run(engine, kernel_id, steps, dt, layer, stream);
  if (dr == nullptr || layer == nullptr) {
    setError("sdg_proxy_run_dr: DR kernels need the layer and its DR faces");
    return SDG_ERR_INVALID;
  }
  const int O = engine->s.order;
  std::vector<double> tp(O), tw(O), coll(O * O), integ(O);
  if (int rc = sdg_gauss_legendre(O, 0.0, dt, tp.data(), tw.data()); rc != SDG_OK) return rc;
  for (int o = 0; o < O; ++o) {
    double fact = 1.0;
    for (int d = 0; d < O; ++d) {
      if (d > 0) fact *= d;
      coll[o * O + d] = std::pow(tp[o], d) / fact;
    }
  }
  sdg_monomial_integrate(O, 0.0, dt, integ.data());
  for (int s = 0; s < steps; ++s) {
    int rc = SDG_OK;
    if (kernel_id == 7) rc = sdg_compute_batched_ader_local(engine, integ.data(), dt, layer, stream);
    if (rc == SDG_OK && (kernel_id == 6 || kernel_id == 7))
      rc = sdg_dr_space_time_interpolation(engine, coll.data(), layer->derivatives, &dr->faces, dr->q_plus,
                                           dr->q_minus, stream);
    if (rc == SDG_OK && (kernel_id == 5 || kernel_id == 7)) {
      rc = sdg_compute_batched_neighbors_integral(engine, layer, stream);
      if (rc == SDG_OK)
        rc = sdg_dr_nodal_flux(engine, layer, dr->dr_mapping, &dr->faces, dr->imposed_plus, dr->imposed_minus,
                               stream);
    }
    if (rc != SDG_OK) return rc;
  }
  return SDG_OK;
}
