// Per-element work estimates of the proxy kernels (the PerformanceEstimate of
// src/Monitoring/Metric.h:16-50 that the proxy sums in src/Proxy/Runner.cpp:120-150).
//
// Non-zero FLOP: every contraction of the element update counted as 2 flops
// per non-zero of the sparse operand, from the sparsity of this library's
// generated global matrices (entries below 1e-13 of the matrix maximum count
// as zero) and the 24 non-zeros of an elastic star matrix.  Hardware FLOP:
// the same contractions with dense operands, times the number of bf16
// products of the split forms.  Bytes: DOF blocks and per-element operators
// each moved once (global matrices excluded); kernel bytes count the face
// traces that the matrix-core forms exchange instead of I.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "global_matrices.h"

namespace sdg {
namespace {

constexpr int kStarNonZeros = 24;

struct Counts {
  double nz = 0, dense = 0;
};

// non-zeros of rows [0, rows) x cols [0, cols) of a row-major block with leading dimension ld
double nonZeros(const double* m, int rows, int cols, int ld) {
  double mx = 0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) mx = std::max(mx, std::fabs(m[r * ld + c]));
  double nz = 0;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) nz += std::fabs(m[r * ld + c]) > 1e-13 * mx ? 1 : 0;
  return nz;
}

struct Work {
  Counts ader, volume, localFlux, neighborFlux;
};

Work elementWork(int order, const std::vector<double>& gm) {
  const GlobalMatrixLayout L = globalMatrixLayout(order);
  const int B = numBasis3(order), B2 = numBasis2(order);
  Work w;
  // CK: dQ_i[B(O-i)][9] = sum_j kDivMT_j[:B(O-i), :B(O-i+1)] dQ_{i-1} star_j, plus I += c_i dQ_i
  for (int i = 1; i < order; ++i) {
    const int rows = numBasis3(order - i), inner = numBasis3(order - i + 1);
    for (int j = 0; j < 3; ++j) {
      w.ader.nz += 2.0 * nonZeros(&gm[L.kDivMT + j * B * B], rows, inner, B) * 9 + 2.0 * kStarNonZeros * rows;
      w.ader.dense += 2.0 * rows * inner * 9 + 2.0 * 81 * rows;
    }
  }
  for (int i = 0; i < order; ++i) {
    w.ader.nz += 2.0 * numBasis3(order - i) * 9;
    w.ader.dense += 2.0 * numBasis3(order - i) * 9;
  }
  // volume: Q += sum_j kDivM_j (I star_j)
  for (int j = 0; j < 3; ++j) {
    w.volume.nz += 2.0 * kStarNonZeros * B + 2.0 * nonZeros(&gm[L.kDivM + j * B * B], B, B, B) * 9;
    w.volume.dense += 2.0 * 81 * B + 2.0 * B * B * 9;
  }
  // local flux of 4 faces: Q += rDivM_f (fMrT_f I) AplusT_f
  for (int f = 0; f < 4; ++f) {
    w.localFlux.nz += 2.0 * nonZeros(&gm[L.fMrT + f * B2 * B], B2, B, B) * 9 + 2.0 * 81 * B2 +
                      2.0 * nonZeros(&gm[L.rDivM + f * B * B2], B, B2, B2) * 9;
    w.localFlux.dense += 2.0 * B2 * B * 9 + 2.0 * 81 * B2 + 2.0 * B * B2 * 9;
  }
  // neighbour flux of 4 faces: Q += rDivM_f fP_h rT_j I_nb AminusT_f, averaged over the
  // 12 (side j, orientation h) relations
  double rt = 0, fp = 0, rdm = 0;
  for (int j = 0; j < 4; ++j) rt += nonZeros(&gm[L.rT + j * B2 * B], B2, B, B) / 4.0;
  for (int h = 0; h < 3; ++h) fp += nonZeros(&gm[L.fP + h * B2 * B2], B2, B2, B2) / 3.0;
  for (int f = 0; f < 4; ++f) rdm += nonZeros(&gm[L.rDivM + f * B * B2], B, B2, B2);
  w.neighborFlux.nz = 4 * (2.0 * rt * 9 + 2.0 * fp * 9 + 2.0 * 81 * B2) + 2.0 * rdm * 9;
  w.neighborFlux.dense = 4 * (2.0 * B2 * B * 9 + 2.0 * B2 * B2 * 9 + 2.0 * 81 * B2 + 2.0 * B * B2 * 9);
  return w;
}

// Dynamic rupture, per DR face: the space-time interpolation of both sides
// at the `order` time points (Taylor sum, V3mTo2n(i, h) Q, rotation by
// TinvT) and the nodal flux of one side (G F, V3mTo2nTWDivM(i, h) (G F));
// V3mTo2n counts averaged over the plus (h = 0) and minus (h = 1..3) forms.
struct DrWork {
  Counts godunovFace, nodalFace;
};

DrWork drWork(int order) {
  const int B = numBasis3(order), np = (order + 1) * (order + 1);
  std::vector<double> v(16LL * np * B), vt(16LL * np * B);
  sdg_dr_face_matrices(order, v.data(), vt.data());
  double nzPlus = 0, nzMinus = 0, nzTwPlus = 0, nzTwMinus = 0;
  for (int i = 0; i < 4; ++i)
    for (int h = 0; h < 4; ++h) {
      const double a = nonZeros(&v[(i * 4 + h) * static_cast<size_t>(np) * B], np, B, B);
      const double b = nonZeros(&vt[(i * 4 + h) * static_cast<size_t>(np) * B], B, np, np);
      (h == 0 ? nzPlus : nzMinus) += h == 0 ? a / 4.0 : a / 12.0;
      (h == 0 ? nzTwPlus : nzTwMinus) += h == 0 ? b / 4.0 : b / 12.0;
    }
  DrWork w;
  double taylor = 0;
  for (int d = 0; d < order; ++d) taylor += 2.0 * 9 * numBasis3(order - d);
  const double rot = 2.0 * np * 81;
  w.godunovFace.nz = order * (2 * taylor + 2.0 * (nzPlus + nzMinus) * 9 + 2 * rot);
  w.godunovFace.dense = order * (2 * taylor + 2.0 * 2 * np * B * 9 + 2 * rot);
  w.nodalFace.nz = rot + 2.0 * 0.5 * (nzTwPlus + nzTwMinus) * 9;
  w.nodalFace.dense = rot + 2.0 * B * np * 9;
  return w;
}

}  // namespace
}  // namespace sdg

using namespace sdg;

extern "C" int sdg_performance_estimate(int order, int arithmetic, int kernel_id, sdg_work_estimate* out) {
  if (out == nullptr || order < kMinOrder || order > kMaxOrder || arithmetic < SDG_FP64 ||
      arithmetic > SDG_F16X3 || kernel_id < 0 || kernel_id > 7 ||
      (kernel_id >= 5 && (order > 7 || (matrixCore(arithmetic) && order < 3)))) {
    setError("sdg_performance_estimate: invalid order, arithmetic, kernel or NULL output");
    return SDG_ERR_INVALID;
  }
  std::vector<double> gm(globalMatrixLayout(order).total);
  if (generateGlobalMatrices(order, gm.data()) != 0) {
    setError("sdg_performance_estimate: global matrices");
    return SDG_ERR_INVALID;
  }
  const Work w = elementWork(order, gm);
  const int B = numBasis3(order), B2 = numBasis2(order);
  const double s = arithmetic == SDG_FP64 ? 8.0 : 4.0;
  const double products = arithmetic == SDG_BF16X6 ? 6.0 : (matrixCore(arithmetic) ? 3.0 : 1.0);
  const bool traces = matrixCore(arithmetic);
  const int B2P = (B2 + 3) / 4 * 4;  // padded face rows of the matrix-core traces
  const double q = 9.0 * B, traceBytes = 4.0 * 9 * B2P * 4;
  Counts c;
  double bytes = 0, kernelBytes = 0;
  auto add = [&](const Counts& x) {
    c.nz += x.nz;
    c.dense += x.dense;
  };
  if (kernel_id >= 5) {
    // the proxy's DR layer: all four faces of every cell are DR faces and
    // there are 4 DR faces per cell (src/Proxy/Allocator.cpp:193, :230)
    const DrWork d = drWork(order);
    const int np = (order + 1) * (order + 1);
    const double dsize = static_cast<double>(sdg_derivatives_size(order));
    const double godunovBytes = 4 * (2 * dsize * s + 81 * 8 + 2.0 * order * 9 * np * 8);
    const double nodalBytes = 2 * q * s + 4 * (9.0 * np * 8 + 81 * 8);
    double localHardware = 0.0, localKernelBytes = 0.0;
    if (kernel_id == 7) {
      add(w.ader);
      add(w.volume);
      bytes += (3 * q + 3 * kStarNonZeros + 4 * 81 + dsize) * s;  // Q in/out, I, derivatives out
      if (traces) {
        // the matrix-core local pass runs its whole product chain (the DR
        // faces' flux operands are zero) and writes traces, not I
        sdg_work_estimate l{};
        if (const int rc = sdg_performance_estimate(order, arithmetic, 1, &l); rc != SDG_OK) return rc;
        localHardware = l.hardware_flop - (w.ader.dense + w.volume.dense);
        localKernelBytes = traceBytes - q * s;
      }
    }
    if (kernel_id == 6 || kernel_id == 7) {
      c.nz += 4 * d.godunovFace.nz;
      c.dense += 4 * d.godunovFace.dense;
      bytes += godunovBytes;
    }
    if (kernel_id == 5 || kernel_id == 7) {
      c.nz += 4 * d.nodalFace.nz;
      c.dense += 4 * d.nodalFace.dense;
      bytes += nodalBytes;
    }
    out->nonzero_flop = c.nz;
    out->hardware_flop = c.dense + localHardware;
    out->bytes = bytes;
    out->kernel_bytes = bytes + localKernelBytes;
    return SDG_OK;
  }
  const bool ader = kernel_id == 0 || kernel_id == 1 || kernel_id == 3;
  const bool local = kernel_id == 0 || kernel_id == 1 || kernel_id == 4;
  const bool neigh = kernel_id == 0 || kernel_id == 2;
  if (ader) add(w.ader);
  if (local) {
    add(w.volume);
    add(w.localFlux);
  }
  if (neigh) add(w.neighborFlux);
  // kernel ids: 0 all, 1 local (ader + integral), 2 neigh, 3 ader, 4 localwoader
  if (kernel_id == 3) bytes = (2 * q + 3 * kStarNonZeros) * s;                    // Q in, I out, star
  if (kernel_id == 1 || kernel_id == 0) bytes += (3 * q + 3 * kStarNonZeros + 4 * 81) * s;  // Q in/out, I out
  if (kernel_id == 4) bytes = (3 * q + 3 * kStarNonZeros + 4 * 81) * s;           // I in, Q in/out
  if (neigh) bytes += (3 * q + 4 * 81) * s;                                        // Q in/out, neighbour I
  kernelBytes = bytes;
  // matrix-core forms: the fused local pass writes the face traces rT_f I
  // instead of I, and the neighbour pass reads the neighbours' traces
  if (traces && (kernel_id == 0 || kernel_id == 1)) kernelBytes += traceBytes - q * s;
  if (traces && neigh) kernelBytes += traceBytes - q * s;
  out->nonzero_flop = c.nz;
  out->hardware_flop = c.dense * products;
  out->bytes = bytes;
  out->kernel_bytes = kernelBytes;
  return SDG_OK;
}
