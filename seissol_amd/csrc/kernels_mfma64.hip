// Host side of the FP64 engines' matrix-core local pass: the FP64 fragments of
// the global matrices and the dispatch over the orders 3-7 (the kernels of one
// order are built by kernels_mfma64_o<O>.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "engine.h"
#include "global_matrices.h"
#include "mfma_common.h"

namespace sdg {
namespace {

// A (rows x cols, 0 outside) in the lane order of v_mfma_f64_16x16x4_f64's A
// operand: lane l of fragment (mt, ks) holds A[16 mt + l % 16][4 ks + l / 16]
void appendFragments64(const std::function<double(int, int)>& A, int rows, int cols, int mtiles, int ksteps,
                       double* out) {
  for (int mt = 0; mt < mtiles; ++mt)
    for (int ks = 0; ks < ksteps; ++ks) {
      double* base = out + (static_cast<int64_t>(mt) * ksteps + ks) * 64;
      for (int l = 0; l < 64; ++l) {
        const int r = 16 * mt + l % 16, c = 4 * ks + l / 16;
        base[l] = (r < rows && c < cols) ? A(r, c) : 0.0;
      }
    }
}

template <int O>
int uploadData64(EngineState& s, const double* gm) {
  using S = mfma64::Shape64<O>;
  const auto g = globalMatrixLayout(O);
  const int B = S::B, B2 = S::B2, B2P = S::B2P;
  std::vector<double> frags(S::total, 0.0);
  for (int j = 0; j < 3; ++j) {
    appendFragments64([&](int r, int c) { return -gm[g.kDivMT + (j * B + r) * B + c]; }, B, B, S::MT, S::KSB,
                      frags.data() + S::ck + static_cast<int64_t>(j) * S::MT * S::KSB * 64);
    appendFragments64([&](int r, int c) { return gm[g.kDivM + (j * B + r) * B + c]; }, B, B, S::MT, S::KSB,
                      frags.data() + S::vol + static_cast<int64_t>(j) * S::MT * S::KSB * 64);
  }
  // the four face operands stacked with B2P rows each: traces (fMrT_f / M2) I
  appendFragments64(
      [&](int r, int c) {
        const int m = r % B2P;
        return m < B2 ? gm[g.fMrT + ((r / B2P) * B2 + m) * B + c] / gm[g.mass2 + m] : 0.0;
      },
      S::FR, B, S::MTF, S::KSB, frags.data() + S::fmrt);
  // rDivM_f diag(M2), one face at a time: [m-tile][face][k-step]
  for (int mt = 0; mt < S::MT; ++mt)
    for (int f = 0; f < 4; ++f)
      for (int ks = 0; ks < S::KSF; ++ks) {
        double* base = frags.data() + S::rdm + ((static_cast<int64_t>(mt) * 4 + f) * S::KSF + ks) * 64;
        for (int l = 0; l < 64; ++l) {
          const int r = 16 * mt + l % 16, m = 4 * ks + l / 16;
          base[l] = (r < B && m < B2) ? gm[g.rDivM + (f * B + r) * B2 + m] * gm[g.mass2 + m] : 0.0;
        }
      }
  // the neighbour pass: rDivM_f per face, and fP_h as rows of at most O non-zeros
  for (int mt = 0; mt < S::MT; ++mt)
    for (int f = 0; f < 4; ++f)
      for (int ks = 0; ks < S::KSF; ++ks) {
        double* base = frags.data() + S::rdmN + ((static_cast<int64_t>(mt) * 4 + f) * S::KSF + ks) * 64;
        for (int l = 0; l < 64; ++l) {
          const int r = 16 * mt + l % 16, m = 4 * ks + l / 16;
          base[l] = (r < B && m < B2) ? gm[g.rDivM + (f * B + r) * B2 + m] : 0.0;
        }
      }
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
        frags[S::tFP + (h * B2 + r) * O + cnt] = v;
        frags[S::tFPCol + (h * B2 + r) * O + cnt] = static_cast<double>(c);
        ++cnt;
      }
    }
  }
  if (s.mfma64Frags == nullptr) SDG_HIP_CHECK(hipMalloc(&s.mfma64Frags, frags.size() * sizeof(double)));
  SDG_HIP_CHECK(hipMemcpy(s.mfma64Frags, frags.data(), frags.size() * sizeof(double), hipMemcpyHostToDevice));
  return SDG_OK;
}

int localOrders64(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream,
                  const mfma64::LtsOut64* lts) {
  switch (e.order) {
    case 3: return mfma64::MfmaOrder64<3>::local(e, coeffs, layer, stream, lts);
    case 4: return mfma64::MfmaOrder64<4>::local(e, coeffs, layer, stream, lts);
    case 5: return mfma64::MfmaOrder64<5>::local(e, coeffs, layer, stream, lts);
    case 6: return mfma64::MfmaOrder64<6>::local(e, coeffs, layer, stream, lts);
    case 7: return mfma64::MfmaOrder64<7>::local(e, coeffs, layer, stream, lts);
    default: setError("the FP64 matrix-core pass supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}

}  // namespace

int64_t mfma64::fragmentCount(int order) {
  switch (order) {
    case 3: return Shape64<3>::total;
    case 4: return Shape64<4>::total;
    case 5: return Shape64<5>::total;
    case 6: return Shape64<6>::total;
    case 7: return Shape64<7>::total;
    default: return 0;
  }
}

int uploadMfma64(EngineState& s, const double* gm) {
  switch (s.order) {
    case 3: return uploadData64<3>(s, gm);
    case 4: return uploadData64<4>(s, gm);
    case 5: return uploadData64<5>(s, gm);
    case 6: return uploadData64<6>(s, gm);
    case 7: return uploadData64<7>(s, gm);
    default: return SDG_OK;  // no FP64 matrix-core pass at this order
  }
}

bool hasMfma64(const EngineState& e) {
  return e.arithmetic == SDG_FP64 && e.mfma64Frags != nullptr && e.kernels == SDG_KERNELS_AUTO;
}

int launchAderLocalMfma64(const EngineState& e, const double* coeffs, const sdg_layer* layer, hipStream_t stream) {
  if (layer->derivatives != nullptr) {
    const mfma64::LtsOut64 lts{nullptr, static_cast<double*>(layer->derivatives), 1, 0};
    return localOrders64(e, coeffs, layer, stream, &lts);
  }
  return localOrders64(e, coeffs, layer, stream, nullptr);
}

int launchNeighborMfma64(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  switch (e.order) {
    case 3: return mfma64::MfmaOrder64<3>::neighbor(e, layer, stream);
    case 4: return mfma64::MfmaOrder64<4>::neighbor(e, layer, stream);
    case 5: return mfma64::MfmaOrder64<5>::neighbor(e, layer, stream);
    case 6: return mfma64::MfmaOrder64<6>::neighbor(e, layer, stream);
    case 7: return mfma64::MfmaOrder64<7>::neighbor(e, layer, stream);
    default: setError("the FP64 matrix-core pass supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}

int launchLocalIntegrationMfma64(const EngineState& e, const double* coeffs, const sdg_layer* layer,
                                 int resetBuffers, hipStream_t stream) {
  const mfma64::LtsOut64 lts{layer->lts_setup, static_cast<double*>(layer->derivatives), resetBuffers, 0};
  return localOrders64(e, coeffs, layer, stream, &lts);
}

}  // namespace sdg
