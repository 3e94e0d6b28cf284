// Host dispatch of the matrix-core element update over the orders 3-7 (the
// kernels of one order are built by kernels_mfma_o<O>.hip).
#include <hip/hip_runtime.h>

#include "engine.h"
#include "mfma_common.h"

namespace sdg {

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

namespace {
int localOrders(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                hipStream_t stream, const mfma::LtsOut* lts) {
  switch (e.order) {
    case 3: return mfma::MfmaOrder<3>::local(e, coeffs, layer, writeIntegrated, stream, lts);
    case 4: return mfma::MfmaOrder<4>::local(e, coeffs, layer, writeIntegrated, stream, lts);
    case 5: return mfma::MfmaOrder<5>::local(e, coeffs, layer, writeIntegrated, stream, lts);
    case 6: return mfma::MfmaOrder<6>::local(e, coeffs, layer, writeIntegrated, stream, lts);
    case 7: return mfma::MfmaOrder<7>::local(e, coeffs, layer, writeIntegrated, stream, lts);
    default: setError("matrix-core path supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}
}  // namespace

int launchLocalMfma(const EngineState& e, const double* coeffs, const sdg_layer* layer, bool writeIntegrated,
                    hipStream_t stream) {
  if (layer->derivatives != nullptr) {
    // the fused ader + local pass with derivatives (every cell stores them)
    const mfma::LtsOut lts{nullptr, static_cast<float*>(layer->derivatives), 1, 0};
    return localOrders(e, coeffs, layer, writeIntegrated, stream, &lts);
  }
  return localOrders(e, coeffs, layer, writeIntegrated, stream, nullptr);
}

int launchLocalIntegrationMfma(const EngineState& e, const double* coeffs, const sdg_layer* layer, int resetBuffers,
                               hipStream_t stream) {
  const mfma::LtsOut lts{layer->lts_setup, static_cast<float*>(layer->derivatives), resetBuffers, 0};
  return localOrders(e, coeffs, layer, layer->time_integrated != nullptr, stream, &lts);
}

int launchNeighborMfma(const EngineState& e, const sdg_layer* layer, hipStream_t stream) {
  switch (e.order) {
    case 3: return mfma::MfmaOrder<3>::neighbor(e, layer, stream);
    case 4: return mfma::MfmaOrder<4>::neighbor(e, layer, stream);
    case 5: return mfma::MfmaOrder<5>::neighbor(e, layer, stream);
    case 6: return mfma::MfmaOrder<6>::neighbor(e, layer, stream);
    case 7: return mfma::MfmaOrder<7>::neighbor(e, layer, stream);
    default: setError("matrix-core path supports orders 3..7"); return SDG_ERR_UNSUPPORTED;
  }
}

}  // namespace sdg
