// Local time stepping (rate 2): the per-cell LTS setup, the cluster-wise
// local and neighbour integration, and the macro step of all clusters,
// enqueued directly or replayed from a captured hipGraph.
//
// Clusters: cluster c advances with dt_c = dt0 * 2^c; its cells are the
// contiguous range [start[c], start[c+1]) of one layer whose arrays hold all
// clusters, so neighbour indices are global.  One macro step (dt of the
// slowest cluster) runs, per cluster step,
//   predict(c, t)  - ader + local integral (TimeCluster::computeLocalIntegration,
//                    TimeCluster.cpp:358-434): buffer (stored or accumulated
//                    over the two sub-steps a slower neighbour sees) and
//                    derivatives of the cells that provide them,
//   step(c-1, t), step(c-1, t + dt_{c-1})  - the two steps of the faster cluster,
//   correct(c, t)  - neighbour integral (TimeCluster::computeNeighboringIntegration,
//                    TimeCluster.cpp:540-628) from buffers or from derivatives
//                    over [t, t + dt_c] (TimeCommon.cpp:40-87).
// A faster cluster reads a slower neighbour's derivatives, expanded at the
// start T of that neighbour's step, over [t - T, t - T + dt_c] (the
// subTimeStart of TimeCluster.cpp:799); the slower cluster reads the faster
// cluster's buffer summed over its two steps.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "engine.h"

using namespace sdg;

namespace {

constexpr uint16_t kNHD0 = 1u << 0;  // neighbour on face 0 provides derivatives
constexpr uint16_t kGTS0 = 1u << 4;  // neighbour on face 0 is in the same cluster
constexpr uint16_t kBuffers = SDG_LTS_HAS_BUFFERS;
constexpr uint16_t kDerivatives = SDG_LTS_HAS_DERIVATIVES;
constexpr uint16_t kAccumulate = SDG_LTS_ACCUMULATE_BUFFERS;

bool ltsLayerOk(const sdg_layer* L, bool neighbors, const char* who) {
  if (L == nullptr || L->num_elements < 0) {
    setError(std::string(who) + ": invalid layer");
    return false;
  }
  if (L->num_elements == 0) return true;
  if (L->dofs == nullptr || L->time_integrated == nullptr || L->star == nullptr || L->aplus == nullptr ||
      L->lts_setup == nullptr || L->derivatives == nullptr) {
    setError(std::string(who) + ": dofs, buffers, derivatives, star, aplus and lts_setup are required");
    return false;
  }
  if (neighbors && (L->face_neighbors == nullptr || L->face_relations == nullptr || L->aminus == nullptr)) {
    setError(std::string(who) + ": neighbour data missing");
    return false;
  }
  return true;
}

// Local-pass variants: the FP64/FP32 engines have the kTileVariants VALU tile
// shapes; a matrix-core engine has its matrix-core pass (variant 0) and the
// FP32 VALU shapes after it.  The neighbour pass is the VALU kernel of the
// engine's real type (FP32 for the matrix-core engines) with kTileVariants
// shapes, then the one-element-per-row kernel (kNeighborRowsVariant).
// The FP64 engine (orders 3-7, kernel family AUTO) likewise has its FP64
// matrix-core pass as variant 0.
bool matrixCoreLocal(const EngineState& e) { return matrixCore(e.arithmetic) || hasMfma64(e); }
// Neighbour-pass variant before any tuning: one element per 16-lane row for
// the FP32 kernel at orders 3-6 and the FP64 kernel at orders 4-5 (where
// the autotuner picked it on every cluster), else tile 0.
int defaultNeighborVariant(const EngineState& e) {
  const bool fp64 = e.arithmetic == SDG_FP64;
  const bool rows = fp64 ? (e.order >= 4 && e.order <= 5) : (e.order >= 3 && e.order <= 6);
  return rows ? kNeighborRowsVariant : 0;
}
int localVariants(const EngineState& e) { return matrixCoreLocal(e) ? 1 + kTileVariants : kTileVariants; }

int runLocal(const EngineState& e, const double* coeffs, const sdg_layer* own, int reset, hipStream_t s, int v) {
  if (!matrixCoreLocal(e)) return launchLocalIntegration(e, coeffs, own, reset, s, v);
  if (v == 0)
    return matrixCore(e.arithmetic) ? launchLocalIntegrationMfma(e, coeffs, own, reset, s)
                                    : launchLocalIntegrationMfma64(e, coeffs, own, reset, s);
  return launchLocalIntegration(e, coeffs, own, reset, s, v - 1);
}

bool ltsArithmetic(const sdg_engine* e, const char* who) {
  if (matrixCore(e->s.arithmetic) && (e->s.order < 3 || e->s.order > 7)) {
    setError(std::string(who) + ": the matrix-core forms run orders 3..7");
    return false;
  }
  return true;
}

}  // namespace

extern "C" int sdg_derive_lts_setups(int64_t n, const int32_t* cluster_ids, const int32_t* neighbors,
                                     const uint8_t* face_types, uint16_t* setups) {
  if (n < 0 || (n > 0 && (cluster_ids == nullptr || neighbors == nullptr || setups == nullptr))) {
    setError("sdg_derive_lts_setups: invalid arguments");
    return SDG_ERR_INVALID;
  }
  for (int64_t e = 0; e < n * 4; ++e)
    if (neighbors[e] < 0 || neighbors[e] >= n) {
      setError("sdg_derive_lts_setups: neighbour index out of range");
      return SDG_ERR_INVALID;
    }
  // getLtsSetup (LtsSetup.cpp:97-182)
  for (int64_t e = 0; e < n; ++e) {
    uint16_t s = 0;
    const int own = cluster_ids[e];
    for (int f = 0; f < 4; ++f) {
      const uint8_t type = face_types != nullptr ? face_types[e * 4 + f] : SDG_FACE_REGULAR;
      if (type == SDG_FACE_DYNAMIC_RUPTURE) {
        s |= (kNHD0 << f) | (kGTS0 << f) | kDerivatives;
        continue;
      }
      const int nc = cluster_ids[neighbors[e * 4 + f]];
      if (nc > own + 1 || own > nc + 1) {
        // the neighbour integral takes the sub-interval of the next cluster only
        setError("sdg_derive_lts_setups: face neighbours must be at most one cluster apart");
        return SDG_ERR_INVALID;
      }
      if (own < nc) s |= (kNHD0 << f) | kAccumulate;
      else if (own == nc) s |= (kGTS0 << f);
      s |= own > nc ? kDerivatives : kBuffers;
    }
    if ((s & kAccumulate) && (s & (0xfu << 4))) s |= kDerivatives;
    setups[e] = s;
  }
  // normalizeLtsSetup (LtsSetup.cpp:205-217): a same-cluster neighbour whose
  // buffer accumulates over several steps provides derivatives instead
  std::vector<uint16_t> first(setups, setups + n);
  for (int64_t e = 0; e < n; ++e)
    for (int f = 0; f < 4; ++f)
      if ((first[e] & (kGTS0 << f)) && (first[neighbors[e * 4 + f]] & kAccumulate)) setups[e] |= (kNHD0 << f);
  return SDG_OK;
}

extern "C" int sdg_compute_local_integration(sdg_engine* engine, const double* coeffs, double /*time_step_width*/,
                                             const sdg_layer* layer, int reset_buffers, void* stream) {
  if (engine == nullptr || coeffs == nullptr) {
    setError("sdg_compute_local_integration: NULL argument");
    return SDG_ERR_INVALID;
  }
  if (!ltsArithmetic(engine, "sdg_compute_local_integration") ||
      !ltsLayerOk(layer, false, "sdg_compute_local_integration") ||
      !layerStridesOk(engine->s, layer, "sdg_compute_local_integration"))
    return SDG_ERR_INVALID;
  return runLocal(engine->s, coeffs, layer, reset_buffers != 0, static_cast<hipStream_t>(stream), 0);
}

extern "C" int sdg_compute_neighboring_integration(sdg_engine* engine, const double* time_coeffs,
                                                   const double* subtime_coeffs, const sdg_layer* layer,
                                                   void* stream) {
  if (engine == nullptr || time_coeffs == nullptr || subtime_coeffs == nullptr) {
    setError("sdg_compute_neighboring_integration: NULL argument");
    return SDG_ERR_INVALID;
  }
  if (!ltsArithmetic(engine, "sdg_compute_neighboring_integration") ||
      !ltsLayerOk(layer, true, "sdg_compute_neighboring_integration") ||
      !layerStridesOk(engine->s, layer, "sdg_compute_neighboring_integration"))
    return SDG_ERR_INVALID;
  return launchNeighboringIntegration(engine->s, time_coeffs, subtime_coeffs, layer,
                                      static_cast<hipStream_t>(stream), defaultNeighborVariant(engine->s));
}

struct sdg_lts {
  sdg_engine* engine = nullptr;
  sdg_layer layer{};
  std::vector<int64_t> start;
  double dt0 = 0.0;
  std::vector<int> tileLocal, tileNeighbor;  // tile variant per cluster
  hipStream_t captureStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// view of cluster c: own arrays offset to the cluster, neighbour sources global
sdg_layer clusterView(const sdg_lts& L, int c, bool neighborPass) {
  const int64_t rs = L.engine->s.arithmetic == SDG_FP64 ? 8 : 4;
  const CellLayout lay = cellLayout(L.engine->s, &L.layer);
  const int64_t q = lay.qStride;
  const int64_t off = L.start[c];
  const sdg_layer& g = L.layer;
  sdg_layer v = g;
  auto shift = [&](const void* p, int64_t bytes) -> char* {
    return p == nullptr ? nullptr : const_cast<char*>(static_cast<const char*>(p)) + off * bytes;
  };
  v.num_elements = L.start[c + 1] - L.start[c];
  v.dofs = shift(g.dofs, q * rs);
  v.star = shift(g.star, lay.starStride * rs);
  v.aplus = shift(g.aplus, lay.aplusStride * rs);
  v.aminus = shift(g.aminus, lay.aminusStride * rs);
  v.face_neighbors = reinterpret_cast<const int32_t*>(shift(g.face_neighbors, 16));
  v.face_relations = reinterpret_cast<const uint8_t*>(shift(g.face_relations, 8));
  v.face_types = reinterpret_cast<const uint8_t*>(shift(g.face_types, 4));
  v.lts_setup = reinterpret_cast<const uint16_t*>(shift(g.lts_setup, 2));
  v.face_traces = nullptr;
  if (!neighborPass) {
    v.time_integrated = shift(g.time_integrated, q * rs);
    v.derivatives = shift(g.derivatives, lay.dsize * rs);
  }
  return v;
}

// one step of cluster c starting at time k * dt0, with its faster clusters nested
int stepCluster(const sdg_lts& L, int c, int64_t k, hipStream_t s) {
  const int order = L.engine->s.order;
  const int C = static_cast<int>(L.start.size()) - 1;
  const double dtc = L.dt0 * static_cast<double>(int64_t(1) << c);
  const int64_t stepIndex = k >> c;  // steps of cluster c since the macro step began
  std::vector<double> timeCoeffs(order), subCoeffs(order);
  sdg_monomial_integrate(order, 0.0, dtc, timeCoeffs.data());
  const sdg_layer own = clusterView(L, c, false);
  int rc = SDG_OK;
  if (own.num_elements > 0)
    rc = runLocal(L.engine->s, timeCoeffs.data(), &own, (stepIndex & 1) == 0, s, L.tileLocal[c]);
  if (rc != SDG_OK) return rc;
  if (c > 0) {
    if ((rc = stepCluster(L, c - 1, k, s)) != SDG_OK) return rc;
    if ((rc = stepCluster(L, c - 1, k + (int64_t(1) << (c - 1)), s)) != SDG_OK) return rc;
  }
  // sub-interval of the slower cluster's step
  if (c + 1 < C) {
    const int64_t slowStart = (k >> (c + 1)) << (c + 1);
    const double sub = L.dt0 * static_cast<double>(k - slowStart);
    sdg_monomial_integrate(order, sub, sub + dtc, subCoeffs.data());
  } else {
    subCoeffs = timeCoeffs;
  }
  const sdg_layer nbv = clusterView(L, c, true);
  if (nbv.num_elements > 0)
    rc = launchNeighboringIntegration(L.engine->s, timeCoeffs.data(), subCoeffs.data(), &nbv, s, L.tileNeighbor[c]);
  return rc;
}

int macroStep(const sdg_lts& L, hipStream_t s) {
  return stepCluster(L, static_cast<int>(L.start.size()) - 2, 0, s);
}

}  // namespace

extern "C" int sdg_lts_create(sdg_lts** out, sdg_engine* engine, const sdg_layer* layer, int num_clusters,
                              const int64_t* cluster_start, double dt0) {
  if (out == nullptr || engine == nullptr || cluster_start == nullptr || num_clusters < 1 || num_clusters > 16 ||
      !(dt0 > 0.0)) {
    setError("sdg_lts_create: invalid arguments (1..16 clusters, dt0 > 0)");
    return SDG_ERR_INVALID;
  }
  if (!ltsArithmetic(engine, "sdg_lts_create") || !ltsLayerOk(layer, true, "sdg_lts_create") ||
      !layerStridesOk(engine->s, layer, "sdg_lts_create"))
    return SDG_ERR_INVALID;
  if (cluster_start[0] != 0 || cluster_start[num_clusters] != layer->num_elements) {
    setError("sdg_lts_create: cluster_start must run from 0 to num_elements");
    return SDG_ERR_INVALID;
  }
  for (int c = 0; c < num_clusters; ++c)
    if (cluster_start[c + 1] < cluster_start[c]) {
      setError("sdg_lts_create: cluster_start must not decrease");
      return SDG_ERR_INVALID;
    }
  auto* L = new sdg_lts;
  L->engine = engine;
  L->layer = *layer;
  L->start.assign(cluster_start, cluster_start + num_clusters + 1);
  L->dt0 = dt0;
  L->tileLocal.assign(num_clusters, 0);
  L->tileNeighbor.assign(num_clusters, defaultNeighborVariant(engine->s));
  *out = L;
  return SDG_OK;
}

// Tile autotuner: times every tile variant of both kernels on every cluster
// (`repeats` launches each, HIP events on `stream`) and keeps the fastest;
// the cells' state is saved before and restored afterwards.
extern "C" int sdg_lts_autotune(sdg_lts* lts, int repeats, void* stream) {
  if (lts == nullptr || repeats < 1) {
    setError("sdg_lts_autotune: invalid arguments");
    return SDG_ERR_INVALID;
  }
  const auto st = static_cast<hipStream_t>(stream);
  const int order = lts->engine->s.order;
  const int64_t n = lts->layer.num_elements;
  const int64_t rs = lts->engine->s.arithmetic == SDG_FP64 ? 8 : 4;
  const CellLayout lay = cellLayout(lts->engine->s, &lts->layer);
  const int64_t qBytes = n * lay.qStride * rs;
  const int64_t dBytes = n * lay.dsize * rs;
  void* saved[3] = {nullptr, nullptr, nullptr};
  void* live[3] = {lts->layer.dofs, lts->layer.time_integrated, lts->layer.derivatives};
  const int64_t bytes[3] = {qBytes, qBytes, dBytes};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = SDG_OK;
  // every failure from here on goes through the one cleanup path below: the
  // state is restored if it was saved, and what was allocated is released
  for (int i = 0; i < 3 && rc == SDG_OK; ++i) {
    hipError_t err = hipMalloc(&saved[i], bytes[i]);
    if (err != hipSuccess) saved[i] = nullptr;
    if (err == hipSuccess) err = hipMemcpyAsync(saved[i], live[i], bytes[i], hipMemcpyDeviceToDevice, st);
    if (err != hipSuccess) rc = hipFail(err, "sdg_lts_autotune: saving the state");
  }
  if (rc == SDG_OK) {
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    if (err != hipSuccess) rc = hipFail(err, "sdg_lts_autotune: events");
  }
  const bool stateSaved = rc == SDG_OK;
  const int C = static_cast<int>(lts->start.size()) - 1;
  for (int c = 0; c < C && rc == SDG_OK; ++c) {
    const double dtc = lts->dt0 * static_cast<double>(int64_t(1) << c);
    std::vector<double> coeffs(order);
    sdg_monomial_integrate(order, 0.0, dtc, coeffs.data());
    const sdg_layer own = clusterView(*lts, c, false), nbv = clusterView(*lts, c, true);
    if (own.num_elements == 0) continue;
    for (int kind = 0; kind < 2 && rc == SDG_OK; ++kind) {
      float best = 0.0f;
      int bestV = 0;
      const int variants = kind == 0 ? localVariants(lts->engine->s) : kNeighborVariants;
      for (int v = 0; v < variants && rc == SDG_OK; ++v) {
        (void)hipEventRecord(e0, st);
        for (int r = 0; r < repeats && rc == SDG_OK; ++r)
          rc = kind == 0 ? runLocal(lts->engine->s, coeffs.data(), &own, 1, st, v)
                         : launchNeighboringIntegration(lts->engine->s, coeffs.data(), coeffs.data(), &nbv, st, v);
        (void)hipEventRecord(e1, st);
        if (rc != SDG_OK) break;
        const hipError_t err = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (err != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = hipFail(err != hipSuccess ? err : hipErrorUnknown, "sdg_lts_autotune timing");
          break;
        }
        if (v == 0 || ms < best) {
          best = ms;
          bestV = v;
        }
      }
      (kind == 0 ? lts->tileLocal : lts->tileNeighbor)[c] = bestV;
    }
  }
  // restore whenever the state was saved, also after a failed tuning launch
  // (the first error is the one reported)
  if (stateSaved)
    for (int i = 0; i < 3; ++i) {
      const hipError_t err = hipMemcpyAsync(live[i], saved[i], bytes[i], hipMemcpyDeviceToDevice, st);
      if (err != hipSuccess && rc == SDG_OK) rc = hipFail(err, "sdg_lts_autotune restore");
    }
  (void)hipStreamSynchronize(st);
  for (int i = 0; i < 3; ++i)
    if (saved[i] != nullptr) (void)hipFree(saved[i]);
  if (e0 != nullptr) (void)hipEventDestroy(e0);
  if (e1 != nullptr) (void)hipEventDestroy(e1);
  // a captured macro step used the previous tiles
  if (lts->exec != nullptr) {
    (void)hipGraphExecDestroy(lts->exec);
    lts->exec = nullptr;
  }
  if (lts->graph != nullptr) {
    (void)hipGraphDestroy(lts->graph);
    lts->graph = nullptr;
  }
  return rc;
}

extern "C" int sdg_lts_set_tile_variant(sdg_lts* lts, int cluster, int kernel, int variant) {
  if (lts == nullptr || cluster < 0 || cluster >= static_cast<int>(lts->tileLocal.size()) || kernel < 0 ||
      kernel > 1 || variant < 0 || variant >= (kernel == 0 ? localVariants(lts->engine->s) : kNeighborVariants)) {
    setError("sdg_lts_set_tile_variant: invalid arguments");
    return SDG_ERR_INVALID;
  }
  (kernel == 0 ? lts->tileLocal : lts->tileNeighbor)[cluster] = variant;
  if (lts->exec != nullptr) {
    (void)hipGraphExecDestroy(lts->exec);
    lts->exec = nullptr;
  }
  if (lts->graph != nullptr) {
    (void)hipGraphDestroy(lts->graph);
    lts->graph = nullptr;
  }
  return SDG_OK;
}

extern "C" int sdg_lts_tile_variant(const sdg_lts* lts, int cluster, int kernel) {
  if (lts == nullptr || cluster < 0 || cluster >= static_cast<int>(lts->tileLocal.size()) || kernel < 0 ||
      kernel > 1) {
    setError("sdg_lts_tile_variant: invalid arguments");
    return SDG_ERR_INVALID;
  }
  return kernel == 0 ? lts->tileLocal[cluster] : lts->tileNeighbor[cluster];
}

extern "C" int sdg_lts_advance(sdg_lts* lts, int macro_steps, void* stream) {
  if (lts == nullptr || macro_steps < 0) {
    setError("sdg_lts_advance: invalid arguments");
    return SDG_ERR_INVALID;
  }
  for (int m = 0; m < macro_steps; ++m) {
    const int rc = macroStep(*lts, static_cast<hipStream_t>(stream));
    if (rc != SDG_OK) return rc;
  }
  return SDG_OK;
}

extern "C" int sdg_lts_graph_advance(sdg_lts* lts, int macro_steps, void* stream) {
  if (lts == nullptr || macro_steps < 0) {
    setError("sdg_lts_graph_advance: invalid arguments");
    return SDG_ERR_INVALID;
  }
  if (lts->exec == nullptr) {
    // capture one macro step on a private stream (default priority)
    if (lts->captureStream == nullptr) SDG_HIP_CHECK(hipStreamCreateWithFlags(&lts->captureStream, hipStreamNonBlocking));
    SDG_HIP_CHECK(hipStreamBeginCapture(lts->captureStream, hipStreamCaptureModeThreadLocal));
    const int rc = macroStep(*lts, lts->captureStream);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(lts->captureStream, &g);
    if (rc != SDG_OK) {
      if (g != nullptr) (void)hipGraphDestroy(g);
      return rc;
    }
    SDG_HIP_CHECK(end);
    lts->graph = g;
    SDG_HIP_CHECK(hipGraphInstantiate(&lts->exec, lts->graph, nullptr, nullptr, 0));
  }
  for (int m = 0; m < macro_steps; ++m) SDG_HIP_CHECK(hipGraphLaunch(lts->exec, static_cast<hipStream_t>(stream)));
  return SDG_OK;
}

extern "C" int sdg_lts_destroy(sdg_lts* lts) {
  if (lts == nullptr) return SDG_OK;
  if (lts->exec != nullptr) (void)hipGraphExecDestroy(lts->exec);
  if (lts->graph != nullptr) (void)hipGraphDestroy(lts->graph);
  if (lts->captureStream != nullptr) (void)hipStreamDestroy(lts->captureStream);
  delete lts;
  return SDG_OK;
}
