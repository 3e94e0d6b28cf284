// C++ host interface of seissol_amd: the reference's kernel and friction
// call surfaces over the C ABI (include/seissol_amd.h).
//
//   seissol_amd::GlobalData                      setGlobalData / GlobalDataInitializer
//   seissol_amd::kernels::Spacetime::computeBatchedAder   src/Kernels/Spacetime.h:52-69
//   seissol_amd::kernels::Time::evaluateBatched           src/Kernels/Time.h:42-61
//   seissol_amd::kernels::Local::computeBatchedIntegral   src/Kernels/Local.h:58-79
//   seissol_amd::kernels::Neighbor::computeBatchedNeighborsIntegral  src/Kernels/Neighbor.h:44-55
//     (with the nodal flux of the layer's dynamic-rupture faces, Neighbor.cpp:93-101)
//   seissol_amd::kernels::DynamicRupture::batchedSpaceTimeInterpolation  src/Kernels/DynamicRupture.h:30-58
//   *::metrics()                                 src/Monitoring/Metric.h:16-50
//   seissol_amd::dr::friction_law::FrictionSolver  src/DynamicRupture/FrictionLaws/FrictionSolver.h:28-108
//
// The batched calls take the layer (an sdg_layer of device arrays) and a
// StreamRuntime holding the hipStream_t, as the reference takes its recorded
// tables and its runtime.  A failing call throws std::runtime_error with the
// library's message, where the reference logs a fatal error.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/seissol_amd.h"

namespace seissol_amd {

inline void check(int rc, const char* what) {
  if (rc != SDG_OK) throw std::runtime_error(std::string(what) + ": " + sdg_last_error());
}

enum class Arithmetic : int {
  FP64 = SDG_FP64,
  FP32 = SDG_FP32,
  BF16x3 = SDG_BF16X3,
  BF16x6 = SDG_BF16X6,
  F16x3 = SDG_F16X3
};

inline Arithmetic arithmeticFromString(const std::string& s) {
  if (s == "fp64") return Arithmetic::FP64;
  if (s == "fp32") return Arithmetic::FP32;
  if (s == "bf16x3") return Arithmetic::BF16x3;
  if (s == "bf16x6") return Arithmetic::BF16x6;
  if (s == "f16x3") return Arithmetic::F16x3;
  throw std::runtime_error("unknown arithmetic " + s + " (fp64, fp32, bf16x3, bf16x6, f16x3)");
}

/// Stream on which the batched kernels are enqueued (the reference's StreamRuntime).
struct StreamRuntime {
  void* stream = nullptr;  // hipStream_t
};

/// Work per element-update of a kernel (PerformanceEstimate).
struct PerformanceEstimate {
  double hardwareFlop = 0, nonzeroFlop = 0, bytes = 0, kernelBytes = 0;
  PerformanceEstimate operator+(const PerformanceEstimate& o) const {
    return {hardwareFlop + o.hardwareFlop, nonzeroFlop + o.nonzeroFlop, bytes + o.bytes,
            kernelBytes + o.kernelBytes};
  }
  PerformanceEstimate operator*(double n) const { return {hardwareFlop * n, nonzeroFlop * n, bytes * n, kernelBytes * n}; }
};

inline PerformanceEstimate estimate(int order, Arithmetic a, int kernelId) {
  sdg_work_estimate e{};
  check(sdg_performance_estimate(order, static_cast<int>(a), kernelId, &e), "sdg_performance_estimate");
  return {e.hardware_flop, e.nonzero_flop, e.bytes, e.kernel_bytes};
}

/// Global matrices of one order resident on the device in one arithmetic.
class GlobalData {
 public:
  GlobalData(int order, Arithmetic arithmetic) : order_(order), arithmetic_(arithmetic) {
    check(sdg_engine_create(&engine_, order, static_cast<int>(arithmetic)), "sdg_engine_create");
  }
  GlobalData(const GlobalData&) = delete;
  GlobalData& operator=(const GlobalData&) = delete;
  ~GlobalData() {
    if (engine_ != nullptr) sdg_engine_destroy(engine_);
  }
  /// Replace the generated matrices by caller data (layout of sdg_global_matrices).
  void setMatrices(const std::vector<double>& flat) {
    check(sdg_engine_set_global_matrices(engine_, flat.data()), "sdg_engine_set_global_matrices");
  }
  /// Storage layout of the cells' arrays (the reference fixes it at build time:
  /// 128-byte alignment gives [9][64] DOF blocks at ORDER 6) and star format
  /// (SDG_STAR_DENSE / SDG_STAR_COMPACT).
  void setLayout(int alignment, int starFormat) {
    check(sdg_engine_set_layout(engine_, alignment, starFormat), "sdg_engine_set_layout");
  }
  /// Kernel family: SDG_KERNELS_AUTO (matrix cores where available) or SDG_KERNELS_VALU.
  void setKernels(int family) { check(sdg_engine_set_kernels(engine_, family), "sdg_engine_set_kernels"); }
  sdg_engine* engine() const { return engine_; }
  int order() const { return order_; }
  Arithmetic arithmetic() const { return arithmetic_; }

 private:
  sdg_engine* engine_ = nullptr;
  int order_;
  Arithmetic arithmetic_;
};

namespace kernels {

class Spacetime {
 public:
  void setGlobalData(const GlobalData& g) { global_ = &g; }
  void computeBatchedAder(const double* coeffs, double timeStepWidth, const sdg_layer& layer, StreamRuntime& rt) {
    check(sdg_compute_batched_ader(global_->engine(), coeffs, timeStepWidth, &layer, rt.stream),
          "Spacetime::computeBatchedAder");
  }
  /// ader fused with the local integral (the proxy's `local` kernel)
  void computeBatchedAderLocal(const double* coeffs, double timeStepWidth, const sdg_layer& layer,
                               StreamRuntime& rt) {
    check(sdg_compute_batched_ader_local(global_->engine(), coeffs, timeStepWidth, &layer, rt.stream),
          "Spacetime::computeBatchedAderLocal");
  }
  PerformanceEstimate metrics() const { return estimate(global_->order(), global_->arithmetic(), 3); }

 private:
  const GlobalData* global_ = nullptr;
};

class Time {
 public:
  void setGlobalData(const GlobalData& g) { global_ = &g; }
  void evaluateBatched(const double* coeffs, const void* timeDerivatives, void* timeEvaluated,
                       std::int64_t numElements, StreamRuntime& rt) {
    check(sdg_evaluate_batched(global_->engine(), coeffs, timeDerivatives, timeEvaluated, numElements, rt.stream),
          "Time::evaluateBatched");
  }

 private:
  const GlobalData* global_ = nullptr;
};

class Local {
 public:
  void setGlobalData(const GlobalData& g) { global_ = &g; }
  void computeBatchedIntegral(const sdg_layer& layer, StreamRuntime& rt) {
    check(sdg_compute_batched_integral(global_->engine(), &layer, rt.stream), "Local::computeBatchedIntegral");
  }
  PerformanceEstimate metrics() const { return estimate(global_->order(), global_->arithmetic(), 4); }

 private:
  const GlobalData* global_ = nullptr;
};

class Neighbor {
 public:
  void setGlobalData(const GlobalData& g) { global_ = &g; }
  void computeBatchedNeighborsIntegral(const sdg_layer& layer, StreamRuntime& rt) {
    check(sdg_compute_batched_neighbors_integral(global_->engine(), &layer, rt.stream),
          "Neighbor::computeBatchedNeighborsIntegral");
  }
  /// regular faces, then the nodal flux of the dynamic-rupture faces from
  /// their imposed states (imposedPlus/Minus [faces][9][points], FP64)
  void computeBatchedNeighborsIntegral(const sdg_layer& layer, const int32_t* drMapping, const sdg_dr_layer& faces,
                                       const double* imposedPlus, const double* imposedMinus, StreamRuntime& rt) {
    computeBatchedNeighborsIntegral(layer, rt);
    check(sdg_dr_nodal_flux(global_->engine(), &layer, drMapping, &faces, imposedPlus, imposedMinus, rt.stream),
          "Neighbor::computeBatchedNeighborsIntegral (nodal flux)");
  }
  PerformanceEstimate metrics() const { return estimate(global_->order(), global_->arithmetic(), 2); }

 private:
  const GlobalData* global_ = nullptr;
};

/// Space-time interpolation of the dynamic-rupture faces (DynamicRupture.h:30-58).
class DynamicRupture {
 public:
  void setGlobalData(const GlobalData& g) { global_ = &g; }
  /// coeffsCollocate [order][order] (c[o][d] = t_o^d / d!), derivatives of the
  /// layer's cells; qPlus/qMinus [faces][order][9][points] FP64
  void batchedSpaceTimeInterpolation(const sdg_dr_layer& faces, const double* coeffsCollocate,
                                     const void* derivatives, double* qPlus, double* qMinus, StreamRuntime& rt) {
    check(sdg_dr_space_time_interpolation(global_->engine(), coeffsCollocate, derivatives, &faces, qPlus, qMinus,
                                          rt.stream),
          "DynamicRupture::batchedSpaceTimeInterpolation");
  }
  /// work per DR face (the proxy's godunov_dr estimate counts 4 faces per cell)
  PerformanceEstimate metrics() const { return estimate(global_->order(), global_->arithmetic(), 6) * 0.25; }

 private:
  const GlobalData* global_ = nullptr;
};

}  // namespace kernels

namespace dr {
namespace friction_law {

/// FrictionSolver of the fast-velocity-weakening rate-and-state law (103)
/// on device data (FrictionSolver.h:38-65).  Trivially copyable, like the
/// reference's solver objects; clone() returns a copy.
class FrictionSolver {
 public:
  FrictionSolver(const sdg_friction_params& params, const double* deviceResample)
      : params_(params), resample_(deviceResample) {}

  static std::vector<double> computeDeltaT(const std::vector<double>& timePoints) {
    std::vector<double> dt(timePoints.size());
    check(sdg_friction_delta_t(static_cast<int>(timePoints.size()), timePoints.data(), dt.data()),
          "FrictionSolver::computeDeltaT");
    return dt;
  }
  void setupLayer(const sdg_fault_layer& layer) { layer_ = layer; }
  /// warm start of the first slip-rate solve of each time point: creeping
  /// points at the closed-form creep root, the others at the MLP's predicted
  /// root (nullptr: no network); (false, nullptr) is plain Newton
  void setWarmStart(bool closedForm, const float* deviceMlpParams) {
    options_.closed_form_start = closedForm ? 1 : 0;
    options_.mlp_params = deviceMlpParams;
  }
  void setSampleRecording(float* deviceSamples, std::int64_t capacity, unsigned long long* deviceCount,
                          int stride) {
    options_.samples = deviceSamples;
    options_.capacity = capacity;
    options_.sample_count = deviceCount;
    options_.sample_stride = stride;
  }
  void evaluate(double fullUpdateTime, const std::vector<double>& deltaT, const std::vector<double>& timeWeights,
                StreamRuntime& rt) {
    if (static_cast<int>(deltaT.size()) != layer_.order || static_cast<int>(timeWeights.size()) != layer_.order)
      throw std::runtime_error("FrictionSolver::evaluate: deltaT and timeWeights need one value per time point");
    check(sdg_friction_evaluate(&params_, &layer_, fullUpdateTime, deltaT.data(), timeWeights.data(), resample_,
                                &options_, rt.stream),
          "FrictionSolver::evaluate");
  }
  static const char* allocationPlace() { return "device"; }
  FrictionSolver clone() const { return *this; }

 private:
  sdg_friction_params params_;
  const double* resample_;
  sdg_fault_layer layer_{};
  sdg_friction_options options_{};
};

}  // namespace friction_law
}  // namespace dr

namespace proxy {

/// Proxy kernel selection by name (src/Proxy/Common.cpp:22-41).
inline int kernelId(const std::string& name) {
  const int id = sdg_proxy_kernel_id(name.c_str());
  if (id < 0) throw std::runtime_error("Unknown kernel string: " + name);
  return id;
}

}  // namespace proxy
}  // namespace seissol_amd
