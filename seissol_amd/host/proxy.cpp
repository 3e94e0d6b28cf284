// seissol_amd_proxy: the SeisSol proxy benchmark on seissol_amd.
//
// Usage (as app/Proxy/Main.cpp:49-66 plus this engine's options):
//   seissol_amd_proxy [--order O] [--arithmetic fp64|fp32|bf16x3|bf16x6|f16x3]
//                     [--mesh proxy|periodic] [--seed S] [-f plain|json]
//                     [--dump FILE] <cells> <timesteps> <kernel[,kernel...]>
// --dump writes the final DOFs ([cells][9][B], FP64, native byte order).
// Kernels: all, local, neigh, ader, localwoader, neigh_dr, godunov_dr, all_dr
// (src/Proxy/Common.cpp:22-41); a comma-separated list runs its kernels in
// sequence in every pass.  With a DR kernel every cell face is a DR face and
// the layer gets the proxy's random DR faces, derivatives and imposed states
// (src/Proxy/Allocator.cpp:186-294; every arithmetic, the DR kernels in the
// engine's real type).
//
// Data as src/Proxy/Allocator.cpp:49-134 (random affine tetrahedra with
// random neighbours and face relations; or a conforming periodic box),
// rho ~ U[2000, 3000], vp 6000, vs 3464, DOFs ~ U[-1, 1], dt from CFL 0.5.
// Timing as src/Proxy/Runner.cpp:57-154: one untimed pass, then `timesteps`
// passes between device synchronisations; the summary follows
// Aux::writeOutput (src/Proxy/Common.cpp:87-160).  The kernels 0-4 run
// through the C++ kernel classes as the reference's device proxy runs them
// (src/Proxy/KernelDevice.cpp:32-124): ader = Spacetime::computeBatchedAder,
// localwoader = Local::computeBatchedIntegral, local = ader fused with the
// local integral, neigh = Neighbor::computeBatchedNeighborsIntegral, all =
// local then neigh.  The matrix-core forms' local pass writes the face traces
// the neighbour pass reads, so there the layer carries no time_integrated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "seissol_amd.hpp"

namespace {

using seissol_amd::Arithmetic;
using seissol_amd::check;

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
T* deviceCopy(const std::vector<T>& host) {
  T* d = nullptr;
  hipCheck(hipMalloc(&d, host.size() * sizeof(T)), "hipMalloc");
  hipCheck(hipMemcpy(d, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return d;
}

void* deviceZeros(size_t bytes) {
  void* d = nullptr;
  hipCheck(hipMalloc(&d, bytes), "hipMalloc");
  hipCheck(hipMemset(d, 0, bytes), "hipMemset");
  return d;
}

// even box dimensions (>= 4) whose 5-tet periodic mesh has about `target` cells
void periodicDims(int64_t target, int dims[3]) {
  const int64_t cubes = std::max<int64_t>(64, target / 5);
  int64_t best = -1;
  for (int nx = 4; nx < 400; nx += 2)
    for (int ny = nx; ny < 400; ny += 2) {
      int64_t nz = cubes / (static_cast<int64_t>(nx) * ny);
      nz -= nz % 2;
      if (nz < ny) break;
      const int64_t got = 5LL * nx * ny * nz;
      if (best < 0 || std::llabs(got - target) < std::llabs(best - target)) {
        best = got;
        dims[0] = nx;
        dims[1] = ny;
        dims[2] = static_cast<int>(nz);
      }
    }
}

std::vector<std::string> split(const std::string& s) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string part;
  while (std::getline(ss, part, ',')) {
    const auto b = part.find_first_not_of(" \t"), e = part.find_last_not_of(" \t");
    out.push_back(b == std::string::npos ? std::string() : part.substr(b, e - b + 1));
  }
  return out;
}

int usage(const char* msg) {
  if (msg != nullptr) std::cerr << msg << "\n";
  std::cerr << "usage: seissol_amd_proxy [--order O] [--arithmetic fp64|fp32|bf16x3|bf16x6|f16x3] "
               "[--mesh proxy|periodic] [--seed S] [-f plain|json] [--dump FILE] <cells> <timesteps> "
               "<kernel[,kernel...]>\n"
               "kernels: all, local, neigh, ader, localwoader, neigh_dr, godunov_dr, all_dr\n";
  return -1;
}

struct Output {
  double time = 0, cycles = 0, nzGflop = 0, hwGflop = 0, gib = 0, kernelGib = 0;
};

void writeOutput(std::ostream& s, const Output& o, const std::string& kernelStr, bool json) {
  const double nzFpc = o.nzGflop * 1e9 / o.cycles, hwFpc = o.hwGflop * 1e9 / o.cycles;
  const double gib = 1024.0 * 1024.0 * 1024.0;
  const double bpc = o.gib * gib / o.cycles, kbpc = o.kernelGib * gib / o.cycles;
  if (!json) {
    s << "=================================================\n";
    s << "===            PERFORMANCE SUMMARY            ===\n";
    s << "=================================================\n";
    s << "seissol proxy mode                  : " << kernelStr << '\n';
    s << "time for seissol proxy              : " << o.time << '\n';
    s << "cycles                              : " << o.cycles << '\n';
    s << '\n';
    s << "GFLOP (non-zero) for seissol proxy  : " << o.nzGflop << '\n';
    s << "GFLOP (hardware) for seissol proxy  : " << o.hwGflop << '\n';
    s << "GiB (estimate) for seissol proxy    : " << o.gib << '\n';
    s << "GiB (kernels) for seissol proxy     : " << o.kernelGib << '\n';
    s << '\n';
    s << "FLOPS/cycle (non-zero)              : " << nzFpc << '\n';
    s << "FLOPS/cycle (hardware)              : " << hwFpc << '\n';
    s << "Bytes/cycle (estimate)              : " << bpc << '\n';
    s << "Bytes/cycle (kernel)                : " << kbpc << '\n';
    s << '\n';
    s << "GFLOPS (non-zero) for seissol proxy : " << o.nzGflop / o.time << '\n';
    s << "GFLOPS (hardware) for seissol proxy : " << o.hwGflop / o.time << '\n';
    s << "GiB/s (estimate) for seissol proxy  : " << o.gib / o.time << '\n';
    s << "GiB/s (kernel) for seissol proxy    : " << o.kernelGib / o.time << '\n';
    s << "=================================================\n\n";
    return;
  }
  s << "{\"name\":\"" << kernelStr << "\",\"time\":" << o.time << ",\"cycles\":" << o.cycles
    << ",\"gflop-nz\":" << o.nzGflop << ",\"gflop-hw\":" << o.hwGflop << ",\"gib\":" << o.gib
    << ",\"gib-kernel\":" << o.kernelGib << ",\"flop-per-cycle-nz\":" << nzFpc << ",\"flop-per-cycle-hw\":" << hwFpc
    << ",\"bytes-per-cycle\":" << bpc << ",\"bytes-per-cycle-kernel\":" << kbpc
    << ",\"gflops-nz\":" << o.nzGflop / o.time << ",\"gflops-hw\":" << o.hwGflop / o.time
    << ",\"gibs\":" << o.gib / o.time << ",\"gibs-kernel\":" << o.kernelGib / o.time << "}\n";
}

int run(int argc, char** argv) {
  int order = 6;
  std::string arith = "fp64", mesh = "proxy", format = "plain", dumpPath;
  uint64_t seed = 1234;
  std::vector<std::string> positional;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* name) -> std::string {
      if (i + 1 >= argc) throw std::invalid_argument(std::string("missing value for ") + name);
      return argv[++i];
    };
    if (a == "--order") order = std::stoi(next("--order"));
    else if (a == "--arithmetic") arith = next("--arithmetic");
    else if (a == "--mesh") mesh = next("--mesh");
    else if (a == "--seed") seed = std::stoull(next("--seed"));
    else if (a == "-f" || a == "--format") format = next("--format");
    else if (a == "--dump") dumpPath = next("--dump");
    else if (a == "-h" || a == "--help") return usage(nullptr);
    else positional.push_back(a);
  }
  if (positional.size() != 3) return usage("expected <cells> <timesteps> <kernel>");
  if (format != "plain" && format != "json") return usage("format must be plain or json");
  if (mesh != "proxy" && mesh != "periodic") return usage("mesh must be proxy or periodic");
  int64_t cells = std::stoll(positional[0]);
  const int64_t timesteps = std::stoll(positional[1]);
  const std::string kernelStr = positional[2];
  if (cells <= 0 || timesteps <= 0) return usage("cells and timesteps must be positive");
  std::vector<int> kernelIds;
  for (const auto& k : split(kernelStr)) {
    const int id = sdg_proxy_kernel_id(k.c_str());
    if (id < 0) {
      std::cerr << "Unknown kernel string: " << k << std::endl;
      return -1;
    }
    kernelIds.push_back(id);
  }
  const Arithmetic arithmetic = seissol_amd::arithmeticFromString(arith);
  const bool withDR = std::any_of(kernelIds.begin(), kernelIds.end(), [](int id) { return id >= 5; });

  // synthetic data (host, FP64)
  std::cerr << "Allocating fake data..." << std::endl;
  int dims[3] = {0, 0, 0};
  if (mesh == "periodic") {
    periodicDims(cells, dims);
    cells = sdg_mesh_periodic_size(dims[0], dims[1], dims[2]);
  }
  const size_t n = static_cast<size_t>(cells);
  const int B = order * (order + 1) * (order + 2) / 6, B2 = order * (order + 1) / 2;
  std::vector<double> vertices(n * 12), material(n * 3), dofs(n * 9 * B);
  std::vector<int32_t> neighbors(n * 4);
  std::vector<uint8_t> relations(n * 8), faceTypes(n * 4, withDR ? SDG_FACE_DYNAMIC_RUPTURE : SDG_FACE_REGULAR);
  if (mesh == "proxy")
    check(sdg_mesh_proxy(cells, seed, vertices.data(), neighbors.data(), relations.data()), "sdg_mesh_proxy");
  else
    check(sdg_mesh_periodic(dims[0], dims[1], dims[2], 100.0, 0.1, seed, vertices.data(), neighbors.data(),
                            relations.data()),
          "sdg_mesh_periodic");
  check(sdg_random_materials(cells, seed + 1, 2000.0, 3000.0, 6000.0, 3464.0, material.data()),
        "sdg_random_materials");
  check(sdg_uniform_dofs(static_cast<int64_t>(dofs.size()), seed + 2, dofs.data()), "sdg_uniform_dofs");
  const double dt = sdg_cfl_timestep(cells, order, 0.5, vertices.data(), material.data());

  // device layer: FP64 per-element operators on the device, then the engine's real type
  seissol_amd::GlobalData global(order, arithmetic);
  const size_t real = arithmetic == Arithmetic::FP64 ? 8 : 4;
  double* dVert = deviceCopy(vertices);
  double* dMat = deviceCopy(material);
  int32_t* dNb = deviceCopy(neighbors);
  uint8_t* dRel = deviceCopy(relations);
  uint8_t* dTypes = deviceCopy(faceTypes);
  double* star64 = static_cast<double*>(deviceZeros(n * 3 * 81 * 8));
  double* aplus64 = static_cast<double*>(deviceZeros(n * 4 * 81 * 8));
  double* aminus64 = static_cast<double*>(deviceZeros(n * 4 * 81 * 8));
  check(sdg_cell_local_matrices(cells, dVert, dMat, dNb, dTypes, star64, aplus64, aminus64, nullptr),
        "sdg_cell_local_matrices");
  void *star = star64, *aplus = aplus64, *aminus = aminus64;
  if (real == 4) {
    star = deviceZeros(n * 3 * 81 * 4);
    aplus = deviceZeros(n * 4 * 81 * 4);
    aminus = deviceZeros(n * 4 * 81 * 4);
    check(sdg_convert_fp64_to_real(global.engine(), star, star64, cells * 3 * 81, nullptr), "convert");
    check(sdg_convert_fp64_to_real(global.engine(), aplus, aplus64, cells * 4 * 81, nullptr), "convert");
    check(sdg_convert_fp64_to_real(global.engine(), aminus, aminus64, cells * 4 * 81, nullptr), "convert");
  }
  void* dDofs = deviceZeros(n * 9 * B * real);
  check(sdg_upload_reals(global.engine(), dDofs, dofs.data(), static_cast<int64_t>(dofs.size()), nullptr),
        "sdg_upload_reals");
  void* dInt = deviceZeros(n * 9 * B * real);
  const bool traces = arithmetic == Arithmetic::BF16x3 || arithmetic == Arithmetic::BF16x6 ||
                      arithmetic == Arithmetic::F16x3;
  // the FP64 engine's matrix-core passes (orders 3-7) exchange FP64 face traces
  const bool traces64 = arithmetic == Arithmetic::FP64 && order >= 3 && order <= 7;
  const int B2P = (B2 + 3) / 4 * 4;  // face rows of the traces, padded to a multiple of 4
  void* dTraces = traces ? deviceZeros(n * 4 * 9 * B2P * 4)
                         : (traces64 ? deviceZeros(n * 4 * 9 * B2P * 8) : nullptr);
  // dynamic rupture: random DR faces, derivatives and imposed states
  void* dDeriv = nullptr;
  std::vector<void*> drBuffers;
  sdg_proxy_dr dr{};
  if (withDR) {
    const int64_t dsize = sdg_derivatives_size(order), nf = 4 * cells, points = (order + 1) * (order + 1);
    std::vector<double> deriv(n * dsize);
    check(sdg_uniform_dofs(static_cast<int64_t>(deriv.size()), seed + 5, deriv.data()), "sdg_uniform_dofs");
    dDeriv = deviceZeros(deriv.size() * real);
    check(sdg_upload_reals(global.engine(), dDeriv, deriv.data(), static_cast<int64_t>(deriv.size()), nullptr),
          "sdg_upload_reals");
    std::vector<int32_t> mapping(n * 4), plus(nf), minus(nf);
    std::vector<uint8_t> info(nf * 3);
    check(sdg_proxy_dr_faces(cells, seed + 3, mapping.data(), plus.data(), minus.data(), info.data()),
          "sdg_proxy_dr_faces");
    std::vector<double> tinv(nf * 81), fluxPlus(nf * 81), fluxMinus(nf * 81), imposed(2 * nf * 9 * points);
    check(sdg_dr_face_data(nf, vertices.data(), material.data(), plus.data(), minus.data(), info.data(), tinv.data(),
                           fluxPlus.data(), fluxMinus.data(), nullptr),
          "sdg_dr_face_data");
    check(sdg_uniform_dofs(static_cast<int64_t>(imposed.size()), seed + 4, imposed.data()), "sdg_uniform_dofs");
    auto keep = [&](void* p) {
      drBuffers.push_back(p);
      return p;
    };
    dr.dr_mapping = static_cast<int32_t*>(keep(deviceCopy(mapping)));
    dr.faces.num_faces = nf;
    dr.faces.points = static_cast<int32_t>(points);
    dr.faces.plus_element = static_cast<int32_t*>(keep(deviceCopy(plus)));
    dr.faces.minus_element = static_cast<int32_t*>(keep(deviceCopy(minus)));
    dr.faces.face_info = static_cast<uint8_t*>(keep(deviceCopy(info)));
    dr.faces.tinv_t = static_cast<double*>(keep(deviceCopy(tinv)));
    dr.faces.flux_solver_plus = static_cast<double*>(keep(deviceCopy(fluxPlus)));
    dr.faces.flux_solver_minus = static_cast<double*>(keep(deviceCopy(fluxMinus)));
    double* dImposed = static_cast<double*>(keep(deviceCopy(imposed)));
    dr.imposed_plus = dImposed;
    dr.imposed_minus = dImposed + nf * 9 * points;
    dr.q_plus = static_cast<double*>(keep(deviceZeros(nf * order * 9 * points * 8)));
    dr.q_minus = static_cast<double*>(keep(deviceZeros(nf * order * 9 * points * 8)));
  }
  hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
  sdg_layer layer{cells, dDofs, dInt, dDeriv, star, aplus, aminus, dNb, dRel, dTypes, dTraces, nullptr};
  // the matrix-core local pass hands face traces to the neighbour pass, not I
  sdg_layer classLayer = layer;
  if (traces) classLayer.time_integrated = nullptr;
  std::cerr << "...done" << std::endl;

  seissol_amd::StreamRuntime runtime{};
  hipStream_t stream;
  hipCheck(hipStreamCreate(&stream), "hipStreamCreate");
  runtime.stream = stream;
  seissol_amd::kernels::Spacetime spacetime;
  seissol_amd::kernels::Local local;
  seissol_amd::kernels::Neighbor neighbor;
  spacetime.setGlobalData(global);
  local.setGlobalData(global);
  neighbor.setGlobalData(global);
  std::vector<double> coeffs(order);
  check(sdg_monomial_integrate(order, 0.0, dt, coeffs.data()), "sdg_monomial_integrate");
  auto run1 = [&](int id, int steps) {
    if (withDR) {
      check(sdg_proxy_run_dr(global.engine(), id, steps, dt, &layer, &dr, runtime.stream), "sdg_proxy_run_dr");
      return;
    }
    for (int s = 0; s < steps; ++s) switch (id) {
        case 0:
          spacetime.computeBatchedAderLocal(coeffs.data(), dt, classLayer, runtime);
          neighbor.computeBatchedNeighborsIntegral(classLayer, runtime);
          break;
        case 1: spacetime.computeBatchedAderLocal(coeffs.data(), dt, classLayer, runtime); break;
        case 2: neighbor.computeBatchedNeighborsIntegral(classLayer, runtime); break;
        case 3: spacetime.computeBatchedAder(coeffs.data(), dt, layer, runtime); break;
        case 4: local.computeBatchedIntegral(layer, runtime); break;
        default: throw std::runtime_error("unknown proxy kernel");
      }
  };
  auto pass = [&](int64_t steps) {
    if (kernelIds.size() == 1) {
      run1(kernelIds[0], static_cast<int>(steps));
      return;
    }
    for (int64_t t = 0; t < steps; ++t)
      for (int id : kernelIds) run1(id, 1);
  };
  pass(1);  // untimed pass
  hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
  const auto t0 = std::chrono::steady_clock::now();
  pass(timesteps);
  hipCheck(hipStreamSynchronize(stream), "hipStreamSynchronize");
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  seissol_amd::PerformanceEstimate per{};
  for (int id : kernelIds) per = per + seissol_amd::estimate(order, arithmetic, id);
  const double work = static_cast<double>(cells) * static_cast<double>(timesteps);
  int clockKhz = 0;
  hipCheck(hipDeviceGetAttribute(&clockKhz, hipDeviceAttributeClockRate, 0), "hipDeviceGetAttribute");
  Output o;
  o.time = sec;
  o.cycles = sec * clockKhz * 1e3;
  o.nzGflop = per.nonzeroFlop * work * 1e-9;
  o.hwGflop = per.hardwareFlop * work * 1e-9;
  o.gib = per.bytes * work / (1024.0 * 1024.0 * 1024.0);
  o.kernelGib = per.kernelBytes * work / (1024.0 * 1024.0 * 1024.0);
  std::cerr << "order " << order << ", " << arith << ", " << cells << " cells (" << mesh << "), dt " << dt << ": "
            << work / sec << " element-updates/s" << std::endl;
  writeOutput(std::cout, o, kernelStr, format == "json");
  std::cout.flush();
  if (!dumpPath.empty()) {
    std::vector<double> out(dofs.size());
    check(sdg_download_reals(global.engine(), out.data(), dDofs, static_cast<int64_t>(out.size()), nullptr),
          "sdg_download_reals");
    std::ofstream f(dumpPath, std::ios::binary);
    f.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * sizeof(double)));
    if (!f) throw std::runtime_error("cannot write " + dumpPath);
  }

  hipStreamDestroy(stream);
  for (void* p : {static_cast<void*>(dVert), static_cast<void*>(dMat), static_cast<void*>(dNb),
                  static_cast<void*>(dRel), static_cast<void*>(dTypes), static_cast<void*>(star64),
                  static_cast<void*>(aplus64), static_cast<void*>(aminus64), dDofs, dInt, dTraces, dDeriv})
    if (p != nullptr) hipFree(p);
  for (void* p : drBuffers) hipFree(p);
  if (real == 4) {
    hipFree(star);
    hipFree(aplus);
    hipFree(aminus);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "seissol_amd_proxy: " << e.what() << std::endl;
    return -1;
  }
}
