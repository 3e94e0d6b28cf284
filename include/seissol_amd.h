/*
 * seissol_amd -- MI355X (gfx950) engine for the SeisSol elastic ADER-DG
 * element update and the rate-and-state friction solve.
 *
 * C ABI: plain pointers and sizes, int status codes (0 = success), the text
 * of the last failure from sdg_last_error().  Device pointers are HIP device
 * memory owned by the caller; `stream` is a hipStream_t (NULL = null stream).
 * Every compute entry point is asynchronous on `stream` and allocates no
 * device memory unless stated.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the SeisSol source tree).
 */
#ifndef SEISSOL_AMD_H_
#define SEISSOL_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDG_OK 0
#define SDG_ERR_INVALID (-1)
#define SDG_ERR_HIP (-2)
#define SDG_ERR_UNSUPPORTED (-3)

/* arithmetic of the element-update kernels */
#define SDG_FP64 0   /* plain FP64: the drop-in for the reference GPU path */
#define SDG_FP32 1   /* FP32 storage and VALU arithmetic                    */
#define SDG_BF16X3 2 /* FP32 storage, MFMA with 2-plane bf16 splits (3 products), FP32 accumulate */
#define SDG_BF16X6 3 /* FP32 storage, MFMA with 3-plane bf16 splits (6 products), FP32 accumulate */

/* face types (src/Initializer/CellLocalInformation.h FaceType subset) */
#define SDG_FACE_REGULAR 0
#define SDG_FACE_DYNAMIC_RUPTURE 1

const char* sdg_last_error(void);
int sdg_version(void);
int sdg_device_count(void);

/* ---------------------------------------------------------------------- */
/* Global matrices (replaces the build-time data codegen/matrices/         */
/* matrices_{B}.xml consumed by codegen/kernels/aderdg.py:37-41 and staged */
/* by GlobalDataInitializer::init, src/Memory/GlobalData.cpp:135-320).     */
/* Flat FP64 buffer, row-major blocks in this order:                       */
/*   kDivM[3][B][B], kDivMT[3][B][B] (not negated), rDivM[4][B][B2],       */
/*   fMrT[4][B2][B], rT[4][B2][B], fP[3][B2][B2], mass[B], mass2[B2].      */
/* ---------------------------------------------------------------------- */
size_t sdg_global_matrices_size(int order);
int sdg_global_matrices(int order, double* out);
/* Modal basis evaluation (src/Numerical/BasisFunction.h:133-143 ordering) */
int sdg_basis_values(int order, double xi, double eta, double zeta, double* out /* [B] */);
int sdg_basis_gradients(int order, double xi, double eta, double zeta, double* out /* [3][B] */);

/* ---------------------------------------------------------------------- */
/* Synthetic inputs of the proxy harness (src/Proxy/Allocator.cpp:49-134)   */
/* Host arrays: vertices [n][4][3], material [n][3] = (rho, lambda, mu),   */
/* neighbors [n][4], face_relations [n][4][2] = (neighbour side, orient.). */
/* ---------------------------------------------------------------------- */
int sdg_mesh_proxy(int64_t n, uint64_t seed, double* vertices, int32_t* neighbors, uint8_t* face_relations);
int64_t sdg_mesh_periodic_size(int nx, int ny, int nz);
int sdg_mesh_periodic(int nx, int ny, int nz, double h, double jitter, uint64_t seed, double* vertices,
                      int32_t* neighbors, uint8_t* face_relations);
int sdg_random_materials(int64_t n, uint64_t seed, double rho_min, double rho_max, double vp, double vs,
                         double* material);
double sdg_cfl_timestep(int64_t n, int order, double cfl, const double* vertices, const double* material);
int sdg_uniform_dofs(int64_t count, uint64_t seed, double* out);
/* Time-basis coefficients of the monomial basis integrated over [t0, t1]
 * (src/Numerical/TimeBasis.h MonomialBasis::integrate): coeffs[i] =
 * (t1^{i+1} - t0^{i+1}) / (i+1)!  */
int sdg_monomial_integrate(int order, double t0, double t1, double* coeffs);

/* ---------------------------------------------------------------------- */
/* Per-element FP64 precompute on the GPU (replaces                        */
/* initializeCellLocalMatrices, src/Initializer/Model/CellLocalMatrices.cpp:66). */
/* Device arrays: vertices [n][4][3], material [n][3], neighbors [n][4],   */
/* face_types [n][4] or NULL (all regular).  Outputs (FP64, 9x9 column-    */
/* major): star [n][3][81], aplus [n][4][81], aminus [n][4][81].           */
/* ---------------------------------------------------------------------- */
int sdg_cell_local_matrices(int64_t n, const double* vertices, const double* material,
                            const int32_t* neighbors, const uint8_t* face_types, double* star, double* aplus,
                            double* aminus, void* stream);

/* ---------------------------------------------------------------------- */
/* Engine: global matrices resident on the device for one order and one    */
/* arithmetic.  Mirrors Spacetime/Time/Local/Neighbor::setGlobalData       */
/* (src/Kernels/LinearCK/{Time,Local,Neighbor}.cpp).                       */
/* ---------------------------------------------------------------------- */
typedef struct sdg_engine sdg_engine;
int sdg_engine_create(sdg_engine** engine, int order, int arithmetic);
/* Replace the generated global matrices by caller-provided ones (layout as
 * sdg_global_matrices). */
int sdg_engine_set_global_matrices(sdg_engine* engine, const double* host_matrices);
int sdg_engine_destroy(sdg_engine* engine);
int64_t sdg_derivatives_size(int order); /* reals per element: sum_i 9*B(O-i) */

/* A layer of cells in device memory.  `real` is double for SDG_FP64 and
 * float otherwise.  DOF-like arrays are [n][9][B] (basis index fastest);
 * derivatives are [n][sdg_derivatives_size] with block i stored [9][B(O-i)]. */
typedef struct {
  int64_t num_elements;
  void* dofs;                     /* [n][9][B]                                */
  void* time_integrated;          /* [n][9][B] (the proxy's buffers)          */
  void* derivatives;              /* optional, NULL if not wanted             */
  const void* star;               /* [n][3][81]                               */
  const void* aplus;              /* [n][4][81]                               */
  const void* aminus;             /* [n][4][81]                               */
  const int32_t* face_neighbors;  /* [n][4], index into time_integrated       */
  const uint8_t* face_relations;  /* [n][4][2]                                */
  const uint8_t* face_types;      /* [n][4] or NULL (all regular)             */
  void* face_traces;              /* [n][4][9][B2]: rT_f I of every face;     */
                                  /* required by SDG_BF16X3/X6 (written by the*/
                                  /* local pass, read by the neighbour pass)  */
} sdg_layer;

/* Spacetime::computeBatchedAder (src/Kernels/LinearCK/Time.cpp:142-217):
 * CK derivatives of dofs, time integral into time_integrated, optional
 * derivatives.  coeffs[order] = time-basis coefficients (host). */
int sdg_compute_batched_ader(sdg_engine* engine, const double* coeffs, double time_step_width,
                             const sdg_layer* layer, void* stream);
/* Time::evaluateBatched (Time.cpp:257-290; HIP taylorSum KernelsAux.cpp:117-184):
 * out[e] = sum_i coeffs[i] * dQ_i[e]. */
int sdg_evaluate_batched(sdg_engine* engine, const double* coeffs, const void* derivatives,
                         void* time_evaluated, int64_t num_elements, void* stream);
/* Local::computeBatchedIntegral (src/Kernels/LinearCK/Local.cpp:255-395):
 * volume + local flux of all non-DR faces, dofs += ... */
int sdg_compute_batched_integral(sdg_engine* engine, const sdg_layer* layer, void* stream);
/* Neighbor::computeBatchedNeighborsIntegral (src/Kernels/LinearCK/Neighbor.cpp:111-188):
 * neighbour flux of all regular faces from the neighbours' time_integrated. */
int sdg_compute_batched_neighbors_integral(sdg_engine* engine, const sdg_layer* layer, void* stream);
/* Fused ader + local integral (the proxy's `local` kernel,
 * src/Proxy/KernelDevice.cpp:69-88): one pass over dofs. */
int sdg_compute_batched_ader_local(sdg_engine* engine, const double* coeffs, double time_step_width,
                                   const sdg_layer* layer, void* stream);

/* Proxy kernel selection (src/Proxy/Common.cpp:22-41, KernelDevice.cpp:144-165):
 * "all", "local", "neigh", "ader", "localwoader".  Runs `steps` repetitions
 * of that kernel on the layer with time-step width dt. */
int sdg_proxy_kernel_id(const char* name);
int sdg_proxy_run(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer, void* stream);

/* Copy FP64 host data into a layer array of the engine's real type (and
 * back); element counts in reals. */
int sdg_upload_reals(sdg_engine* engine, void* device_dst, const double* host_src, int64_t count, void* stream);
int sdg_download_reals(sdg_engine* engine, double* host_dst, const void* device_src, int64_t count, void* stream);
int sdg_convert_fp64_to_real(sdg_engine* engine, void* device_dst, const double* device_src, int64_t count,
                             void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SEISSOL_AMD_H_ */
