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
#define SDG_F16X3 4  /* FP32 storage, MFMA with 2-plane fp16 splits of power-of-two scaled operands
                        (3 products), FP32 accumulate */

/* storage of the per-cell star matrices (sdg_engine_set_layout) */
#define SDG_STAR_DENSE 0   /* [3][81]: three 9x9 column-major matrices                          */
#define SDG_STAR_COMPACT 1 /* [3][24]: the 24 non-zeros of each, column by column with rows
                              ascending (the sparse memory layout of the reference's star tensor:
                              codegen/matrices/star.xml, codegen/config/gpu/sparse.xml:26,
                              LocalIntegrationData::starMatrices, src/Initializer/Typedefs.h:189) */

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
/* Kernel family of an engine (no reference counterpart: the reference picks
 * its generated kernels at build time).  SDG_KERNELS_AUTO (default): the
 * matrix-core passes where the engine has them (the split-16-bit forms; the
 * FP64 engine's fused ader + local pass at orders 3-7 on v_mfma_f64_16x16x4);
 * SDG_KERNELS_VALU: the FP64/FP32 VALU kernels only (FP64 and FP32 engines). */
#define SDG_KERNELS_AUTO 0
#define SDG_KERNELS_VALU 1
int sdg_engine_set_kernels(sdg_engine* engine, int family);
int sdg_engine_destroy(sdg_engine* engine);
int64_t sdg_derivatives_size(int order); /* reals per element, packed: sum_i 9*B(O-i) */

/* Storage layout of the cells' arrays for every later call on this engine
 * (the reference fixes it at build time):
 *  alignment (bytes, 0 = packed, else a power of two that is a multiple of
 *    sizeof(real), at most 1024): every DOF-like block (dofs, time_integrated /
 *    LTS buffers, sdg_evaluate_batched's output) is [9][ld] per cell with
 *    ld = B rounded up to a multiple of alignment / sizeof(real), the
 *    quantities' padding rows never read or written -- tensor::Q and
 *    tensor::I with alignStride (codegen/kernels/aderdg.py:58-65); and
 *    derivative block i is [9][ld_i], ld_i = B(O-i) rounded up likewise, the
 *    blocks one after another -- tensor::dQ(i) (the size check of the HIP
 *    taylorSum, src/Kernels/LinearCK/DeviceAux/cudahip/KernelsAux.cpp:53-66).
 *    The reference's AMD GPU build uses 128 (VECTORSIZE/ALIGNMENT,
 *    cmake/process_users_input.cmake:402-404): ld = 64 at ORDER 6.
 *  star_format: SDG_STAR_DENSE or SDG_STAR_COMPACT.
 * The per-cell strides of star/aplus/aminus are fields of each sdg_layer. */
int sdg_engine_set_layout(sdg_engine* engine, int alignment, int star_format);
/* Sizes under the engine's layout: ld of a DOF-like block, reals per cell of
 * a DOF-like block (9 ld) and of the derivatives. Any output may be NULL. */
int sdg_layout_sizes(const sdg_engine* engine, int32_t* dofs_ld, int64_t* dofs_reals, int64_t* derivatives_reals);

/* A layer of cells in device memory.  `real` is double for SDG_FP64 and
 * float otherwise.  DOF-like arrays are [n][9][ld] (basis index fastest,
 * ld = B unless the engine's layout aligns it); derivatives are
 * [n][derivatives_reals] with block i stored [9][ld_i] (sdg_engine_set_layout,
 * sdg_layout_sizes).  The per-cell operators of cell e start at
 * e * stride reals (0 = the dense default: star 243 or 72, aplus and aminus
 * 324), so the fields of an array of structs such as the reference's
 * LocalIntegrationData {starMatrices[3][24], nApNm1[4][81], ...} and
 * NeighboringIntegrationData {nAmNm1[4][81], ...} (src/Initializer/Typedefs.h:186-203)
 * are passed as they lie, with stride = sizeof(struct) / sizeof(real). */
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
  void* face_traces;              /* [n][4][9][B2P]: rT_f I of every face, B2P = B2 rounded up to a
                                     multiple of 4 (padding rows zero).  Required by the matrix-core
                                     arithmetics SDG_BF16X3/X6, SDG_F16X3 (FP32; written by the local
                                     pass, read by the neighbour pass).  Optional for the FP64 engine
                                     at orders 3-7 (FP64): its local pass then also writes them and
                                     the proxy kernels all/neigh read them; its batched neighbour
                                     integral always reads time_integrated */
  const uint16_t* lts_setup;      /* [n] LTS bits (SDG_LTS_*) or NULL: global */
                                  /* time stepping                            */
  int64_t star_stride;            /* reals between cells' star matrices, 0: dense */
  int64_t aplus_stride;           /* reals between cells' AplusT blocks, 0: 324 */
  int64_t aminus_stride;          /* reals between cells' AminusT blocks, 0: 324 */
} sdg_layer;

/* Local time stepping bits of a cell (src/Initializer/LtsSetup.h:18-110). */
#define SDG_LTS_NEIGHBOR_HAS_DERIVATIVES(face) (1u << (face))      /* neighbour gives derivatives */
#define SDG_LTS_NEIGHBOR_GTS_RELATION(face) (1u << (4 + (face)))   /* same cluster: time coeffs   */
#define SDG_LTS_HAS_BUFFERS (1u << 8)         /* cell stores its time-integrated buffer          */
#define SDG_LTS_HAS_DERIVATIVES (1u << 9)     /* cell stores its derivatives                     */
#define SDG_LTS_ACCUMULATE_BUFFERS (1u << 10) /* buffer sums the steps of a larger neighbour step */

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

/* ---------------------------------------------------------------------- */
/* Local time stepping (rate 2).  Every engine: the matrix-core engines    */
/* (orders 3-7) run the local pass on the matrix cores; the neighbour pass  */
/* runs on the VALU kernel of the engine's real type (FP32 for the split   */
/* forms, FP64 for the FP64 engine).                                        */
/* ---------------------------------------------------------------------- */
/* deriveLtsSetups (src/Initializer/InitProcedure/Internal/LtsSetup.cpp:97-330,
 * getLtsSetup + normalizeLtsSetup), host: LTS bits of every cell from the
 * cluster ids (larger id = larger time step) of the cell and its neighbours;
 * face neighbours must be at most one cluster apart. */
int sdg_derive_lts_setups(int64_t n, const int32_t* cluster_ids, const int32_t* neighbors,
                          const uint8_t* face_types, uint16_t* setups);
/* TimeCluster::computeLocalIntegration (src/Solver/TimeStepping/TimeCluster.cpp:358-434):
 * ader + local integral of the layer's cells; `time_integrated` receives the
 * buffer of cells with SDG_LTS_HAS_BUFFERS (added to it for cells with
 * SDG_LTS_ACCUMULATE_BUFFERS unless reset_buffers), `derivatives` those of
 * cells with SDG_LTS_HAS_DERIVATIVES.  Requires lts_setup and derivatives. */
int sdg_compute_local_integration(sdg_engine* engine, const double* coeffs, double time_step_width,
                                  const sdg_layer* layer, int reset_buffers, void* stream);
/* TimeCluster::computeNeighboringIntegration (TimeCluster.cpp:540-628) with
 * TimeCommon::computeIntegrals (src/Kernels/TimeCommon.cpp:40-87): for each
 * regular face the neighbour's buffer, or its derivatives integrated with
 * time_coeffs (same cluster) or subtime_coeffs (larger neighbour step).
 * `time_integrated`, `derivatives` are the arrays of all cells (indexed by
 * face_neighbors); the other arrays are the layer's own. */
int sdg_compute_neighboring_integration(sdg_engine* engine, const double* time_coeffs,
                                        const double* subtime_coeffs, const sdg_layer* layer, void* stream);
/* Rate-2 clusters of one layer (the TimeManager's clusters, src/Solver/TimeStepping):
 * cluster c = cells [cluster_start[c], cluster_start[c+1]) advancing with
 * dt0 * 2^c; the layer holds all clusters (neighbour indices global). */
typedef struct sdg_lts sdg_lts;
int sdg_lts_create(sdg_lts** lts, sdg_engine* engine, const sdg_layer* layer, int num_clusters,
                   const int64_t* cluster_start, double dt0);
/* Advance every cluster by `macro_steps` steps of the slowest cluster. */
int sdg_lts_advance(sdg_lts* lts, int macro_steps, void* stream);
/* Same, replaying a hipGraph of one macro step captured on first use
 * (StreamRuntime::runGraph, src/Parallel/Runtime/Stream.h:163-190). */
int sdg_lts_graph_advance(sdg_lts* lts, int macro_steps, void* stream);
/* Tile autotuner (no reference counterpart: tile shapes are the engine's
 * own): times the tile variants of both kernels of every cluster on the
 * cluster's data, keeps the fastest and restores the cells' state.
 * Allocates temporaries for the saved state; synchronises `stream`. */
int sdg_lts_autotune(sdg_lts* lts, int repeats, void* stream);
/* Tile variant in use for `cluster` and kernel 0 (local) / 1 (neighbour).
 * VALU variants 0-3: bit 0 = one more basis row per lane, bit 1 = two waves
 * per block; neighbour variant 4 = one element per 16-lane row (the default
 * of the FP32 neighbour pass at orders 3-6 and of the FP64 one at orders
 * 4-5, else 0).  Matrix-core engines,
 * local kernel: 0 = the matrix-core pass, 1 + v = the FP32 VALU variant v. */
int sdg_lts_tile_variant(const sdg_lts* lts, int cluster, int kernel);
int sdg_lts_set_tile_variant(sdg_lts* lts, int cluster, int kernel, int variant);
int sdg_lts_destroy(sdg_lts* lts);

/* Proxy kernel selection (src/Proxy/Common.cpp:22-41, KernelDevice.cpp:144-165):
 * "all" 0, "local" 1, "neigh" 2, "ader" 3, "localwoader" 4, "neigh_dr" 5,
 * "godunov_dr" 6, "all_dr" 7 (the enum of src/Proxy/Common.h:18).
 * sdg_proxy_run runs `steps` repetitions of kernels 0-4 on the layer with
 * time-step width dt; the DR kernels run through sdg_proxy_run_dr. */
int sdg_proxy_kernel_id(const char* name);
int sdg_proxy_run(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer, void* stream);
/* No counterpart in the reference.  sdg_proxy_run without any device
 * allocation or caller workspace, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops): the matrix-core forms' face traces
 * live in one library-owned 64 MiB array per GPU (e.g. 19,418 elements at
 * ORDER 6) and layer->face_traces is ignored; larger layers are rejected.
 * Kernels 0, 1, 3, 4 (kernel 2 alone is rejected for matrix-core engines: it
 * would read the traces of an earlier call); the FP64/FP32 engines use no
 * scratch and run as sdg_proxy_run.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU race. */
int sdg_proxy_run_static(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer,
                         void* stream);

/* Work of one element-update of a proxy kernel (PerformanceEstimate,
 * src/Monitoring/Metric.h:16-50, summed by src/Proxy/Runner.cpp:120-150),
 * per element: non-zero and hardware FLOP, bytes (DOF blocks and
 * per-element operators moved once) and kernel bytes (what the kernels of
 * this arithmetic move, face traces included).  The DR kernels 5-7 (every
 * arithmetic; the DR parts in the engine's real type) count the proxy's
 * layout of 4 DR faces per cell, all faces DR
 * (DynamicRupture::metrics, src/Kernels/DynamicRupture.cpp:170-195, and the
 * nodalFlux estimate of Neighbor.cpp:198-214). */
typedef struct {
  double nonzero_flop;
  double hardware_flop;
  double bytes;
  double kernel_bytes;
} sdg_work_estimate;
int sdg_performance_estimate(int order, int arithmetic, int kernel_id, sdg_work_estimate* out);

/* Copy FP64 host data into a layer array of the engine's real type (and
 * back); element counts in reals. */
int sdg_upload_reals(sdg_engine* engine, void* device_dst, const double* host_src, int64_t count, void* stream);
int sdg_download_reals(sdg_engine* engine, double* host_dst, const void* device_src, int64_t count, void* stream);
int sdg_convert_fp64_to_real(sdg_engine* engine, void* device_dst, const double* device_src, int64_t count,
                             void* stream);

/* ---------------------------------------------------------------------- */
/* Rate-and-state fast-velocity-weakening friction (friction law 103).      */
/* ---------------------------------------------------------------------- */

/* Subset of DRParameters (src/Initializer/Parameters/DRParameters.h:55-76). */
typedef struct {
  double rs_sr0;                     /* V0                                   */
  double rs_slip_rate_tolerance;     /* Newton |g| tolerance (1e-8)          */
  double rs_state_tolerance;         /* fixed-point tolerance (1e-8)         */
  int32_t rs_max_slip_rate_updates;  /* Newton iterations (60)               */
  int32_t rs_state_variable_updates; /* fixed-point iterations (10)          */
  double eta_damp;
  double eta_damp_end;
  int32_t nucleation_count;          /* 0 or 1                               */
  int32_t pad_;
  double t0;                         /* nucleation rise time                 */
  double s0;                         /* nucleation start time                */
} sdg_friction_params;

/* One layer of dynamic-rupture faces in device memory (FP64).  Per-point
 * arrays are [faces][points] with `points` >= `num_points` real Gauss points
 * (padding ignored).  Fields follow the friction-law storage of
 * src/DynamicRupture/FrictionLaws/GpuImpl/FrictionSolverInterface.h:18-90. */
typedef struct {
  int64_t num_faces;
  int32_t points;                    /* stride of per-point arrays           */
  int32_t num_points;                /* real Gauss points per face (<= 64)   */
  int32_t order;                     /* time points per step                 */
  int32_t pad_;
  const double* imp_and_eta;         /* [f][11] zp zs zpNeig zsNeig etaP etaS invEtaS invZp invZs invZpNeig invZsNeig */
  const double* q_interpolated_plus; /* [f][order][9][points]                */
  const double* q_interpolated_minus;
  double* initial_stress;            /* [f][6][points], nucleation adds to it */
  double* initial_pressure;          /* [f][points]                          */
  const double* nucleation_stress;   /* [f][6][points] or NULL               */
  const double* nucleation_pressure; /* [f][points] or NULL                  */
  const double* rs_a;                /* [f][points] each                     */
  const double* rs_sl0;
  const double* rs_f0;
  const double* rs_muw;
  const double* rs_b;
  const double* rs_srw;
  double* state_variable;            /* [f][points] each                     */
  double* slip_rate1;
  double* slip_rate2;
  double* slip_rate_magnitude;
  double* mu;
  double* traction1;
  double* traction2;
  double* slip1;
  double* slip2;
  double* accumulated_slip;
  double* rupture_time;
  double* peak_slip_rate;
  double* dyn_stress_time;
  uint8_t* rupture_time_pending;
  uint8_t* dyn_stress_time_pending;
  uint8_t* convergence_inner;
  uint8_t* convergence_outer;
  double* imposed_state_plus;        /* [f][9][points]                       */
  double* imposed_state_minus;
  int32_t* newton_iterations;        /* optional [f][points]: residual evaluations g(V) of the slip-rate solves, accumulated */
  int32_t* slip_rate_updates;        /* optional [f][points]: Newton steps (the reference's "slip-rate updates",
                                        rsMaxNumberSlipRateUpdates bounds them per solve), accumulated */
} sdg_fault_layer;

/* Options of sdg_friction_evaluate: the warm start of the first slip-rate
 * solve of each time point, and the recording of (features, target) training
 * samples from the solves of a run without the MLP.
 *   closed_form_start != 0: a point whose closed-form creep root
 *     V_c = 2 V0 exp(-psi/a) sinh(S/N)  (S = invEtaS tau, N = invEtaS |sigma_n| a;
 *     the root of g(V) + V = 0, so g(V_c) = -V_c) lies below half the slip-rate
 *     tolerance starts its Newton solve there: its first residual test passes.
 *   mlp_params != NULL: every other point takes the 6-32-32-1 MLP's predicted
 *     root (FP32 parameters, layout of sdg_mlp_param_count()) when it lies
 *     more than a decade from the previous slip rate.
 * Either way the residual test and the tolerances are unchanged; g is
 * strictly decreasing, so a converged solve returns the plain root up to the
 * tolerance. */
typedef struct {
  const float* mlp_params;           /* device, or NULL                      */
  float* samples;                    /* device [capacity][7] or NULL         */
  int64_t capacity;
  unsigned long long* sample_count;  /* device counter (atomic)              */
  int32_t sample_stride;             /* keep 1 of `stride` steady solves     */
  int32_t closed_form_start;         /* 0: plain start at the previous slip rate */
} sdg_friction_options;

/* dr::friction_law::FrictionSolver::evaluate for the FVW law on the GPU
 * (src/DynamicRupture/FrictionLaws/GpuImpl/BaseFrictionSolver.h:253-267,
 * CPU semantics of CpuImpl/RateAndState.h: face-wide convergence).
 * delta_t, time_weights: host [order]; resample: device [num_points]^2. */
int sdg_friction_evaluate(const sdg_friction_params* params, const sdg_fault_layer* layer,
                          double full_update_time, const double* delta_t, const double* time_weights,
                          const double* resample, const sdg_friction_options* options, void* stream);
/* FrictionSolver::computeDeltaT (src/DynamicRupture/FrictionLaws/FrictionSolver.cpp:20-35) */
int sdg_friction_delta_t(int order, const double* time_points, double* delta_t);
/* Gauss-Legendre points/weights on [t0, t1] (src/Numerical/Quadrature.h:35-90) */
int sdg_gauss_legendre(int n, double t0, double t1, double* points, double* weights);
/* Stroud (collapsed Gauss-Jacobi) rule on the reference triangle with
 * (order+1)^2 points and the state-variable resampling matrix (the data of
 * codegen/matrices/dr_stroud_matrices_{order}.json): points [n][2],
 * weights [n], resample [n][n], n = (order+1)^2. */
int sdg_dr_quadrature(int order, double* points, double* weights, double* resample);
/* A layer of dynamic-rupture faces (DynamicRupture::FaceInformation and
 * DRGodunovData, src/Initializer/Typedefs.h:213-236).  Element indices refer
 * to the element layer; 9x9 matrices are column-major FP64. */
typedef struct {
  int64_t num_faces;
  int32_t points;                   /* padded Gauss points per face, >= (order+1)^2 */
  int32_t pad_;
  const int32_t* plus_element;      /* [f]                                  */
  const int32_t* minus_element;     /* [f]                                  */
  const uint8_t* face_info;         /* [f][3]: plus side, minus side, faceRelation (orientation + 1) */
  const double* tinv_t;             /* [f][81] TinvT                        */
  const double* flux_solver_plus;   /* [f][81]                              */
  const double* flux_solver_minus;  /* [f][81]                              */
} sdg_dr_layer;

/* Per DR face (DynamicRuptureMatrices.cpp:150-470), host, FP64: TinvT from
 * the plus side's face frame, the flux solvers fluxScale * A0 * T^T with
 * fluxScale = -2 area / (6 V_plus) and +2 area / (6 V_minus), and optionally
 * the impedances imp_and_eta [f][11] of the friction law (layout of
 * sdg_fault_layer).  vertices [n][4][3], material [n][3] of the cells. */
int sdg_dr_face_data(int64_t num_faces, const double* vertices, const double* material,
                     const int32_t* plus_element, const int32_t* minus_element, const uint8_t* face_info,
                     double* tinv_t, double* flux_solver_plus, double* flux_solver_minus, double* imp_and_eta);
/* DynamicRupture::batchedSpaceTimeInterpolation (src/Kernels/DynamicRupture.cpp:118-168):
 * both sides' derivatives (`derivatives` [n][sdg_derivatives_size], the
 * engine's real type) at the `order` time points (coeffs [order][order]:
 * c[o][d] = t_o^d / d!) and the Gauss points, rotated to the face frame;
 * q_plus, q_minus [f][order][9][points] FP64 (the friction law's
 * q_interpolated_plus/minus).  Every engine; the matrix-core engines
 * interpolate their FP32 derivatives in FP32. */
int sdg_dr_space_time_interpolation(sdg_engine* engine, const double* coeffs, const void* derivatives,
                                    const sdg_dr_layer* layer, double* q_plus, double* q_minus, void* stream);
/* Nodal flux of the dynamic-rupture faces (Neighbor.cpp:93-101; spec
 * dynamic_rupture.py:174-190): for every cell face with dr_mapping [n][4] >= 0,
 * dofs += V3mTo2nTWDivM(face, h) * imposed state * flux solver of the cell's
 * side (imposed_plus/minus [f][9][points] FP64, e.g. from sdg_friction_evaluate). */
int sdg_dr_nodal_flux(sdg_engine* engine, const sdg_layer* layer, const int32_t* dr_mapping,
                      const sdg_dr_layer* dr_layer, const double* imposed_plus, const double* imposed_minus,
                      void* stream);
/* Dynamic-rupture data of the proxy's DR kernels (the DR storage of
 * src/Proxy/Allocator.cpp:186-294). */
typedef struct {
  const int32_t* dr_mapping;     /* [num_elements][4] DR face of each cell face, -1 none */
  sdg_dr_layer faces;
  const double* imposed_plus;    /* [faces][9][points]                     */
  const double* imposed_minus;   /* [faces][9][points]                     */
  double* q_plus;                /* [faces][order][9][points] outputs      */
  double* q_minus;
} sdg_proxy_dr;
/* The proxy's random DR faces (Allocator.cpp:257-294), host: 4 n faces
 * between random cells (plus_element, minus_element [4n]) with random sides
 * and orientations (face_info [4n][3]), every cell face mapped to a random
 * DR face (dr_mapping [n][4]).  Deterministic in `seed`. */
int sdg_proxy_dr_faces(int64_t n, uint64_t seed, int32_t* dr_mapping, int32_t* plus_element,
                       int32_t* minus_element, uint8_t* face_info);
/* Proxy kernels 0-7 (KernelDevice.cpp:109-123, KernelHost.cpp:219-244):
 * neigh_dr = neighbour integral + nodal flux of the DR faces; godunov_dr =
 * space-time interpolation at the Gauss-Legendre points of [0, dt] from
 * layer->derivatives; all_dr = ader + local integral (writing the
 * derivatives), godunov_dr, neigh_dr.  Every engine (a matrix-core engine
 * runs ader + local on the matrix cores and the DR kernels in FP32). */
int sdg_proxy_run_dr(sdg_engine* engine, int kernel_id, int steps, double dt, const sdg_layer* layer,
                     const sdg_proxy_dr* dr, void* stream);
/* Face evaluation matrices of dynamic rupture (V3mTo2n(i, h) and
 * V3mTo2nTWDivM(i, h), codegen/kernels/dynamic_rupture.py:108-190; the data of
 * codegen/matrices/dr_stroud_matrices_{order}.json), host:
 * v3m [4][4][n][B] (face i, h = 0 plus side, h = orientation + 1 minus side),
 * v3mtw [4][4][B][n] or NULL; n = (order+1)^2 Stroud points. */
int sdg_dr_face_matrices(int order, double* v3m, double* v3mtw);
/* RateAndStateFastVelocityInitializer::computeInitialStateAndFriction
 * (src/DynamicRupture/Initializer/RateAndStateInitializer.cpp:154-180), host. */
int sdg_rs_initial_state(double traction1, double traction2, double pressure, double rs_a, double rs_sr0,
                         double initial_slip_rate, double* state_variable, double* friction_coefficient);

/* MLP training on the GPU (no counterpart in the reference): parameters
 * [sdg_mlp_param_count()] FP32 = normalisation (in_mean[6], in_scale[6],
 * out_mean, out_scale), W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3. */
int sdg_mlp_param_count(void);
/* mean squared error and its gradient over samples [n][7] (features, target),
 * both in normalised units; grads [param_count] device (overwritten). */
int sdg_mlp_loss_grad(const float* params, const float* samples, int64_t n, float* grads, float* loss,
                      void* stream);
/* one Adam step on params with moments m, v (device, param_count each). */
int sdg_mlp_adam(float* params, const float* grads, float* m, float* v, int step, float lr, void* stream);
/* forward pass: out[n] = predicted log10 slip rate for features [n][stride] */
int sdg_mlp_forward(const float* params, const float* features, int64_t n, int stride, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SEISSOL_AMD_H_ */
