/*
 * dgx — MI355X-native DG right-hand side and RK4 engine: C ABI.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in launchers with the reference's names and signatures
 *     (marchdf/dg code/src/{dg/kernels.h, physics/physics.h, bc/boundaries.h,
 *     time_integration/rk_kernels.h, limiting/limiting_kernels.h,
 *     sensors/sensor_kernels.h}).  Pointers are device buffers in the
 *     reference's layouts; the launch runs on the null stream and returns when
 *     the work is done, like the reference's launches under its timers.  The
 *     reference fixes D, N_F, physics and flux at compile time; here they come
 *     from the configuration set with dgx_set_config().
 *
 *  2. Stream-aware fused entry points used by the C++ host classes
 *     (DG_SOLVER / RK / Limiting / SENSOR).  They return a dgx_status and take
 *     a hipStream_t passed as void*.
 *
 * Plus the setup builder (structured meshes and the reference's DG tables).
 */
#ifndef DGX_H
#define DGX_H
#ifdef __cplusplus
extern "C" {
#endif

typedef int dgx_status; /* 0 ok, <0 error; dgx_last_error() holds the message */
#ifdef __cplusplus
typedef bool dgx_bool;
#else
typedef _Bool dgx_bool;
#endif

#define DGX_OK 0
#define DGX_EINVAL -1
#define DGX_EUNSUPPORTED -2
#define DGX_EDEVICE -3
#define DGX_EDT -4 /* Dt below 1e-14 or NaN (rk.cc:147-148) */

#define DGX_SINGLEFLUID 0
#define DGX_MULTIFLUID 1
#define DGX_PASSIVE 2   /* rho, momenta, E, rho*phic (conservative), phinc (non-conservative); gamma = 1.4 */
#define DGX_STIFFENED 3 /* rho, momenta, E, alpha = 1/(gamma-1), beta = gamma*pinf/(gamma-1) */
#define DGX_RUSANOV 0
#define DGX_ROE 1

typedef struct {
  int dim;     /* D: 1 or 2 */
  int order;   /* p: 1..3 */
  int physics; /* DGX_SINGLEFLUID | DGX_MULTIFLUID (GAMNCON) | DGX_PASSIVE | DGX_STIFFENED; no mass fractions */
  int flux;    /* DGX_RUSANOV | DGX_ROE */
} dgx_config;

const char* dgx_last_error(void);
int dgx_nfields(const dgx_config* cfg); /* N_F implied by the configuration */

/* ------------------------------------------------------------------ setup */
typedef struct dgx_mesh dgx_mesh;

/* Cartesian mesh on [x0,x1] (x [y0,y1] in 2D) with nx (x ny) elements.
 * bc: 1D {left,right}; 2D {left,right,bottom,top}; 1 periodic, 2 farfield,
 * 3 reflective wall.  Builds every table of main.cc:312-740 in the
 * reference's layouts. */
dgx_status dgx_mesh_build(int dim, int order, int nx, int ny, double x0, double x1, double y0, double y1,
                          const int* bc, int nfields, dgx_mesh** out);
void dgx_mesh_free(dgx_mesh* m);
/* Mesh from a Gmsh 2.2 ASCII file, as simpleMesh::load reads it
 * (simpleMesh.cc:141-204): quadrilaterals (2D) or lines (1D) of order 1..3
 * in file order, node index = node id - 1, boundary elements with physical
 * tags 1 periodic (paired by geometry, simpleMesh.cc:21-91), 2 farfield,
 * 3 wall, 4 (as farfield); faces built as BuildInterfaces does
 * (simpleMesh.cc:770-989).  Elements must be affine (one invJac each). */
dgx_status dgx_mesh_read_gmsh(const char* path, int nfields, dgx_mesh** out);
/* Writes a mesh as Gmsh 2.2 ASCII, boundary faces as tagged boundary
 * elements; dgx_mesh_read_gmsh of the file gives the same tables. */
dgx_status dgx_mesh_write_gmsh(const dgx_mesh* m, const char* path);
/* 1 when solvers on this mesh run the fused tile kernel by default (uniform
 * Cartesian grid, element counts multiples of the tile), else 0.  Host only. */
int dgx_mesh_tileable(const dgx_mesh* m);
/* sizes[0..15] = D, order, N_s, N_G, M_s, M_G, N_N, N_E, M_T, M_B, N_F, N_s1D, nx, ny, rflctiveIdx, 0 */
void dgx_mesh_sizes(const dgx_mesh* m, int* sizes);
double dgx_mesh_dx(const dgx_mesh* m); /* minimum cell size (simpleMesh.cc:1140-1190) */
/* Named host arrays: phi dphi phi_w dphi_w psi psi_w weight points pointsF weightF ref_nodes
 * xyz J invJac JF normals Minv Lag2Mono Mono2Lag Lag2MonoX MonoX2MonoY MonoY2Lag Lag2MonoY MonoX2Lag
 * (double) and map invmap boundaryMap neighbors face_elem face_closure face_tag closures (int).
 * Returns NULL for an unknown name. */
const double* dgx_mesh_f64(const dgx_mesh* m, const char* name, long* count);
const int* dgx_mesh_i32(const dgx_mesh* m, const char* name, long* count);

/* ------------------------------------------------- drop-in launcher group */
dgx_status dgx_set_config(const dgx_config* cfg);

/* dg/kernels.h:14-18 */
void LmapToFace(int M_s, int M_T, int N_s, int* map, double* U, double* UF);
void LmapToElement(int N_s, int N_E, int M_s, int N_N, int* invmap, double* Q, double* q);
void Lredistribute_sf(int N_G, int N_E, double* sJ, double* fJ, double* s, double* f, double* J, double* invJac);
void Lredistribute_q(int M_G, int M_T, double* qJ, double* q, double* JF);
void LaddSFQ(int N_s, int N_E, double* A, double* S, double* F, double* Q);
/* physics/physics.h:15-16 */
void Levaluate_sf(int N_G, int N_E, double* s, double* f, double* Ug, double* dUg, double* invJac);
void Levaluate_q(int M_G, int M_T, double* q, double* UgF, double* normals);
/* bc/boundaries.h:15 */
void LrflctiveBoundary(int M_s, int M_B, int* boundaryMap, double* normals, int start, double* UF);
/* time_integration/rk_kernels.h:14,16 */
void Lsolver(int N_s, int N_E, double Dt, double* Minv, double* f, double* DU);
void LfindUPA(const int N_s, const int N_E, double* U, double* UPA);
/* limiting/limiting_kernels.h:30-31 */
void Lhrl1D(int N_s, int N_E, int Nfields, int N_N, int slicenum, int* neighbors, int offxy, double* A, double* Alim);
void Lhri1D(int N_s, int N_E, int N_N, int* neighbors, int N_s1D, int slicenum, int offxy, double* Lag2Mono,
            double* Mono2Lag, int* sensors, double* U, double* Unew);
/* limiting/limiting_kernels.h:21-39: the helpers of the m2, primitive-variable and p0 limiters
 * (Limiting::M2limiting, PRlimiting, M2Ilimiting, P0Ilimiting, limiting.cc:102-744).  Lm2i1D takes two-fluid
 * states (DGX_MULTIFLUID); LCons2Prim / LPrim2Cons take every physics with an alpha field (not single fluid). */
void Lstridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB, double* A,
                  double* B);
void Lreconstruct_energy(int N_s, int N_E, int slicenum, double* rhoeLim, double* KLim, double* EMono, double* ELim);
void Linternal_energy_multifluid(int N_s, int N_E, int slicenum, double* p, double* g, double* rhoe);
void Linternal_energy_stiffened(int N_s, int N_E, int slicenum, double* p, double* g, double* b, double* rhoe);
void Lm2i1D(int N_s, int N_E, int N_N, int* neighbors, int N_s1D, int slicenum, int offxy, double* Lag2Mono,
            double* Mono2Lag, int* sensors, double* U, double* Unew);
void Lp0i(int N_s, int N_E, int N_G, double refArea, int* sensors, double* phi_w, double* U, double* Unew);
void LChangeBasis(int size1, int size2, int N_E, double* Transform, double* U, double* Unew);
void LPrim2Cons(int N_s, int N_E, double* U);
void LCons2Prim(int N_s, int N_E, double* U);
/* physics/physics.h:18-20 */
void Lpressure(int N_s, int N_E, double* U, double* p);
void Lkinetic_energy1D(int N_s, int N_E, double* rho, double* rhou, double* K);
void Lkinetic_energy2D(int N_s, int N_E, double* rho, double* rhou, double* rhov, double* K);
/* time_integration/rk_kernels.h:15 */
void Laverage_cell_p0(const int N_s, const int N_E, double* DU);
/* printer/printer_kernels.h:14-15 (the reference's inverse defaults to false in C++) */
void Lformater(int N_s, int N_E, double* U, double* output, dgx_bool inverse);
void Lformat_sensor(int N_s, int N_E, int* sensor, double* output);
/* sensors/sensor_kernels.h:17-18 (flags follow the sequential CPU order) */
void Lcalc_sensors(int N_E, int N_N, dgx_bool sensor1, double thresh1, dgx_bool sensor2, double thresh2, dgx_bool sensor3,
                   double thresh3, int* neighbors, double* Uavg, int* sensors);
void Lcopy_detected(int N_s, int N_E, int* sensors, double* Uold, double* U);
/* No counterpart in the reference: Lcalc_sensors for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and flags as
 * Lcalc_sensors; its scratch is a fixed per-device array (40 MiB) instead of
 * hipMalloc, so N_E is limited to 2^21 elements (DGX_EINVAL beyond, with
 * dgx_last_error()).  Calls that overlap on one GPU share that array and race.
 * Synchronous, null stream, like the reference launchers. */
dgx_status dgx_calc_sensors_noalloc(int N_E, int N_N, dgx_bool sensor1, double thresh1, dgx_bool sensor2,
                                    double thresh2, dgx_bool sensor3, double thresh3, int* neighbors, double* Uavg,
                                    int* sensors);
/* BLAS macro layer (utils/scalar_def.h:81-87): column-major, device pointers */
void dgx_blasDgemm(char transa, char transb, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc);
void dgx_blasDaxpy(int n, double alpha, const double* x, int incx, double* y, int incy);
void dgx_blasDcopy(int n, const double* x, int incx, double* y, int incy);
int dgx_blasIdamax(int n, const double* x, int incx); /* 1-based like BLAS */

/* ------------------------------------------------------ fused engine group */
typedef struct dgx_solver dgx_solver;

/* Uploads the tables of `mesh` for `cfg` (DG_SOLVER ctor, dg_solver.h:104-246;
 * Limiting ctors, limiting.h:123-408; SENSOR ctor, sensor.h:42-77). */
dgx_status dgx_solver_create(const dgx_config* cfg, const dgx_mesh* mesh, dgx_solver** out);
void dgx_solver_free(dgx_solver* s);
long dgx_solver_state_size(const dgx_solver* s); /* N_E * N_F * N_s doubles */

/* f_rk = S + F + Q for U (DG_SOLVER::dg_solver, dg_solver.h:292-364). */
dgx_status dgx_rhs(dgx_solver* s, const double* U, double* f_rk, void* stream);

/* Sensor thresholds (deck "#sensor thresholds", sensor.h:49-53); n = 0..3. */
dgx_status dgx_set_sensor_thresholds(dgx_solver* s, int n, const double* thresholds);

/* limiter (the reference's codes, main.cc:167-175): 0 none, 1 hrl (global HR, Limiting::HRlimiting),
 * 2 m2l (two-fluid: HR of rho, momenta, 1/(gamma-1) and pressure, energy rebuilt from the limited internal and
 * kinetic energies; Limiting::M2limiting, limiting.cc:123-594), 3 hri (sensor-gated HR), 4 m2i (two-fluid: the
 * m2 limiting on the sensor's cells only, element by element; Limiting::M2Ilimiting, limiting.cc:682-720), 5 prl
 * (two-fluid: HR of the primitive variables; Limiting::PRlimiting, limiting.cc:102-121), 6 pri (two-fluid: sensor
 * on the conserved state, hri sweeps on the primitive variables; Limiting::PRIlimiting, limiting.cc:636-662), 7 p0i
 * (two-fluid sensor: the flagged elements set to their cell averages; Limiting::P0Ilimiting, limiting.cc:723-744);
 * and 8, HR gated by the learned indicator (no reference code). */
dgx_status dgx_limit(dgx_solver* s, int limiter, double* U, void* stream);
/* Out = limited In (In unchanged; In != Out, both state-sized device arrays).
 * hrl on a 2D Cartesian tile mesh runs both sweeps in one launch (the same
 * arithmetic as dgx_limit, bit for bit); otherwise In is copied to Out and
 * limited in place.  Same limiter codes as dgx_limit. */
dgx_status dgx_limit_into(dgx_solver* s, int limiter, const double* In, double* Out, void* stream);
/* Solution output and restart.
 * dgx_format replaces Lformater (printer_kernels.h:14; printer_{oned,twod}_*.cu:20-110): inverse 0 fills
 * out (device, state-sized) with the printed fields node by node (rho, ux[, uy][, G], p), inverse 1 fills
 * U from out; the reference's arithmetic bit for bit.
 * dgx_pos_write / dgx_pos_read replace simpleMesh::writeSolution / readSolution (simpleMesh.cc:322-459)
 * with the printed fields in host memory (out[(e*N_F+f)*N_s+i]): one "<prefix><fname><step %010d>.pos"
 * file per field (rho, ux[, uy][, g], p), Gmsh ElementNodeData with the mesh's element ids; reading checks
 * the element count, ids and node counts and returns DGX_EINVAL on a mismatch (the reference exits). */
dgx_status dgx_format(dgx_solver* s, double* U, double* out, int inverse, void* stream);
dgx_status dgx_pos_write(const dgx_mesh* m, int physics, const double* out, const char* prefix, int step,
                         double time);
dgx_status dgx_pos_read(const dgx_mesh* m, int physics, const char* prefix, int step, double* time, double* out);

/* DG_SOLVER::conservation (dg_solver.h:367-389): totals (device double[N_F]) = the integral of each conserved
 * field over the mesh, sum over elements and Gauss points of (phi*U)_g * J_e * w_g.  Each term has the
 * reference's bits; the terms are summed in double-double in a fixed order (deterministic, within about one
 * rounding of their exact sum; the reference adds them one after another). */
dgx_status dgx_conservation(dgx_solver* s, const double* U, double* totals, void* stream);

/* SENSOR::sensing: flags (int[N_E], device) and Uavg (double[N_E*N_F], device; may be NULL). */
dgx_status dgx_sensing(dgx_solver* s, const double* U, int* sensors, double* Uavg, void* stream);

/* Dt = CFL / max(|u|+a) over nodes, written to dt_dev (device double). */
dgx_status dgx_dt_from_cfl(dgx_solver* s, const double* U, double cfl, double* dt_dev, void* stream);

/* One RK4 step (rk.cc:163-209) with Dt read from device memory dt_dev.
 * U is updated in place; limiter as in dgx_limit. */
dgx_status dgx_rk4_step(dgx_solver* s, double* U, const double* dt_dev, int limiter, void* stream);

/* RK::RK_integration (rk.cc:11-247) without output/particles.
 * mode 0: fixed Dt = dt_or_cfl for nsteps steps (deck CFL < 0).
 * mode 1: Dt from CFL = dt_or_cfl (already scaled by Dx/(2p+1)) until tf,
 *         or for at most nsteps steps when nsteps > 0; the initial condition
 *         is limited first (rk.cc:100-114).
 * Returns the step count in *steps and the final time in *t_final.
 * In mode 1 with limiter 0 or 1 the time control (Dt, its stop and the clamp
 * to tf) runs on the device and the steps are enqueued in chunks, with one
 * read-back per chunk instead of one per step (dgx_solver_set_device_time);
 * the same steps, times and bits as the per-step loop. */
dgx_status dgx_rk4_integrate(dgx_solver* s, double* U, int mode, double dt_or_cfl, int nsteps, double tf,
                             int limiter, int* steps, double* t_final, void* stream);
/* RK::RK_integration's CFL branch with its output times (rk.cc:100-233, rk.h:58-83): out_times[0..n_out-1]
 * are the reference's _output_time_array[1..] (entry 0 is the initial time), strictly increasing, the last
 * being the final time.  Dt = CFL/max(|u|+a) is clamped to the final time (rk.cc:149) and to the next output
 * time (rk.cc:150); after a clamped step on_output (may be NULL) is called with the output count (1-based, as
 * the reference's count), the step count, the time and U, with the stream synchronised: it stands for
 * PRINTER::print (rk.cc:215-233) and may call dgx_format / dgx_pos_write.  A non-zero return stops the run
 * with that status. */
typedef dgx_status (*dgx_output_fn)(void* user, int count, int step, double time, double* U, void* stream);
dgx_status dgx_rk4_integrate_out(dgx_solver* s, double* U, double cfl, const double* out_times, int n_out,
                                 int limiter, dgx_output_fn on_output, void* user, int* steps, double* t_final,
                                 void* stream);

/* Kernel path.  0 auto: the fused tile kernel (one launch per RK stage)
 * on uniform Cartesian meshes whose element counts are multiples of the tile
 * (8x4 elements in 2D, 128 in 1D), else the face + element kernels;
 * 1 face + element kernels; 2 tile kernel (DGX_EUNSUPPORTED when the mesh
 * does not qualify).  Both paths compute the same right-hand side as
 * DG_SOLVER::dg_solver (dg_solver.h:292-364).  dgx_solver_path reports the
 * path in use (1 or 2). */
dgx_status dgx_solver_set_path(dgx_solver* s, int path);
/* Precision of the tile kernel's volume contractions: 0 FP64 (default);
 * 1 the map to the Gauss points and the integration against the test
 * functions in FP32 on MFMA with FP32 accumulation; 2 the same two
 * contractions as split-BF16 products (x = hi + lo in BF16, hi*hi + hi*lo +
 * lo*hi on v_mfma_f32_16x16x16_bf16) with FP32 accumulation.  The fluxes,
 * the Riemann solver, Dt*Minv, the RK4 accumulation and the limiter stay
 * FP64 (2D p=2 on the tile path only; choosing the face + element path
 * resets it to 0). */
dgx_status dgx_solver_set_precision(dgx_solver* s, int precision);
int dgx_solver_path(const dgx_solver* s);
/* 1 when the 2D global HR limiter runs as one launch (dgx_limit_into, and
 * the RK driver's limiter): 2D p <= 2 Cartesian grids of single-fluid or
 * two-fluid states whose neighbour table follows the grid and whose
 * Lagrange/Taylor transforms are the standard ones of the Gmsh element (the
 * kernel carries them as constants; every mesh the library builds or reads
 * has them); else 0. */
int dgx_solver_hr_one_launch(const dgx_solver* s);

/* ------------------------------------------------- batched 1D ensemble */
/* Many independent 1D problems advanced together: each runs the loop of
 * RK::RK_integration (rk.cc:132-247) with its own Dt from CFL (rk.cc:250-272),
 * the DG right-hand side (dg_solver.h:292-364) and, optionally, the HR
 * limiter (limiting.cc:11-27).  No reference counterpart batches problems;
 * each problem's result equals a separate run of the reference loop. */
typedef struct dgx_batch dgx_batch;

/* n problems on [x0, x1]: problem k has n_elem[k] >= 2 elements of order
 * order[k] (1..3); cfg->dim must be 1 (cfg->order is ignored); bc = {left,
 * right} tags, 2 (farfield) or 3 (wall). */
dgx_status dgx_batch_create(const dgx_config* cfg, int n, const int* n_elem, const int* order, double x0,
                            double x1, const int* bc, dgx_batch** out);
void dgx_batch_free(dgx_batch* b);
/* Offset (doubles) of problem k in the packed state, k = 0..n (n: total size).
 * Each problem's slice uses the usual layout U[(e*N_F+f)*N_s+i]. */
long dgx_batch_state_offset(const dgx_batch* b, int k);
/* sizes[0] = workgroup size, sizes[1] = workgroups of a launch, sizes[2] =
 * problems cut into chunks.  Without a cut problem every problem has its own
 * workgroup, largest first.  A problem whose work (N_E times a measured cost
 * per element of order p) is well above
 * the ensemble's average per resident workgroup, and that has at least two
 * chunks' elements (dgx_batch_set_min_chunk), is cut into chunks of
 * consecutive elements on consecutive workgroups; the other problems then
 * share the remaining resident workgroups, longest first to the least
 * loaded.  Cut or not, every problem gets the same bits. */
dgx_status dgx_batch_block_sizes(dgx_batch* b, int* sizes);
/* Elements per chunk at least (default 64; 0: no problem is cut; else >= 8). */
dgx_status dgx_batch_set_min_chunk(dgx_batch* b, int min_chunk);
/* sizes[0] = workgroup size, a multiple of 64 in [64, 1024]; the rest is ignored. */
dgx_status dgx_batch_set_block_sizes(dgx_batch* b, const int* sizes);
/* Times ten steps for each candidate workgroup size (64 to 1024; the
 * fastest of two timed runs after a warm-up run) on scratch copies of U (U is
 * not modified) and keeps the fastest size. */
dgx_status dgx_batch_autotune(dgx_batch* b, const double* U, double cfl, int limiter, void* stream);
/* nsteps RK4 steps of every problem in one graph launch.  Problem k uses
 * Dt = cfl*Dx_k/(2p_k+1)/max(|u|+a) (deck CFL scaling, main.cc:896-897);
 * limiter 0 none or 1 hrl; limit_initial limits the initial state first
 * (rk.cc:100-114).  T (double[n], device) accumulates each problem's time;
 * status (int[n], device, zero on entry) becomes 1 for a problem whose Dt
 * fell below 1e-14 or was NaN (rk.cc:147-148); that problem stops.  Status 2
 * marks a cut problem whose chunks stopped waiting for each other (a bounded
 * wait that timed out; not expected when the device runs nothing else). */
dgx_status dgx_batched_rk4(dgx_batch* b, double* U, double cfl, int nsteps, int limiter, int limit_initial,
                           double* T, int* status, void* stream);

/* ------------------------------------------------ learned indicator */
/* A 49 -> 64 -> 64 -> 1 MLP troubled-cell indicator for 2D p=2 two-fluid
 * states, trained on the flags of calc_sensors (sensor_kernels.cu:19-163).
 * No counterpart in the reference.  Parameters (dgx_mlp_param_count() floats):
 * W1[64][64] (inputs 49..63 zero), b1[64], W2[64][64], b2[64], w3[64], b3.
 * Features: float[64] per element (dgx_mlp_features).  BF16 weights and
 * activations, FP32 accumulation. */
typedef struct dgx_mlp dgx_mlp;
int dgx_mlp_param_count(void);
dgx_status dgx_mlp_create(unsigned long long seed, dgx_mlp** out); /* He-uniform init from seed */
void dgx_mlp_free(dgx_mlp* m);
dgx_status dgx_mlp_set_params(dgx_mlp* m, const float* params); /* host array; resets the optimizer */
dgx_status dgx_mlp_get_params(const dgx_mlp* m, float* params);
/* logits z[n] for n feature rows X[n*64] (device pointers) */
dgx_status dgx_mlp_forward(dgx_mlp* m, const float* X, long n, float* z, void* stream);
/* mean gradient of the weighted cross-entropy -[w y log s(z) + (1-y) log(1-s(z))]
 * over the n rows, into grad (device float[param_count + 1], last = loss) */
dgx_status dgx_mlp_gradient(dgx_mlp* m, const float* X, const float* y, long n, float pos_weight, float* grad,
                            void* stream);
/* Adam (lr, 0.9, 0.999, 1e-8) over shuffled minibatches (seeded); epoch_loss
 * (host float[epochs]) receives the mean loss of each epoch */
dgx_status dgx_mlp_train(dgx_mlp* m, const float* X, const float* y, long n, int epochs, int batch, float lr,
                         float pos_weight, unsigned long long seed, float* epoch_loss, void* stream);
/* Attach (not owned) an indicator to a solver; limiter 8 in dgx_limit and the
 * RK drivers then flags cells with z > threshold and applies the HR sweeps
 * of hri to them (limiting.cc:596-634 with the flags in place of the sensor). */
dgx_status dgx_solver_set_indicator(dgx_solver* s, dgx_mlp* m, float threshold);
/* With limiter 8 on the tile kernel (2D p=2 two-fluid), each RK stage also
 * writes the cell means of rho, p and 1/(gamma-1) the indicator reads, and
 * the indicator skips its own pass over the state (on by default; 0 turns
 * it off, the same flags and bits either way). */
dgx_status dgx_solver_set_fused_means(dgx_solver* s, int on);
/* 1 (default): dgx_rk4_integrate's CFL runs with limiter 0 or 1 keep their
 * time control on the device; 0: the host loop reads every step's Dt back
 * (as DtFromCFL, rk.cc:250-272).  No counterpart in the reference: a switch
 * between two schedules of the same computation. */
dgx_status dgx_solver_set_device_time(dgx_solver* s, int on);
/* Limiter 8 and dgx_indicator_flags also flag every cell of a neighbour pair
 * for which the sensor's shock test (calc_sensors test 2, the same arithmetic
 * and cell averages, sensor_kernels.cu:125-136) fires, so that no cell the
 * sensor marks as a shock goes unflagged, whatever the network says.  Needs
 * the second sensor threshold (dgx_set_sensor_thresholds with n >= 2).  On by
 * default; 0 leaves the flags to the network alone. */
dgx_status dgx_solver_set_shock_guard(dgx_solver* s, int on);
dgx_status dgx_mlp_features(dgx_solver* s, const double* U, float* X, void* stream);
/* flags (int[N_E], device) and optionally logits z (float[N_E], device) */
dgx_status dgx_indicator_flags(dgx_solver* s, const double* U, int* flags, float* z, void* stream);

/* Per-kernel timing with HIP events recorded on the launch stream (the
 * reference's TIMERS use std::clock and synchronise after every launch,
 * timers.cc:71-90).  kernel: "face", "elem", "limit", "dt", "stage".  Reading
 * synchronises the recorded events and resets the counters. */
dgx_status dgx_profile_enable(dgx_solver* s, int enable);
dgx_status dgx_profile_read(dgx_solver* s, const char* kernel, double* total_ms, long* launches);

#ifdef __cplusplus
}
#endif
#endif
