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
#define DGX_RUSANOV 0
#define DGX_ROE 1

typedef struct {
  int dim;     /* D: 1 or 2 */
  int order;   /* p: 1..3 */
  int physics; /* DGX_SINGLEFLUID | DGX_MULTIFLUID (GAMNCON, no mass fractions) */
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
/* sensors/sensor_kernels.h:17-18 (flags follow the sequential CPU order) */
void Lcalc_sensors(int N_E, int N_N, dgx_bool sensor1, double thresh1, dgx_bool sensor2, double thresh2, dgx_bool sensor3,
                   double thresh3, int* neighbors, double* Uavg, int* sensors);
void Lcopy_detected(int N_s, int N_E, int* sensors, double* Uold, double* U);
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

/* limiter: 0 none, 1 hrl (global HR), 3 hri (sensor-gated HR) */
dgx_status dgx_limit(dgx_solver* s, int limiter, double* U, void* stream);
/* SENSOR::sensing: flags (int[N_E], device) and Uavg (double[N_E*N_F], device; may be NULL). */
dgx_status dgx_sensing(dgx_solver* s, const double* U, int* sensors, double* Uavg, void* stream);

/* Dt = CFL / max(|u|+a) over nodes, written to dt_dev (device double). */
dgx_status dgx_dt_from_cfl(dgx_solver* s, const double* U, double cfl, double* dt_dev, void* stream);

/* One RK4 step (rk.cc:163-209) with Dt read from device memory dt_dev.
 * U is updated in place; limiter as in dgx_limit. */
dgx_status dgx_rk4_step(dgx_solver* s, double* U, const double* dt_dev, int limiter, void* stream);

/* RK::RK_integration (rk.cc:11-247) without output/particles.
 * mode 0: fixed Dt = dt_or_cfl for nsteps steps (deck CFL < 0).
 * mode 1: Dt from CFL = dt_or_cfl (already scaled by Dx/(2p+1)) until tf;
 *         the initial condition is limited first (rk.cc:100-114).
 * Returns the step count in *steps and the final time in *t_final. */
dgx_status dgx_rk4_integrate(dgx_solver* s, double* U, int mode, double dt_or_cfl, int nsteps, double tf,
                             int limiter, int* steps, double* t_final, void* stream);

/* Kernel path.  0 auto: the fused tile kernel (one launch per RK stage)
 * on uniform Cartesian meshes whose element counts are multiples of the tile
 * (8x4 elements in 2D, 128 in 1D), else the face + element kernels;
 * 1 face + element kernels; 2 tile kernel (DGX_EUNSUPPORTED when the mesh
 * does not qualify).  Both paths compute the same right-hand side as
 * DG_SOLVER::dg_solver (dg_solver.h:292-364).  dgx_solver_path reports the
 * path in use (1 or 2). */
dgx_status dgx_solver_set_path(dgx_solver* s, int path);
int dgx_solver_path(const dgx_solver* s);

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
