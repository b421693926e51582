/* CPU oracle for the DG right-hand side, RK4, HR limiter and sensors.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain FP64 restatement of the
 * reference's CPU path (marchdf/dg, ARCH=USE_CPU), used as the checker for
 * the HIP engine by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  Nothing in the product links or calls it.
 *
 * The reference fixes dimension, physics, flux and field count at compile
 * time (code/Makefile:147-156); here they are runtime fields of or_cfg.
 * Supported: D=1,2; SINGLEFLUID and MULTIFLUID (GAMNCON, no mass fractions);
 * RUS and ROE fluxes.  Every function names the reference lines it follows.
 */
#ifndef DG_ORACLE_H
#define DG_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int D, N_F, physics, flux; /* physics: 0 single fluid, 1 multifluid; flux: 0 Rusanov, 1 Roe */
  int N_s, N_G, M_s, M_G, N_N, N_E, M_T, N_s1D;
  double gx, gy;             /* gravity (constants::GLOBAL_GX/GY), zero in every config here */
} or_cfg;

typedef struct {
  const double *phi, *dphi, *phi_w, *dphi_w, *psi, *psi_w;
  const double *J, *invJac, *JF, *normals, *Minv;
  const int *map, *invmap, *boundaryMap, *neighbors;
  int rflctiveIdx;
  const double *Lag2Mono, *Mono2Lag;                                   /* 1D */
  const double *Lag2MonoX, *MonoX2MonoY, *MonoY2Lag, *Lag2MonoY, *MonoX2Lag; /* 2D */
} or_tables;

/* netlib dgemm restatement, column-major, alpha/beta as in BLAS */
void or_dgemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc);

/* pieces of DG_SOLVER::dg_solver (reference launcher semantics) */
void or_mapToFace(const or_cfg* c, const int* map, const double* U, double* UF);
void or_rflctive(const or_cfg* c, int M_B, const int* boundaryMap, const double* normals, int start, double* UF);
void or_evaluate_sf(const or_cfg* c, double* s, double* f, const double* Ug, const double* dUg, const double* invJac);
void or_evaluate_q(const or_cfg* c, double* q, const double* UgF, const double* normals);
void or_redistribute_sf(const or_cfg* c, double* sJ, double* fJ, const double* s, const double* f,
                        const double* J, const double* invJac);
void or_redistribute_q(const or_cfg* c, double* qJ, const double* q, const double* JF);
void or_mapToElement(const or_cfg* c, const int* invmap, double* Q, const double* Qtcj);
void or_addSFQ(const or_cfg* c, double* A, const double* S, const double* F, const double* Q);
void or_solve(const or_cfg* c, double Dt, const double* Minv, const double* f, double* DU);
void or_findUPA(const or_cfg* c, const double* U, double* UPA);

/* whole right-hand side: f_rk = S + F + Q (dg_solver.h:292-364) */
void or_rhs(const or_cfg* c, const or_tables* t, const double* U, double* f_rk);

/* Dt = CFL / max|u|+a over nodes (rk.cc:250-272); CFL already scaled by Dx/(2p+1) */
double or_dt_from_cfl(const or_cfg* c, const double* U, double CFL);

/* limiter pieces */
void or_hrl1D(int N_s, int N_E, int Nfields, int N_N, int slicenum, const int* neighbors, int offxy,
              const double* A, double* Alim);
void or_limit_hr(const or_cfg* c, const or_tables* t, double* U);      /* Limiting::HRlimiting */
void or_sensing(const or_cfg* c, const int* neighbors, const double* U, const int* on, const double* thresh,
                int* sensors, double* Uavg);                           /* SENSOR::sensing */
void or_calc_sensors(const or_cfg* c, const int* on, const double* thresh, const int* neighbors,
                     const double* Uavg, int* sensors);
void or_hri1D(const or_cfg* c, const int* neighbors, int N_s1D, int slicenum, int offxy, const double* L2M,
              const double* M2L, const int* sensors, const double* U, double* Unew);
void or_copy_detected(const or_cfg* c, const int* sensors, const double* Uold, double* U);
/* HR sweeps of the sensor-gated limiter on given flags (no sensor). */
void or_limit_flagged(const or_cfg* c, const or_tables* t, const int* flags, double* U);
void or_limit_hri(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors); /* Limiting::HRIlimiting (sensors returned for inspection) */

/* Limiting::M2limiting (limiting.cc:123-594), two-fluid GAMNCON, 1D and 2D Cartesian: per-field HR of rho,
 * rho u (rho v), 1/(gamma-1) and the pressure, then the energy rebuilt from the limited internal energy
 * (pressure x 1/(gamma-1), internal_energy_multifluid, limiting_kernels.cu:160-200) and kinetic energy
 * (kinetic_energy1D/2D, physics.cu:814-890), with the mean kept (reconstruct_energy,
 * limiting_kernels.cu:111-158); in 2D an x sweep (back to Lagrange) then a y sweep. */
void or_limit_m2(const or_cfg* c, const or_tables* t, double* U);
/* Limiting::PRlimiting (limiting.cc:102-121): Cons2Prim, HRlimiting, Prim2Cons (limiting_kernels.cu:957-1068). */
void or_limit_pr(const or_cfg* c, const or_tables* t, double* U);
/* Limiting::PRIlimiting (limiting.cc:636-662): sensing, Cons2Prim, hri sweeps, Prim2Cons. */
void or_limit_pri(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors);

/* m2i1D (limiting_kernels.cu:431-722), two-fluid GAMNCON: on the flagged elements, one direction (offxy 0 or 2)
 * of the m2 limiting done element by element in Lagrange form; Unew receives the flagged elements. */
void or_m2i1D(const or_cfg* c, const int* neighbors, int N_s1D, int slicenum, int offxy, const double* L2M,
              const double* M2L, const int* sensors, const double* U, double* Unew);
/* Limiting::M2Ilimiting (limiting.cc:682-720): sensing, then m2i1D + copy_detected per direction. */
void or_limit_m2i(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors);

/* p0i (limiting_kernels.cu:723-768): flagged elements set to their cell averages (phi_w quadrature / refArea). */
void or_p0i(const or_cfg* c, const double* phi_w, double refArea, const int* sensors, const double* U,
            double* Unew);
double or_ref_area(const or_cfg* c); /* 4 for quadrilaterals (main.cc:286), 2 for segments (unset in the reference) */
/* Limiting::P0Ilimiting (limiting.cc:723-744): sensing, p0i, copy_detected. */
void or_limit_p0i(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors);

/* Launcher-level restatements of the reference's helper kernels (their CPU loops):
 * stridedcopy, reconstruct_energy, internal_energy_multifluid (b NULL) / _stiffened, pressure (every physics),
 * kinetic_energy1D (rhov NULL) / 2D, average_cell_p0, ChangeBasis, Cons2Prim / Prim2Cons (to_prim 1 / 0),
 * format_sensor. */
void or_stridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB, const double* A,
                    double* B);
void or_reconstruct_energy(int N_s, int N_E, int slicenum, const double* rhoeLim, const double* KLim,
                           const double* EMono, double* ELim);
void or_internal_energy(int N_s, int N_E, int slicenum, const double* p, const double* g, const double* b,
                        double* rhoe);
void or_pressure(const or_cfg* c, const double* U, double* p);
void or_kinetic_energy(int N_s, int N_E, const double* rho, const double* rhou, const double* rhov, double* K);
void or_average_cell_p0(const or_cfg* c, double* DU);
void or_change_basis(const or_cfg* c, int size1, int size2, const double* T, const double* U, double* Unew);
void or_cons_prim(const or_cfg* c, double* U, int to_prim);
void or_format_sensor(int N_s, int N_E, const int* sensor, double* output);

/* DG_SOLVER::conservation (dg_solver.h:367-389): I[fc] = sum_e sum_g (phi*U)[(e,fc),g] * J[e] * weight[g],
 * accumulated one term after another in that order. */
void or_conservation(const or_cfg* c, const double* phi, const double* J, const double* weight, const double* U,
                     double* I);

/* PRINTER formatter (Lformater, printer_{oned,twod}_{singlefluid,multifluid}.cu:20-110, GAMNCON):
 * inverse 0: out = primitive output fields of U per node (rho, ux[, uy][, gamma], p);
 * inverse 1: U = conserved fields from out (the restart direction). */
void or_format(const or_cfg* c, double* U, double* out, int inverse);

/* RK4 time integration (rk.cc:132-247).
 * mode 0: fixed Dt for nsteps steps; mode 1: CFL-based Dt until T reaches tf (Dt clamped),
 * or for at most nsteps steps when nsteps > 0 (a cap the reference does not have).
 * limiter: 0 none, 1 hrl, 2 m2l, 3 hri, 5 prl, 6 pri.  Returns number of steps taken, or -1 on Dt error.
 * *t_out receives the final time. dt_hist (may be NULL) receives up to max_hist Dt values. */
int or_rk4(const or_cfg* c, const or_tables* t, double* U, int mode, double dt_or_cfl, int nsteps, double tf,
           int limiter, const int* on, const double* thresh, double* t_out, double* dt_hist, int max_hist);
/* Same with the reference's output times (rk.cc:149-150, 215-233): in CFL mode Dt is clamped to the next output
 * time as well as to the final time out_times[n_out-1]; out_steps[k] (may be NULL) receives the step count after
 * which output k happened.  out_times are _output_time_array[1..] (entry 0, the initial time, left out). */
int or_rk4_out(const or_cfg* c, const or_tables* t, double* U, int mode, double dt_or_cfl, int nsteps,
               const double* out_times, int n_out, int limiter, const int* on, const double* thresh, double* t_out,
               double* dt_hist, int max_hist, int* out_steps);

#ifdef __cplusplus
}
#endif
#endif
