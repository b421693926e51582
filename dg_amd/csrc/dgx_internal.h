// Internal declarations shared by the HIP kernels and the host library.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/dgx.h"
#include "mesh.h"

struct dgx_mlp;

namespace dgx {

void set_error(const std::string& msg);
bool hip_ok(hipError_t e, const char* what);

// Device copies of the tables the fused kernels use.  Per-element geometry
// is stored once per element; when every element carries bit-identical
// values (uniform Cartesian meshes) the stride is 0 and one copy is kept.
struct DevTables {
  // reference-element tables (column-major, as the reference)
  double *phi = nullptr, *dphi = nullptr, *phi_w = nullptr, *dphi_w = nullptr;
  double *psi = nullptr, *psi_w = nullptr;
  // faces
  int* face_elem = nullptr;     // [t*2+d]
  int* face_closure = nullptr;  // [t*2+d]
  int* face_tag = nullptr;      // [t]
  double* normals = nullptr;    // [t*D+a]
  double* JF = nullptr;         // [t*2+d]
  // elements
  double* J = nullptr;          // [e*geo_stride]
  double* invJac = nullptr;     // [(e*geo_stride)*D*D + alpha*D + a]
  double* Minv = nullptr;       // [(e*minv_stride*N_s + k)*N_s + i]
  int geo_stride = 1, minv_stride = 1;
  // minv_stride == 0: one matrix, Minv_e = Minv * (J0 / J_e).  The mass
  // matrix of an affine element is J_e times the reference one, so this holds
  // to rounding for every element the mesh builder makes; it is verified when
  // the solver is created.
  double J0 = 1.0;
  // tile kernel face geometry: [face*4 + {n_x, n_y, JF side 0, JF side 1}]
  double* face_geo_x = nullptr;  // faces normal to x, index ey*(nx+1) + fx
  double* face_geo_y = nullptr;  // faces normal to y, index fy*nx + ex
  // learned indicator (2D p=2): tensor-Legendre modes from nodal values,
  // Vinv[q*N_s + i]; row 0 gives the cell mean
  double* mlp_vinv = nullptr;
  double* weight = nullptr;     // volume quadrature weights [g] (conservation totals)
  // limiter / sensor
  int* neighbors = nullptr;     // [e*N_N + {L,R,D,U}]
  double *Lag2Mono = nullptr, *Mono2Lag = nullptr;
  double *Lag2MonoX = nullptr, *MonoX2MonoY = nullptr, *MonoY2Lag = nullptr, *Lag2MonoY = nullptr,
         *MonoX2Lag = nullptr;
};

// Cartesian meshes whose faces follow the structured pattern (element
// e = ey*nx + ex, side 0 = left/lower element) run the fused tile kernel of
// stage.hip, which finds faces from the tile structure and reads their
// normals and Jacobians from per-face arrays in grid order.
struct Cart {
  bool on = false;
  int nx = 0, ny = 1, tx = 1, ty = 1;
  int bc[4] = {0, 0, 0, 0};
  std::vector<double> gx, gy;  // host copies of DevTables::face_geo_x / face_geo_y
};

enum KernelPath { PATH_AUTO = 0, PATH_GENERAL = 1, PATH_TILED = 2 };

// Sensor relaxation sweeps queued per host read-back (the 1024x512
// shock-bubble needs about three).
constexpr int kRelaxBatch = 4;

struct Solver {
  dgx_config cfg{};
  int D = 1, P = 1, NF = 3, N_s = 0, N_G = 0, M_s = 0, M_G = 0, N_N = 0, N_E = 0, M_T = 0, N1 = 0;
  double Dx = 0;
  DevTables t;
  Cart cart;
  int path = PATH_AUTO;
  int precision = 0;  // tile kernel: 0 FP64, 1 FP32-MFMA map/integrate (2D p=2)
  bool tiled() const { return cart.on && path != PATH_GENERAL; }
  bool hr2d = false;  // 2D tile layout whose neighbour table follows it: one-launch HR limiter (k_hr2d)
  bool hr_const = false;  // 2D p <= 2 with the HR transforms of hr_tables.h: constant-table sweeps
  // work buffers
  double* Qe = nullptr;     // face contributions, element-major [((e*N_N+k)*NF+f)*M_s+j]
  double* Uacc = nullptr;   // RK accumulator
  double* Ustar = nullptr;  // stage state
  double* Ustar2 = nullptr; // second stage state (tile path, see rk4_step)
  double* work = nullptr;   // limiter scratch (state size)
  double* work2 = nullptr;  // limiter scratch (state size)
  double* Uavg = nullptr;   // [e*NF+f]
  int* sensors = nullptr;   // [e]
  int* sens_tgt = nullptr;  // sensor resolution scratch [e*2]
  int* sens_act = nullptr;  // [e*2] (ping-pong)
  int* flag = nullptr;      // device flag / counters: [0] sensor relaxation, [2] flagged-element count
  int* sens_changed = nullptr;  // [kRelaxBatch] change flags of a batch of sensor relaxation sweeps
  int* flag_list = nullptr; // [N_E] flagged elements (launch_limit_flagged)
  double* m2buf = nullptr;  // m2 limiting: 15 per-field arrays of N_E*N_s (allocated on first use)
  double* dscratch = nullptr;  // device scalars: [1] Dt, [2] the time of a device-timed CFL run
  int* time_state = nullptr;   // device-timed CFL run: stop, last step taken, Dt error, steps (k_dt_step)
  const int* gate = nullptr;   // while steps of such a run are enqueued: time_state, read by their kernels
  bool device_time = true;     // CFL runs without outputs keep their time control on the device
  double* cons_partial = nullptr;  // conservation totals: per-block double-double partial sums
  double* dt_partial = nullptr;  // per-block maxima of |u|+a (ceil(N_E*N_s/256)), then node 0's |u|+a
  int n_thresh = 0;
  double thresh[3] = {0, 0, 0};
  std::vector<void*> allocations;
  // learned indicator (limiter 8): network (not owned), threshold, scratch
  dgx_mlp* mlp = nullptr;
  float mlp_threshold = 0.0f;
  double* mlp_avg = nullptr;  // [e*3 + {rho, p, alpha}] cell means
  // the state whose cell means mlp_avg holds, set by the RK driver between a
  // stage that emitted them and the indicator that reads them (else null)
  const double* means_of = nullptr;
  bool fuse_means = true;      // RK stages emit the indicator's cell means (dgx_solver_set_fused_means)
  bool shock_guard = true;     // the indicator also flags the sensor's shock pairs (dgx_solver_set_shock_guard)
  double* mlp_psens = nullptr; // [e] the sensor's pressure from the cell averages (shock guard)
  int* mlp_guard = nullptr;    // [e] 1 where the sensor's shock test fires for a pair of e and a neighbour
  // the guard needs the sensor's shock test switched on (a second threshold)
  bool shock_guard_on() const { return shock_guard && n_thresh >= 2; }
  float* mlp_jumps = nullptr;  // [e*4 + n] neighbour jumps L, R, D, U (-1: boundary)
  float* mlp_x = nullptr;     // [e*64 + k] features
  float* mlp_z = nullptr;     // [e] logits
  // event timing (dgx_profile_*): [0] face, [1] elem, [2] limit, [3] dt, [4] stage
  bool prof_on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof[5];
  std::vector<hipEvent_t> ev_pool;
  long state_size() const { return (long)N_E * NF * N_s; }
};

// kernel launch entry points (defined in rhs.hip / limiter.hip)
hipError_t launch_face(const Solver& s, const double* U, hipStream_t st);
// mode 0: f_rk = S+F+Q to out; mode 1: RK stage `stage` update (see rhs.hip)
hipError_t launch_elem(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                       double* Uout, const double* dt_dev, double* f_out, hipStream_t st);
// fused tile kernel (Cartesian meshes): same modes as launch_elem
bool stage_tiles(int D, int nx, int ny, int* tx, int* ty);
// means: also write the output tile's cell means of rho, p and alpha there
// (2D p=2 two-fluid RK stages), and with uavg and psens the sensor's cell
// averages and pressures;
// *means_done tells whether the kernel did
hipError_t launch_stage(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                        double* Uout, const double* dt_dev, double* f_out, hipStream_t st, double* means = nullptr,
                        double* uavg = nullptr, double* psens = nullptr, bool* means_done = nullptr);
hipError_t launch_upa_max(const Solver& s, const double* U, double cfl, double* dt_dev, hipStream_t st);
// one step's Dt on the device with the stop, final-time clamp and step count
// of a CFL run (k_dt_step): state = Solver::time_state, T_dev its time
hipError_t launch_dt_step(const Solver& s, const double* U, double cfl, double tf, double* dt_dev, double* T_dev,
                          int* state, hipStream_t st);
hipError_t launch_limit_hr(Solver& s, double* U, hipStream_t st);
// Out = HR(In), both 2D sweeps in one launch; In != Out, needs s.hr2d
hipError_t launch_limit_hr_oop(Solver& s, const double* In, double* Out, hipStream_t st);
// with list: the flagged elements also appended to list[0 .. *count)
hipError_t launch_sensing(Solver& s, const double* U, int* sensors, double* Uavg, hipStream_t st, int* list = nullptr,
                          int* count = nullptr);
// the sensor's cell averages alone (k_uavg), [e*NF + f]
hipError_t launch_uavg(const Solver& s, const double* U, double* Uavg, hipStream_t st);
hipError_t launch_limit_hri(Solver& s, double* U, hipStream_t st);
hipError_t launch_limit_m2(Solver& s, double* U, hipStream_t st);  // m2l, two-fluid
hipError_t launch_limit_pr(Solver& s, double* U, hipStream_t st);  // prl, two-fluid
hipError_t launch_limit_pri(Solver& s, double* U, hipStream_t st); // pri, two-fluid
hipError_t launch_limit_m2i(Solver& s, double* U, hipStream_t st); // m2i, two-fluid
hipError_t launch_limit_p0i(Solver& s, double* U, hipStream_t st); // p0i (sensor: not single fluid)
// HR sweeps on the cells flagged in s.sensors (the HRI sweeps without the sensor)
hipError_t launch_limit_flagged(Solver& s, double* U, hipStream_t st);
// the same sweeps over a list already in s.flag_list / s.flag[2]
hipError_t launch_limit_listed(Solver& s, double* U, hipStream_t st);
// learned indicator (mlp.hip): features of U into X, and flags into s.sensors
hipError_t mlp_features(Solver& s, const double* U, float* X, hipStream_t st);
// with_list: also list the flagged elements in s.flag_list, their count in s.flag[2]
hipError_t mlp_indicator_flags(Solver& s, const double* U, hipStream_t st, bool with_list = false);
// solution output and restart (output.hip): PRINTER formatter, .pos files
hipError_t launch_format(const Solver& s, double* U, double* out, int inverse, hipStream_t st);
// DG_SOLVER::conservation totals of U into totals (device double[NF]); scratch holds
// conservation_scratch_doubles() doubles
hipError_t launch_conservation(const Solver& s, const double* U, double* totals, double* scratch, hipStream_t st);
size_t conservation_scratch_doubles();
std::string pos_write(const Tables& t, int physics, const double* out, const std::string& prefix, int step,
                      double time);
std::string pos_read(const Tables& t, int physics, const std::string& prefix, int step, double* time, double* out);

// drop-in compat: process-wide configuration for the reference-named launchers
const dgx_config& current_config();

}  // namespace dgx
