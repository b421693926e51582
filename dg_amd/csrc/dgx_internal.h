// Internal declarations shared by the HIP kernels and the host library.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/dgx.h"
#include "mesh.h"

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
  // limiter / sensor
  int* neighbors = nullptr;     // [e*N_N + {L,R,D,U}]
  double *Lag2Mono = nullptr, *Mono2Lag = nullptr;
  double *Lag2MonoX = nullptr, *MonoX2MonoY = nullptr, *MonoY2Lag = nullptr, *Lag2MonoY = nullptr,
         *MonoX2Lag = nullptr;
};

struct Solver {
  dgx_config cfg{};
  int D = 1, P = 1, NF = 3, N_s = 0, N_G = 0, M_s = 0, M_G = 0, N_N = 0, N_E = 0, M_T = 0, N1 = 0;
  double Dx = 0;
  DevTables t;
  // work buffers
  double* Qe = nullptr;     // face contributions, element-major [((e*N_N+k)*NF+f)*M_s+j]
  double* Uacc = nullptr;   // RK accumulator
  double* Ustar = nullptr;  // stage state
  double* work = nullptr;   // limiter scratch (state size)
  double* work2 = nullptr;  // limiter scratch (state size)
  double* Uavg = nullptr;   // [e*NF+f]
  int* sensors = nullptr;   // [e]
  int* sens_tgt = nullptr;  // sensor resolution scratch [e*2]
  int* sens_act = nullptr;  // [e*2] (ping-pong)
  int* flag = nullptr;      // device flag / counters
  double* dscratch = nullptr;  // device scalars: [0] max UPA bits, [1] Dt, [2] T
  int n_thresh = 0;
  double thresh[3] = {0, 0, 0};
  std::vector<void*> allocations;
  // event timing (dgx_profile_*): [0] face, [1] elem, [2] limit, [3] dt
  bool prof_on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> prof[4];
  std::vector<hipEvent_t> ev_pool;
  long state_size() const { return (long)N_E * NF * N_s; }
};

// kernel launch entry points (defined in rhs.hip / limiter.hip)
hipError_t launch_face(const Solver& s, const double* U, hipStream_t st);
// mode 0: f_rk = S+F+Q to out; mode 1: RK stage `stage` update (see rhs.hip)
hipError_t launch_elem(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                       double* Uout, const double* dt_dev, double* f_out, hipStream_t st);
hipError_t launch_upa_max(const Solver& s, const double* U, double cfl, double* dt_dev, hipStream_t st);
hipError_t launch_limit_hr(Solver& s, double* U, hipStream_t st);
hipError_t launch_sensing(Solver& s, const double* U, int* sensors, double* Uavg, hipStream_t st);
hipError_t launch_limit_hri(Solver& s, double* U, hipStream_t st);

// drop-in compat: process-wide configuration for the reference-named launchers
const dgx_config& current_config();

}  // namespace dgx
