// Host side of the C ABI (include/dgx.h): setup tables, solver lifecycle,
// and the RK4 driver (RK::RK_integration, rk.cc:11-247).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dgx_internal.h"
#include "hr_tables.h"

namespace dgx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

}  // namespace dgx

using namespace dgx;

struct dgx_mesh {
  Tables t;
};

struct dgx_solver {
  Solver s;
  std::vector<double> ref_nodes;  // reference-element node coordinates [i*D + a]
};

namespace {

template <typename T>
bool upload(Solver& s, T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  if (!hip_ok(hipMalloc(&p, n * sizeof(T)), "hipMalloc")) return false;
  s.allocations.push_back(p);
  *dst = static_cast<T*>(p);
  if (src) return hip_ok(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return true;
}

template <typename T>
bool alloc(Solver& s, T** dst, size_t n) {
  return upload<T>(s, dst, nullptr, n);
}

void release_events(Solver& s) {
  for (auto& v : s.prof)
    for (auto& pr : v) {
      s.ev_pool.push_back(pr.first);
      s.ev_pool.push_back(pr.second);
    }
  for (auto& v : s.prof) v.clear();
}

void free_solver(Solver& s) {
  release_events(s);
  for (hipEvent_t e : s.ev_pool) (void)hipEventDestroy(e);
  s.ev_pool.clear();
  for (void* p : s.allocations) (void)hipFree(p);
  s.allocations.clear();
}

hipEvent_t take_event(Solver& s) {
  if (!s.ev_pool.empty()) {
    hipEvent_t e = s.ev_pool.back();
    s.ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events when profiling is on.
struct KernelTimer {
  Solver& s;
  int slot;
  hipStream_t st;
  hipEvent_t a = nullptr;
  KernelTimer(Solver& s_, int slot_, hipStream_t st_) : s(s_), slot(slot_), st(st_) {
    if (s.prof_on && (a = take_event(s))) (void)hipEventRecord(a, st);
  }
  ~KernelTimer() {
    if (!a) return;
    hipEvent_t b = take_event(s);
    if (!b) return;
    (void)hipEventRecord(b, st);
    s.prof[slot].emplace_back(a, b);
  }
};

hipStream_t as_stream(void* st) { return reinterpret_cast<hipStream_t>(st); }

dgx_status fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}


// Checks that the mesh is a Cartesian grid whose faces follow the structured
// pattern the tile kernel assumes (element e = ey*nx + ex; side 0 is the
// left/lower element, or the boundary element of a boundary face; closures
// and periodic partners as built by build_tables; normals along the axes)
// and gathers each face's normal and Jacobians into grid order.  Any
// mismatch leaves the general path.
bool cartesian_layout(const Tables& t, Cart& out) {
  Cart c;
  const int D = t.D, nx = t.nx, ny = (D == 2) ? t.ny : 1;
  if (nx * ny != t.N_E) return false;
  if (!stage_tiles(D, nx, ny, &c.tx, &c.ty)) return false;
  c.nx = nx;
  c.ny = ny;
  for (int k = 0; k < 4; ++k) c.bc[k] = t.bc[k];
  c.gx.assign((size_t)(nx + 1) * ny * 4, 0.0);
  c.gy.assign(D == 2 ? (size_t)nx * (ny + 1) * 4 : 4, 0.0);
  std::vector<char> seen_x((size_t)(nx + 1) * ny, 0), seen_y(D == 2 ? (size_t)nx * (ny + 1) : 0, 0);
  const bool perx = c.bc[0] == BC_PERIODIC, pery = D == 2 && c.bc[2] == BC_PERIODIC;
  for (int f = 0; f < t.M_T; ++f) {
    const int e0 = t.face_elem[2 * f], c0 = t.face_closure[2 * f];
    const int e1 = t.face_elem[2 * f + 1], c1 = t.face_closure[2 * f + 1], tag = t.face_tag[f];
    const int ex = e0 % nx, ey = e0 / nx;
    int we1, wc1, wtag;
    double wn[2] = {0, 0};
    size_t slot;
    bool xface;
    if (D == 1) {
      xface = true;
      if (c0 == 1 && ex < nx - 1) { wtag = 0; we1 = e0 + 1; wc1 = 0; wn[0] = 1; slot = ex + 1; }
      else if (c0 == 1) { wtag = c.bc[1]; wn[0] = 1; we1 = perx ? 0 : e0; wc1 = perx ? 0 : 1; slot = nx; }
      else if (c0 == 0 && ex == 0) { wtag = c.bc[0]; wn[0] = -1; we1 = perx ? nx - 1 : e0; wc1 = perx ? 1 : 0; slot = 0; }
      else return false;
    } else {
      xface = (c0 == 1 || c0 == 3);
      const size_t sx = (size_t)ey * (nx + 1), sy = (size_t)ey * nx + ex;
      if (c0 == 1 && ex < nx - 1) { wtag = 0; we1 = e0 + 1; wc1 = 7; wn[0] = 1; slot = sx + ex + 1; }
      else if (c0 == 1) { wtag = c.bc[1]; wn[0] = 1; we1 = perx ? ey * nx : e0; wc1 = perx ? 7 : 1; slot = sx + nx; }
      else if (c0 == 3 && ex == 0) { wtag = c.bc[0]; wn[0] = -1; we1 = perx ? ey * nx + nx - 1 : e0; wc1 = perx ? 5 : 3; slot = sx; }
      else if (c0 == 2 && ey < ny - 1) { wtag = 0; we1 = e0 + nx; wc1 = 4; wn[1] = 1; slot = sy + nx; }
      else if (c0 == 2) { wtag = c.bc[3]; wn[1] = 1; we1 = pery ? ex : e0; wc1 = pery ? 4 : 2; slot = sy + nx; }
      else if (c0 == 0 && ey == 0) { wtag = c.bc[2]; wn[1] = -1; we1 = pery ? (ny - 1) * nx + ex : e0; wc1 = pery ? 6 : 0; slot = sy; }
      else return false;
    }
    if (tag != wtag || e1 != we1 || c1 != wc1) return false;
    // normals of a Cartesian grid: the axis direction up to rounding
    for (int a = 0; a < D; ++a)
      if (!(std::fabs(t.normals[(size_t)f * D + a] - wn[a]) <= 1e-6)) return false;
    std::vector<char>& seen = xface ? seen_x : seen_y;
    if (seen[slot]) return false;
    seen[slot] = 1;
    double* g = (xface ? c.gx.data() : c.gy.data()) + slot * 4;
    g[0] = t.normals[(size_t)f * D];
    g[1] = (D == 2) ? t.normals[(size_t)f * D + 1] : 0.0;
    g[2] = t.JF[2 * f];
    g[3] = t.JF[2 * f + 1];
  }
  for (char v : seen_x)
    if (!v) return false;
  for (char v : seen_y)
    if (!v) return false;
  c.on = true;
  out = std::move(c);
  return true;
}

}  // namespace

extern "C" {

const char* dgx_last_error(void) { return g_err.c_str(); }

int dgx_nfields(const dgx_config* c) { return 2 + c->dim + (c->physics == 0 ? 0 : (c->physics == 1 ? 1 : 2)); }

// ------------------------------------------------------------------ mesh
dgx_status dgx_mesh_build(int dim, int order, int nx, int ny, double x0, double x1, double y0, double y1,
                          const int* bc, int nfields, dgx_mesh** out) {
  if (!out) return fail(DGX_EINVAL, "dgx_mesh_build: null output");
  MeshSpec sp;
  sp.dim = dim;
  sp.order = order;
  sp.nx = nx;
  sp.ny = ny;
  sp.x0 = x0; sp.x1 = x1; sp.y0 = y0; sp.y1 = y1;
  for (int k = 0; k < 4; ++k) sp.bc[k] = bc ? bc[k < 2 * dim ? k : 0] : 2;
  sp.nfields = nfields;
  dgx_mesh* m = new dgx_mesh;
  std::string err = build_tables(sp, m->t);
  if (!err.empty()) {
    delete m;
    return fail(DGX_EINVAL, "dgx_mesh_build: " + err);
  }
  *out = m;
  return DGX_OK;
}

dgx_status dgx_mesh_read_gmsh(const char* path, int nfields, dgx_mesh** out) {
  if (!path || !out) return fail(DGX_EINVAL, "dgx_mesh_read_gmsh: null argument");
  GmshMesh g;
  std::string err = read_gmsh(path, g);
  dgx_mesh* m = new dgx_mesh;
  if (err.empty()) err = build_tables_gmsh(g, nfields, m->t);
  if (!err.empty()) {
    delete m;
    return fail(DGX_EINVAL, std::string("dgx_mesh_read_gmsh: ") + err);
  }
  *out = m;
  return DGX_OK;
}

dgx_status dgx_mesh_write_gmsh(const dgx_mesh* m, const char* path) {
  if (!m || !path) return fail(DGX_EINVAL, "dgx_mesh_write_gmsh: null argument");
  std::string err = write_gmsh(m->t, path);
  return err.empty() ? DGX_OK : fail(DGX_EINVAL, "dgx_mesh_write_gmsh: " + err);
}

void dgx_mesh_free(dgx_mesh* m) { delete m; }

int dgx_mesh_tileable(const dgx_mesh* m) {
  Cart c;
  return (m && cartesian_layout(m->t, c)) ? 1 : 0;
}

void dgx_mesh_sizes(const dgx_mesh* m, int* z) {
  const Tables& t = m->t;
  int v[16] = {t.D, t.order, t.N_s, t.N_G, t.M_s, t.M_G, t.N_N, t.N_E, t.M_T, t.M_B, t.N_F, t.N_s1D, t.nx, t.ny,
               t.rflctiveIdx, 0};
  std::memcpy(z, v, sizeof(v));
}

double dgx_mesh_dx(const dgx_mesh* m) { return m->t.Dx; }

const double* dgx_mesh_f64(const dgx_mesh* m, const char* name, long* count) {
  const Tables& t = m->t;
  const std::map<std::string, const std::vector<double>*> tab = {
      {"phi", &t.phi}, {"dphi", &t.dphi}, {"phi_w", &t.phi_w}, {"dphi_w", &t.dphi_w}, {"psi", &t.psi},
      {"psi_w", &t.psi_w}, {"weight", &t.weight}, {"points", &t.points}, {"pointsF", &t.pointsF},
      {"weightF", &t.weightF}, {"ref_nodes", &t.ref_nodes}, {"xyz", &t.xyz}, {"J", &t.J}, {"invJac", &t.invJac},
      {"JF", &t.JF}, {"normals", &t.normals}, {"Minv", &t.Minv}, {"Lag2Mono", &t.Lag2Mono},
      {"Mono2Lag", &t.Mono2Lag}, {"Lag2MonoX", &t.Lag2MonoX}, {"MonoX2MonoY", &t.MonoX2MonoY},
      {"MonoY2Lag", &t.MonoY2Lag}, {"Lag2MonoY", &t.Lag2MonoY}, {"MonoX2Lag", &t.MonoX2Lag}};
  auto it = tab.find(name);
  if (it == tab.end()) return nullptr;
  if (count) *count = (long)it->second->size();
  return it->second->data();
}

const int* dgx_mesh_i32(const dgx_mesh* m, const char* name, long* count) {
  const Tables& t = m->t;
  const std::map<std::string, const std::vector<int>*> tab = {
      {"map", &t.map}, {"invmap", &t.invmap}, {"boundaryMap", &t.boundaryMap}, {"neighbors", &t.neighbors},
      {"face_elem", &t.face_elem}, {"face_closure", &t.face_closure}, {"face_tag", &t.face_tag}, {"elem_id", &t.elem_id},
      {"closures", &t.closures}};
  auto it = tab.find(name);
  if (it == tab.end()) return nullptr;
  if (count) *count = (long)it->second->size();
  return it->second->data();
}

// ---------------------------------------------------------------- solver
dgx_status dgx_solver_create(const dgx_config* cfg, const dgx_mesh* mesh, dgx_solver** out) {
  if (!cfg || !mesh || !out) return fail(DGX_EINVAL, "dgx_solver_create: null argument");
  const Tables& t = mesh->t;
  if (cfg->dim != t.D || cfg->order != t.order)
    return fail(DGX_EINVAL, "dgx_solver_create: configuration dim/order do not match the mesh");
  if (cfg->physics < 0 || cfg->physics > 3 || cfg->flux < 0 || cfg->flux > 1)
    return fail(DGX_EINVAL, "dgx_solver_create: unknown physics or flux");
  dgx_solver* h = new dgx_solver;
  h->ref_nodes = mesh->t.ref_nodes;
  Solver& s = h->s;
  s.cfg = *cfg;
  s.D = t.D; s.P = t.order; s.NF = dgx_nfields(cfg);
  s.N_s = t.N_s; s.N_G = t.N_G; s.M_s = t.M_s; s.M_G = t.M_G; s.N_N = t.N_N; s.N_E = t.N_E; s.M_T = t.M_T;
  s.N1 = t.order + 1;
  s.Dx = t.Dx;
  const int D = s.D, Ns = s.N_s, NN = s.N_N, NE = s.N_E;

  // per-element geometry, compacted (affine elements: values at g=0)
  std::vector<double> J(NE), ij((size_t)NE * D * D);
  bool uniform_geo = true;
  for (int e = 0; e < NE; ++e) {
    J[e] = t.J[e];
    for (int k = 0; k < D * D; ++k) ij[(size_t)e * D * D + k] = t.invJac[(size_t)e * t.N_G * D * D + k];
    if (e > 0 && (std::memcmp(&J[e], &J[0], 8) != 0 ||
                  std::memcmp(&ij[(size_t)e * D * D], &ij[0], 8 * D * D) != 0))
      uniform_geo = false;
  }
  // one inverse mass matrix when Minv_e = Minv_0 * J_0 / J_e for every
  // element to rounding (affine elements), else one per element
  bool uniform_minv = true;
  {
    double mmax = 0.0;
    for (int k = 0; k < Ns * Ns; ++k) mmax = std::max(mmax, std::fabs(t.Minv[k]));
    for (int e = 1; e < NE && uniform_minv; ++e) {
      const double r = t.J[0] / t.J[e];
      const double* Me = &t.Minv[(size_t)e * Ns * Ns];
      for (int k = 0; k < Ns * Ns; ++k)
        if (!(std::fabs(Me[k] - t.Minv[k] * r) <= 1e-13 * mmax * std::max(1.0, r))) {
          uniform_minv = false;
          break;
        }
    }
  }
  s.t.geo_stride = uniform_geo ? 0 : 1;
  s.t.minv_stride = uniform_minv ? 0 : 1;
  s.t.J0 = t.J[0];
  cartesian_layout(t, s.cart);

  // every element edge must receive exactly one contributing face side
  std::vector<int> ef((size_t)NE * NN, -1);
  for (int f = 0; f < t.M_T; ++f)
    for (int d = 0; d < 2; ++d) {
      if (d == 1 && t.face_tag[f] != 0) continue;
      int e = t.face_elem[f * 2 + d], cl = t.face_closure[f * 2 + d];
      int k = (D == 1) ? cl : cl % 4;
      int rev = (D == 2 && cl >= 4) ? 1 : 0;
      if (ef[(size_t)e * NN + k] != -1) {
        free_solver(s);
        delete h;
        return fail(DGX_EINVAL, "dgx_solver_create: element edge shared by two contributing faces");
      }
      ef[(size_t)e * NN + k] = ((f * 2 + d) << 1) | rev;
    }
  for (size_t k = 0; k < ef.size(); ++k)
    if (ef[k] < 0) {
      free_solver(s);
      delete h;
      return fail(DGX_EINVAL, "dgx_solver_create: element edge without a face");
    }

  bool ok = upload(s, &s.t.phi, t.phi.data(), t.phi.size()) && upload(s, &s.t.dphi, t.dphi.data(), t.dphi.size()) &&
            upload(s, &s.t.phi_w, t.phi_w.data(), t.phi_w.size()) &&
            upload(s, &s.t.dphi_w, t.dphi_w.data(), t.dphi_w.size()) &&
            upload(s, &s.t.psi, t.psi.data(), t.psi.size()) && upload(s, &s.t.psi_w, t.psi_w.data(), t.psi_w.size()) &&
            upload(s, &s.t.face_elem, t.face_elem.data(), t.face_elem.size()) &&
            upload(s, &s.t.face_closure, t.face_closure.data(), t.face_closure.size()) &&
            upload(s, &s.t.face_tag, t.face_tag.data(), t.face_tag.size()) &&
            upload(s, &s.t.normals, t.normals.data(), t.normals.size()) &&
            upload(s, &s.t.JF, t.JF.data(), t.JF.size()) &&
            upload(s, &s.t.J, J.data(), uniform_geo ? 1 : J.size()) &&
            upload(s, &s.t.invJac, ij.data(), uniform_geo ? (size_t)D * D : ij.size()) &&
            upload(s, &s.t.Minv, t.Minv.data(), uniform_minv ? (size_t)Ns * Ns : t.Minv.size()) &&
            upload(s, &s.t.neighbors, t.neighbors.data(), t.neighbors.size()) &&
            upload(s, &s.t.weight, t.weight.data(), t.weight.size());
  if (ok && s.cart.on)
    ok = upload(s, &s.t.face_geo_x, s.cart.gx.data(), s.cart.gx.size()) &&
         upload(s, &s.t.face_geo_y, s.cart.gy.data(), s.cart.gy.size());
  if (ok && D == 1)
    ok = upload(s, &s.t.Lag2Mono, t.Lag2Mono.data(), t.Lag2Mono.size()) &&
         upload(s, &s.t.Mono2Lag, t.Mono2Lag.data(), t.Mono2Lag.size());
  if (ok && D == 2)
    ok = upload(s, &s.t.Lag2MonoX, t.Lag2MonoX.data(), t.Lag2MonoX.size()) &&
         upload(s, &s.t.MonoX2MonoY, t.MonoX2MonoY.data(), t.MonoX2MonoY.size()) &&
         upload(s, &s.t.MonoY2Lag, t.MonoY2Lag.data(), t.MonoY2Lag.size()) &&
         upload(s, &s.t.Lag2MonoY, t.Lag2MonoY.data(), t.Lag2MonoY.size()) &&
         upload(s, &s.t.MonoX2Lag, t.MonoX2Lag.data(), t.MonoX2Lag.size());
  s.hr_const = D == 2 && s.P <= 2 && hr_tables_match(s.P, t.Lag2MonoX.data(), t.MonoX2MonoY.data(), t.MonoY2Lag.data());
  // k_hr2d's tiles are 8 x 8 or 8 x 4 elements whatever the stage tile
  if (s.cart.on && s.hr_const && s.NF <= 5 && s.cart.nx % 8 == 0 && s.cart.ny % 4 == 0) {
    // the one-launch HR limiter assumes e -/+ 1 and e -/+ nx inside a tile,
    // takes the rows below and above it and the ends of every row from the
    // table, and takes the x neighbours of a halo-row element to be the halo
    // elements of the x neighbours (p = 3 would not fit its LDS: two sweeps);
    // it carries the three transforms as compile-time constants, so the
    // solver's tables must equal them
    const int nx = s.cart.nx;
    auto nbr = [&](int e, int k) { return t.neighbors[(size_t)e * NN + k]; };
    s.hr2d = true;
    for (int e = 0; e < NE && s.hr2d; ++e) {
      const int ex = e % nx, ey = e / nx;
      if (ex > 0 && nbr(e, 0) != e - 1) s.hr2d = false;
      if (ex < nx - 1 && nbr(e, 1) != e + 1) s.hr2d = false;
      if (ey > 0 && nbr(e, 2) != e - nx) s.hr2d = false;
      if (ey < s.cart.ny - 1 && nbr(e, 3) != e + nx) s.hr2d = false;
      for (int d = 2; d < 4 && s.hr2d; ++d) {
        const int h = nbr(e, d);
        if (h < 0) continue;
        if (ex > 0 && nbr(h, 0) != nbr(e - 1, d)) s.hr2d = false;
        if (ex < nx - 1 && nbr(h, 1) != nbr(e + 1, d)) s.hr2d = false;
      }
    }
  }
  const size_t S = (size_t)s.state_size();
  ok = ok && alloc(s, &s.Qe, (size_t)NE * NN * s.NF * s.M_s) && alloc(s, &s.Uacc, S) && alloc(s, &s.Ustar, S) &&
       alloc(s, &s.work, S) && alloc(s, &s.work2, S) && alloc(s, &s.Uavg, (size_t)NE * s.NF) &&
       alloc(s, &s.sensors, (size_t)NE) && alloc(s, &s.sens_tgt, (size_t)2 * NE) &&
       alloc(s, &s.sens_act, (size_t)3 * NE) && alloc(s, &s.flag, 4) && alloc(s, &s.sens_changed, kRelaxBatch) && alloc(s, &s.flag_list, (size_t)NE) &&
       alloc(s, &s.dscratch, 8) && alloc(s, &s.time_state, 4) &&
       alloc(s, &s.cons_partial, conservation_scratch_doubles()) &&
       alloc(s, &s.dt_partial, ((size_t)NE * Ns + 255) / 256 + 1) &&
       ((!s.cart.on && D != 1) || alloc(s, &s.Ustar2, S));
  if (ok) ok = hip_ok(hipMemset(s.sensors, 0, sizeof(int) * NE), "hipMemset");
  if (!ok) {
    free_solver(s);
    delete h;
    return DGX_EDEVICE;
  }
  *out = h;
  return DGX_OK;
}

void dgx_solver_free(dgx_solver* h) {
  if (!h) return;
  free_solver(h->s);
  delete h;
}

long dgx_solver_state_size(const dgx_solver* h) { return h ? h->s.state_size() : 0; }

dgx_status dgx_rhs(dgx_solver* h, const double* U, double* f_rk, void* stream) {
  if (!h || !U || !f_rk) return fail(DGX_EINVAL, "dgx_rhs: null argument");
  Solver& s = h->s;
  hipStream_t st = as_stream(stream);
  if (s.tiled())
    return hip_ok(launch_stage(s, 0, 0, U, U, nullptr, nullptr, nullptr, f_rk, st), "dgx_rhs tile kernel")
               ? DGX_OK
               : DGX_EDEVICE;
  if (!hip_ok(launch_face(s, U, st), "dgx_rhs face kernel")) return DGX_EDEVICE;
  if (!hip_ok(launch_elem(s, 0, 0, U, U, nullptr, nullptr, nullptr, f_rk, st), "dgx_rhs element kernel"))
    return DGX_EDEVICE;
  return DGX_OK;
}

dgx_status dgx_set_sensor_thresholds(dgx_solver* h, int n, const double* th) {
  if (!h || n < 0 || n > 3 || (n > 0 && !th)) return fail(DGX_EINVAL, "dgx_set_sensor_thresholds: bad arguments");
  h->s.n_thresh = n;
  for (int k = 0; k < 3; ++k) h->s.thresh[k] = k < n ? th[k] : 0.0;
  return DGX_OK;
}

static dgx_status do_limit(Solver& s, int limiter, double* U, hipStream_t st) {
  if (limiter == 0) return DGX_OK;
  if (limiter == 1) return hip_ok(launch_limit_hr(s, U, st), "HR limiter") ? DGX_OK : DGX_EDEVICE;
  if (limiter == 2 || limiter == 4 || limiter == 5 || limiter == 6) {
    if (s.cfg.physics != DGX_MULTIFLUID)
      return fail(DGX_EUNSUPPORTED,
                  "m2l, m2i, prl and pri limit two-fluid states (the reference implements them for GAMNCON)");
    if (limiter == 2) return hip_ok(launch_limit_m2(s, U, st), "m2 limiter") ? DGX_OK : DGX_EDEVICE;
    if (limiter == 4) return hip_ok(launch_limit_m2i(s, U, st), "m2i limiter") ? DGX_OK : DGX_EDEVICE;
    if (limiter == 6) return hip_ok(launch_limit_pri(s, U, st), "primitive HRI limiter") ? DGX_OK : DGX_EDEVICE;
    return hip_ok(launch_limit_pr(s, U, st), "primitive HR limiter") ? DGX_OK : DGX_EDEVICE;
  }
  if (limiter == 3 || limiter == 7) {
    if (s.cfg.physics == DGX_SINGLEFLUID)
      return fail(DGX_EUNSUPPORTED, "hri and p0i limiting need the sensor, which reads past Uavg for single-fluid "
                                    "states in the reference (sensor_kernels.cu:55-60)");
    if (limiter == 7) return hip_ok(launch_limit_p0i(s, U, st), "p0i limiter") ? DGX_OK : DGX_EDEVICE;
    return hip_ok(launch_limit_hri(s, U, st), "HRI limiter") ? DGX_OK : DGX_EDEVICE;
  }
  if (limiter == 8) {
    if (!s.mlp) return fail(DGX_EUNSUPPORTED, "limiter 8 needs an indicator (dgx_solver_set_indicator)");
    // the indicator lists the flagged elements itself (no separate compaction pass)
    if (!hip_ok(mlp_indicator_flags(s, U, st, true), "indicator")) return DGX_EDEVICE;
    return hip_ok(launch_limit_listed(s, U, st), "flagged HR limiter") ? DGX_OK : DGX_EDEVICE;
  }
  return fail(DGX_EUNSUPPORTED,
              "limiter must be 0 (none), 1 (hrl), 2 (m2l), 3 (hri), 4 (m2i), 5 (prl), 6 (pri), 7 (p0i) or 8 "
              "(learned indicator + HR)");
}

dgx_status dgx_limit(dgx_solver* h, int limiter, double* U, void* stream) {
  if (!h || !U) return fail(DGX_EINVAL, "dgx_limit: null argument");
  return do_limit(h->s, limiter, U, as_stream(stream));
}

dgx_status dgx_limit_into(dgx_solver* h, int limiter, const double* In, double* Out, void* stream) {
  if (!h || !In || !Out) return fail(DGX_EINVAL, "dgx_limit_into: null argument");
  if (In == Out) return fail(DGX_EINVAL, "dgx_limit_into: In and Out must differ");
  Solver& s = h->s;
  hipStream_t st = as_stream(stream);
  if (limiter == 1 && (s.hr2d || s.D == 1))
    return hip_ok(launch_limit_hr_oop(s, In, Out, st), "HR limiter") ? DGX_OK : DGX_EDEVICE;
  if (!hip_ok(hipMemcpyAsync(Out, In, sizeof(double) * s.state_size(), hipMemcpyDeviceToDevice, st), "copy"))
    return DGX_EDEVICE;
  return do_limit(s, limiter, Out, st);
}

dgx_status dgx_format(dgx_solver* h, double* U, double* out, int inverse, void* stream) {
  if (!h || !U || !out) return fail(DGX_EINVAL, "dgx_format: null argument");
  return hip_ok(launch_format(h->s, U, out, inverse ? 1 : 0, as_stream(stream)), "formatter") ? DGX_OK : DGX_EDEVICE;
}

dgx_status dgx_pos_write(const dgx_mesh* m, int physics, const double* out, const char* prefix, int step,
                         double time) {
  if (!m || !out || !prefix) return fail(DGX_EINVAL, "dgx_pos_write: null argument");
  if (physics < DGX_SINGLEFLUID || physics > DGX_STIFFENED) return fail(DGX_EINVAL, "dgx_pos_write: physics");
  const std::string err = pos_write(m->t, physics, out, prefix, step, time);
  return err.empty() ? DGX_OK : fail(DGX_EINVAL, "dgx_pos_write: " + err);
}

dgx_status dgx_pos_read(const dgx_mesh* m, int physics, const char* prefix, int step, double* time, double* out) {
  if (!m || !out || !prefix) return fail(DGX_EINVAL, "dgx_pos_read: null argument");
  if (physics < DGX_SINGLEFLUID || physics > DGX_STIFFENED) return fail(DGX_EINVAL, "dgx_pos_read: physics");
  const std::string err = pos_read(m->t, physics, prefix, step, time, out);
  return err.empty() ? DGX_OK : fail(DGX_EINVAL, "dgx_pos_read: " + err);
}

dgx_status dgx_conservation(dgx_solver* h, const double* U, double* totals, void* stream) {
  if (!h || !U || !totals) return fail(DGX_EINVAL, "dgx_conservation: null argument");
  const Solver& s = h->s;
  if (s.N_G > 25 || s.N_s > 16) return fail(DGX_EUNSUPPORTED, "dgx_conservation: element larger than p=3");
  return hip_ok(launch_conservation(s, U, totals, s.cons_partial, as_stream(stream)), "conservation") ? DGX_OK
                                                                                                      : DGX_EDEVICE;
}

dgx_status dgx_sensing(dgx_solver* h, const double* U, int* sensors, double* Uavg, void* stream) {
  if (!h || !U || !sensors) return fail(DGX_EINVAL, "dgx_sensing: null argument");
  return hip_ok(launch_sensing(h->s, U, sensors, Uavg, as_stream(stream)), "sensing") ? DGX_OK : DGX_EDEVICE;
}

dgx_status dgx_dt_from_cfl(dgx_solver* h, const double* U, double cfl, double* dt_dev, void* stream) {
  if (!h || !U || !dt_dev) return fail(DGX_EINVAL, "dgx_dt_from_cfl: null argument");
  return hip_ok(launch_upa_max(h->s, U, cfl, dt_dev, as_stream(stream)), "dt kernel") ? DGX_OK : DGX_EDEVICE;
}

static dgx_status timed_limit(Solver& s, int limiter, double* U, hipStream_t st) {
  if (limiter == 0) return DGX_OK;
  KernelTimer kt(s, 2, st);
  return do_limit(s, limiter, U, st);
}

#ifndef DGX_RK_HR2D
// 1: on grids where the one-launch HR kernel applies (Solver::hr2d) the RK
// driver limits out of place with it; 0: the two sweep launches in place.
// On the 1024x512 p=2 two-fluid grid: 166 us against 212 us per application
// (the kernel computes each element's Taylor-x coefficients once per tile
// and keeps the intermediate state in LDS).
#define DGX_RK_HR2D 1
#endif

static dgx_status timed_hr_oop(Solver& s, const double* In, double* Out, hipStream_t st) {
  KernelTimer kt(s, 2, st);
  return hip_ok(launch_limit_hr_oop(s, In, Out, st), "HR limiter") ? DGX_OK : DGX_EDEVICE;
}

static dgx_status rk4_step(Solver& s, double* U, const double* dt, int limiter, hipStream_t st) {
  dgx_status r;
  if (DGX_RK_HR2D && limiter == 1 && ((s.tiled() && s.hr2d) || s.D == 1)) {
    // One-launch HR limiter (out of place; 2D tiles, or 1D): every stage
    // writes Ustar, the limiter takes Ustar to Ustar2 (the next stage's input)
    // and, after the last stage, to U.  U keeps the step's initial state
    // until then.
    const double* Uin = U;
    for (int k = 0; k < 4; ++k) {
      if (k > 0) {
        if ((r = timed_hr_oop(s, s.Ustar, s.Ustar2, st)) != DGX_OK) return r;
        Uin = s.Ustar2;
      }
      if (s.tiled()) {
        KernelTimer kt(s, 4, st);
        if (!hip_ok(launch_stage(s, 1, k, Uin, U, s.Uacc, s.Ustar, dt, nullptr, st), "tile kernel"))
          return DGX_EDEVICE;
        continue;
      }
      {
        KernelTimer kt(s, 0, st);
        if (!hip_ok(launch_face(s, Uin, st), "face kernel")) return DGX_EDEVICE;
      }
      KernelTimer kt(s, 1, st);
      if (!hip_ok(launch_elem(s, 1, k, Uin, U, s.Uacc, s.Ustar, dt, nullptr, st), "element kernel"))
        return DGX_EDEVICE;
    }
    return timed_hr_oop(s, s.Ustar, U, st);
  }
  // The tile kernel reads halo traces of neighbouring tiles while other
  // blocks write their stage output, so its stages ping-pong between two
  // buffers (U -> Ustar -> Ustar2 -> Ustar -> U); the face + element path
  // consumes all traces in the face kernel and updates Ustar in place.
  const bool tiled = s.tiled();
  // with the learned indicator the tile kernel also writes its output's cell
  // means, and the indicator that follows skips its own pass over the state
  double* means = (limiter == 8 && s.mlp && s.fuse_means) ? s.mlp_avg : nullptr;
  bool means_ready = false;
  auto limit = [&](double* V) {
    s.means_of = means_ready ? V : nullptr;
    const dgx_status q = timed_limit(s, limiter, V, st);
    s.means_of = nullptr;
    return q;
  };
  double* cur = U;
  for (int k = 0; k < 4; ++k) {
    const double* Uin = cur;
    if (k > 0 && (r = limit(cur)) != DGX_OK) return r;
    double* Uout = (k == 3) ? U : ((tiled && k == 1) ? s.Ustar2 : s.Ustar);
    cur = Uout;
    if (tiled) {
      KernelTimer kt(s, 4, st);
      if (!hip_ok(launch_stage(s, 1, k, Uin, U, s.Uacc, Uout, dt, nullptr, st, means,
                               means && s.shock_guard_on() ? s.Uavg : nullptr, s.mlp_psens, &means_ready),
                  "tile kernel"))
        return DGX_EDEVICE;
      continue;
    }
    means_ready = false;
    {
      KernelTimer kt(s, 0, st);
      if (!hip_ok(launch_face(s, Uin, st), "face kernel")) return DGX_EDEVICE;
    }
    {
      KernelTimer kt(s, 1, st);
      if (!hip_ok(launch_elem(s, 1, k, Uin, U, s.Uacc, Uout, dt, nullptr, st), "element kernel")) return DGX_EDEVICE;
    }
  }
  return limit(U);
}

dgx_status dgx_rk4_step(dgx_solver* h, double* U, const double* dt_dev, int limiter, void* stream) {
  if (!h || !U || !dt_dev) return fail(DGX_EINVAL, "dgx_rk4_step: null argument");
  return rk4_step(h->s, U, dt_dev, limiter, as_stream(stream));
}

// The CFL loop of a run without outputs, its time control on the device:
// steps are enqueued in chunks without waiting.  k_dt_step computes each
// step's Dt, stops the run on a tiny or NaN Dt and clamps the last step to the
// final time; once the run has stopped, every kernel of the steps still
// queued returns at once (Solver::gate).  The host reads the state back once
// per chunk; the next chunk covers the time left at the last Dt.
static dgx_status integrate_cfl_device(Solver& s, double* U, double cfl, int nsteps, double tf, int limiter,
                                       int* n_out_steps, double* T_out, hipStream_t st) {
  double* dt_dev = s.dscratch + 1;
  double* T_dev = s.dscratch + 2;
  struct Ungate {
    Solver& s;
    ~Ungate() { s.gate = nullptr; }
  } ungate{s};
  int hs[4] = {0, 0, 0, 0};
  double hd[2] = {0, 0};  // Dt, T
  auto finish = [&]() {
    *n_out_steps = hs[3];
    *T_out = hd[1];
  };
  if (!hip_ok(hipMemsetAsync(s.time_state, 0, sizeof(int) * 4, st), "time state") ||
      !hip_ok(hipMemsetAsync(T_dev, 0, sizeof(double), st), "time"))
    return DGX_EDEVICE;
  dgx_status r;
  int queued = 0, chunk = 4;
  for (;;) {
    const int k = nsteps > 0 ? std::min(chunk, nsteps - queued) : chunk;
    if (k <= 0) break;
    s.gate = s.time_state;
    for (int i = 0; i < k; ++i) {
      if (!hip_ok(launch_dt_step(s, U, cfl, tf, dt_dev, T_dev, s.time_state, st), "dt step")) return DGX_EDEVICE;
      if ((r = rk4_step(s, U, dt_dev, limiter, st)) != DGX_OK) return r;
    }
    s.gate = nullptr;
    queued += k;
    if (!hip_ok(hipMemcpyAsync(hs, s.time_state, sizeof(hs), hipMemcpyDeviceToHost, st), "time state readback") ||
        !hip_ok(hipMemcpyAsync(hd, dt_dev, sizeof(hd), hipMemcpyDeviceToHost, st), "time readback") ||
        !hip_ok(hipStreamSynchronize(st), "time readback"))
      return DGX_EDEVICE;
    if (hs[2]) {
      finish();
      return fail(DGX_EDT, "time step too small or NaN (rk.cc:147-148)");
    }
    if (hs[0] || hs[1]) break;
    const double left = (tf - hd[1]) / hd[0];
    chunk = left < 1 ? 1 : (int)std::min(256.0, std::ceil(left) + 1);
  }
  finish();
  return DGX_OK;
}

// RK::RK_integration's CFL branch (rk.cc:100-233): the initial limit, then
// per step Dt from CFL (one 8-byte read-back, as DtFromCFL's, rk.cc:265-268),
// the stop on a tiny or NaN Dt (rk.cc:147-148), the clamps to the final time
// and to the next output time (rk.cc:149-150), the step, and the output call
// after a clamped step (rk.cc:215-233).  out_times are the reference's
// _output_time_array[1..]; the last one is the final time.
static dgx_status integrate_cfl(Solver& s, double* U, double cfl, int nsteps, const double* out_times, int n_out,
                                int limiter, dgx_output_fn on_output, void* user, int* n_out_steps, double* T_out,
                                hipStream_t st) {
  double* dt_dev = s.dscratch + 1;
  dgx_status r;
  int n = 0, count = 0;
  double T = 0;
  const double tf = out_times[n_out - 1];
  auto finish = [&]() {
    *n_out_steps = n;
    *T_out = T;
  };
  if ((r = do_limit(s, limiter, U, st)) != DGX_OK) return r;
  if (s.device_time && !on_output && n_out == 1 && (limiter == 0 || limiter == 1))
    return integrate_cfl_device(s, U, cfl, nsteps, tf, limiter, n_out_steps, T_out, st);
  bool done = false;
  while (!done && (nsteps <= 0 || n < nsteps)) {
    {
      KernelTimer kt(s, 3, st);
      if (!hip_ok(launch_upa_max(s, U, cfl, dt_dev, st), "dt kernel")) return DGX_EDEVICE;
    }
    double dt = 0;
    if (!hip_ok(hipMemcpyAsync(&dt, dt_dev, sizeof(double), hipMemcpyDeviceToHost, st), "dt readback") ||
        !hip_ok(hipStreamSynchronize(st), "dt readback"))
      return DGX_EDEVICE;
    if (dt < 1e-14 || dt != dt) {
      finish();
      return fail(DGX_EDT, "time step too small or NaN (rk.cc:147-148)");
    }
    bool output = false;
    const double tout = out_times[count < n_out ? count : n_out - 1];
    if (dt > (tf - T)) {
      dt = tf - T;
      output = done = true;
    } else if (dt > (tout - T)) {
      dt = tout - T;
      output = true;
    }
    if (output && (!hip_ok(hipMemcpyAsync(dt_dev, &dt, sizeof(double), hipMemcpyHostToDevice, st), "dt clamp") ||
                   !hip_ok(hipStreamSynchronize(st), "dt clamp")))
      return DGX_EDEVICE;
    if ((r = rk4_step(s, U, dt_dev, limiter, st)) != DGX_OK) return r;
    T = T + dt;
    ++n;
    if (output) {
      ++count;
      if (on_output && count <= n_out) {
        if (!hip_ok(hipStreamSynchronize(st), "rk4 output")) return DGX_EDEVICE;
        if ((r = on_output(user, count, n, T, U, st)) != DGX_OK) {
          finish();
          return r;
        }
      }
    }
  }
  finish();
  return DGX_OK;
}

dgx_status dgx_rk4_integrate(dgx_solver* h, double* U, int mode, double dt_or_cfl, int nsteps, double tf, int limiter,
                             int* steps, double* t_final, void* stream) {
  if (!h || !U) return fail(DGX_EINVAL, "dgx_rk4_integrate: null argument");
  Solver& s = h->s;
  hipStream_t st = as_stream(stream);
  double* dt_dev = s.dscratch + 1;
  dgx_status r;
  int n = 0;
  double T = 0;
  if (mode == 0) {
    if (!hip_ok(hipMemcpyAsync(dt_dev, &dt_or_cfl, sizeof(double), hipMemcpyHostToDevice, st), "dt upload") ||
        !hip_ok(hipStreamSynchronize(st), "dt upload"))
      return DGX_EDEVICE;
    for (; n < nsteps; ++n) {
      if ((r = rk4_step(s, U, dt_dev, limiter, st)) != DGX_OK) return r;
      T = T + dt_or_cfl;
    }
  } else if (mode == 1) {
    r = integrate_cfl(s, U, dt_or_cfl, nsteps, &tf, 1, limiter, nullptr, nullptr, &n, &T, st);
    if (r != DGX_OK) {
      if (steps) *steps = n;
      if (t_final) *t_final = T;
      return r;
    }
  } else {
    return fail(DGX_EINVAL, "dgx_rk4_integrate: mode must be 0 (fixed Dt) or 1 (CFL)");
  }
  if (!hip_ok(hipStreamSynchronize(st), "rk4 integrate")) return DGX_EDEVICE;
  if (steps) *steps = n;
  if (t_final) *t_final = T;
  return DGX_OK;
}

dgx_status dgx_rk4_integrate_out(dgx_solver* h, double* U, double cfl, const double* out_times, int n_out,
                                 int limiter, dgx_output_fn on_output, void* user, int* steps, double* t_final,
                                 void* stream) {
  if (!h || !U || !out_times || n_out < 1) return fail(DGX_EINVAL, "dgx_rk4_integrate_out: bad arguments");
  for (int k = 0; k < n_out; ++k)
    if (!(out_times[k] > (k ? out_times[k - 1] : 0.0)))
      return fail(DGX_EINVAL, "dgx_rk4_integrate_out: output times must increase from above 0");
  hipStream_t st = as_stream(stream);
  int n = 0;
  double T = 0;
  dgx_status r = integrate_cfl(h->s, U, cfl, 0, out_times, n_out, limiter, on_output, user, &n, &T, st);
  if (r == DGX_OK && !hip_ok(hipStreamSynchronize(st), "rk4 integrate")) r = DGX_EDEVICE;
  if (steps) *steps = n;
  if (t_final) *t_final = T;
  return r;
}

// Tensor-Legendre transform of the 2D p=2 element: V[i][a*3+b] =
// P_a(xi_i) P_b(eta_i) at the element's nodes, inverted by Gauss-Jordan with
// partial pivoting; Vinv[q][i] maps nodal values to modal coefficients.
static bool legendre_vinv(const std::vector<double>& nodes, int Ns, std::vector<double>& vinv) {
  auto P = [](int a, double x) { return a == 0 ? 1.0 : (a == 1 ? x : 0.5 * (3 * x * x - 1)); };
  std::vector<double> A((size_t)Ns * 2 * Ns, 0.0);
  for (int i = 0; i < Ns; ++i) {
    for (int q = 0; q < Ns; ++q) A[(size_t)i * 2 * Ns + q] = P(q / 3, nodes[i * 2]) * P(q % 3, nodes[i * 2 + 1]);
    A[(size_t)i * 2 * Ns + Ns + i] = 1.0;
  }
  for (int c = 0; c < Ns; ++c) {
    int piv = c;
    for (int r = c + 1; r < Ns; ++r)
      if (std::fabs(A[(size_t)r * 2 * Ns + c]) > std::fabs(A[(size_t)piv * 2 * Ns + c])) piv = r;
    if (std::fabs(A[(size_t)piv * 2 * Ns + c]) < 1e-12) return false;
    for (int k = 0; k < 2 * Ns; ++k) std::swap(A[(size_t)c * 2 * Ns + k], A[(size_t)piv * 2 * Ns + k]);
    const double d = A[(size_t)c * 2 * Ns + c];
    for (int k = 0; k < 2 * Ns; ++k) A[(size_t)c * 2 * Ns + k] /= d;
    for (int r = 0; r < Ns; ++r) {
      if (r == c) continue;
      const double f = A[(size_t)r * 2 * Ns + c];
      for (int k = 0; k < 2 * Ns; ++k) A[(size_t)r * 2 * Ns + k] -= f * A[(size_t)c * 2 * Ns + k];
    }
  }
  // inverse of V (nodes x modes) is modes x nodes: Vinv[q][i] = A[q][Ns + i]
  vinv.assign((size_t)Ns * Ns, 0.0);
  for (int q = 0; q < Ns; ++q)
    for (int i = 0; i < Ns; ++i) vinv[(size_t)q * Ns + i] = A[(size_t)q * 2 * Ns + Ns + i];
  return true;
}

dgx_status dgx_solver_set_indicator(dgx_solver* h, dgx_mlp* mlp, float threshold) {
  if (!h) return fail(DGX_EINVAL, "dgx_solver_set_indicator: null solver");
  Solver& s = h->s;
  if (mlp && !(s.D == 2 && s.P == 2 && s.cfg.physics == DGX_MULTIFLUID))
    return fail(DGX_EUNSUPPORTED, "dgx_solver_set_indicator: the indicator takes 2D p=2 two-fluid states");
  if (mlp && !s.mlp_x) {
    std::vector<double> vinv;
    if (!legendre_vinv(h->ref_nodes, s.N_s, vinv)) return fail(DGX_EINVAL, "dgx_solver_set_indicator: singular transform");
    if (!upload(s, &s.t.mlp_vinv, vinv.data(), vinv.size()) || !alloc(s, &s.mlp_avg, (size_t)s.N_E * 3) ||
        !alloc(s, &s.mlp_x, (size_t)s.N_E * 64) || !alloc(s, &s.mlp_z, (size_t)s.N_E) ||
        !alloc(s, &s.mlp_jumps, (size_t)s.N_E * 4) || !alloc(s, &s.mlp_psens, (size_t)s.N_E) ||
        !alloc(s, &s.mlp_guard, (size_t)s.N_E))
      return DGX_EDEVICE;
  }
  s.mlp = mlp;
  s.mlp_threshold = threshold;
  return DGX_OK;
}

dgx_status dgx_solver_set_shock_guard(dgx_solver* h, int on) {
  if (!h) return fail(DGX_EINVAL, "dgx_solver_set_shock_guard: null solver");
  h->s.shock_guard = on != 0;
  return DGX_OK;
}

dgx_status dgx_solver_set_device_time(dgx_solver* h, int on) {
  if (!h) return fail(DGX_EINVAL, "dgx_solver_set_device_time: null solver");
  h->s.device_time = on != 0;
  return DGX_OK;
}

dgx_status dgx_solver_set_fused_means(dgx_solver* h, int on) {
  if (!h) return fail(DGX_EINVAL, "dgx_solver_set_fused_means: null solver");
  h->s.fuse_means = on != 0;
  return DGX_OK;
}

dgx_status dgx_mlp_features(dgx_solver* h, const double* U, float* X, void* stream) {
  if (!h || !U || !X) return fail(DGX_EINVAL, "dgx_mlp_features: null argument");
  Solver& s = h->s;
  if (!s.mlp_x) return fail(DGX_EUNSUPPORTED, "dgx_mlp_features: attach an indicator first");
  return hip_ok(mlp_features(s, U, X, as_stream(stream)), "features") ? DGX_OK : DGX_EDEVICE;
}

dgx_status dgx_indicator_flags(dgx_solver* h, const double* U, int* flags, float* z, void* stream) {
  if (!h || !U || !flags) return fail(DGX_EINVAL, "dgx_indicator_flags: null argument");
  Solver& s = h->s;
  if (!s.mlp) return fail(DGX_EUNSUPPORTED, "dgx_indicator_flags: no indicator attached");
  hipStream_t st = as_stream(stream);
  if (!hip_ok(mlp_indicator_flags(s, U, st), "indicator")) return DGX_EDEVICE;
  bool ok = hip_ok(hipMemcpyAsync(flags, s.sensors, sizeof(int) * s.N_E, hipMemcpyDeviceToDevice, st), "flags copy");
  if (ok && z) ok = hip_ok(hipMemcpyAsync(z, s.mlp_z, sizeof(float) * s.N_E, hipMemcpyDeviceToDevice, st), "z copy");
  return ok ? DGX_OK : DGX_EDEVICE;
}

dgx_status dgx_solver_set_path(dgx_solver* h, int path) {
  if (!h || path < 0 || path > 2) return fail(DGX_EINVAL, "dgx_solver_set_path: path must be 0, 1 or 2");
  if (path == PATH_TILED && !h->s.cart.on)
    return fail(DGX_EUNSUPPORTED, "dgx_solver_set_path: the tile kernel needs a uniform Cartesian mesh whose "
                                  "element counts are multiples of the tile shape");
  h->s.path = path;
  if (!h->s.tiled()) h->s.precision = 0;  // the face + element kernels are FP64
  return DGX_OK;
}

dgx_status dgx_solver_set_precision(dgx_solver* h, int precision) {
  if (!h || precision < 0 || precision > 2)
    return fail(DGX_EINVAL, "dgx_solver_set_precision: precision must be 0, 1 or 2");
  Solver& s = h->s;
  if (precision > 0 && !(s.tiled() && s.D == 2 && s.P == 2 && s.cfg.physics <= DGX_MULTIFLUID))
    return fail(DGX_EUNSUPPORTED, "dgx_solver_set_precision: FP32 and split-BF16 contractions need the tile "
                                  "kernel, 2D, p=2, single-fluid or two-fluid states");
  s.precision = precision;
  return DGX_OK;
}

int dgx_solver_path(const dgx_solver* h) { return (h && h->s.tiled()) ? PATH_TILED : PATH_GENERAL; }
int dgx_solver_hr_one_launch(const dgx_solver* h) { return (h && h->s.hr2d) ? 1 : 0; }

dgx_status dgx_profile_enable(dgx_solver* h, int enable) {
  if (!h) return fail(DGX_EINVAL, "dgx_profile_enable: null solver");
  h->s.prof_on = enable != 0;
  return DGX_OK;
}

dgx_status dgx_profile_read(dgx_solver* h, const char* kernel, double* total_ms, long* launches) {
  if (!h || !kernel) return fail(DGX_EINVAL, "dgx_profile_read: null argument");
  static const char* names[5] = {"face", "elem", "limit", "dt", "stage"};
  int slot = -1;
  for (int k = 0; k < 5; ++k)
    if (std::strcmp(kernel, names[k]) == 0) slot = k;
  if (slot < 0) return fail(DGX_EINVAL, "dgx_profile_read: kernel must be face, elem, limit, dt or stage");
  Solver& s = h->s;
  double tot = 0;
  for (auto& pr : s.prof[slot]) {
    if (!hip_ok(hipEventSynchronize(pr.second), "hipEventSynchronize")) return DGX_EDEVICE;
    float ms = 0;
    if (!hip_ok(hipEventElapsedTime(&ms, pr.first, pr.second), "hipEventElapsedTime")) return DGX_EDEVICE;
    tot += ms;
  }
  if (total_ms) *total_ms = tot;
  if (launches) *launches = (long)s.prof[slot].size();
  for (auto& pr : s.prof[slot]) {
    s.ev_pool.push_back(pr.first);
    s.ev_pool.push_back(pr.second);
  }
  s.prof[slot].clear();
  return DGX_OK;
}

}  // extern "C"
