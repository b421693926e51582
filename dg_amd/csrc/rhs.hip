// Fused DG right-hand side and RK4 stage kernels (FP64).
//
// The reference evaluates f(U) = S + F + Q with fifteen launches and
// quadrature-resolution temporaries in HBM (dg_solver.h:292-364).  Here one
// RHS is two kernels:
//
//   k_face  one thread per interface: gathers both traces straight from U
//           through the face closures (mapToFace, kernels.cu:41-75), applies
//           the wall reflection (boundaries.cu:20-67), interpolates to the face
//           Gauss points (psi), evaluates the Riemann solver (evaluate_q,
//           physics.cu:367-811), scales by the face Jacobian (redistribute_q)
//           and integrates against psi_w.  Writes each contributing side's
//           nodal contributions into an element-major buffer Qe (the
//           reference's Qtcj values, placed where mapToElement sends them).
//   k_elem  a block of elements: maps U to the volume Gauss points, evaluates
//           the volume flux and alpha source (evaluate_sf), applies the metric
//           terms (redistribute_sf), integrates against phi_w / dphi_w, adds
//           its Qe contributions (mapToElement: at most two per node, so the
//           sum is order independent; no atomics), forms S+F+Q (addSFQ) and,
//           in RK mode, applies Dt*Minv (solve, rk_kernels.cu:22-56) and the
//           RK4 stage update (rk.cc:164-198).
#include <type_traits>

#include "dgx_internal.h"
#include "physics.cuh"
#include "volume.cuh"

namespace dgx {

struct FaceArgs {
  const double* __restrict__ U;
  const int* __restrict__ face_elem;
  const int* __restrict__ face_closure;
  const int* __restrict__ face_tag;
  const double* __restrict__ normals;
  const double* __restrict__ JF;
  const double* __restrict__ psi;
  const double* __restrict__ psi_w;
  double* __restrict__ Qe;
  int M_T;
  const int* gate;  // set and nonzero: the launch does nothing (a step skipped by the device-side time control)
};

template <int D, int P, int PH, int FL>
__global__ __launch_bounds__(256) void k_face(FaceArgs a) {
  using S = Shape<D, P>;
  constexpr int NF = Fields<D, PH>::NF, Ms = S::Ms, MG = S::MG, Ns = S::Ns;
  if (a.gate && *a.gate) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.M_T) return;
  const int e0 = a.face_elem[2 * t], e1 = a.face_elem[2 * t + 1];
  const int c0 = a.face_closure[2 * t], c1 = a.face_closure[2 * t + 1];
  double uf[2][NF][Ms];
#pragma unroll
  for (int j = 0; j < Ms; ++j) {
    const int n0 = closure_node<D, P>(c0, j), n1 = closure_node<D, P>(c1, j);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      uf[0][f][j] = a.U[(e0 * NF + f) * Ns + n0];
      uf[1][f][j] = a.U[(e1 * NF + f) * Ns + n1];
    }
  }
  double nx = a.normals[t * D + 0];
  double ny = (D == 2) ? a.normals[t * D + 1] : 0.0;
  if (a.face_tag[t] == BC_REFLECTIVE) {
#pragma unroll
    for (int j = 0; j < Ms; ++j) {
      if constexpr (D == 1) {
        uf[1][1][j] = -uf[1][1][j];
      } else {
        double nx2 = nx * nx, ny2 = ny * ny, m2nxny = -2 * nx * ny, inv = 1.0 / (nx2 + ny2);
        double vxL = uf[0][1][j], vyL = uf[0][2][j];
        uf[1][1][j] = inv * ((-nx2 + ny2) * vxL + m2nxny * vyL);
        uf[1][2][j] = inv * (m2nxny * vxL + (nx2 - ny2) * vyL);
      }
    }
  }
  const double jf0 = a.JF[2 * t], jf1 = a.JF[2 * t + 1];
  double qt[2][NF][Ms];
#pragma unroll
  for (int d = 0; d < 2; ++d)
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int j = 0; j < Ms; ++j) qt[d][f][j] = 0.0;
#pragma unroll
  for (int g = 0; g < MG; ++g) {
    double ul[NF], ur[NF], F[NF], nc[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double sl = 0.0, sr = 0.0;
#pragma unroll
      for (int j = 0; j < Ms; ++j) {
        sl += uf[0][f][j] * a.psi[g + j * MG];
        sr += uf[1][f][j] * a.psi[g + j * MG];
      }
      ul[f] = sl;
      ur[f] = sr;
    }
    riemann<D, PH, FL>(ul, ur, nx, ny, F, nc);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double q0 = (-F[f] + nc[f]) * jf0;
      const double q1 = (F[f] + nc[f]) * jf1;
#pragma unroll
      for (int j = 0; j < Ms; ++j) {
        qt[0][f][j] += a.psi_w[g + j * MG] * q0;
        qt[1][f][j] += a.psi_w[g + j * MG] * q1;
      }
    }
  }
  // element-major output: Qe[((e*NN + k)*NF + f)*Ms + j] in the forward
  // closure order of local edge k; side 1 of a boundary face contributes
  // nothing (dg_functions.cc:225-233)
  const int tag = a.face_tag[t];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    if (d == 1 && tag != 0) break;
    const int e = d ? e1 : e0, cl = d ? c1 : c0;
    const int k = (D == 1) ? cl : (cl & 3);
    const bool rev = (D == 2) && cl >= 4;
    double* q = a.Qe + (size_t)(e * S::NN + k) * NF * Ms;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int j = 0; j < Ms; ++j) q[f * Ms + j] = qt[d][f][rev ? reversed_pos<P>(j) : j];
  }
}

struct ElemArgs {
  const double* __restrict__ phi;
  const double* __restrict__ dphi;
  const double* __restrict__ phi_w;
  const double* __restrict__ dphi_w;
  const double* __restrict__ J;
  const double* __restrict__ invJac;
  const double* __restrict__ Minv;
  int geo_stride, minv_stride;
  double J0;  // minv_stride == 0: Minv_e = Minv * J0 / J_e
  const double* __restrict__ Qe;
  int N_E;
  int stage;
  const double* Uin;
  const double* Un;
  double* Uacc;
  double* Uout;
  const double* __restrict__ dt;
  double* fout;
  const int* gate;  // as FaceArgs::gate
};

// One block of 256 threads handles EPB elements.  Phase 1: lane = (element,
// quadrature point); phase 2: lane = (element, node); phase 3: flat over the
// block's contiguous state (coalesced).  Element state, face contributions,
// per-point fluxes and the reference tables are staged in LDS.
template <int D, int P, int PH, int FL, int MODE>
__global__ __launch_bounds__(256) void k_elem(ElemArgs a) {
  if (a.gate && *a.gate) return;
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, NU = V::NU, NQE = V::NQE, NQ = V::NQ;
  constexpr int TPE = V::TPE, EPB = V::EPB;
  __shared__ double sU[EPB * NU];
  __shared__ double sQ[EPB * NQE];
  __shared__ double sF[EPB * NG * NQ];
  __shared__ double sPhi[V::NTAB_PHI];
  __shared__ double sDphi[PH ? V::NTAB_D : 1];
  __shared__ double sDw[V::NTAB_D];
  __shared__ double sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns];
  __shared__ double sScale[EPB];
  const VolTables<D, P, PH> tb{sPhi, sDphi, sDw, sPw, sMinv};

  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  const int n_el = min(EPB, a.N_E - e0);
  const bool uniform_minv = a.minv_stride == 0;
  load_vol_tables(tb, a.phi, a.dphi, a.phi_w, a.dphi_w, a.Minv, MODE == 1 && uniform_minv, tid);
  {  // every load in flight before the first LDS store
    constexpr int RU = (EPB * NU + 255) / 256, RQ = (EPB * NQE + 255) / 256;
    const double* src = a.Uin + (size_t)e0 * NU;
    const double* q = a.Qe + (size_t)e0 * NQE;
    double ru[RU], rq[RQ];
#pragma unroll
    for (int i = 0; i < RU; ++i) ru[i] = tid + 256 * i < n_el * NU ? src[tid + 256 * i] : 0.0;
#pragma unroll
    for (int i = 0; i < RQ; ++i) rq[i] = tid + 256 * i < n_el * NQE ? q[tid + 256 * i] : 0.0;
#pragma unroll
    for (int i = 0; i < RU; ++i)
      if (tid + 256 * i < n_el * NU) sU[tid + 256 * i] = ru[i];
#pragma unroll
    for (int i = 0; i < RQ; ++i)
      if (tid + 256 * i < n_el * NQE) sQ[tid + 256 * i] = rq[i];
  }
  __syncthreads();

  {  // phase 1: fluxes at the quadrature points
    const int el = tid / TPE, g = tid % TPE;
    if (el < n_el && g < NG) {
      const int eg = (e0 + el) * a.geo_stride;
      double ij[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) ij[k] = a.invJac[eg * D * D + k];
      vol_phase1<D, P, PH>(tb, g, sU + el * NU, a.J[eg], ij, sF + (el * NG + g) * NQ);
    }
  }
  __syncthreads();
  {  // phase 2: integrate against the test functions, add face terms
    const int el = tid / TPE, i = tid % TPE;
    double rhs[NF];
    const bool act = el < n_el && i < Ns;
    if (act) vol_phase2<D, P, PH>(tb, i, sF + el * NG * NQ, sQ + el * NQE, rhs);
    if (MODE == 1 && uniform_minv && act && i == 0) sScale[el] = a.J0 / a.J[(e0 + el) * a.geo_stride];
    __syncthreads();  // sU is reused for the right-hand side
    if (act)
#pragma unroll
      for (int f = 0; f < NF; ++f) sU[el * NU + f * Ns + i] = rhs[f];
  }
  __syncthreads();

  // phase 3: coalesced output
  const size_t base = (size_t)e0 * NU;
  if constexpr (MODE == 0) {
    for (int k = tid; k < n_el * NU; k += 256) a.fout[base + k] = sU[k];
  } else {
    const double dt = a.dt[0];
    for (int k = tid; k < n_el * NU; k += 256) {
      const int el = k / NU;
      const double dte = uniform_minv ? dt * sScale[el] : dt;
      rk_update<D, P, PH>(tb, a.Minv + (size_t)(e0 + el) * Ns * Ns, uniform_minv, k % NU, sU + el * NU, dte, a.stage,
                          base + k, a.Un, a.Uacc, a.Uout);
    }
  }
}

// The same for 2D p=2 single-fluid and two-fluid states with the element
// phases on the FP64 matrix cores (vol_mfma_wave): a block of 256 threads
// handles 32 consecutive elements; with the shared mass-matrix inverse the
// wave also forms Dt*Minv*R.  Padded per-element LDS strides (a multiple of
// 4 doubles apart for neighbouring elements: distinct bank groups).
constexpr bool elem_bank_ok(int n) { return (2 * n) % 32 >= 4 && (2 * n) % 32 <= 28; }
constexpr int elem_stride(int n) { return elem_bank_ok(n) ? n : elem_stride(n + 1); }

template <int PH, int FL, int MODE>
__global__ __launch_bounds__(256) void k_elem_mfma(ElemArgs a) {
  if (a.gate && *a.gate) return;
  constexpr int D = 2, P = 2, NT = 32, NGEO = 1 + D * D;
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NU = V::NU, NQE = V::NQE;
  constexpr int SU = elem_stride(NU), SQ = elem_stride(NQE);
  __shared__ double sU[NT * SU];
  __shared__ double sQ[NT * SQ];
  __shared__ double sW[4 * vol_wave_lds<double, PH>()];
  __shared__ double sPhi[V::NTAB_PHI];
  __shared__ double sDphi[PH ? V::NTAB_D : 1];
  __shared__ double sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns];
  __shared__ double sGeo[NT * NGEO];
  __shared__ double sDte[NT];
  __shared__ int sNodeQ[18];
  const VolTables<D, P, PH> tb{sPhi, sDphi, nullptr, sPw, sMinv};

  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * NT;
  const int n_el = min(NT, a.N_E - e0);
  const bool uniform_minv = a.minv_stride == 0;
  VolTableFetch<D, P, PH, false> tf;
  tf.fetch(a.phi, a.dphi, a.phi_w, a.dphi_w, a.Minv, MODE == 1 && uniform_minv, tid);
  constexpr int RU = (NT * NU + 255) / 256, RQ = (NT * NQE + 255) / 256, RG = (NT * NGEO + 255) / 256;
  double ru[RU], rq[RQ], rg[RG];
  // elements past the end of the mesh (last block) repeat the block's first element
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int k = tid + 256 * r, el = k / NU;
    if (k < NT * NU) ru[r] = a.Uin[(size_t)(e0 + (el < n_el ? el : 0)) * NU + k % NU];
  }
#pragma unroll
  for (int r = 0; r < RQ; ++r) {
    const int k = tid + 256 * r, el = k / NQE;
    if (k < NT * NQE) rq[r] = a.Qe[(size_t)(e0 + (el < n_el ? el : 0)) * NQE + k % NQE];
  }
#pragma unroll
  for (int r = 0; r < RG; ++r) {
    const int k = tid + 256 * r, el = k / NGEO, q = k % NGEO;
    if (k < NT * NGEO) {
      const int eg = (e0 + (el < n_el ? el : 0)) * a.geo_stride;
      rg[r] = q == 0 ? a.J[eg] : a.invJac[eg * D * D + q - 1];
    }
  }
  tf.store(tb, MODE == 1 && uniform_minv, tid);
  node_q_table<NF>(sNodeQ, tid);
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int k = tid + 256 * r;
    if (k < NT * NU) sU[(k / NU) * SU + k % NU] = ru[r];
  }
#pragma unroll
  for (int r = 0; r < RQ; ++r) {
    const int k = tid + 256 * r;
    if (k < NT * NQE) sQ[(k / NQE) * SQ + k % NQE] = rq[r];
  }
#pragma unroll
  for (int r = 0; r < RG; ++r) {
    const int k = tid + 256 * r;
    if (k < NT * NGEO) sGeo[k] = rg[r];
  }
  __syncthreads();
  const double dt = MODE == 1 ? a.dt[0] : 0.0;
  if (MODE == 1 && tid < NT) sDte[tid] = uniform_minv ? dt * (a.J0 / sGeo[tid * NGEO]) : dt;
  // RK operands of this thread's output entries, in flight during the element phases
  double pre_un[RU], pre_acc[RU];
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    pre_un[r] = pre_acc[r] = 0.0;
    const int k = tid + 256 * r;
    if (MODE == 1 && k < n_el * NU) {
      if (a.stage < 3) pre_un[r] = a.Un[(size_t)e0 * NU + k];
      if (a.stage > 0) pre_acc[r] = a.Uacc[(size_t)e0 * NU + k];
    }
  }
  __syncthreads();
  const bool wave_minv = MODE == 1 && uniform_minv;
  vol_mfma_wave<double, PH, NT, SU, SQ, NGEO>(tb, a.dphi_w, sU, sGeo, sQ, sW, sNodeQ, tid,
                                              wave_minv ? tb.minv : nullptr, sDte);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RU; ++r) {
    const int k = tid + 256 * r;
    if (k >= n_el * NU) break;
    const int el = k / NU, rr = k % NU;
    const size_t gi = (size_t)e0 * NU + k;
    if constexpr (MODE == 0) {
      a.fout[gi] = sU[el * SU + rr];
    } else if (wave_minv) {
      rk_combine(a.stage, sU[el * SU + rr], pre_un[r], pre_acc[r], a.Uacc + gi, a.Uout + gi);
    } else {
      const double* M = a.Minv + (size_t)(e0 + el) * Ns * Ns;
      const int f = rr / Ns, i = rr % Ns;
      const double* R = sU + el * SU + f * Ns;
      double sol = 0.0;
#pragma unroll
      for (int ii = 0; ii < Ns; ++ii) sol += M[ii * Ns + i] * R[ii];
      rk_combine(a.stage, sDte[el] * sol, pre_un[r], pre_acc[r], a.Uacc + gi, a.Uout + gi);
    }
  }
}

// ---------------------------------------------------------------------------
// Dt from CFL: max over nodes of |u|+a, then Dt = cfl / max (rk.cc:250-272).
// Two levels: per-block maxima, then one block reduces them (no atomics).
// The reference takes the node with BLAS idamax, which starts from node 0 and
// moves on only to a strictly larger |x|: a NaN (negative pressure, upa.h)
// after node 0 is passed over, a NaN at node 0 is kept.  Here the maxima pass
// over NaNs (b > a is false for a NaN b; every |u|+a is >= 0, so starting
// from 0 changes nothing) and node 0's own value, stored after the per-block
// maxima, turns Dt into NaN when it is NaN: the same Dt as the reference's.
// ---------------------------------------------------------------------------
__device__ inline double skip_nan_max(double a, double b) { return b > a ? b : a; }

__device__ inline double block_max(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = skip_nan_max(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double m = red[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = skip_nan_max(m, red[k]);
  return m;
}

template <int D, int PH>
__global__ __launch_bounds__(256) void k_upa_max(const double* __restrict__ U, int N_E, int Ns,
                                                 double* __restrict__ partial) {
  constexpr int NF = Fields<D, PH>::NF;
  __shared__ double red[4];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (n < N_E * Ns) {
    const int e = n / Ns, i = n % Ns;
    double u[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) u[f] = U[(e * NF + f) * Ns + i];
    v = upa<D, PH>(u);
  }
  if (n == 0) partial[gridDim.x] = v;  // node 0, where idamax starts
  const double m = block_max(v, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

// partial[0..n-1]: per-block maxima; partial[n]: |u|+a of node 0.
__global__ __launch_bounds__(256) void k_dt_finalize(const double* __restrict__ partial, int n, double cfl,
                                                     double* dt) {
  __shared__ double red[4];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v = skip_nan_max(v, partial[k]);
  const double m = block_max(v, red);
  const double first = partial[n];
  if (threadIdx.x == 0) dt[0] = cfl / (first != first ? first : m);
}

// One step's time control on the device, for CFL-stepped runs without
// outputs (rk.cc:140-150): Dt from the partial maxima exactly as
// k_dt_finalize, then the stop on a tiny or NaN Dt, the clamp to the final
// time and the step count.  state[0]: stop (every gated kernel of later
// steps returns at once), [1]: the clamped last step was taken, [2]: Dt too
// small or NaN, [3]: steps taken; T[0]: the time reached.
__global__ __launch_bounds__(256) void k_dt_step(const double* __restrict__ partial, int n, double cfl, double tf,
                                                 double* dt, double* T, int* state) {
  if (state[0]) return;
  __syncthreads();  // every thread has read state[0] before thread 0 may set it
  if (state[1]) {
    if (threadIdx.x == 0) state[0] = 1;
    return;
  }
  __shared__ double red[4];
  double v = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) v = skip_nan_max(v, partial[k]);
  const double m = block_max(v, red);
  const double first = partial[n];
  if (threadIdx.x == 0) {
    double d = cfl / (first != first ? first : m);
    if (d < 1e-14 || d != d) {
      state[2] = 1;
      state[0] = 1;
      return;
    }
    const double t = T[0];
    if (d > (tf - t)) {
      d = tf - t;
      state[1] = 1;
    }
    dt[0] = d;
    T[0] = t + d;
    state[3] = state[3] + 1;
  }
}

// ---------------------------------------------------------------------------
// host-side dispatch
// ---------------------------------------------------------------------------
template <typename Fn>
static hipError_t dispatch(const dgx_config& c, Fn&& fn) {
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  auto by_flux = [&](auto d, auto p, auto ph) -> hipError_t {
    return c.flux == 0 ? fn(d, p, ph, I0{}) : fn(d, p, ph, I1{});
  };
  auto by_phys = [&](auto d, auto p) -> hipError_t {
    switch (c.physics) {
      case 0: return by_flux(d, p, I0{});
      case 1: return by_flux(d, p, I1{});
      case 2: return by_flux(d, p, I2{});
      default: return by_flux(d, p, I3{});
    }
  };
  auto by_order = [&](auto d) -> hipError_t {
    switch (c.order) {
      case 1: return by_phys(d, I1{});
      case 2: return by_phys(d, I2{});
      default: return by_phys(d, I3{});
    }
  };
  return c.dim == 1 ? by_order(I1{}) : by_order(I2{});
}

static inline unsigned blocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_face(const Solver& s, const double* U, hipStream_t st) {
  FaceArgs a{U,        s.t.face_elem, s.t.face_closure, s.t.face_tag, s.t.normals, s.t.JF, s.t.psi, s.t.psi_w,
             s.Qe,     s.M_T,         s.gate};
  return dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value, PH = decltype(ph)::value, FL = decltype(fl)::value;
    hipLaunchKernelGGL((k_face<D, P, PH, FL>), dim3(blocks_for(s.M_T, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_elem(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                       double* Uout, const double* dt_dev, double* f_out, hipStream_t st) {
  ElemArgs a{s.t.phi, s.t.dphi, s.t.phi_w, s.t.dphi_w, s.t.J, s.t.invJac, s.t.Minv, s.t.geo_stride,
             s.t.minv_stride, s.t.J0, s.Qe, s.N_E, stage, Uin, Un, Uacc, Uout, dt_dev, f_out, s.gate};
  return dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value, PH = decltype(ph)::value, FL = decltype(fl)::value;
    if constexpr (D == 2 && P == 2 && PH <= 1) {
      const unsigned nbm = blocks_for(s.N_E, 32);
      if (mode == 0)
        hipLaunchKernelGGL((k_elem_mfma<PH, FL, 0>), dim3(nbm), dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((k_elem_mfma<PH, FL, 1>), dim3(nbm), dim3(256), 0, st, a);
      return hipGetLastError();
    }
    constexpr int EPB = Vol<D, P, PH>::EPB;
    const unsigned nb = blocks_for(s.N_E, EPB);
    if (mode == 0)
      hipLaunchKernelGGL((k_elem<D, P, PH, FL, 0>), dim3(nb), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((k_elem<D, P, PH, FL, 1>), dim3(nb), dim3(256), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_upa_max(const Solver& s, const double* U, double cfl, double* dt_dev, hipStream_t st) {
  const long n = (long)s.N_E * s.N_s;
  const unsigned nb = blocks_for(n, 256);
  hipError_t err = dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, PH = decltype(ph)::value;
    (void)p; (void)fl;
    hipLaunchKernelGGL((k_upa_max<D, PH>), dim3(nb), dim3(256), 0, st, U, s.N_E, s.N_s, s.dt_partial);
    return hipGetLastError();
  });
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(256), 0, st, s.dt_partial, (int)nb, cfl, dt_dev);
  return hipGetLastError();
}

hipError_t launch_dt_step(const Solver& s, const double* U, double cfl, double tf, double* dt_dev, double* T_dev,
                          int* state, hipStream_t st) {
  const long n = (long)s.N_E * s.N_s;
  const unsigned nb = blocks_for(n, 256);
  hipError_t err = dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, PH = decltype(ph)::value;
    (void)p; (void)fl;
    hipLaunchKernelGGL((k_upa_max<D, PH>), dim3(nb), dim3(256), 0, st, U, s.N_E, s.N_s, s.dt_partial);
    return hipGetLastError();
  });
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_dt_step, dim3(1), dim3(256), 0, st, s.dt_partial, (int)nb, cfl, tf, dt_dev, T_dev, state);
  return hipGetLastError();
}

}  // namespace dgx
