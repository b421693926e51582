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
};

template <int D, int P, int PH, int FL>
__global__ __launch_bounds__(256) void k_face(FaceArgs a) {
  using S = Shape<D, P>;
  constexpr int NF = Fields<D, PH>::NF, Ms = S::Ms, MG = S::MG, Ns = S::Ns;
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
  const double* __restrict__ Qe;
  int N_E;
  int stage;
  const double* Uin;
  const double* Un;
  double* Uacc;
  double* Uout;
  const double* __restrict__ dt;
  double* fout;
};

__device__ constexpr double kRkBeta[4] = {0.0, 0.5, 0.5, 1.0};
__device__ constexpr double kRkGamma[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};

constexpr int pow2ceil(int n) { return n <= 1 ? 1 : 2 * pow2ceil((n + 1) / 2); }

// One block of 256 threads handles EPB elements.  Phase 1: lane = (element,
// quadrature point); phase 2: lane = (element, node); phase 3: flat over the
// block's contiguous state (coalesced).  Element state, face contributions,
// per-point fluxes and the reference tables are staged in LDS.
template <int D, int P, int PH, int FL, int MODE>
__global__ __launch_bounds__(256) void k_elem(ElemArgs a) {
  using S = Shape<D, P>;
  constexpr int NF = Fields<D, PH>::NF, Ns = S::Ns, NG = S::NG, Ms = S::Ms, NN = S::NN;
  constexpr int NFX = NF - PH;           // fields that carry a volume flux (alpha has none)
  constexpr int NU = NF * Ns;            // state doubles per element
  constexpr int NQE = NN * NF * Ms;      // face-contribution doubles per element
  constexpr int NQ = NFX * D + PH;       // per-point values handed to phase 2
  constexpr int TPE = pow2ceil(NG);      // lanes per element in phase 1
  constexpr int EPB = 256 / TPE;         // elements per block

  __shared__ double sU[EPB * NU];
  __shared__ double sQ[EPB * NQE];
  __shared__ double sF[EPB * NG * NQ];
  __shared__ double sPhi[NG * Ns];
  __shared__ double sDphi[PH ? NG * D * Ns : 1];
  __shared__ double sDw[NG * D * Ns];   // transposed: [(g*D+al)*Ns + i]
  __shared__ double sPw[PH ? NG * Ns : 1];  // transposed: [g*Ns + i]
  __shared__ double sMinv[Ns * Ns];

  const int tid = threadIdx.x;
  const int e0 = blockIdx.x * EPB;
  const int n_el = min(EPB, a.N_E - e0);

  for (int k = tid; k < NG * Ns; k += 256) sPhi[k] = a.phi[k];
  for (int k = tid; k < NG * D * Ns; k += 256) {
    const int i = k / (NG * D), r = k % (NG * D);
    sDw[r * Ns + i] = a.dphi_w[k];
    if constexpr (PH == 1) sDphi[k] = a.dphi[k];
  }
  if constexpr (PH == 1)
    for (int k = tid; k < NG * Ns; k += 256) sPw[(k % NG) * Ns + k / NG] = a.phi_w[k];
  if (MODE == 1 && a.minv_stride == 0)
    for (int k = tid; k < Ns * Ns; k += 256) sMinv[k] = a.Minv[k];
  {
    const double* src = a.Uin + (size_t)e0 * NU;
    for (int k = tid; k < n_el * NU; k += 256) sU[k] = src[k];
    const double* q = a.Qe + (size_t)e0 * NQE;
    for (int k = tid; k < n_el * NQE; k += 256) sQ[k] = q[k];
  }
  __syncthreads();

  // ---- phase 1: fluxes at the quadrature points
  {
    const int el = tid / TPE, g = tid % TPE;
    if (el < n_el && g < NG) {
      const double* u = sU + el * NU;
      double ug[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < Ns; ++i) s += u[f * Ns + i] * sPhi[g + i * NG];
        ug[f] = s;
      }
      double dua[D];
#pragma unroll
      for (int al = 0; al < D; ++al) {
        double s = 0.0;
        if constexpr (PH == 1) {
#pragma unroll
          for (int i = 0; i < Ns; ++i) s += u[(NF - 1) * Ns + i] * sDphi[(g * D + al) + i * NG * D];
        }
        dua[al] = s;
      }
      const int eg = (e0 + el) * a.geo_stride;
      const double Je = a.J[eg];
      double ij[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) ij[k] = a.invJac[eg * D * D + k];
      double fx[NF * D];
      double salpha;
      volume_flux<D, PH>(ug, dua, ij, fx, salpha);
      double* out = sF + (el * NG + g) * NQ;
#pragma unroll
      for (int f = 0; f < NFX; ++f)
#pragma unroll
        for (int al = 0; al < D; ++al) {
          double sol = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) sol += ij[al * D + k] * fx[f * D + k] * Je;
          out[f * D + al] = sol;
        }
      if constexpr (PH == 1) out[NFX * D] = salpha * Je;
    }
  }
  __syncthreads();

  // ---- phase 2: integrate against the test functions, add face terms
  {
    const int el = tid / TPE, i = tid % TPE;
    double rhs[NF];
    if (el < n_el && i < Ns) {
      double acc[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) acc[f] = 0.0;
      const double* fe = sF + el * NG * NQ;
#pragma unroll 4
      for (int g = 0; g < NG; ++g) {
#pragma unroll
        for (int al = 0; al < D; ++al) {
          const double w = sDw[(g * D + al) * Ns + i];
#pragma unroll
          for (int f = 0; f < NFX; ++f) acc[f] += w * fe[g * NQ + f * D + al];
        }
        if constexpr (PH == 1) acc[NF - 1] += sPw[g * Ns + i] * fe[g * NQ + NFX * D];
      }
      const double* qe = sQ + el * NQE;
      // face contributions landing on node i: at most two (corners), so
      // the sum does not depend on the order
      double q[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) q[f] = 0.0;
#pragma unroll
      for (int kk = 0; kk < NN; ++kk)
#pragma unroll
        for (int jj = 0; jj < Ms; ++jj)
          if (closure_node<D, P>(kk, jj) == i)
#pragma unroll
            for (int f = 0; f < NF; ++f) q[f] += qe[(kk * NF + f) * Ms + jj];
#pragma unroll
      for (int f = 0; f < NF; ++f) rhs[f] = acc[f] + q[f];
    }
    __syncthreads();  // sU is reused below for the right-hand side
    if (el < n_el && i < Ns)
#pragma unroll
      for (int f = 0; f < NF; ++f) sU[el * NU + f * Ns + i] = rhs[f];
  }
  __syncthreads();

  // ---- phase 3: coalesced output
  const size_t base = (size_t)e0 * NU;
  if constexpr (MODE == 0) {
    for (int k = tid; k < n_el * NU; k += 256) a.fout[base + k] = sU[k];
  } else {
    const double dt = a.dt[0];
    const int st = a.stage;
    const double gk = kRkGamma[st];
    const double bk = (st < 3) ? kRkBeta[st + 1] : 0.0;
    for (int k = tid; k < n_el * NU; k += 256) {
      const int el = k / NU, r = k % NU, f = r / Ns, i = r % Ns;
      const double* R = sU + el * NU + f * Ns;
      double sol = 0.0;
      if (a.minv_stride == 0) {
#pragma unroll
        for (int ii = 0; ii < Ns; ++ii) sol += sMinv[ii * Ns + i] * R[ii];
      } else {
        const double* Mi = a.Minv + (size_t)(e0 + el) * Ns * Ns;
#pragma unroll
        for (int ii = 0; ii < Ns; ++ii) sol += Mi[ii * Ns + i] * R[ii];
      }
      const double du = dt * sol;
      const size_t idx = base + k;
      if (st == 0) {
        const double un = a.Un[idx];
        a.Uacc[idx] = un + gk * du;
        a.Uout[idx] = un + bk * du;
      } else if (st < 3) {
        a.Uacc[idx] = a.Uacc[idx] + gk * du;
        a.Uout[idx] = a.Un[idx] + bk * du;
      } else {
        a.Uout[idx] = a.Uacc[idx] + gk * du;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Dt from CFL: max over nodes of |u|+a, then Dt = cfl / max (rk.cc:250-272).
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

template <int D, int PH>
__global__ __launch_bounds__(256) void k_upa_max(const double* __restrict__ U, int N_E, int Ns,
                                                 unsigned long long* out) {
  constexpr int NF = Fields<D, PH>::NF;
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (n < N_E * Ns) {
    const int e = n / Ns, i = n % Ns;
    double u[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) u[f] = U[(e * NF + f) * Ns + i];
    v = upa<D, PH>(u);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off));
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = dmax(m, red[k]);
    atomicMax(out, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_dt_finalize(const unsigned long long* mx, double cfl, double* dt) {
  dt[0] = cfl / __longlong_as_double((long long)mx[0]);
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
    return c.physics == 0 ? by_flux(d, p, I0{}) : by_flux(d, p, I1{});
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
  FaceArgs a{U, s.t.face_elem, s.t.face_closure, s.t.face_tag, s.t.normals, s.t.JF, s.t.psi, s.t.psi_w, s.Qe, s.M_T};
  return dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value, PH = decltype(ph)::value, FL = decltype(fl)::value;
    hipLaunchKernelGGL((k_face<D, P, PH, FL>), dim3(blocks_for(s.M_T, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_elem(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                       double* Uout, const double* dt_dev, double* f_out, hipStream_t st) {
  ElemArgs a{s.t.phi, s.t.dphi, s.t.phi_w, s.t.dphi_w, s.t.J, s.t.invJac, s.t.Minv, s.t.geo_stride,
             s.t.minv_stride, s.Qe, s.N_E, stage, Uin, Un, Uacc, Uout, dt_dev, f_out};
  return dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value, PH = decltype(ph)::value, FL = decltype(fl)::value;
    constexpr int EPB = 256 / pow2ceil(Shape<D, P>::NG);
    const unsigned nb = blocks_for(s.N_E, EPB);
    if (mode == 0)
      hipLaunchKernelGGL((k_elem<D, P, PH, FL, 0>), dim3(nb), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((k_elem<D, P, PH, FL, 1>), dim3(nb), dim3(256), 0, st, a);
    return hipGetLastError();
  });
}

hipError_t launch_upa_max(const Solver& s, const double* U, double cfl, double* dt_dev, hipStream_t st) {
  unsigned long long* mx = reinterpret_cast<unsigned long long*>(s.dscratch);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, mx, 0ull);
  long n = (long)s.N_E * s.N_s;
  hipError_t err = dispatch(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value, PH = decltype(ph)::value;
    (void)p; (void)fl;
    hipLaunchKernelGGL((k_upa_max<D, PH>), dim3(blocks_for(n, 256)), dim3(256), 0, st, U, s.N_E, s.N_s, mx);
    return hipGetLastError();
  });
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, st, mx, cfl, dt_dev);
  return hipGetLastError();
}

}  // namespace dgx
