// Element-local DG work shared by the element kernel (general meshes) and
// the fused stage kernel (Cartesian meshes): volume fluxes at the Gauss
// points (evaluate_sf + redistribute_sf), their integration against the test
// functions plus the face contributions (the GEMMs of dg_solver.h:336-340,
// mapToElement, addSFQ), and Dt*Minv with the RK4 stage update
// (rk_kernels.cu:22-56, rk.cc:164-198).  All data is in LDS; one 256-thread
// block processes batches of EPB elements.
#pragma once
#include "physics.cuh"

namespace dgx {

__device__ constexpr double kRkBeta[4] = {0.0, 0.5, 0.5, 1.0};
__device__ constexpr double kRkGamma[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};

constexpr int pow2ceil(int n) { return n <= 1 ? 1 : 2 * pow2ceil((n + 1) / 2); }

template <int D, int P, int PH>
struct Vol {
  using S = Shape<D, P>;
  static constexpr int NF = Fields<D, PH>::NF, Ns = S::Ns, NG = S::NG, Ms = S::Ms, NN = S::NN;
  static constexpr int NFX = NF - PH;       // fields that carry a volume flux (alpha has none)
  static constexpr int NU = NF * Ns;        // state doubles per element
  static constexpr int NQE = NN * NF * Ms;  // face-contribution doubles per element
  static constexpr int NQ = NFX * D + PH;   // per-point values handed from phase 1 to phase 2
  static constexpr int TPE = pow2ceil(NG);  // lanes per element in phase 1
  static constexpr int EPB = 256 / TPE;     // elements per 256-thread batch
  static constexpr int NTAB_PHI = NG * Ns, NTAB_D = NG * D * Ns;
};

// Reference-element tables staged in LDS (derivative tables transposed so
// that lanes over nodes read consecutive banks).
template <int D, int P, int PH>
struct VolTables {
  double* phi;    // [g + i*NG]
  double* dphi;   // [(g*D+al) + i*NG*D]           (multifluid only)
  double* dw;     // [(g*D+al)*Ns + i]  = dphi_w transposed
  double* pw;     // [g*Ns + i]         = phi_w transposed (multifluid only)
  double* minv;   // [ii*Ns + i]        (uniform Minv only)
};

template <int D, int P, int PH>
__device__ inline void load_vol_tables(const VolTables<D, P, PH>& t, const double* phi, const double* dphi,
                                       const double* phi_w, const double* dphi_w, const double* Minv,
                                       bool uniform_minv, int tid, int nthr = 256) {
  using V = Vol<D, P, PH>;
  for (int k = tid; k < V::NTAB_PHI; k += nthr) t.phi[k] = phi[k];
  for (int k = tid; k < V::NTAB_D; k += nthr) {
    const int i = k / (V::NG * D), r = k % (V::NG * D);
    t.dw[r * V::Ns + i] = dphi_w[k];
    if constexpr (PH == 1) t.dphi[k] = dphi[k];
  }
  if constexpr (PH == 1)
    for (int k = tid; k < V::NTAB_PHI; k += nthr) t.pw[(k % V::NG) * V::Ns + k / V::NG] = phi_w[k];
  if (uniform_minv && Minv)
    for (int k = tid; k < V::Ns * V::Ns; k += nthr) t.minv[k] = Minv[k];
}

// Phase 1 for lane (element u, Gauss point g): fluxes with metric terms.
template <int D, int P, int PH>
__device__ inline void vol_phase1(const VolTables<D, P, PH>& t, int g, const double* u, double Je, const double* ij,
                                  double* out) {
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, NFX = V::NFX;
  double ug[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < Ns; ++i) s += u[f * Ns + i] * t.phi[g + i * NG];
    ug[f] = s;
  }
  double dua[D];
#pragma unroll
  for (int al = 0; al < D; ++al) {
    double s = 0.0;
    if constexpr (PH == 1) {
#pragma unroll
      for (int i = 0; i < Ns; ++i) s += u[(NF - 1) * Ns + i] * t.dphi[(g * D + al) + i * NG * D];
    }
    dua[al] = s;
  }
  double fx[NF * D];
  double salpha;
  volume_flux<D, PH>(ug, dua, ij, fx, salpha);
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

// Phase 2 for lane (element, node i): S+F from the per-point values fe
// (NG x NQ) plus the face contributions qe (N_N x N_F x M_s, forward closure
// order per local edge).  At most two face terms land on a node, so their
// sum does not depend on the order.
template <int D, int P, int PH>
__device__ inline void vol_phase2(const VolTables<D, P, PH>& t, int i, const double* fe, const double* qe,
                                  double* rhs) {
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, NFX = V::NFX, NQ = V::NQ, NN = V::NN, Ms = V::Ms;
  double acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = 0.0;
#pragma unroll 4
  for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int al = 0; al < D; ++al) {
      const double w = t.dw[(g * D + al) * Ns + i];
#pragma unroll
      for (int f = 0; f < NFX; ++f) acc[f] += w * fe[g * NQ + f * D + al];
    }
    if constexpr (PH == 1) acc[NF - 1] += t.pw[g * Ns + i] * fe[g * NQ + NFX * D];
  }
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

// RK4 stage combination for one state entry (rk.cc:164-198 with DU fused):
// un = U^n, acc = accumulator (read for stages 1-3), du = Dt*Minv*f.
__device__ inline void rk_combine(int st, double du, double un, double acc, double* acc_out, double* out) {
  const double gk = kRkGamma[st];
  if (st == 0) {
    *acc_out = un + gk * du;
    *out = un + kRkBeta[1] * du;
  } else if (st < 3) {
    *acc_out = acc + gk * du;
    *out = un + kRkBeta[st + 1] * du;
  } else {
    *out = acc + gk * du;
  }
}

// Phase 3 for one state entry r = f*Ns + i of an element whose right-hand
// side R (NF x Ns) is in LDS: du = Dt * Minv * R, then the RK4 stage update
// at global index idx.
template <int D, int P, int PH>
__device__ inline void rk_update(const VolTables<D, P, PH>& t, const double* Mglob, bool uniform_minv, int r,
                                 const double* R, double dt, int st, size_t idx, const double* Un, double* Uacc,
                                 double* Uout) {
  using V = Vol<D, P, PH>;
  constexpr int Ns = V::Ns;
  const int f = r / Ns, i = r % Ns;
  const double* Rf = R + f * Ns;
  double sol = 0.0;
  if (uniform_minv) {
#pragma unroll
    for (int ii = 0; ii < Ns; ++ii) sol += t.minv[ii * Ns + i] * Rf[ii];
  } else {
#pragma unroll
    for (int ii = 0; ii < Ns; ++ii) sol += Mglob[ii * Ns + i] * Rf[ii];
  }
  const double un = st < 3 ? Un[idx] : 0.0;
  const double acc = st > 0 ? Uacc[idx] : 0.0;
  rk_combine(st, dt * sol, un, acc, Uacc + idx, Uout + idx);
}

}  // namespace dgx
