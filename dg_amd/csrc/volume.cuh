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

namespace dgx {

typedef float f32x4_t __attribute__((ext_vector_type(4)));

// FP32-MFMA volume terms of a tile of NT 2D p=2 elements (256 threads), for
// the mixed-precision mode: the map U -> Gauss points (and the alpha
// gradient) and the integration against the weighted derivative tables run
// as v_mfma_f32_16x16x4f32 contractions with FP32 accumulation; the
// pointwise flux, the face sum and everything after stay FP64.  Rows of the
// MFMA tiles are (element, field) pairs, columns Gauss points (map) or nodes
// (integration).  Operand layouts (16x16x4 f32): A lane l = row l%16,
// k = l/16; B lane l = column l%16, k = l/16; D lane l = rows 4*(l/16)+v,
// column l%16.  Writes S+F+Q of every element, in FP64, into sU.
template <int PH, int NT, int SU, int SQ, int NGEO>
__device__ void vol_mfma_tile(const VolTables<2, 2, PH>& t, double* sU, const double* sGeo, const double* sQ,
                              float* sM, int tid) {
  constexpr int NF = 4 + PH, NFX = 4, Ns = 9, NG = 16, D = 2, Ms = 3;
  constexpr int NFP = NF + 2;          // per (element, point): fields, then d(alpha)/dxi, d(alpha)/deta
  constexpr int KP = NG * D + 1;       // padded row of the flux operand
  constexpr int NGP = NG + 1;          // padded row of the alpha-source operand
  float* sUg = sM;                     // [(el*NG + g)*NFP + f]
  float* sFe = sM;                     // [(el*NFX + f)*KP + g*D + al]  (after the flux phase)
  float* sS = sM + NT * NFX * KP;      // [el*NGP + g]
  const int w = tid >> 6, l = tid & 63, c = l & 15, q = l >> 4;

  // ---- map: Ug[(el,f)][g] = sum_i U[el][f][i] phi[g][i]
  float bphi[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int k = 4 * s + q;
    bphi[s] = k < Ns ? (float)t.phi[c + k * NG] : 0.0f;
  }
  constexpr int RT = NT * NF / 16;
  for (int rt = w; rt < RT; rt += 4) {
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const int r = 16 * rt + c, el = r / NF, f = r % NF;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int k = 4 * s + q;
      const float a = k < Ns ? (float)sU[el * SU + f * Ns + k] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bphi[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = 16 * rt + 4 * q + v;
      sUg[((rr / NF) * NG + c) * NFP + rr % NF] = acc[v];
    }
  }
  if constexpr (PH == 1) {  // d(alpha)/dxi_al at the points: rows el, columns g for direction al
    const int rt = w >> 1, al = w & 1;
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
    const int el = 16 * rt + c;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int k = 4 * s + q;
      const float a = k < Ns ? (float)sU[el * SU + (NF - 1) * Ns + k] : 0.0f;
      const float b = k < Ns ? (float)t.dphi[(c * D + al) + k * NG * D] : 0.0f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) sUg[((16 * rt + 4 * q + v) * NG + c) * NFP + NF + al] = acc[v];
  }
  __syncthreads();

  // ---- pointwise flux with metric terms (FP64), two (element, point) pairs per lane
  double fe[2][NFX * D + PH];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = tid + h * 256, el = p / NG, g = p % NG;
    double ug[NF], dua[D];
#pragma unroll
    for (int f = 0; f < NF; ++f) ug[f] = sUg[(el * NG + g) * NFP + f];
#pragma unroll
    for (int al = 0; al < D; ++al) dua[al] = PH ? (double)sUg[(el * NG + g) * NFP + NF + al] : 0.0;
    const double Je = sGeo[el * NGEO];
    const double* ij = sGeo + el * NGEO + 1;
    double fx[NF * D], salpha;
    volume_flux<2, PH>(ug, dua, ij, fx, salpha);
#pragma unroll
    for (int f = 0; f < NFX; ++f)
#pragma unroll
      for (int al = 0; al < D; ++al) {
        double sol = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) sol += ij[al * D + k] * fx[f * D + k] * Je;
        fe[h][f * D + al] = sol;
      }
    if constexpr (PH == 1) fe[h][NFX * D] = salpha * Je;
  }
  __syncthreads();  // the flux operand overwrites the point values
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = tid + h * 256, el = p / NG, g = p % NG;
#pragma unroll
    for (int f = 0; f < NFX; ++f)
#pragma unroll
      for (int al = 0; al < D; ++al) sFe[(el * NFX + f) * KP + g * D + al] = (float)fe[h][f * D + al];
    if constexpr (PH == 1) sS[el * NGP + g] = (float)fe[h][NFX * D];
  }
  __syncthreads();

  // face terms of node i, field f of element el (at most two, summed in edge order)
  auto qsum = [&](int el, int f, int i) {
    double s = 0.0;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int jj = 0; jj < Ms; ++jj)
        if (closure_node<2, 2>(kk, jj) == i) s += sQ[el * SQ + (kk * NF + f) * Ms + jj];
    return s;
  };
  // ---- integration: F[(el,f)][i] = sum_{g,al} fe[(el,f)][g,al] dphi_w[(g,al)][i]
  float bdw[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) bdw[s] = c < Ns ? (float)t.dw[(4 * s + q) * Ns + c] : 0.0f;
  constexpr int RTI = NT * NFX / 16;
  for (int rt = w; rt < RTI; rt += 4) {
    f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; ++s)
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sFe[(16 * rt + c) * KP + 4 * s + q], bdw[s], acc, 0, 0, 0);
    if (c < Ns) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rr = 16 * rt + 4 * q + v, el = rr / NFX, f = rr % NFX;
        sU[el * SU + f * Ns + c] = (double)acc[v] + qsum(el, f, c);
      }
    }
  }
  if constexpr (PH == 1) {  // alpha source: S[el][i] = sum_g s[el][g] phi_w[g][i]
    if (w < 2) {
      f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int k = 4 * s + q;
        const float b = c < Ns ? (float)t.pw[k * Ns + c] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(sS[(16 * w + c) * NGP + k], b, acc, 0, 0, 0);
      }
      if (c < Ns) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int el = 16 * w + 4 * q + v;
          sU[el * SU + (NF - 1) * Ns + c] = (double)acc[v] + qsum(el, NF - 1, c);
        }
      }
    }
  }
  __syncthreads();
}

}  // namespace dgx

namespace dgx {

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Element phases of a tile of NT single-fluid 2D p=2 elements in FP64 on the
// matrix cores (v_mfma_f64_16x16x4f64), same arithmetic precision as the
// VALU path.  Wave w owns elements 4w..4w+3 of every group of 16: the map
// (rows = element x field, K = 9 nodes padded to 12, columns = 16 Gauss
// points) leaves in each lane the 4 fields of one element at one point
// (D lane l, register v = row l/16 + 4v, column l%16), so the flux is evaluated in
// registers; the metric-scaled fluxes go through this wave's LDS rows into
// the integration (rows = element x field, K = 32 (point, direction),
// columns = 9 nodes padded to 16), whose D fragments plus the face sums are
// the right-hand side, written to sU.  No block barrier is needed: every
// exchange stays inside the wave.  sF needs 4 * 16 * KP doubles.
template <int NT, int SU, int SQ, int NGEO>
__device__ void vol_mfma64_sf(const VolTables<2, 2, 0>& t, double* sU, const double* sGeo, const double* sQ,
                              double* sF, const int* sNodeQ, int tid) {
  constexpr int NF = 4, Ns = 9, NG = 16, D = 2, KP = NG * D + 1;
  const int w = tid >> 6, l = tid & 63, c = l & 15, q = l >> 4;
  double* F = sF + w * 16 * KP;
  double bphi[3], bdw[8];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int k = 4 * s + q;
    bphi[s] = k < Ns ? t.phi[c + k * NG] : 0.0;
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) bdw[s] = c < Ns ? t.dw[(4 * s + q) * Ns + c] : 0.0;
#pragma unroll
  for (int b = 0; b < NT; b += 16) {
    const int e_base = b + 4 * w;  // this wave's four elements
    // ---- map: Ug[(el,f)][g]. The f64 16x16x4 result holds row (l/16) + 4v of
    // column l%16 in register v, so A row c is (element c%4, field c/4): lane
    // (q, g) then receives the four fields of element q at point g.
    f64x4_t acc = {0.0, 0.0, 0.0, 0.0};
    {
      const int el = e_base + (c & 3), f = c >> 2;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int k = 4 * s + q;
        const double a = k < Ns ? sU[el * SU + f * Ns + k] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bphi[s], acc, 0, 0, 0);
      }
    }
    // ---- flux of element e_base + q at point g = c, metric terms, into LDS
    {
      const int el = e_base + q;
      double ug[NF] = {acc[0], acc[1], acc[2], acc[3]};
      double dua[D] = {0.0, 0.0};
      const double Je = sGeo[el * NGEO];
      const double* ij = sGeo + el * NGEO + 1;
      double fx[NF * D], salpha;
      volume_flux<2, 0>(ug, dua, ij, fx, salpha);
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int al = 0; al < D; ++al) {
          double sol = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) sol += ij[al * D + k] * fx[f * D + k] * Je;
          F[(f * 4 + q) * KP + c * D + al] = sol;  // row 4f + el, see the map
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---- integration: R[(el,f)][i] = sum_k F[(el,f)][k] dphi_w[k][i]
    f64x4_t r = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 8; ++s) r = __builtin_amdgcn_mfma_f64_16x16x4f64(F[c * KP + 4 * s + q], bdw[s], r, 0, 0, 0);
    __builtin_amdgcn_wave_barrier();  // F is rewritten by the next group
    if (c < Ns) {
      const int el = e_base + q;
      const int q0 = sNodeQ[2 * c], q1 = sNodeQ[2 * c + 1];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double qs = 0.0;
        if (q0 >= 0) qs += sQ[el * SQ + q0 + v * 3];
        if (q1 >= 0) qs += sQ[el * SQ + q1 + v * 3];
        sU[el * SU + v * Ns + c] = r[v] + qs;
      }
    }
  }
}

// Offsets (kk*NF*Ms + jj) of the at most two face slots that land on node i
// (closure_node(kk, jj) == i, edges in order), -1 when absent; field f adds f*Ms.
template <int NF>
__device__ inline void node_q_table(int* t, int tid) {
  if (tid < 9) {
    int n = 0, o[2] = {-1, -1};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj)
        if (closure_node<2, 2>(kk, jj) == tid && n < 2) o[n++] = kk * NF * 3 + jj;
    t[2 * tid] = o[0];
    t[2 * tid + 1] = o[1];
  }
}

}  // namespace dgx
