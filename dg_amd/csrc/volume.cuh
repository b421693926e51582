// Element-local DG work shared by the element kernel (general meshes) and
// the fused stage kernel (Cartesian meshes): volume fluxes at the Gauss
// points (evaluate_sf + redistribute_sf), their integration against the test
// functions plus the face contributions (the GEMMs of dg_solver.h:336-340,
// mapToElement, addSFQ), and Dt*Minv with the RK4 stage update
// (rk_kernels.cu:22-56, rk.cc:164-198).  All data is in LDS; one 256-thread
// block processes batches of EPB elements.
#pragma once
#include <type_traits>

#include "physics.cuh"

namespace dgx {

__device__ constexpr double kRkBeta[4] = {0.0, 0.5, 0.5, 1.0};
__device__ constexpr double kRkGamma[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};

constexpr int pow2ceil(int n) { return n <= 1 ? 1 : 2 * pow2ceil((n + 1) / 2); }

template <int D, int P, int PH>
struct Vol {
  using S = Shape<D, P>;
  static constexpr int NF = Fields<D, PH>::NF, Ns = S::Ns, NG = S::NG, Ms = S::Ms, NN = S::NN;
  static constexpr int NNC = Fields<D, PH>::NNC;  // non-conservative fields (the last NNC; no volume flux)
  static constexpr int NFX = NF - NNC;            // fields that carry a volume flux
  static constexpr int NU = NF * Ns;        // state doubles per element
  static constexpr int NQE = NN * NF * Ms;  // face-contribution doubles per element
  static constexpr int NQ = NFX * D + NNC;  // per-point values handed from phase 1 to phase 2
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
                                       bool uniform_minv, int tid, int nthr = 256, bool with_dw = true) {
  using V = Vol<D, P, PH>;
  for (int k = tid; k < V::NTAB_PHI; k += nthr) t.phi[k] = phi[k];
  for (int k = tid; k < V::NTAB_D; k += nthr) {
    const int i = k / (V::NG * D), r = k % (V::NG * D);
    if (with_dw) t.dw[r * V::Ns + i] = dphi_w[k];
    if constexpr (V::NNC > 0) t.dphi[k] = dphi[k];
  }
  if constexpr (V::NNC > 0)
    for (int k = tid; k < V::NTAB_PHI; k += nthr) t.pw[(k % V::NG) * V::Ns + k / V::NG] = phi_w[k];
  if (uniform_minv && Minv)
    for (int k = tid; k < V::Ns * V::Ns; k += nthr) t.minv[k] = Minv[k];
}

// The same tables for 256-thread blocks in two steps: fetch() issues every
// global load into registers and store() writes them to LDS, so that a
// caller can issue its own loads in between and wait once for all of them.
template <int D, int P, int PH, bool DW>
struct VolTableFetch {
  using V = Vol<D, P, PH>;
  static constexpr int RP = (V::NTAB_PHI + 255) / 256, RD = (V::NTAB_D + 255) / 256;
  static constexpr int RM = (V::Ns * V::Ns + 255) / 256;
  static constexpr bool NC = V::NNC > 0;
  double phi[RP], pw[NC ? RP : 1], dw[DW ? RD : 1], dphi[NC ? RD : 1], minv[RM];
  __device__ inline void fetch(const double* phi_g, const double* dphi_g, const double* phi_w, const double* dphi_w,
                               const double* Minv, bool with_minv, int tid) {
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int k = tid + r * 256;
      if (k < V::NTAB_PHI) {
        phi[r] = phi_g[k];
        if constexpr (NC) pw[r] = phi_w[k];
      }
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int k = tid + r * 256;
      if (k < V::NTAB_D) {
        if constexpr (DW) dw[r] = dphi_w[k];
        if constexpr (NC) dphi[r] = dphi_g[k];
      }
    }
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      const int k = tid + r * 256;
      if (with_minv && k < V::Ns * V::Ns) minv[r] = Minv[k];
    }
  }
  __device__ inline void store(const VolTables<D, P, PH>& t, bool with_minv, int tid) const {
#pragma unroll
    for (int r = 0; r < RP; ++r) {
      const int k = tid + r * 256;
      if (k < V::NTAB_PHI) {
        t.phi[k] = phi[r];
        if constexpr (NC) t.pw[(k % V::NG) * V::Ns + k / V::NG] = pw[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const int k = tid + r * 256;
      if (k < V::NTAB_D) {
        const int i = k / (V::NG * D), q = k % (V::NG * D);
        if constexpr (DW) t.dw[q * V::Ns + i] = dw[r];
        if constexpr (NC) t.dphi[k] = dphi[r];
      }
    }
#pragma unroll
    for (int r = 0; r < RM; ++r) {
      const int k = tid + r * 256;
      if (with_minv && k < V::Ns * V::Ns) t.minv[k] = minv[r];
    }
  }
};

// Phase 1 for lane (element u, Gauss point g): fluxes with metric terms.
template <int D, int P, int PH>
__device__ inline void vol_phase1(const VolTables<D, P, PH>& t, int g, const double* u, double Je, const double* ij,
                                  double* out) {
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, NFX = V::NFX, NNC = V::NNC;
  double ug[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < Ns; ++i) s += u[f * Ns + i] * t.phi[g + i * NG];
    ug[f] = s;
  }
  double dua[(NNC ? NNC : 1) * D];
#pragma unroll
  for (int m = 0; m < NNC; ++m)
#pragma unroll
    for (int al = 0; al < D; ++al) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < Ns; ++i) s += u[(NFX + m) * Ns + i] * t.dphi[(g * D + al) + i * NG * D];
      dua[m * D + al] = s;
    }
  double fx[NF * D];
  double snc[NNC ? NNC : 1];
  volume_flux<D, PH>(ug, dua, ij, fx, snc);
#pragma unroll
  for (int f = 0; f < NFX; ++f)
#pragma unroll
    for (int al = 0; al < D; ++al) {
      double sol = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) sol += ij[al * D + k] * fx[f * D + k] * Je;
      out[f * D + al] = sol;
    }
#pragma unroll
  for (int m = 0; m < NNC; ++m) out[NFX * D + m] = snc[m] * Je;
}

// Phase 2 for lane (element, node i): S+F from the per-point values fe
// (NG x NQ) plus the face contributions qe (N_N x N_F x M_s, forward closure
// order per local edge).  At most two face terms land on a node, so their
// sum does not depend on the order.
template <int D, int P, int PH>
__device__ inline void vol_phase2(const VolTables<D, P, PH>& t, int i, const double* fe, const double* qe,
                                  double* rhs) {
  using V = Vol<D, P, PH>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, NFX = V::NFX, NQ = V::NQ, NN = V::NN, Ms = V::Ms, NNC = V::NNC;
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
#pragma unroll
    for (int m = 0; m < NNC; ++m) acc[NFX + m] += t.pw[g * Ns + i] * fe[g * NQ + NFX * D + m];
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
// The values rk_combine stores: the new accumulator (stages 0-2) and the
// stage output.
__device__ inline void rk_values(int st, double du, double un, double acc, double& acc_new, double& out) {
  const double gk = kRkGamma[st];
  if (st == 0) {
    acc_new = un + gk * du;
    out = un + kRkBeta[1] * du;
  } else if (st < 3) {
    acc_new = acc + gk * du;
    out = un + kRkBeta[st + 1] * du;
  } else {
    acc_new = acc;
    out = acc + gk * du;
  }
}

__device__ inline double rk_combine(int st, double du, double un, double acc, double* acc_out, double* out) {
  double an, o;
  rk_values(st, du, un, acc, an, o);
  if (st < 3) *acc_out = an;
  *out = o;
  return o;  // the stage output written
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
typedef double f64x4_t __attribute__((ext_vector_type(4)));

// 16x16x4 matrix-core step in T, and where its result rows land.  Operands
// (both precisions): A lane l = row l%16, k = l/16; B lane l = column l%16,
// k = l/16.  Result register v of lane l holds column l%16 of row
//   v_mfma_f64_16x16x4f64: l/16 + 4v        v_mfma_f32_16x16x4f32: 4(l/16) + v
// Rows are (element, field) pairs of four elements x four fields, numbered
// by row(el, f) so that in both cases register v of lane l is field v of
// element l/16.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  using V4 = f64x4_t;
  static __device__ inline int row(int el, int f) { return el + 4 * f; }
  static __device__ inline int el(int r) { return r & 3; }
  static __device__ inline int field(int r) { return r >> 2; }
  static __device__ inline V4 mma(double a, double b, V4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma16<float> {
  using V4 = f32x4_t;
  static __device__ inline int row(int el, int f) { return 4 * el + f; }
  static __device__ inline int el(int r) { return r >> 2; }
  static __device__ inline int field(int r) { return r & 3; }
  static __device__ inline V4 mma(float a, float b, V4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Split-BF16 operands for v_mfma_f32_16x16x16_bf16: x ~ hi + lo, both BF16
// (round to nearest even); a product is taken as hi*hi + hi*lo + lo*hi with
// FP32 accumulation (the lo*lo term, ~2^-16 relative, is dropped).
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
__device__ inline unsigned bf16_rne(unsigned u) { return (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u; }
__device__ inline void bf16_split(float x, short& hi, short& lo) {
  const unsigned h = bf16_rne(__float_as_uint(x));
  const unsigned l = bf16_rne(__float_as_uint(x - __uint_as_float(h)));
  hi = (short)(h >> 16);
  lo = (short)(l >> 16);
}
__device__ inline f32x4_t mma_bf16x3(const bf16x4_t& ah, const bf16x4_t& al, const bf16x4_t& bh, const bf16x4_t& bl,
                                     f32x4_t c) {
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
}

// Doubles of per-wave LDS that vol_mfma_wave needs: the alpha source of four
// elements (multifluid, FP64), then 16 flux rows of one direction in T
// (16 points, padded).
template <typename T, int PH>
constexpr int vol_wave_lds() { return (PH ? 4 * 16 : 0) + (16 * 17 * (int)sizeof(T) + 7) / 8; }


// Element phases of a tile of NT 2D p=2 elements (256 threads, NT a multiple
// of 16) on the matrix cores, contractions in T: double for the FP64 path,
// float for the mixed-precision mode (FP32 inputs and accumulation; the
// pointwise flux, the alpha terms, the face sums and everything after stay
// FP64).
//   map        Ug[(el,f)][g]  = sum_i U[el][f][i] phi[g][i]         fields 0..3
//   flux       fe[(el,f)][g,al] = J_e sum_k invJac[al][k] f_k(Ug)    in registers
//   integrate  R[(el,f)][i]   = sum_{al} sum_g fe[(el,f)][g,al] dphi_w[g,al][i]
// Wave w owns elements 4w..4w+3 of every group of 16.  After the map each
// lane holds the four conserved fields of one element at one Gauss point, so
// the flux needs no exchange; the scaled fluxes pass through the wave's own
// LDS rows, one direction at a time, into the integration, whose result
// registers plus the face sums are the right-hand side, written to sU.
// The integration operand (dphi_w, [i*NG*D + g*D + al]) is read from global
// memory into registers once.  Multifluid adds alpha and its reference
// gradient at the point (FP64 sums over the 9 nodes) and the alpha source
// integrated against phi_w (FP64).  With minv (the shared mass-matrix
// inverse in LDS, [ii*Ns + i]) and dte (Dt per element), the right-hand side
// is then replaced by du = dte * Minv * R, again on the FP64 matrix cores
// (A = R rows read back from sU, B = Minv; alpha by FP64 sums).  No block
// barrier: every exchange stays inside the wave.  sWork needs
// 4 * vol_wave_lds<T, PH>() doubles.
// SPLIT (T = TM = float): the map and the integration as split-BF16
// products on v_mfma_f32_16x16x16_bf16 (K = 16 per instruction: A lane l =
// row l%16, k = 4(l/16)+0..3; B lane l = column l%16, same k; D as the FP32
// 16x16x4 layout), three instructions per contraction block.
template <typename T, int PH, int NT, int SU, int SQ, int NGEO, typename TM = T, bool SPLIT = false>
__device__ void vol_mfma_wave(const VolTables<2, 2, PH>& t, const double* __restrict__ dphi_w, double* sU,
                              const double* sGeo, const double* sQ, double* sWork, const int* sNodeQ, int tid,
                              const double* minv = nullptr, const double* dte = nullptr) {
  static_assert(!SPLIT || (std::is_same<T, float>::value && std::is_same<TM, float>::value),
                "split-BF16 accumulates in FP32");
  // TM: type of the map (T unless set), T: type of the integration.  On gfx950
  // the FP32 16x16x4 integration measured slower than the FP64 one (volume
  // phase of a 2D p=2 tile: 11.0k against 7.3k cycles), the FP32 map about
  // as fast as the FP64 map.
  using M = Mfma16<T>;
  using MM = Mfma16<TM>;
  constexpr int NF = 4 + PH, Ns = 9, NG = 16, D = 2, KP = NG + 1, Ms = 3;
  static_assert(NT % 16 == 0, "groups of 16 elements, four per wave");
  static_assert(PH <= 1, "single-fluid and two-fluid states only");
  const int w = tid >> 6, l = tid & 63, c = l & 15, q = l >> 4;
  double* S = sWork + w * vol_wave_lds<T, PH>();
  T* F = reinterpret_cast<T*>(S + (PH ? 4 * NG : 0));
  TM bphi[3];
  T bdw[8];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int k = 4 * s + q;
    bphi[s] = k < Ns ? (TM)t.phi[c + k * NG] : TM(0);
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) {  // direction s/4, points 4(s%4)+q
    const int g = 4 * (s & 3) + q, al = s >> 2;
    bdw[s] = c < Ns ? (T)dphi_w[c * NG * D + g * D + al] : T(0);
  }
  const int a_el = MM::el(c), a_f = MM::field(c);
  // split-BF16 B operands: phi[g = c][i = 4q+j] for the map, dphi_w[i = c][g = 4q+j][al] for the integration
  bf16x4_t sphi_h = {0, 0, 0, 0}, sphi_l = {0, 0, 0, 0}, sdw_h[2], sdw_l[2];
  if constexpr (SPLIT) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 4 * q + j;
      short h, lo;
      bf16_split(k < Ns ? (float)t.phi[c + k * NG] : 0.0f, h, lo);
      sphi_h[j] = h;
      sphi_l[j] = lo;
#pragma unroll
      for (int al = 0; al < D; ++al) {
        bf16_split(c < Ns ? (float)dphi_w[c * NG * D + k * D + al] : 0.0f, h, lo);
        sdw_h[al][j] = h;
        sdw_l[al][j] = lo;
      }
    }
  }
  // Minv operand: column c (output node), rows k = 4s + q (input node)
  double bminv[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int k = 4 * s + q;
    bminv[s] = (minv && k < Ns && c < Ns) ? minv[k * Ns + c] : 0.0;
  }
#pragma unroll 1
  for (int b = 0; b < NT; b += 16) {
    const int e_base = b + 4 * w, el = e_base + q;
    // ---- map
    typename MM::V4 acc = {TM(0), TM(0), TM(0), TM(0)};
    if constexpr (SPLIT) {
      bf16x4_t ah, alo;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = 4 * q + j;
        short h, lo;
        bf16_split(k < Ns ? (float)sU[(e_base + a_el) * SU + a_f * Ns + k] : 0.0f, h, lo);
        ah[j] = h;
        alo[j] = lo;
      }
      acc = mma_bf16x3(ah, alo, sphi_h, sphi_l, acc);
    } else {
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int k = 4 * s + q;
        const TM a = k < Ns ? (TM)sU[(e_base + a_el) * SU + a_f * Ns + k] : TM(0);
        acc = MM::mma(a, bphi[s], acc);
      }
    }
    double ug[NF], dua[D] = {0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) ug[v] = (double)acc[v];
    if constexpr (PH == 1) {
      const double* ua = sU + el * SU + 4 * Ns;
      double al = 0, d0 = 0, d1 = 0;
#pragma unroll
      for (int k = 0; k < Ns; ++k) {
        const double uk = ua[k];
        al += t.phi[c + k * NG] * uk;
        d0 += t.dphi[(c * D + 0) + k * NG * D] * uk;
        d1 += t.dphi[(c * D + 1) + k * NG * D] * uk;
      }
      ug[4] = al;
      dua[0] = d0;
      dua[1] = d1;
    }
    // ---- flux of element el at point g = c with metric terms
    T fe[4][D];
    {
      const double Je = sGeo[el * NGEO];
      const double* ij = sGeo + el * NGEO + 1;
      double fx[NF * D], salpha;
      volume_flux<2, PH>(ug, dua, ij, fx, &salpha);
#pragma unroll
      for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int al = 0; al < D; ++al) {
          double sol = 0.0;
#pragma unroll
          for (int k = 0; k < D; ++k) sol += ij[al * D + k] * fx[f * D + k] * Je;
          fe[f][al] = (T)sol;
        }
      if constexpr (PH == 1) S[q * NG + c] = salpha * Je;
    }
    // ---- integration, one direction at a time through this wave's rows
    typename M::V4 r = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int al = 0; al < D; ++al) {
#pragma unroll
      for (int f = 0; f < 4; ++f) F[M::row(q, f) * KP + c] = fe[f][al];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if constexpr (SPLIT) {
        bf16x4_t ah, alo;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          short h, lo;
          bf16_split((float)F[c * KP + 4 * q + j], h, lo);
          ah[j] = h;
          alo[j] = lo;
        }
        r = mma_bf16x3(ah, alo, sdw_h[al], sdw_l[al], r);
      } else {
#pragma unroll
        for (int s = 0; s < 4; ++s) r = M::mma(F[c * KP + 4 * s + q], bdw[4 * al + s], r);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();  // the rows are rewritten next
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    double sa = 0;
    if constexpr (PH == 1) {
      if (c < Ns) {
#pragma unroll
        for (int g = 0; g < NG; ++g) sa += S[q * NG + g] * t.pw[g * Ns + c];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // F and S are rewritten by the next group
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (c < Ns) {
      const int q0 = sNodeQ[2 * c], q1 = sNodeQ[2 * c + 1];
      auto face = [&](int f) {
        double qs = 0.0;
        if (q0 >= 0) qs += sQ[el * SQ + q0 + f * Ms];
        if (q1 >= 0) qs += sQ[el * SQ + q1 + f * Ms];
        return qs;
      };
#pragma unroll
      for (int v = 0; v < 4; ++v) sU[el * SU + v * Ns + c] = (double)r[v] + face(v);
      if constexpr (PH == 1) sU[el * SU + 4 * Ns + c] = sa + face(4);
    }
    if (minv) {
      // du = dte * Minv R: rows (el, f) = el + 4f of the four elements, columns = nodes
      using M64 = Mfma16<double>;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();  // the right-hand side of the four elements is in sU
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      M64::V4 du = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int k = 4 * s + q;
        const double a = k < Ns ? sU[(e_base + M64::el(c)) * SU + M64::field(c) * Ns + k] : 0.0;
        du = M64::mma(a, bminv[s], du);
      }
      double dua5 = 0.0;
      if constexpr (PH == 1) {
        if (c < Ns) {
#pragma unroll
          for (int k = 0; k < Ns; ++k) dua5 += minv[k * Ns + c] * sU[el * SU + 4 * Ns + k];
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();  // every lane has read the right-hand side
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (c < Ns) {
        const double d = dte[el];
#pragma unroll
        for (int v = 0; v < 4; ++v) sU[el * SU + v * Ns + c] = d * du[v];
        if constexpr (PH == 1) sU[el * SU + 4 * Ns + c] = d * dua5;
      }
    }
  }
}

// Offsets (kk*NF*Ms + jj) of the at most two face slots that land on node i
// (closure_node(kk, jj) == i, edges in order), -1 when absent; field f adds f*Ms.
// Built at compile time; each block copies it into LDS.
template <int NF>
struct NodeQ {
  int v[18];
  constexpr NodeQ() : v{} {
    for (int i = 0; i < 9; ++i) {
      int n = 0, o[2] = {-1, -1};
      for (int kk = 0; kk < 4; ++kk)
        for (int jj = 0; jj < 3; ++jj)
          if (closure_node<2, 2>(kk, jj) == i && n < 2) o[n++] = kk * NF * 3 + jj;
      v[2 * i] = o[0];
      v[2 * i + 1] = o[1];
    }
  }
};
template <int NF>
__device__ constexpr NodeQ<NF> kNodeQ{};

template <int NF>
__device__ inline void node_q_table(int* t, int tid) {
  if (tid < 18) t[tid] = kNodeQ<NF>.v[tid];
}

}  // namespace dgx
