// Batched ensemble: many independent 1D problems of uneven size and order,
// each advanced by the RK4 loop of RK::RK_integration (rk.cc:132-247) with
// its own CFL time step (rk.cc:250-272), the HR limiter (limiting.cc:11-27)
// and the DG right-hand side of DG_SOLVER::dg_solver (dg_solver.h:292-364).
//
// One persistent kernel for every order: a workgroup owns one problem and
// runs all its steps, stages and limiter sweeps, separated by workgroup
// barriers only (the problems never exchange data).  Problems are dispatched
// largest first.  Per step: Dt from the workgroup maximum of |u|+a; per
// stage: a face pass (one lane per interface, Riemann flux integrated into
// element-major edge slots), an element pass (one lane per element, volume
// terms accumulated point by point with the element-local helpers of
// volume.cuh, Dt*Minv and the RK4 combination), and the HR limiter
// (Lagrange->Taylor, per-column limiting with the neighbours' Taylor
// coefficients, Taylor->Lagrange, without FMA contraction as in the CPU
// path).  The launch is captured once in a hipGraph and replayed; the
// workgroup size (up to 1024 threads, so that the few largest problems, which
// set the time, keep four waves on every SIMD of their CU) comes from a
// timing autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "dgx_internal.h"
#include "hr.cuh"
#include "physics.cuh"
#include "volume.cuh"

namespace dgx {

struct BatchTables {
  // reference-element tables of one order
  const double* phi;
  const double* dphi;
  const double* phi_w;
  const double* dphi_w;
  const double* psi_w;  // 1 x 1 face weight
  const double* L2M;    // Lag2Mono, column-major N_s x N_s
  const double* M2L;    // Mono2Lag
};

struct BatchArgs {
  BatchTables tab[3];       // orders 1..3
  const int* wg_prob;       // [workgroup] its problem (largest first)
  const int* order;         // [problem] 1..3
  const int* n_elem;        // [problem]
  const long* state_off;    // [problem] doubles into the packed state
  const long* elem_off;     // [problem] elements into the packed per-element arrays
  const double* J;          // [element]
  const double* ij;         // [element]
  const double* normals;    // [elem_off + problem + face]
  const double* minv;       // [problem*16 + k]: Minv of the problem's element 0 (N_s x N_s)
  const double* J0;         // [problem] Jacobian of that element
  const double* cfl;        // [problem] CFL * Dx / (2p+1)
  int bc[2];
  double* U;
  double* Uacc;
  double* S;    // stage state
  double* A;    // limiter: Taylor coefficients
  double* Qe;   // [element][edge][field]
  double* T;    // [problem] accumulated time
  int* status;  // [problem] 0 ok, 1 Dt too small or NaN
  int nsteps, limiter, limit_initial;
};

// max that keeps a NaN: a problem whose state holds a NaN gets a NaN Dt and
// stops with status 1 at once (the reference's idamax skips NaNs that are
// not first and stops only once its Dt turns NaN, rk.cc:147-148)
__device__ inline double nanmax(double a, double b) { return (b != b || b > a) ? b : a; }

template <int P, int PH>
struct BatchShape {
  static constexpr int NF = Fields<1, PH>::NF, Ns = P + 1, NG = P + 1, NU = NF * Ns;
};

// Taylor coefficients of the problem's columns (the first half of the HR
// limiter, limiting.cc:11-27 in 1D): A = L2M U per (element, field) column.
// KL columns per thread at a time, their loads issued together (a loop that
// stores one column before it loads the next waits a memory round trip per
// column).
template <int P, int PH>
__device__ void batch_limit_l2m(const double* U, double* A, int e0, int e1, const double* sL2M) {
#pragma clang fp contract(off)
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, KL = 4;
  const int end = e1 * NF, stride = blockDim.x;
  for (int c0 = e0 * NF + threadIdx.x; c0 < end; c0 += KL * stride) {
    double x[KL][Ns];
#pragma unroll
    for (int k = 0; k < KL; ++k) {
      const int col = c0 + k * stride;
#pragma unroll
      for (int l = 0; l < Ns; ++l) x[k][l] = col < end ? U[col * Ns + l] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < KL; ++k) {
      const int col = c0 + k * stride;
      if (col >= end) break;
      double y[Ns];
#pragma unroll
      for (int r = 0; r < Ns; ++r) y[r] = 0.0;
#pragma unroll
      for (int l = 0; l < Ns; ++l)
#pragma unroll
        for (int r = 0; r < Ns; ++r) y[r] = y[r] + x[k][l] * sL2M[r + l * Ns];
#pragma unroll
      for (int r = 0; r < Ns; ++r) A[col * Ns + r] = y[r];
    }
  }
}

// The second half: per column, hrl1D with the neighbours' Taylor
// coefficients, then U = M2L Alim; KA columns per thread at a time, each with
// its neighbours' coefficients loaded together (KA = 2 spilled registers in the
// order-3 body).
template <int P, int PH>
__device__ void batch_limit_apply(const BatchArgs& a, double* U, const double* A, int NE, int e0, int e1,
                                  const double* sM2L) {
#pragma clang fp contract(off)
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, KA = 1;
  const int end = e1 * NF, stride = blockDim.x;
  for (int c0 = e0 * NF + threadIdx.x; c0 < end; c0 += KA * stride) {
    double al[KA][Ns], ac[KA][Ns], ar[KA][Ns];
    int physical[KA];
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int col = c0 + k * stride;
      const int e = col / NF, fc = col % NF;
      const int left = e > 0 ? e - 1 : -a.bc[0];
      const int right = e < NE - 1 ? e + 1 : -a.bc[1];
      int ph = 0;
      if (left < 0) ph = -left;
      if (right < 0) ph = -right;
      physical[k] = ph;
      const bool in = col < end, inner = in && ph == 0;
#pragma unroll
      for (int n = 0; n < Ns; ++n) {
        ac[k][n] = in ? A[col * Ns + n] : 0.0;
        al[k][n] = inner ? A[(left * NF + fc) * Ns + n] : 0.0;
        ar[k][n] = inner ? A[(right * NF + fc) * Ns + n] : 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < KA; ++k) {
      const int col = c0 + k * stride;
      if (col >= end) break;
      const int fc = col % NF;
      double lim[Ns];
#pragma unroll
      for (int n = 0; n < Ns; ++n) lim[n] = 0.0;
      if (physical[k] == 2 || (physical[k] == 3 && fc != 1)) {
        average_monomial_d<P>(ac[k], lim);
      } else if (physical[k] == 3) {
#pragma unroll
        for (int n = 0; n < Ns; ++n) lim[n] = ac[k][n];
      } else {
        limit_monomial_d<P>(al[k], ac[k], ar[k], lim);
      }
      double y[Ns];
#pragma unroll
      for (int r = 0; r < Ns; ++r) y[r] = 0.0;
#pragma unroll
      for (int l = 0; l < Ns; ++l)
#pragma unroll
        for (int r = 0; r < Ns; ++r) y[r] = y[r] + lim[l] * sM2L[r + l * Ns];
#pragma unroll
      for (int r = 0; r < Ns; ++r) U[col * Ns + r] = y[r];
    }
  }
}

// One workgroup's problem (order P): every step, stage and limiter
// application, separated by workgroup barriers.  PF: the element pass loads
// the next element's state ahead (the 512-thread instance, whose register
// budget holds it).
template <int P, int PH, int FL, bool PF>
__device__ void batch_problem(const BatchArgs& a, int pid) {
  using B = BatchShape<P, PH>;
  using V = Vol<1, P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, NG = B::NG, NU = B::NU, NQ = V::NQ;
  __shared__ double sPhi[V::NTAB_PHI], sDphi[PH ? V::NTAB_D : 1], sDw[V::NTAB_D], sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns], sL2M[Ns * Ns], sM2L[Ns * Ns];
  __shared__ double sRed[16];
  const VolTables<1, P, PH> tb{sPhi, sDphi, sDw, sPw, sMinv};
  const BatchTables& bt = a.tab[P - 1];

  const int NE = a.n_elem[pid];
  const long so = a.state_off[pid], eo = a.elem_off[pid];
  double* U = a.U + so;
  double* Uacc = a.Uacc + so;
  double* S = a.S + so;
  double* A = a.A + so;
  double* Qe = a.Qe + eo * 2 * NF;
  const double* J = a.J + eo;
  const double* ij = a.ij + eo;
  const double* nrm = a.normals + eo + pid;
  const double cfl = a.cfl[pid], J0 = a.J0[pid], psiw = bt.psi_w[0];
  const int tid = threadIdx.x, nthr = blockDim.x;

  load_vol_tables(tb, bt.phi, bt.dphi, bt.phi_w, bt.dphi_w, nullptr, false, tid, nthr);
  for (int k = tid; k < Ns * Ns; k += nthr) {
    sMinv[k] = a.minv[(long)pid * 16 + k];
    sL2M[k] = bt.L2M[k];
    sM2L[k] = bt.M2L[k];
  }
  __syncthreads();
  if (a.status[pid] != 0) return;
  auto limit = [&](double* V) {
    batch_limit_l2m<P, PH>(V, A, 0, NE, sL2M);
    __syncthreads();  // the neighbours' Taylor coefficients
    batch_limit_apply<P, PH>(a, V, A, NE, 0, NE, sM2L);
    __syncthreads();
  };
  if (a.limiter && a.limit_initial) limit(U);

  for (int step = 0; step < a.nsteps; ++step) {
    // ---- Dt = cfl / max over nodes of |u|+a (rk.cc:250-272)
    // (KD nodes per thread at a time, loads issued together; max is exact in
    // any order)
    double m = 0.0;
    constexpr int KD = 4;
    for (int n0 = tid; n0 < NE * Ns; n0 += KD * nthr) {
      double u[KD][NF];
#pragma unroll
      for (int k = 0; k < KD; ++k) {
        const int n = n0 + k * nthr;
        const int e = n / Ns, i = n % Ns;
#pragma unroll
        for (int f = 0; f < NF; ++f) u[k][f] = n < NE * Ns ? U[(e * NF + f) * Ns + i] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < KD; ++k)
        if (n0 + k * nthr < NE * Ns) m = nanmax(m, upa<1, PH>(u[k]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) sRed[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double mx = sRed[0];
      for (int k = 1; k < (nthr + 63) / 64; ++k) mx = nanmax(mx, sRed[k]);
      sRed[0] = cfl / mx;
    }
    __syncthreads();
    const double dt = sRed[0];
    __syncthreads();  // sRed is reused by the next step
    if (!(dt >= 1e-14)) {
      if (tid == 0) a.status[pid] = 1;
      return;
    }
    if (tid == 0) a.T[pid] = a.T[pid] + dt;

    double* cur = U;
    for (int k = 0; k < 4; ++k) {
      if (k > 0 && a.limiter) limit(cur);
      // ---- faces: interface j between elements j-1 and j; boundary faces
      // contribute to the boundary element only (dg_functions.cc:225-233).
      // KF faces per thread at a time, their loads issued together.
      constexpr int KF = 2;
      for (int j0 = tid; j0 <= NE; j0 += KF * nthr) {
        double ul[KF][NF], ur[KF][NF];
#pragma unroll
        for (int q = 0; q < KF; ++q) {
          const int j = j0 + q * nthr;
          const bool in = j <= NE, lb = j == 0, rb = j == NE;
          const int ea = lb ? 0 : j - 1, na = lb ? 0 : 1;  // side 0: element and node
          const int eb = (lb || rb) ? ea : j, nb = (lb || rb) ? na : 0;
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            ul[q][f] = in ? cur[(ea * NF + f) * Ns + na] : 0.0;
            ur[q][f] = in ? cur[(eb * NF + f) * Ns + nb] : 0.0;
          }
        }
#pragma unroll
        for (int q = 0; q < KF; ++q) {
          const int j = j0 + q * nthr;
          if (j > NE) break;
          const bool lb = j == 0, rb = j == NE;
          const int ea = lb ? 0 : j - 1, na = lb ? 0 : 1;
          const int eb = (lb || rb) ? ea : j, nb = (lb || rb) ? na : 0;
          const int tag = lb ? a.bc[0] : (rb ? a.bc[1] : 0);
          double F[NF], nc[NF];
          if (tag == BC_REFLECTIVE) ur[q][1] = -ur[q][1];  // boundaries.cu:20-35
          riemann<1, PH, FL>(ul[q], ur[q], nrm[j], 0.0, F, nc);
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            Qe[(ea * 2 + na) * NF + f] = psiw * (-F[f] + nc[f]);
            if (!lb && !rb) Qe[(eb * 2 + nb) * NF + f] = psiw * (F[f] + nc[f]);
          }
        }
      }
      __syncthreads();
      // ---- elements: volume terms, face terms, Dt*Minv and the RK4 update.
      // The element's RK operands are loaded as batches issued together (a
      // store to `out`, which may be U, would otherwise order every later
      // load after it): with PF all of them before the volume terms, and the
      // next element's state too when it fits the registers (PFU); else half
      // the fields at a time after the volume terms.
      constexpr bool PFU = PF && NU <= 12;
      constexpr int FB = PF ? NF : (NF + 1) / 2;  // fields per operand batch
      double* out = (k == 3) ? U : S;
      double u[NU];
      if (PFU && tid < NE) {
#pragma unroll
        for (int r = 0; r < NU; ++r) u[r] = cur[tid * NU + r];
      }
      for (int e = tid; e < NE; e += nthr) {
        // vol_phase1 at each Gauss point, accumulated at once into every
        // node's sum in vol_phase2's order (point by point), so that the
        // per-point values need no registers of their own
        constexpr int NFX = V::NFX, NNC = V::NNC;
        double R[NU], pun[FB * Ns], pacc[FB * Ns];
        auto operands = [&](int f0) {
#pragma unroll
          for (int r = 0; r < FB * Ns; ++r) {
            const long idx = (long)e * NU + f0 * Ns + r;
            const bool in = f0 * Ns + r < NU;
            pun[r] = (in && k < 3) ? U[idx] : 0.0;
            pacc[r] = (in && k > 0) ? Uacc[idx] : 0.0;
          }
        };
        if (!PFU) {
#pragma unroll
          for (int r = 0; r < NU; ++r) u[r] = cur[e * NU + r];
        }
        const double Je = J[e], ije = ij[e];
        if (PF) operands(0);
        double u_next[PFU ? NU : 1];
        if (PFU) {
          const int en = e + nthr;
#pragma unroll
          for (int r = 0; r < NU; ++r) u_next[r] = en < NE ? cur[en * NU + r] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < NU; ++r) R[r] = 0.0;
#pragma unroll 1
        for (int g = 0; g < NG; ++g) {
          double feg[NQ];
          vol_phase1<1, P, PH>(tb, g, u, Je, &ije, feg);
#pragma unroll
          for (int i = 0; i < Ns; ++i) {
            const double wd = tb.dw[g * Ns + i];
#pragma unroll
            for (int f = 0; f < NFX; ++f) R[f * Ns + i] += wd * feg[f];
#pragma unroll
            for (int mm = 0; mm < NNC; ++mm) R[(NFX + mm) * Ns + i] += tb.pw[g * Ns + i] * feg[NFX + mm];
          }
        }
        // face terms: edge 0 lands on node 0, edge 1 on node 1 (1D
        // closures); vol_phase2 adds a zero face sum to the interior nodes
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          R[f * Ns + 0] = R[f * Ns + 0] + (0.0 + Qe[(e * 2 + 0) * NF + f]);
          R[f * Ns + 1] = R[f * Ns + 1] + (0.0 + Qe[(e * 2 + 1) * NF + f]);
#pragma unroll
          for (int i = 2; i < Ns; ++i) R[f * Ns + i] = R[f * Ns + i] + 0.0;
        }
        const double dte = dt * (J0 / Je);
#pragma unroll
        for (int f0 = 0; f0 < NF; f0 += FB) {
          if (!PF) operands(f0);
#pragma unroll
          for (int r = f0 * Ns; r < (f0 + FB) * Ns && r < NU; ++r) {
            const int f = r / Ns, i = r % Ns;
            double sol = 0.0;
#pragma unroll
            for (int ii = 0; ii < Ns; ++ii) sol += sMinv[ii * Ns + i] * R[f * Ns + ii];
            const long idx = (long)e * NU + r;
            rk_combine(k, dte * sol, pun[r - f0 * Ns], pacc[r - f0 * Ns], Uacc + idx, out + idx);
          }
        }
        if (PFU) {
#pragma unroll
          for (int r = 0; r < NU; ++r) u[r] = u_next[r];
        }
      }
      __syncthreads();
      cur = out;
    }
    if (a.limiter) limit(U);
  }
}

// LB: the launch bound (512, or 1024 with four waves on every SIMD of the CU)
template <int PH, int FL, int LB>
__global__ __launch_bounds__(LB) void k_batch(BatchArgs a) {
  const int pid = a.wg_prob[blockIdx.x];
  switch (a.order[pid]) {
    case 1: batch_problem<1, PH, FL, LB <= 512>(a, pid); break;
    case 2: batch_problem<2, PH, FL, LB <= 512>(a, pid); break;
    default: batch_problem<3, PH, FL, LB <= 512>(a, pid); break;
  }
}

__global__ void k_fill_zero_batch(double* T, int* status, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    T[k] = 0.0;
    status[k] = 0;
  }
}

// ------------------------------------------------------------------- host
template <typename Fn>
static hipError_t dispatch_batch(const dgx_config& c, Fn&& fn) {
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  auto by_flux = [&](auto ph) -> hipError_t { return c.flux == 0 ? fn(ph, I0{}) : fn(ph, I1{}); };
  return c.physics == 0 ? by_flux(I0{}) : by_flux(I1{});
}

}  // namespace dgx

using namespace dgx;

struct dgx_batch {
  dgx_config cfg{};
  int n = 0;
  int bc[2] = {2, 2};
  std::vector<int> ne, order;
  std::vector<long> state_off, elem_off;
  BatchTables tab[3] = {};
  int block = 512;   // workgroup size
  std::vector<int> order_by_work;  // problems, largest first (workgroup w advances order_by_work[w])
  int *d_ne = nullptr, *d_order = nullptr, *d_wg_prob = nullptr;
  long *d_state_off = nullptr, *d_elem_off = nullptr;
  double *d_J = nullptr, *d_ij = nullptr, *d_normals = nullptr, *d_minv = nullptr, *d_J0 = nullptr,
         *d_dx = nullptr, *d_cfl = nullptr;
  double *Uacc = nullptr, *S = nullptr, *A = nullptr, *Qe = nullptr;
  double *tuneU = nullptr, *tuneT = nullptr;
  int* tune_status = nullptr;
  std::vector<void*> allocs;
  // graph cache
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  hipStream_t cs = nullptr;
  double* g_U = nullptr;
  double* g_T = nullptr;
  int* g_status = nullptr;
  double g_cfl = 0;
  int g_nsteps = -1, g_limiter = -1, g_init = -1;
  std::vector<double> dx;
};

namespace {

template <typename T>
bool bupload(dgx_batch* b, T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  if (!hip_ok(hipMalloc(&p, n * sizeof(T)), "hipMalloc")) return false;
  b->allocs.push_back(p);
  *dst = static_cast<T*>(p);
  if (src) return hip_ok(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return true;
}

void drop_graph(dgx_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void free_batch(dgx_batch* b) {
  drop_graph(b);
  if (b->cs) (void)hipStreamDestroy(b->cs);
  b->cs = nullptr;
  for (void* p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
}

dgx_status bfail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

BatchArgs make_args(dgx_batch* b, double* U, int nsteps, int limiter, int limit_initial, double* T, int* status) {
  BatchArgs a{};
  for (int q = 0; q < 3; ++q) a.tab[q] = b->tab[q];
  a.wg_prob = b->d_wg_prob; a.order = b->d_order;
  a.n_elem = b->d_ne; a.state_off = b->d_state_off; a.elem_off = b->d_elem_off;
  a.J = b->d_J; a.ij = b->d_ij; a.normals = b->d_normals; a.minv = b->d_minv; a.J0 = b->d_J0; a.cfl = b->d_cfl;
  a.bc[0] = b->bc[0]; a.bc[1] = b->bc[1];
  a.U = U; a.Uacc = b->Uacc; a.S = b->S; a.A = b->A; a.Qe = b->Qe; a.T = T; a.status = status;
  a.nsteps = nsteps; a.limiter = limiter; a.limit_initial = limit_initial;
  return a;
}

// one workgroup per problem; the 1024-thread instance for the largest workgroups
hipError_t launch_batch(dgx_batch* b, const BatchArgs& a, int block, hipStream_t st) {
  return dispatch_batch(b->cfg, [&](auto ph, auto fl) -> hipError_t {
    constexpr int PH = decltype(ph)::value, FL = decltype(fl)::value;
    if (block <= 512) hipLaunchKernelGGL((k_batch<PH, FL, 512>), dim3(b->n), dim3(block), 0, st, a);
    else hipLaunchKernelGGL((k_batch<PH, FL, 1024>), dim3(b->n), dim3(block), 0, st, a);
    return hipGetLastError();
  });
}

// per-problem CFL step: deck CFL * Dx / (2p+1) (main.cc:896-897)
bool upload_cfl(dgx_batch* b, double cfl) {
  std::vector<double> c(b->n);
  for (int k = 0; k < b->n; ++k) c[k] = cfl * b->dx[k] / (2 * b->order[k] + 1);
  return hip_ok(hipMemcpy(b->d_cfl, c.data(), sizeof(double) * b->n, hipMemcpyHostToDevice), "cfl upload");
}

}  // namespace

extern "C" {

dgx_status dgx_batch_create(const dgx_config* cfg, int n, const int* n_elem, const int* order, double x0, double x1,
                            const int* bc, dgx_batch** out) {
  if (!cfg || !n_elem || !order || !bc || !out || n <= 0) return bfail(DGX_EINVAL, "dgx_batch_create: bad arguments");
  if (cfg->dim != 1) return bfail(DGX_EUNSUPPORTED, "dgx_batch_create: the ensemble runs 1D problems");
  if (cfg->physics < 0 || cfg->physics > 1 || cfg->flux < 0 || cfg->flux > 1)
    return bfail(DGX_EINVAL, "dgx_batch_create: unknown physics or flux");
  for (int k = 0; k < 2; ++k)
    if (bc[k] != BC_FARFIELD && bc[k] != BC_REFLECTIVE)
      return bfail(DGX_EUNSUPPORTED, "dgx_batch_create: boundary tags must be 2 (farfield) or 3 (wall)");
  dgx_batch* b = new dgx_batch;
  b->cfg = *cfg;
  b->n = n;
  b->bc[0] = bc[0];
  b->bc[1] = bc[1];
  const int NF = 2 + 1 + cfg->physics;
  std::vector<double> J, ij, nrm, minv((size_t)n * 16, 0.0), J0(n), dx(n);
  long so = 0, eo = 0;
  auto bail = [&](int code, const std::string& msg) {
    free_batch(b);
    delete b;
    return msg.empty() ? code : bfail(code, msg);
  };
  for (int k = 0; k < n; ++k) {
    if (n_elem[k] < 2 || order[k] < 1 || order[k] > 3)
      return bail(DGX_EINVAL, "dgx_batch_create: problems need >= 2 elements and order 1..3");
    MeshSpec ms;
    ms.dim = 1;
    ms.order = order[k];
    ms.nx = n_elem[k];
    ms.x0 = x0;
    ms.x1 = x1;
    ms.bc[0] = bc[0];
    ms.bc[1] = bc[1];
    ms.nfields = NF;
    Tables t;
    std::string err = build_tables(ms, t);
    if (!err.empty()) return bail(DGX_EINVAL, "dgx_batch_create: " + err);
    const int Ns = t.N_s, NE = t.N_E;
    b->ne.push_back(NE);
    b->order.push_back(order[k]);
    b->state_off.push_back(so);
    b->elem_off.push_back(eo);
    so += (long)NE * NF * Ns;
    eo += NE;
    for (int e = 0; e < NE; ++e) {
      J.push_back(t.J[e]);
      ij.push_back(t.invJac[(size_t)e * t.N_G]);
    }
    // normals in geometric face order: face j between elements j-1 and j
    std::vector<double> nk(NE + 1, 0.0);
    std::vector<char> seen(NE + 1, 0);
    for (int f = 0; f < t.M_T; ++f) {
      const int e = t.face_elem[2 * f], c = t.face_closure[2 * f];
      const int j = (c == 1) ? e + 1 : e;
      if (!(c == 0 && e == 0) && c != 1) return bail(DGX_EINVAL, "dgx_batch_create: unexpected face layout");
      nk[j] = t.normals[f];
      seen[j] = 1;
    }
    for (char v : seen)
      if (!v) return bail(DGX_EINVAL, "dgx_batch_create: unexpected face layout");
    nrm.insert(nrm.end(), nk.begin(), nk.end());
    // one inverse mass matrix per problem, scaled by J_0/J_e (affine elements)
    double mmax = 0.0;
    for (int q = 0; q < Ns * Ns; ++q) mmax = std::max(mmax, std::fabs(t.Minv[q]));
    for (int e = 1; e < NE; ++e) {
      const double r = t.J[0] / t.J[e];
      for (int q = 0; q < Ns * Ns; ++q)
        if (!(std::fabs(t.Minv[(size_t)e * Ns * Ns + q] - t.Minv[q] * r) <= 1e-13 * mmax * std::max(1.0, r)))
          return bail(DGX_EUNSUPPORTED, "dgx_batch_create: element mass matrices are not J-scaled copies");
    }
    for (int q = 0; q < Ns * Ns; ++q) minv[(size_t)k * 16 + q] = t.Minv[q];
    J0[k] = t.J[0];
    dx[k] = t.Dx;
    BatchTables& g = b->tab[order[k] - 1];
    if (!g.phi) {
      double *phi, *dphi, *phi_w, *dphi_w, *psi_w, *L2M, *M2L;
      const bool ok = bupload(b, &phi, t.phi.data(), t.phi.size()) && bupload(b, &dphi, t.dphi.data(), t.dphi.size()) &&
                      bupload(b, &phi_w, t.phi_w.data(), t.phi_w.size()) &&
                      bupload(b, &dphi_w, t.dphi_w.data(), t.dphi_w.size()) &&
                      bupload(b, &psi_w, t.psi_w.data(), t.psi_w.size()) &&
                      bupload(b, &L2M, t.Lag2Mono.data(), t.Lag2Mono.size()) &&
                      bupload(b, &M2L, t.Mono2Lag.data(), t.Mono2Lag.size());
      if (!ok) return bail(DGX_EDEVICE, "");
      g = BatchTables{phi, dphi, phi_w, dphi_w, psi_w, L2M, M2L};
    }
  }
  b->state_off.push_back(so);
  b->elem_off.push_back(eo);
  b->dx = dx;
  bool ok = bupload(b, &b->d_ne, b->ne.data(), n) && bupload(b, &b->d_order, b->order.data(), n) &&
            bupload(b, &b->d_state_off, b->state_off.data(), n + 1) &&
            bupload(b, &b->d_elem_off, b->elem_off.data(), n + 1) && bupload(b, &b->d_J, J.data(), J.size()) &&
            bupload(b, &b->d_ij, ij.data(), ij.size()) && bupload(b, &b->d_normals, nrm.data(), nrm.size()) &&
            bupload(b, &b->d_J0, J0.data(), n) && bupload<double>(b, &b->d_cfl, nullptr, n) &&
            bupload<double>(b, &b->Uacc, nullptr, so) && bupload<double>(b, &b->S, nullptr, so) &&
            bupload<double>(b, &b->A, nullptr, so) && bupload<double>(b, &b->Qe, nullptr, (size_t)eo * 2 * NF) &&
            bupload<double>(b, &b->tuneU, nullptr, so) && bupload<double>(b, &b->tuneT, nullptr, n) &&
            bupload<int>(b, &b->tune_status, nullptr, n);
  if (ok) ok = bupload(b, &b->d_minv, minv.data(), minv.size());
  // problems by descending DOFs: the largest, which set the time, start first
  b->order_by_work.resize(n);
  std::iota(b->order_by_work.begin(), b->order_by_work.end(), 0);
  std::stable_sort(b->order_by_work.begin(), b->order_by_work.end(), [&](int x, int y) {
    return (long)b->ne[x] * (b->order[x] + 1) > (long)b->ne[y] * (b->order[y] + 1);
  });
  if (ok) ok = bupload(b, &b->d_wg_prob, b->order_by_work.data(), n);
  if (ok) ok = hip_ok(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking), "hipStreamCreate");
  if (!ok) return bail(DGX_EDEVICE, "");
  *out = b;
  return DGX_OK;
}

void dgx_batch_free(dgx_batch* b) {
  if (!b) return;
  free_batch(b);
  delete b;
}

long dgx_batch_state_offset(const dgx_batch* b, int k) {
  if (!b || k < 0 || k > b->n) return -1;
  return b->state_off[k];
}

dgx_status dgx_batch_block_sizes(const dgx_batch* b, int* sizes) {
  if (!b || !sizes) return bfail(DGX_EINVAL, "dgx_batch_block_sizes: null argument");
  sizes[0] = b->block;
  sizes[1] = b->n;
  sizes[2] = 0;
  return DGX_OK;
}

dgx_status dgx_batch_set_block_sizes(dgx_batch* b, const int* sizes) {
  if (!b || !sizes) return bfail(DGX_EINVAL, "dgx_batch_set_block_sizes: null argument");
  if (sizes[0] < 64 || sizes[0] > 1024 || sizes[0] % 64)
    return bfail(DGX_EINVAL, "dgx_batch_set_block_sizes: workgroup size a multiple of 64 in [64, 1024]");
  b->block = sizes[0];
  drop_graph(b);
  return DGX_OK;
}

dgx_status dgx_batch_autotune(dgx_batch* b, const double* U, double cfl, int limiter, void* stream) {
  if (!b || !U) return bfail(DGX_EINVAL, "dgx_batch_autotune: null argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!upload_cfl(b, cfl)) return DGX_EDEVICE;
  hipEvent_t t0, t1;
  if (!hip_ok(hipEventCreate(&t0), "hipEventCreate") || !hip_ok(hipEventCreate(&t1), "hipEventCreate"))
    return DGX_EDEVICE;
  const int sizes[3] = {256, 512, 1024};
  float best = 1e30f;
  int best_b = b->block;
  dgx_status r = DGX_OK;
  for (int bs : sizes) {
    float tot = 0.0f;
    for (int rep = 0; rep < 2 && r == DGX_OK; ++rep) {
      // scratch copies: the caller's state and times are left untouched
      if (!hip_ok(hipMemcpyAsync(b->tuneU, U, sizeof(double) * b->state_off[b->n], hipMemcpyDeviceToDevice, st),
                  "autotune copy")) { r = DGX_EDEVICE; break; }
      hipLaunchKernelGGL(k_fill_zero_batch, dim3((b->n + 255) / 256), dim3(256), 0, st, b->tuneT, b->tune_status,
                         b->n);
      BatchArgs a = make_args(b, b->tuneU, 1, limiter, 0, b->tuneT, b->tune_status);
      (void)hipEventRecord(t0, st);
      if (!hip_ok(launch_batch(b, a, bs, st), "autotune launch")) { r = DGX_EDEVICE; break; }
      (void)hipEventRecord(t1, st);
      if (!hip_ok(hipEventSynchronize(t1), "autotune sync")) { r = DGX_EDEVICE; break; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, t0, t1);
      if (rep == 1) tot += ms;  // first run warms the code object and caches
    }
    if (r != DGX_OK) break;
    if (tot < best) {
      best = tot;
      best_b = bs;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  b->block = best_b;
  drop_graph(b);
  return r;
}

dgx_status dgx_batched_rk4(dgx_batch* b, double* U, double cfl, int nsteps, int limiter, int limit_initial, double* T,
                           int* status, void* stream) {
  if (!b || !U || !T || !status || nsteps < 0) return bfail(DGX_EINVAL, "dgx_batched_rk4: bad arguments");
  if (limiter != 0 && limiter != 1) return bfail(DGX_EUNSUPPORTED, "dgx_batched_rk4: limiter must be 0 or 1 (hrl)");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const bool same = b->exec && b->g_U == U && b->g_T == T && b->g_status == status && b->g_cfl == cfl &&
                    b->g_nsteps == nsteps && b->g_limiter == limiter && b->g_init == limit_initial;
  if (!same) {
    drop_graph(b);
    if (!upload_cfl(b, cfl)) return DGX_EDEVICE;
    // the kernel, captured on the internal stream
    if (!hip_ok(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return DGX_EDEVICE;
    BatchArgs a = make_args(b, U, nsteps, limiter, limit_initial, T, status);
    const bool ok = hip_ok(launch_batch(b, a, b->block, b->cs), "batched launch");
    hipGraph_t graph = nullptr;
    const bool ended = hip_ok(hipStreamEndCapture(b->cs, &graph), "hipStreamEndCapture");
    if (!ok || !ended) {
      if (graph) (void)hipGraphDestroy(graph);
      return DGX_EDEVICE;
    }
    b->graph = graph;
    if (!hip_ok(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) {
      drop_graph(b);
      return DGX_EDEVICE;
    }
    b->g_U = U; b->g_T = T; b->g_status = status; b->g_cfl = cfl;
    b->g_nsteps = nsteps; b->g_limiter = limiter; b->g_init = limit_initial;
  }
  if (!hip_ok(hipGraphLaunch(b->exec, st), "hipGraphLaunch")) return DGX_EDEVICE;
  return DGX_OK;
}

}  // extern "C"
