// Batched ensemble: many independent 1D problems of uneven size and order,
// each advanced by the RK4 loop of RK::RK_integration (rk.cc:132-247) with
// its own CFL time step (rk.cc:250-272), the HR limiter (limiting.cc:11-27)
// and the DG right-hand side of DG_SOLVER::dg_solver (dg_solver.h:292-364).
//
// One persistent kernel for every order.  The host splits each problem into
// contiguous element chunks of about equal work (DOFs) and gives each chunk
// a workgroup, so that a 4000-element p = 3 problem runs on several CUs while
// a 200-element p = 1 problem takes one workgroup; the grid never exceeds
// what the device holds at once.  The workgroups of one problem meet at a
// barrier of their own (problem_sync: an arrival counter per problem and
// agent-scope fences) between the phases that read a neighbouring chunk's
// data: the Dt reduction, the face pass, and the two halves of the limiter.
// Per step: Dt from the problem's maximum of |u|+a; per stage: a face pass
// (one lane per interface, Riemann flux integrated into element-major edge
// slots), an element pass (one lane per element, volume terms in registers
// with the element-local helpers of volume.cuh, Dt*Minv and the RK4
// combination), and the HR limiter (Lagrange->Taylor, per-column limiting
// with the neighbours' Taylor coefficients, Taylor->Lagrange, without FMA
// contraction as in the CPU path).  The launch is captured once in a hipGraph
// (after a fill node that zeroes the barrier counters) and replayed; the
// workgroup size and the number of chunks come from a timing autotuner.
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
  // work split: workgroup w advances elements [wg_e0[w], wg_e1[w]) of problem wg_prob[w]
  const int* wg_prob;
  const int* wg_e0;
  const int* wg_e1;
  const int* prob_wg0;      // [problem] its first workgroup
  const int* prob_nwg;      // [problem] its number of workgroups
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
  int* status;  // [problem] 0 ok, 1 Dt too small or NaN, 2 a barrier wait timed out
  unsigned* bar;             // [problem] barrier arrivals (zero at launch)
  double* dtp;               // [workgroup] partial maxima of |u|+a
  long long sync_ticks;      // wall-clock ticks after which a barrier wait gives up
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

// Barrier of the nwg workgroups that advance one problem.  Arrivals are
// counted in *ctr, which starts at zero and is never reset within a launch:
// the g-th barrier completes when the count reaches g*nwg.  Every wave
// fences at agent scope before the arrival (its global writes reach the
// other workgroups, whatever XCD they run on) and after the wait (stale
// cached lines are dropped).  A wait longer than a.sync_ticks gives up and
// returns false; the grid is sized so that every workgroup is resident, so
// that only happens when something else holds the device.
__device__ bool problem_sync(unsigned* ctr, int nwg, unsigned& gen, long long ticks, int* s_ok) {
  if (nwg == 1) {
    __syncthreads();
    return true;
  }
  __threadfence();
  __syncthreads();
  gen += 1u;
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned target = gen * (unsigned)nwg;
    const unsigned prev = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev + 1u < target) {
      const long long t0 = wall_clock64();
      while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        __builtin_amdgcn_s_sleep(2);
        if (wall_clock64() - t0 > ticks) {
          ok = 0;
          break;
        }
      }
    }
    *s_ok = ok;
  }
  __syncthreads();
  __threadfence();
  return *s_ok != 0;
}

// Taylor coefficients of the chunk's columns (the first half of the HR
// limiter, limiting.cc:11-27 in 1D): A = L2M U per (element, field) column.
template <int P, int PH>
__device__ void batch_limit_l2m(const double* U, double* A, int e0, int e1, const double* sL2M) {
#pragma clang fp contract(off)
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns;
  for (int col = e0 * NF + threadIdx.x; col < e1 * NF; col += blockDim.x) {
    double x[Ns], y[Ns];
#pragma unroll
    for (int l = 0; l < Ns; ++l) x[l] = U[col * Ns + l];
#pragma unroll
    for (int r = 0; r < Ns; ++r) y[r] = 0.0;
#pragma unroll
    for (int l = 0; l < Ns; ++l)
#pragma unroll
      for (int r = 0; r < Ns; ++r) y[r] = y[r] + x[l] * sL2M[r + l * Ns];
#pragma unroll
    for (int r = 0; r < Ns; ++r) A[col * Ns + r] = y[r];
  }
}

// The second half: per column of the chunk, hrl1D with the neighbours'
// Taylor coefficients (another chunk's at the chunk ends), then U = M2L Alim.
template <int P, int PH>
__device__ void batch_limit_apply(const BatchArgs& a, double* U, const double* A, int NE, int e0, int e1,
                                  const double* sM2L) {
#pragma clang fp contract(off)
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns;
  for (int col = e0 * NF + threadIdx.x; col < e1 * NF; col += blockDim.x) {
    const int e = col / NF, fc = col % NF;
    const int left = e > 0 ? e - 1 : -a.bc[0];
    const int right = e < NE - 1 ? e + 1 : -a.bc[1];
    int physical = 0;
    if (left < 0) physical = -left;
    if (right < 0) physical = -right;
    double lim[Ns], ac[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) { ac[n] = A[col * Ns + n]; lim[n] = 0.0; }
    if (physical == 2 || (physical == 3 && fc != 1)) {
      average_monomial_d<P>(ac, lim);
    } else if (physical == 3) {
#pragma unroll
      for (int n = 0; n < Ns; ++n) lim[n] = ac[n];
    } else {
      double al[Ns], ar[Ns];
#pragma unroll
      for (int n = 0; n < Ns; ++n) {
        al[n] = A[(left * NF + fc) * Ns + n];
        ar[n] = A[(right * NF + fc) * Ns + n];
      }
      limit_monomial_d<P>(al, ac, ar, lim);
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

// One workgroup's share of its problem (order P): every step, stage and
// limiter application over elements [e0, e1), meeting the problem's other
// workgroups at problem_sync.
template <int P, int PH, int FL>
__device__ void batch_chunk(const BatchArgs& a, int w) {
  using B = BatchShape<P, PH>;
  using V = Vol<1, P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, NG = B::NG, NU = B::NU, NQ = V::NQ;
  __shared__ double sPhi[V::NTAB_PHI], sDphi[PH ? V::NTAB_D : 1], sDw[V::NTAB_D], sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns], sL2M[Ns * Ns], sM2L[Ns * Ns];
  __shared__ double sRed[8];
  __shared__ int sOk;
  const VolTables<1, P, PH> tb{sPhi, sDphi, sDw, sPw, sMinv};
  const BatchTables& bt = a.tab[P - 1];

  const int pid = a.wg_prob[w];
  const int NE = a.n_elem[pid];
  const int e0 = a.wg_e0[w], e1 = a.wg_e1[w];
  const int w0 = a.prob_wg0[pid], nwg = a.prob_nwg[pid];
  const bool lead = w == w0;
  const long so = a.state_off[pid], eo = a.elem_off[pid];
  double* U = a.U + so;
  double* Uacc = a.Uacc + so;
  double* S = a.S + so;
  double* A = a.A + so;
  double* Qe = a.Qe + eo * 2 * NF;
  const double* J = a.J + eo;
  const double* ij = a.ij + eo;
  const double* nrm = a.normals + eo + pid;
  unsigned* bar = a.bar + pid;
  const double cfl = a.cfl[pid], J0 = a.J0[pid], psiw = bt.psi_w[0];
  const int tid = threadIdx.x, nthr = blockDim.x;
  unsigned gen = 0;
  auto sync = [&]() { return problem_sync(bar, nwg, gen, a.sync_ticks, &sOk); };
  auto give_up = [&]() {
    if (tid == 0) a.status[pid] = 2;
  };

  load_vol_tables(tb, bt.phi, bt.dphi, bt.phi_w, bt.dphi_w, nullptr, false, tid, nthr);
  for (int k = tid; k < Ns * Ns; k += nthr) {
    sMinv[k] = a.minv[(long)pid * 16 + k];
    sL2M[k] = bt.L2M[k];
    sM2L[k] = bt.M2L[k];
  }
  __syncthreads();
  if (a.status[pid] != 0) return;  // every workgroup of the problem reads the same value
  auto limit = [&](double* V) {
    batch_limit_l2m<P, PH>(V, A, e0, e1, sL2M);
    if (!sync()) return false;  // the neighbouring chunks' Taylor coefficients
    batch_limit_apply<P, PH>(a, V, A, NE, e0, e1, sM2L);
    __syncthreads();  // the chunk's limited columns, before this workgroup reads them by node
    return true;
  };
  if (a.limiter && a.limit_initial && !limit(U)) return give_up();

  for (int step = 0; step < a.nsteps; ++step) {
    // ---- Dt = cfl / max over the problem's nodes of |u|+a (rk.cc:250-272)
    double m = 0.0;
    for (int n = e0 * Ns + tid; n < e1 * Ns; n += nthr) {
      const int e = n / Ns, i = n % Ns;
      double u[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) u[f] = U[(e * NF + f) * Ns + i];
      m = nanmax(m, upa<1, PH>(u));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) sRed[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double mx = sRed[0];
      for (int k = 1; k < (nthr + 63) / 64; ++k) mx = nanmax(mx, sRed[k]);
      a.dtp[w] = mx;
    }
    // the chunks' maxima: every workgroup reduces them in the same order
    if (!sync()) return give_up();
    if (tid == 0) {
      double mx = a.dtp[w0];
      for (int k = 1; k < nwg; ++k) mx = nanmax(mx, a.dtp[w0 + k]);
      sRed[0] = cfl / mx;
    }
    __syncthreads();
    const double dt = sRed[0];
    __syncthreads();  // sRed is reused by the next step
    if (!(dt >= 1e-14)) {
      if (lead && tid == 0) a.status[pid] = 1;
      return;
    }
    if (lead && tid == 0) a.T[pid] = a.T[pid] + dt;

    double* cur = U;
    for (int k = 0; k < 4; ++k) {
      // the previous stage's output (and, at k = 0, the previous step's
      // limited state) of the neighbouring chunks is complete here
      if (k > 0 && a.limiter && !limit(cur)) return give_up();
      if (k > 0 && !sync()) return give_up();
      // ---- faces: interface j between elements j-1 and j; boundary faces
      // contribute to the boundary element only (dg_functions.cc:225-233).
      // The chunk owns interfaces e0 .. e1-1, the last chunk also NE.
      const int nif = e1 - e0 + (e1 == NE ? 1 : 0);
      for (int jj = tid; jj < nif; jj += nthr) {
        const int j = e0 + jj;
        const bool lb = j == 0, rb = j == NE;
        const int ea = lb ? 0 : j - 1, na = lb ? 0 : 1;  // side 0: element and node
        const int eb = (lb || rb) ? ea : j, nb = (lb || rb) ? na : 0;
        const int tag = lb ? a.bc[0] : (rb ? a.bc[1] : 0);
        double ul[NF], ur[NF], F[NF], nc[NF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          ul[f] = cur[(ea * NF + f) * Ns + na];
          ur[f] = cur[(eb * NF + f) * Ns + nb];
        }
        if (tag == BC_REFLECTIVE) ur[1] = -ur[1];  // boundaries.cu:20-35
        riemann<1, PH, FL>(ul, ur, nrm[j], 0.0, F, nc);
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          Qe[(ea * 2 + na) * NF + f] = psiw * (-F[f] + nc[f]);
          if (!lb && !rb) Qe[(eb * 2 + nb) * NF + f] = psiw * (F[f] + nc[f]);
        }
      }
      // the left neighbour chunk wrote edge 1 of this chunk's first element
      if (!sync()) return give_up();
      // ---- elements: volume terms, face terms, Dt*Minv and the RK4 update
      double* out = (k == 3) ? U : S;
      for (int e = e0 + tid; e < e1; e += nthr) {
        // vol_phase1 at each Gauss point, accumulated at once into every
        // node's sum in vol_phase2's order (point by point), so that the
        // per-point values need no registers of their own
        constexpr int NFX = V::NFX, NNC = V::NNC;
        double u[NU], R[NU];
#pragma unroll
        for (int r = 0; r < NU; ++r) {
          u[r] = cur[e * NU + r];
          R[r] = 0.0;
        }
        const double Je = J[e], ije = ij[e];
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
            for (int m = 0; m < NNC; ++m) R[(NFX + m) * Ns + i] += tb.pw[g * Ns + i] * feg[NFX + m];
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
        for (int r = 0; r < NU; ++r) {
          const int f = r / Ns, i = r % Ns;
          double sol = 0.0;
#pragma unroll
          for (int ii = 0; ii < Ns; ++ii) sol += sMinv[ii * Ns + i] * R[f * Ns + ii];
          const long idx = (long)e * NU + r;
          const double un = k < 3 ? U[idx] : 0.0;
          const double acc = k > 0 ? Uacc[idx] : 0.0;
          rk_combine(k, dte * sol, un, acc, Uacc + idx, out + idx);
        }
      }
      __syncthreads();  // the chunk's stage output, before this workgroup reads it by column or node
      cur = out;
    }
    if (a.limiter && !limit(U)) return give_up();
    // the next step's Dt reads only this chunk's nodes; its barrier orders
    // the limited state before the next face pass reads a neighbour's
  }
}

template <int PH, int FL>
__global__ __launch_bounds__(256) void k_batch(BatchArgs a) {
  const int w = blockIdx.x;
  switch (a.order[a.wg_prob[w]]) {
    case 1: batch_chunk<1, PH, FL>(a, w); break;
    case 2: batch_chunk<2, PH, FL>(a, w); break;
    default: batch_chunk<3, PH, FL>(a, w); break;
  }
}

__global__ void k_fill_zero_batch(double* T, int* status, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    T[k] = 0.0;
    status[k] = 0;
  }
}

// the barrier counters start every launch at zero (a kernel node, not a
// memset node, in the captured graph)
__global__ void k_zero_u32(unsigned* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

// ------------------------------------------------------------------- host
template <typename Fn>
static hipError_t dispatch_batch(const dgx_config& c, Fn&& fn) {
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  auto by_flux = [&](auto ph) -> hipError_t { return c.flux == 0 ? fn(ph, I0{}) : fn(ph, I1{}); };
  return c.physics == 0 ? by_flux(I0{}) : by_flux(I1{});
}

// Workgroups of one problem: its element chunks.
struct BatchPlan {
  int block = 0, nwg = 0;
  std::vector<int> wg_prob, wg_e0, wg_e1, prob_wg0, prob_nwg;
};

// Splits the problems over at most `target` workgroups (at least one each):
// repeatedly the problem with the most DOFs per workgroup gets one more, as
// long as its chunks keep >= kMinChunk elements; each problem's elements are
// then cut into equal chunks.  target <= the problems' count: one each.
constexpr int kMinChunk = 32;
BatchPlan make_plan(const std::vector<int>& ne, const std::vector<int>& order, int nf, int block, int target) {
  const int n = (int)ne.size();
  BatchPlan pl;
  pl.block = block;
  std::vector<int> k(n, 1);
  std::vector<double> work(n);
  for (int p = 0; p < n; ++p) work[p] = (double)ne[p] * nf * (order[p] + 1);
  int total = n;
  while (total < target) {
    int best = -1;
    double bw = 0.0;
    for (int p = 0; p < n; ++p)
      if (ne[p] / (k[p] + 1) >= kMinChunk && work[p] / k[p] > bw) {
        bw = work[p] / k[p];
        best = p;
      }
    if (best < 0) break;
    ++k[best];
    ++total;
  }
  pl.prob_wg0.resize(n);
  pl.prob_nwg = k;
  for (int p = 0; p < n; ++p) {
    pl.prob_wg0[p] = (int)pl.wg_prob.size();
    for (int c = 0; c < k[p]; ++c) {
      pl.wg_prob.push_back(p);
      pl.wg_e0.push_back((int)((long)ne[p] * c / k[p]));
      pl.wg_e1.push_back((int)((long)ne[p] * (c + 1) / k[p]));
    }
  }
  pl.nwg = (int)pl.wg_prob.size();
  return pl;
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
  int block = 256;   // workgroup size
  int target = 0;    // workgroups to split the problems over (0: as many as the device holds at once)
  BatchPlan plan;    // the split in use (uploaded)
  int wg_cap = 0;    // capacity of the per-workgroup device arrays
  int *d_ne = nullptr, *d_order = nullptr;
  int *d_wg_prob = nullptr, *d_wg_e0 = nullptr, *d_wg_e1 = nullptr, *d_prob_wg0 = nullptr, *d_prob_nwg = nullptr;
  long *d_state_off = nullptr, *d_elem_off = nullptr;
  double *d_J = nullptr, *d_ij = nullptr, *d_normals = nullptr, *d_minv = nullptr, *d_J0 = nullptr,
         *d_dx = nullptr, *d_cfl = nullptr;
  double *Uacc = nullptr, *S = nullptr, *A = nullptr, *Qe = nullptr;
  double *tuneU = nullptr, *tuneT = nullptr, *d_dtp = nullptr;
  unsigned* d_bar = nullptr;
  int* tune_status = nullptr;
  long long sync_ticks = 0;
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

// Workgroups of `block` threads the device holds at once (the kernel's
// occupancy times the CU count): the most a split may use, since the
// workgroups of a split problem wait for each other.
int resident_workgroups(const dgx_batch* b, int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  hipError_t e = dispatch_batch(b->cfg, [&](auto ph, auto fl) -> hipError_t {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch<decltype(ph)::value, decltype(fl)::value>,
                                                        block, 0);
  });
  return e == hipSuccess ? per_cu * cus : 0;
}

// Makes and uploads the split for (block, target); target 0: the resident limit.
bool set_plan(dgx_batch* b, int block, int target) {
  const int cap = resident_workgroups(b, block);
  if (cap <= 0) return hip_ok(hipErrorInvalidValue, "dgx_batch: occupancy query");
  const int t = std::min(target > 0 ? target : cap, cap);
  const int nf = 2 + 1 + b->cfg.physics;
  BatchPlan pl = make_plan(b->ne, b->order, nf, block, t);
  if (pl.nwg > b->wg_cap) {
    const int c = std::max(pl.nwg, 2 * b->wg_cap);
    if (!bupload<int>(b, &b->d_wg_prob, nullptr, c) || !bupload<int>(b, &b->d_wg_e0, nullptr, c) ||
        !bupload<int>(b, &b->d_wg_e1, nullptr, c) || !bupload<double>(b, &b->d_dtp, nullptr, c))
      return false;
    b->wg_cap = c;
  }
  const size_t nw = (size_t)pl.nwg, np = (size_t)b->n;
  if (!hip_ok(hipMemcpy(b->d_wg_prob, pl.wg_prob.data(), nw * 4, hipMemcpyHostToDevice), "plan upload") ||
      !hip_ok(hipMemcpy(b->d_wg_e0, pl.wg_e0.data(), nw * 4, hipMemcpyHostToDevice), "plan upload") ||
      !hip_ok(hipMemcpy(b->d_wg_e1, pl.wg_e1.data(), nw * 4, hipMemcpyHostToDevice), "plan upload") ||
      !hip_ok(hipMemcpy(b->d_prob_wg0, pl.prob_wg0.data(), np * 4, hipMemcpyHostToDevice), "plan upload") ||
      !hip_ok(hipMemcpy(b->d_prob_nwg, pl.prob_nwg.data(), np * 4, hipMemcpyHostToDevice), "plan upload"))
    return false;
  b->plan = std::move(pl);
  b->block = block;
  b->target = target;
  drop_graph(b);
  return true;
}

BatchArgs make_args(dgx_batch* b, double* U, int nsteps, int limiter, int limit_initial, double* T, int* status) {
  BatchArgs a{};
  for (int q = 0; q < 3; ++q) a.tab[q] = b->tab[q];
  a.wg_prob = b->d_wg_prob; a.wg_e0 = b->d_wg_e0; a.wg_e1 = b->d_wg_e1;
  a.prob_wg0 = b->d_prob_wg0; a.prob_nwg = b->d_prob_nwg; a.order = b->d_order;
  a.n_elem = b->d_ne; a.state_off = b->d_state_off; a.elem_off = b->d_elem_off;
  a.J = b->d_J; a.ij = b->d_ij; a.normals = b->d_normals; a.minv = b->d_minv; a.J0 = b->d_J0; a.cfl = b->d_cfl;
  a.bc[0] = b->bc[0]; a.bc[1] = b->bc[1];
  a.U = U; a.Uacc = b->Uacc; a.S = b->S; a.A = b->A; a.Qe = b->Qe; a.T = T; a.status = status;
  a.bar = b->d_bar; a.dtp = b->d_dtp; a.sync_ticks = b->sync_ticks;
  a.nsteps = nsteps; a.limiter = limiter; a.limit_initial = limit_initial;
  return a;
}

// the counter fill and the ensemble kernel
hipError_t launch_batch(dgx_batch* b, const BatchArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u32, dim3((b->n + 255) / 256), dim3(256), 0, st, b->d_bar, b->n);
  return dispatch_batch(b->cfg, [&](auto ph, auto fl) -> hipError_t {
    hipLaunchKernelGGL((k_batch<decltype(ph)::value, decltype(fl)::value>), dim3(b->plan.nwg), dim3(b->plan.block), 0,
                       st, a);
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
            bupload<int>(b, &b->tune_status, nullptr, n) && bupload<unsigned>(b, &b->d_bar, nullptr, n) &&
            bupload<int>(b, &b->d_prob_wg0, nullptr, n) && bupload<int>(b, &b->d_prob_nwg, nullptr, n);
  if (ok) ok = bupload(b, &b->d_minv, minv.data(), minv.size());
  // a barrier wait gives up after 2 s of the device's constant-rate clock
  int dev = 0, khz = 0;
  if (ok)
    ok = hip_ok(hipGetDevice(&dev), "hipGetDevice") &&
         hip_ok(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
  b->sync_ticks = 2000LL * (khz > 0 ? khz : 100000);
  if (ok) ok = hip_ok(hipStreamCreateWithFlags(&b->cs, hipStreamNonBlocking), "hipStreamCreate");
  if (ok) ok = set_plan(b, 256, 0);
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
  sizes[0] = b->plan.block;
  sizes[1] = b->plan.nwg;
  sizes[2] = b->target;
  return DGX_OK;
}

dgx_status dgx_batch_set_block_sizes(dgx_batch* b, const int* sizes) {
  if (!b || !sizes) return bfail(DGX_EINVAL, "dgx_batch_set_block_sizes: null argument");
  if (sizes[0] < 64 || sizes[0] > 256 || sizes[0] % 64 || sizes[1] < 0)
    return bfail(DGX_EINVAL, "dgx_batch_set_block_sizes: workgroup size 64, 128, 192 or 256; workgroups >= 0");
  return set_plan(b, sizes[0], sizes[1]) ? DGX_OK : DGX_EDEVICE;
}

dgx_status dgx_batch_autotune(dgx_batch* b, const double* U, double cfl, int limiter, void* stream) {
  if (!b || !U) return bfail(DGX_EINVAL, "dgx_batch_autotune: null argument");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!upload_cfl(b, cfl)) return DGX_EDEVICE;
  hipEvent_t t0, t1;
  if (!hip_ok(hipEventCreate(&t0), "hipEventCreate") || !hip_ok(hipEventCreate(&t1), "hipEventCreate"))
    return DGX_EDEVICE;
  // workgroup sizes x splits: one workgroup per problem, or a half or all of
  // the workgroups the device holds at once
  const int sizes[3] = {64, 128, 256};
  float best = 1e30f;
  int best_b = b->block, best_t = b->target;
  dgx_status r = DGX_OK;
  for (int bs : sizes) {
    const int cap = resident_workgroups(b, bs);
    const int targets[3] = {b->n, cap / 2, 0};
    for (int tg : targets) {
      if (tg > 0 && tg < b->n) continue;
      if (!set_plan(b, bs, tg)) { r = DGX_EDEVICE; break; }
      float tot = 0.0f;
      for (int rep = 0; rep < 2 && r == DGX_OK; ++rep) {
        // scratch copies: the caller's state and times are left untouched
        if (!hip_ok(hipMemcpyAsync(b->tuneU, U, sizeof(double) * b->state_off[b->n], hipMemcpyDeviceToDevice, st),
                    "autotune copy")) { r = DGX_EDEVICE; break; }
        hipLaunchKernelGGL(k_fill_zero_batch, dim3((b->n + 255) / 256), dim3(256), 0, st, b->tuneT, b->tune_status,
                           b->n);
        BatchArgs a = make_args(b, b->tuneU, 1, limiter, 0, b->tuneT, b->tune_status);
        (void)hipEventRecord(t0, st);
        if (!hip_ok(launch_batch(b, a, st), "autotune launch")) { r = DGX_EDEVICE; break; }
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
        best_t = tg;
      }
    }
    if (r != DGX_OK) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (r == DGX_OK && !set_plan(b, best_b, best_t)) r = DGX_EDEVICE;
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
    // the counter fill and the kernel, captured on the internal stream
    if (!hip_ok(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return DGX_EDEVICE;
    BatchArgs a = make_args(b, U, nsteps, limiter, limit_initial, T, status);
    const bool ok = hip_ok(launch_batch(b, a, b->cs), "batched launch");
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
