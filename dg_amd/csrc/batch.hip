// Batched ensemble: many independent 1D problems of uneven size and order,
// each advanced by the RK4 loop of RK::RK_integration (rk.cc:132-247) with
// its own CFL time step (rk.cc:250-272), the HR limiter (limiting.cc:11-27)
// and the DG right-hand side of DG_SOLVER::dg_solver (dg_solver.h:292-364).
//
// One persistent kernel for every order.  The host cuts the ensemble into
// segments of about equal work, one list of segments per workgroup, every
// workgroup resident at once: a problem much larger than the average is
// cut into chunks of consecutive elements on consecutive workgroups, and the
// smaller problems are packed several to a workgroup (longest first), which
// runs them one after the other.  A workgroup runs all steps, stages and
// limiter sweeps of its segment behind workgroup barriers.  Per step: Dt from
// the maximum of |u|+a; per stage: a face pass (one lane per interface,
// Riemann flux integrated into element-major edge slots), an element pass
// (one lane per element, volume terms accumulated point by point with the
// element-local helpers of volume.cuh, Dt*Minv and the RK4 combination), and
// the HR limiter (Lagrange->Taylor, per-column limiting with the neighbours'
// Taylor coefficients, Taylor->Lagrange, without FMA contraction as in the CPU
// path).
//
// The chunks of a problem exchange, once per limiter application (or per
// stage without the limiter), the Taylor coefficients (or states) of their two
// first and two last elements with the neighbouring chunks, and each limits
// the two elements just outside it itself (the same arithmetic on the same
// bits as their owner): its faces at the chunk ends then need nothing more.
// The chunks' maxima of |u|+a are gathered once per step.  The exchange
// goes through agent-scope relaxed atomics (written with exchanges whose
// returned values are consumed before the chunk raises its epoch flag, read
// after the neighbour's flag shows the epoch), so that no agent-scope fence
// writes back or invalidates the L2 of the XCDs; every wait is bounded, and a
// chunk that times out marks its problem (status 2) and ends.  The launch is
// captured once in a hipGraph and replayed; the workgroup size comes from a
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

// A workgroup's segment: the elements [c0, c1) of problem pid; left/right:
// the workgroups of the neighbouring chunks (-1 at the problem's ends); the
// problem's chunks are the workgroups wg0 .. wg0+nch-1 (nch = 1: the whole
// problem, no exchange).
struct BatchSeg {
  int pid, c0, c1, left, right, wg0, nch;
};

constexpr int kXA = 64;  // exchanged doubles per chunk and epoch: 4 elements x N_F x N_s (<= 4 x 4 x 4)

struct BatchArgs {
  BatchTables tab[3];       // orders 1..3
  const BatchSeg* seg;      // segments
  const int* wg_seg;        // [workgroup]: its segments seg[wg_seg[w] .. wg_seg[w+1])
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
  int* status;  // [problem] 0 ok, 1 Dt too small or NaN, 2 a chunk exchange timed out
  // chunk exchange (zeroed before every launch)
  double* xA;        // [workgroup][2][kXA]: boundary elements, by epoch parity
  double* xD;        // [workgroup][2]: the chunk's maximum of |u|+a, by step parity
  unsigned* flagA;   // [workgroup]: last epoch published to xA
  unsigned* flagD;   // [workgroup]: last step published to xD
  long spin_max;     // bound of every wait (polls)
  int nsteps, limiter, limit_initial;
};

// Agent-scope relaxed accesses of the exchange.  A value is written with an
// exchange whose returned old value the wave waits for, so the write has
// reached the coherence point before the epoch flag that follows it.
__device__ inline void x_put(double* p, double v) {
  const unsigned long long old =
      __hip_atomic_exchange(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("" ::"v"(old));
}
__device__ inline double x_get(double* p) {
  return __longlong_as_double(
      (long long)__hip_atomic_load(reinterpret_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void x_flag(unsigned* f, unsigned v) {
  const unsigned old = __hip_atomic_exchange(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("" ::"v"(old));
}
// true once *f >= v; false after spin_max polls
__device__ inline bool x_wait(unsigned* f, unsigned v, long spin_max) {
  for (long n = 0; __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v; ++n) {
    if (n >= spin_max) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

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

// One column's limited nodal values (the second half of the HR limiter):
// hrl1D with the neighbours' Taylor coefficients (or the boundary rule),
// then out = M2L Alim.
template <int P>
__device__ inline void hr_column(int physical, int fc, const double* al, const double* ac, const double* ar,
                                 const double* sM2L, double* out) {
#pragma clang fp contract(off)
  constexpr int Ns = P + 1;
  double lim[Ns];
#pragma unroll
  for (int n = 0; n < Ns; ++n) lim[n] = 0.0;
  if (physical == 2 || (physical == 3 && fc != 1)) {
    average_monomial_d<P>(ac, lim);
  } else if (physical == 3) {
#pragma unroll
    for (int n = 0; n < Ns; ++n) lim[n] = ac[n];
  } else {
    limit_monomial_d<P>(al, ac, ar, lim);
  }
#pragma unroll
  for (int r = 0; r < Ns; ++r) out[r] = 0.0;
#pragma unroll
  for (int l = 0; l < Ns; ++l)
#pragma unroll
    for (int r = 0; r < Ns; ++r) out[r] = out[r] + lim[l] * sM2L[r + l * Ns];
}

__device__ inline int batch_physical(const BatchArgs& a, int e, int NE, int* left, int* right) {
  *left = e > 0 ? e - 1 : -a.bc[0];
  *right = e < NE - 1 ? e + 1 : -a.bc[1];
  int ph = 0;
  if (*left < 0) ph = -*left;
  if (*right < 0) ph = -*right;
  return ph;
}

// The columns of elements [e0, e1): U = M2L hrl(A).  The Taylor
// coefficients of element e0-1 and e1 come from gl and gr when given (the
// neighbouring chunks' elements, [f*Ns + n]), else from A.  The column's own
// and its neighbours' coefficients are loaded together.
template <int P, int PH>
__device__ void batch_limit_apply(const BatchArgs& a, double* U, const double* A, int NE, int e0, int e1,
                                  const double* sM2L, const double* gl, const double* gr) {
#pragma clang fp contract(off)
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns;
  const int end = e1 * NF;
  for (int col = e0 * NF + threadIdx.x; col < end; col += blockDim.x) {
    const int e = col / NF, fc = col % NF;
    int left, right;
    const int physical = batch_physical(a, e, NE, &left, &right);
    const double* pl = (gl && left == e0 - 1) ? gl + fc * Ns : A + (left * NF + fc) * Ns;
    const double* pr = (gr && right == e1) ? gr + fc * Ns : A + (right * NF + fc) * Ns;
    double al[Ns], ac[Ns], ar[Ns], y[Ns];
#pragma unroll
    for (int n = 0; n < Ns; ++n) {
      ac[n] = A[col * Ns + n];
      al[n] = physical == 0 ? pl[n] : 0.0;
      ar[n] = physical == 0 ? pr[n] : 0.0;
    }
    hr_column<P>(physical, fc, al, ac, ar, sM2L, y);
#pragma unroll
    for (int r = 0; r < Ns; ++r) U[col * Ns + r] = y[r];
  }
}

// The chunk exchange: publish elements c0, c0+1, c1-2, c1-1 of X (NU values
// each) for this epoch, wait for the neighbouring chunks' epoch, read their
// elements c0-2, c0-1 (left) and c1, c1+1 (right) into sGA.  False when a
// wait timed out.
template <int NU>
__device__ __forceinline__ bool seg_exchange(const BatchArgs& a, const BatchSeg sg, int wg, unsigned epoch,
                                             const double* X, double* sGA, int* sOk) {
  const int tid = threadIdx.x, nthr = blockDim.x, c0 = sg.c0, c1 = sg.c1;
  double* mine = a.xA + ((size_t)wg * 2 + (epoch & 1)) * kXA;
  for (int t = tid; t < 4 * NU; t += nthr) {
    const int k = t / NU, r = t % NU;
    const int e = k < 2 ? c0 + k : c1 - 4 + k;
    x_put(mine + t, X[(long)e * NU + r]);
  }
  __syncthreads();
  if (tid == 0) x_flag(a.flagA + wg, epoch);
  if (tid < 2) {
    const int nb = tid == 0 ? sg.left : sg.right;
    sOk[tid] = nb < 0 || x_wait(a.flagA + nb, epoch, a.spin_max);
  }
  __syncthreads();
  if (!(sOk[0] && sOk[1])) return false;
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  for (int t = tid; t < 4 * NU; t += nthr) {
    const int k = t / NU, r = t % NU;
    const int nb = k < 2 ? sg.left : sg.right;
    if (nb >= 0) sGA[t] = x_get(a.xA + ((size_t)nb * 2 + (epoch & 1)) * kXA + (k < 2 ? k + 2 : k - 2) * NU + r);
  }
  __syncthreads();
  return true;
}

// The limiter of a segment's state V (or, without the limiter, only the
// exchange of a chunk): V limited, sG the limited (or plain) states of the
// elements just outside the segment.  False when an exchange timed out.
template <int P, int PH>
__device__ __forceinline__ bool seg_sync(const BatchArgs& a, const BatchSeg sg, int wg, unsigned* epoch, int NE,
                                         double* V, double* A, const double* sL2M, const double* sM2L, double* sGA,
                                         double* sG, int* sOk) {
  using B = BatchShape<P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, NU = B::NU;
  const int c0 = sg.c0, c1 = sg.c1, tid = threadIdx.x;
  const bool chunked = sg.nch > 1;
  if (a.limiter) {
    batch_limit_l2m<P, PH>(V, A, c0, c1, sL2M);
    __syncthreads();  // the neighbours' Taylor coefficients
  }
  if (chunked && !seg_exchange<NU>(a, sg, wg, ++*epoch, a.limiter ? A : V, sGA, sOk)) return false;
  if (a.limiter) {
    batch_limit_apply<P, PH>(a, V, A, NE, c0, c1, sM2L, chunked && sg.left >= 0 ? sGA + NU : nullptr,
                             chunked && sg.right >= 0 ? sGA + 2 * NU : nullptr);
    // the two elements outside, as their owners limit them
    if (chunked && tid < 2 * NF) {
      const int side = tid / NF, fc = tid % NF;
      const int e = side ? c1 : c0 - 1;
      if (side ? sg.right >= 0 : sg.left >= 0) {
        int left, right;
        const int physical = batch_physical(a, e, NE, &left, &right);
        const double* gc = sGA + (side ? 2 : 1) * NU + fc * Ns;
        const double* pl = side ? A + ((c1 - 1) * NF + fc) * Ns : sGA + fc * Ns;
        const double* pr = side ? sGA + 3 * NU + fc * Ns : A + (c0 * NF + fc) * Ns;
        double al[Ns], ac[Ns], ar[Ns], y[Ns];
#pragma unroll
        for (int n = 0; n < Ns; ++n) {
          ac[n] = gc[n];
          al[n] = physical == 0 ? pl[n] : 0.0;
          ar[n] = physical == 0 ? pr[n] : 0.0;
        }
        hr_column<P>(physical, fc, al, ac, ar, sM2L, y);
#pragma unroll
        for (int n = 0; n < Ns; ++n) sG[side * NU + fc * Ns + n] = y[n];
      }
    }
  } else if (chunked) {
    for (int t = tid; t < 2 * NU; t += blockDim.x) sG[t] = sGA[(t < NU ? 1 : 2) * NU + t % NU];
  }
  __syncthreads();
  return true;
}

// One workgroup's segment (order P): every step, stage and limiter
// application, separated by workgroup barriers, with the chunk exchange when
// the segment is a chunk of a problem.  Returns false when an exchange timed
// out.
template <int P, int PH, int FL>
__device__ __forceinline__ bool batch_segment(const BatchArgs& a, const BatchSeg sg, int wg) {
  using B = BatchShape<P, PH>;
  using V = Vol<1, P, PH>;
  constexpr int NF = B::NF, Ns = B::Ns, NG = B::NG, NU = B::NU, NQ = V::NQ;
  static_assert(4 * NU <= kXA, "exchange slot");
  __shared__ double sPhi[V::NTAB_PHI], sDphi[PH ? V::NTAB_D : 1], sDw[V::NTAB_D], sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns], sL2M[Ns * Ns], sM2L[Ns * Ns];
  __shared__ double sRed[16];
  __shared__ double sGA[4 * NU];  // exchanged elements c0-2, c0-1, c1, c1+1 (Taylor coefficients or states)
  __shared__ double sG[2 * NU];   // limited (or plain) states of elements c0-1 and c1
  __shared__ int sOk[2];
  const VolTables<1, P, PH> tb{sPhi, sDphi, sDw, sPw, sMinv};
  const BatchTables& bt = a.tab[P - 1];

  const int pid = sg.pid, c0 = sg.c0, c1 = sg.c1;
  const bool chunked = sg.nch > 1, first = c0 == 0;
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
  if (a.status[pid] != 0) return true;

  unsigned epoch = 0;  // this chunk's exchanges in this launch (every chunk of the problem counts alike)
  // a timed-out exchange marks the problem and ends the workgroup's segment
#define DGX_SEG_SYNC(V)                                                                                   \
  if (!seg_sync<P, PH>(a, sg, wg, &epoch, NE, V, A, sL2M, sM2L, sGA, sG, sOk)) {                          \
    if (tid == 0) a.status[pid] = 2;                                                                      \
    return false;                                                                                         \
  }
  if (a.limiter && a.limit_initial) {
    DGX_SEG_SYNC(U)
  } else if (chunked) {
    // the states just outside, from U as the launch finds it: no chunk
    // writes U before every chunk has passed the first step's Dt gather
    for (int t = tid; t < 2 * NU; t += nthr) {
      const int side = t / NU;
      if (side ? sg.right >= 0 : sg.left >= 0) sG[t] = U[(long)(side ? c1 : c0 - 1) * NU + t % NU];
    }
    __syncthreads();
  }

  for (int step = 0; step < a.nsteps; ++step) {
    // ---- Dt = cfl / max over nodes of |u|+a (rk.cc:250-272); KD nodes per
    // thread at a time, loads issued together; max is exact in any order
    double m = 0.0;
    constexpr int KD = 4;
    for (int n0 = c0 * Ns + tid; n0 < c1 * Ns; n0 += KD * nthr) {
      double u[KD][NF];
#pragma unroll
      for (int k = 0; k < KD; ++k) {
        const int n = n0 + k * nthr;
        const int e = n / Ns, i = n % Ns;
#pragma unroll
        for (int f = 0; f < NF; ++f) u[k][f] = n < c1 * Ns ? U[(e * NF + f) * Ns + i] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < KD; ++k)
        if (n0 + k * nthr < c1 * Ns) m = nanmax(m, upa<1, PH>(u[k]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) sRed[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      double mx = sRed[0];
      for (int k = 1; k < (nthr + 63) / 64; ++k) mx = nanmax(mx, sRed[k]);
      sRed[0] = mx;
    }
    __syncthreads();
    if (chunked) {
      // gather the chunks' maxima of this step
      if (tid == 0) {
        x_put(a.xD + (size_t)wg * 2 + (step & 1), sRed[0]);
        x_flag(a.flagD + wg, (unsigned)step + 1);
      }
      if (tid == 0) sOk[0] = 1;
      __syncthreads();
      for (int k = tid; k < sg.nch; k += nthr)
        if (!x_wait(a.flagD + sg.wg0 + k, (unsigned)step + 1, a.spin_max)) sOk[0] = 0;
      __syncthreads();
      if (!sOk[0]) {
        if (tid == 0) a.status[pid] = 2;
        return false;
      }
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      if (tid == 0) {
        double mx = sRed[0];
        for (int k = 0; k < sg.nch; ++k) mx = nanmax(mx, x_get(a.xD + (size_t)(sg.wg0 + k) * 2 + (step & 1)));
        sRed[1] = mx;
      }
      __syncthreads();
    } else if (tid == 0) {
      sRed[1] = sRed[0];
    }
    if (tid == 0) sRed[0] = cfl / sRed[1];
    __syncthreads();
    const double dt = sRed[0];
    __syncthreads();  // sRed is reused by the next step
    if (!(dt >= 1e-14)) {
      if (tid == 0 && first) a.status[pid] = 1;
      return true;
    }
    if (tid == 0 && first) a.T[pid] = a.T[pid] + dt;

    double* cur = U;
    for (int k = 0; k < 4; ++k) {
      if (k > 0 && (a.limiter || chunked)) {
        DGX_SEG_SYNC(cur)
      }
      // ---- faces j in [c0, c1]: interface j between elements j-1 and j;
      // boundary faces contribute to the boundary element only
      // (dg_functions.cc:225-233); at a chunk end the element outside comes
      // from sG.  KF faces per thread at a time, their loads issued together.
      constexpr int KF = 2;
      for (int j0 = c0 + tid; j0 <= c1; j0 += KF * nthr) {
        double ul[KF][NF], ur[KF][NF];
#pragma unroll
        for (int q = 0; q < KF; ++q) {
          const int j = j0 + q * nthr;
          const bool in = j <= c1, lb = j == 0, rb = j == NE;
          const int ea = lb ? 0 : j - 1, na = lb ? 0 : 1;  // side 0: element and node
          const int eb = (lb || rb) ? ea : j, nb = (lb || rb) ? na : 0;
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            ul[q][f] = !in ? 0.0 : (ea < c0 ? sG[f * Ns + 1] : cur[(ea * NF + f) * Ns + na]);
            ur[q][f] = !in ? 0.0 : (eb >= c1 ? sG[NU + f * Ns] : cur[(eb * NF + f) * Ns + nb]);
          }
        }
#pragma unroll
        for (int q = 0; q < KF; ++q) {
          const int j = j0 + q * nthr;
          if (j > c1) break;
          const bool lb = j == 0, rb = j == NE;
          const int ea = lb ? 0 : j - 1, na = lb ? 0 : 1;
          const int eb = (lb || rb) ? ea : j, nb = (lb || rb) ? na : 0;
          const int tag = lb ? a.bc[0] : (rb ? a.bc[1] : 0);
          double F[NF], nc[NF];
          if (tag == BC_REFLECTIVE) ur[q][1] = -ur[q][1];  // boundaries.cu:20-35
          riemann<1, PH, FL>(ul[q], ur[q], nrm[j], 0.0, F, nc);
#pragma unroll
          for (int f = 0; f < NF; ++f) {
            if (ea >= c0) Qe[(ea * 2 + na) * NF + f] = psiw * (-F[f] + nc[f]);
            if (!lb && !rb && eb < c1) Qe[(eb * 2 + nb) * NF + f] = psiw * (F[f] + nc[f]);
          }
        }
      }
      __syncthreads();
      // ---- elements: volume terms, face terms, Dt*Minv and the RK4 update.
      // The element's RK operands are loaded half the fields at a time, each
      // batch issued together (a store to `out`, which may be U, would
      // otherwise order every later load after it).
      constexpr int FB = (NF + 1) / 2;  // fields per operand batch
      double* out = (k == 3) ? U : S;
      for (int e = c0 + tid; e < c1; e += nthr) {
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
          double pun[FB * Ns], pacc[FB * Ns];
#pragma unroll
          for (int r = 0; r < FB * Ns; ++r) {
            const long idx = (long)e * NU + f0 * Ns + r;
            const bool in = f0 * Ns + r < NU;
            pun[r] = (in && k < 3) ? U[idx] : 0.0;
            pacc[r] = (in && k > 0) ? Uacc[idx] : 0.0;
          }
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
      }
      __syncthreads();
      cur = out;
    }
    if (a.limiter || chunked) {
      DGX_SEG_SYNC(U)
    }
  }
#undef DGX_SEG_SYNC
  return true;
}

// LB: the launch bound (512, or 1024 with four waves on every SIMD of the
// CU), and W the waves per SIMD the registers must allow: 256-thread
// workgroups are built for three (168 registers, a few spilled), so that
// three of them share a CU where the unbounded build's 238 registers allow
// two.  The arguments live in device memory: a by-value argument block of
// this size was copied to scratch, and its pointers took the registers of
// the element pass.
#ifndef DGX_BATCH_W256
#define DGX_BATCH_W256 3
#endif
template <int PH, int FL, int LB, int W = 1>
__global__ __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(W))) void k_batch(
    const BatchArgs* __restrict__ pa) {
  const BatchArgs& a = *pa;
  const int wg = blockIdx.x;
  for (int k = a.wg_seg[wg]; k < a.wg_seg[wg + 1]; ++k) {
    const BatchSeg sg = a.seg[k];
    bool ok;
    switch (a.order[sg.pid]) {
      case 1: ok = batch_segment<1, PH, FL>(a, sg, wg); break;
      case 2: ok = batch_segment<2, PH, FL>(a, sg, wg); break;
      default: ok = batch_segment<3, PH, FL>(a, sg, wg); break;
    }
    if (!ok) return;  // a timed-out chunk: its problem is marked, the workgroup ends
  }
}

__global__ void k_zero_flags(unsigned* f, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) f[k] = 0u;
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
  int block = 512;      // workgroup size
  int min_chunk = 64;   // elements per chunk at least (0: no problem is cut)
  // the partition into workgroup segments, for part_block and part_min_chunk
  int part_block = -1, part_min_chunk = -1, G = 0, nsplit = 0;
  BatchSeg* d_seg = nullptr;
  int* d_wg_seg = nullptr;
  double *xA = nullptr, *xD = nullptr;
  unsigned* flags = nullptr;  // flagA [G], then flagD [G]
  std::vector<void*> part_allocs;
  BatchArgs* d_args = nullptr;  // the kernel's arguments (device memory)
  int *d_ne = nullptr, *d_order = nullptr;
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

void free_partition(dgx_batch* b) {
  for (void* p : b->part_allocs) (void)hipFree(p);
  b->part_allocs.clear();
  b->d_seg = nullptr;
  b->d_wg_seg = nullptr;
  b->xA = b->xD = nullptr;
  b->flags = nullptr;
  b->part_block = -1;
}

void free_batch(dgx_batch* b) {
  drop_graph(b);
  free_partition(b);
  if (b->cs) (void)hipStreamDestroy(b->cs);
  b->cs = nullptr;
  for (void* p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
}

dgx_status bfail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

template <typename T>
bool palloc(dgx_batch* b, T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  if (!hip_ok(hipMalloc(&p, n * sizeof(T)), "hipMalloc")) return false;
  b->part_allocs.push_back(p);
  *dst = static_cast<T*>(p);
  if (src) return hip_ok(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  return true;
}

// Workgroups of `block` threads that the device holds at once.
int resident_workgroups(dgx_batch* b, int block) {
  int per_cu = 0, dev = 0, cus = 0;
  const hipError_t e = dispatch_batch(b->cfg, [&](auto ph, auto fl) -> hipError_t {
    constexpr int PH = decltype(ph)::value, FL = decltype(fl)::value;
    const void* k = block <= 256   ? reinterpret_cast<const void*>(&k_batch<PH, FL, 256, DGX_BATCH_W256>)
                    : block <= 512 ? reinterpret_cast<const void*>(&k_batch<PH, FL, 512>)
                                   : reinterpret_cast<const void*>(&k_batch<PH, FL, 1024>);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, 0);
  });
  if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  return per_cu * cus;
}

// Work per element and step by order, relative to order 1 (per-workgroup
// times of the 256-problem two-fluid ensemble on MI355X: about 1 : 1.75 :
// 2.7; the flux and limiter work per element does not grow as (p+1)^2 at
// these sizes, where latency sets the time).
// (overridable in tuning builds; at 128-thread workgroups 1.5 : 2.2 measured the
// same, 2.0 : 3.2 and 1.75 : 3.5 4-5 % slower)
#ifndef DGX_BATCH_COST2
#define DGX_BATCH_COST2 1.75
#endif
#ifndef DGX_BATCH_COST3
#define DGX_BATCH_COST3 2.7
#endif
constexpr double kOrderCost[3] = {1.0, DGX_BATCH_COST2, DGX_BATCH_COST3};

// The segments of every workgroup for workgroups of `block` threads.  Work
// of a problem: N_E kOrderCost[p-1].  With R workgroups resident at once
// and the average load L (a little above W/R until the cut problems leave
// room), a problem of more than L and at least two chunks' elements is cut
// into min(ceil(w/L), N_E/min_chunk) chunks on consecutive workgroups, the
// first ones of the launch, so that all of them are resident together (at
// most R-1 of them); every other problem has a workgroup of its own after
// them, largest first, which the device starts as soon as a slot is free.
// (Counting rounds of one element per thread instead of elements, and
// packing the uncut problems into static bins, both measured slower.)
bool make_partition(dgx_batch* b, int block) {
  if (b->part_block == block && b->part_min_chunk == b->min_chunk && b->d_seg) return true;
  drop_graph(b);
  free_partition(b);
  const int n = b->n;
  std::vector<double> w(n);
  double W = 0;
  for (int k = 0; k < n; ++k) W += (w[k] = (double)b->ne[k] * kOrderCost[b->order[k] - 1]);
  std::vector<int> chunks(n, 1);
  const int R = b->min_chunk > 0 ? resident_workgroups(b, block) : 0;
  bool split = false;
  if (R > 1 && b->min_chunk > 0) {
    for (double f = 1.0; f < 4.0 && !split; f *= 1.1) {
      const double L = f * W / R;
      long S = 0;
      double Wsmall = 0;
      for (int k = 0; k < n; ++k) {
        const long c = std::min((long)std::ceil(w[k] / L), (long)(b->ne[k] / b->min_chunk));
        chunks[k] = (w[k] > L && c >= 2) ? (int)c : 1;
        if (chunks[k] > 1) S += chunks[k];
        else Wsmall += w[k];
      }
      if (S > 0 && S < R && Wsmall <= (R - S) * L) split = true;
    }
    if (!split) std::fill(chunks.begin(), chunks.end(), 1);
  }
  std::vector<BatchSeg> seg;
  std::vector<int> wg_seg{0};
  int nsplit = 0, wg = 0;
  std::vector<int> whole;
  for (int k = 0; k < n; ++k) {
    if (chunks[k] == 1) {
      whole.push_back(k);
      continue;
    }
    const int c = chunks[k], ne = b->ne[k];
    ++nsplit;
    for (int i = 0; i < c; ++i) {
      seg.push_back(BatchSeg{k, (int)((long)i * ne / c), (int)((long)(i + 1) * ne / c), i > 0 ? wg + i - 1 : -1,
                             i < c - 1 ? wg + i + 1 : -1, wg, c});
      wg_seg.push_back((int)seg.size());
    }
    wg += c;
  }
  std::stable_sort(whole.begin(), whole.end(), [&](int x, int y) { return w[x] > w[y]; });
  for (int k : whole) {
    seg.push_back(BatchSeg{k, 0, b->ne[k], -1, -1, wg++, 1});
    wg_seg.push_back((int)seg.size());
  }
  const int G = (int)wg_seg.size() - 1;
  const bool ok = palloc(b, &b->d_seg, seg.data(), seg.size()) && palloc(b, &b->d_wg_seg, wg_seg.data(), G + 1) &&
                  palloc<double>(b, &b->xA, nullptr, (size_t)G * 2 * kXA) &&
                  palloc<double>(b, &b->xD, nullptr, (size_t)G * 2) &&
                  palloc<unsigned>(b, &b->flags, nullptr, (size_t)2 * G);
  if (!ok) {
    free_partition(b);
    return false;
  }
  b->G = G;
  b->nsplit = nsplit;
  b->part_block = block;
  b->part_min_chunk = b->min_chunk;
  return true;
}

BatchArgs make_args(dgx_batch* b, double* U, int nsteps, int limiter, int limit_initial, double* T, int* status) {
  BatchArgs a{};
  for (int q = 0; q < 3; ++q) a.tab[q] = b->tab[q];
  a.seg = b->d_seg; a.wg_seg = b->d_wg_seg; a.order = b->d_order;
  a.n_elem = b->d_ne; a.state_off = b->d_state_off; a.elem_off = b->d_elem_off;
  a.J = b->d_J; a.ij = b->d_ij; a.normals = b->d_normals; a.minv = b->d_minv; a.J0 = b->d_J0; a.cfl = b->d_cfl;
  a.bc[0] = b->bc[0]; a.bc[1] = b->bc[1];
  a.U = U; a.Uacc = b->Uacc; a.S = b->S; a.A = b->A; a.Qe = b->Qe; a.T = T; a.status = status;
  a.xA = b->xA; a.xD = b->xD; a.flagA = b->flags; a.flagD = b->flags + b->G;
  a.spin_max = 200000;
  a.nsteps = nsteps; a.limiter = limiter; a.limit_initial = limit_initial;
  return a;
}

// The kernel's arguments into device memory, once the stream's earlier
// launches (which read the previous ones) are done.
bool upload_args(dgx_batch* b, const BatchArgs& a, hipStream_t st) {
  return hip_ok(hipStreamSynchronize(st), "batch args") &&
         hip_ok(hipMemcpy(b->d_args, &a, sizeof(BatchArgs), hipMemcpyHostToDevice), "batch args");
}

// the exchange flags cleared (a kernel, not a memset: the launch is captured
// in a graph), then every workgroup's segments, with the uploaded arguments
hipError_t launch_batch(dgx_batch* b, int block, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_flags, dim3((2 * b->G + 255) / 256), dim3(256), 0, st, b->flags, 2 * b->G);
  return dispatch_batch(b->cfg, [&](auto ph, auto fl) -> hipError_t {
    constexpr int PH = decltype(ph)::value, FL = decltype(fl)::value;
    const BatchArgs* pa = b->d_args;
    if (block <= 256) hipLaunchKernelGGL((k_batch<PH, FL, 256, DGX_BATCH_W256>), dim3(b->G), dim3(block), 0, st, pa);
    else if (block <= 512) hipLaunchKernelGGL((k_batch<PH, FL, 512>), dim3(b->G), dim3(block), 0, st, pa);
    else hipLaunchKernelGGL((k_batch<PH, FL, 1024>), dim3(b->G), dim3(block), 0, st, pa);
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
  if (ok) ok = bupload(b, &b->d_minv, minv.data(), minv.size()) && bupload<BatchArgs>(b, &b->d_args, nullptr, 1);
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

dgx_status dgx_batch_block_sizes(dgx_batch* b, int* sizes) {
  if (!b || !sizes) return bfail(DGX_EINVAL, "dgx_batch_block_sizes: null argument");
  if (!make_partition(b, b->block)) return DGX_EDEVICE;
  sizes[0] = b->block;
  sizes[1] = b->G;
  sizes[2] = b->nsplit;
  return DGX_OK;
}

dgx_status dgx_batch_set_min_chunk(dgx_batch* b, int min_chunk) {
  if (!b) return bfail(DGX_EINVAL, "dgx_batch_set_min_chunk: null argument");
  if (min_chunk != 0 && min_chunk < 8) return bfail(DGX_EINVAL, "dgx_batch_set_min_chunk: 0 or at least 8 elements");
  b->min_chunk = min_chunk;
  drop_graph(b);
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
  const int sizes[5] = {64, 128, 256, 512, 1024};
  float best = 1e30f;
  int best_b = b->block;
  dgx_status r = DGX_OK;
  // per size: a warm-up run, then the fastest of two timed runs of kTuneSteps
  // steps (shorter trials picked 512 over 256 threads on the 256-problem
  // ensemble, which is 10 % slower over 20 steps)
  constexpr int kTuneSteps = 10;
  for (int bs : sizes) {
    float tot = 1e30f;
    for (int rep = 0; rep < 3 && r == DGX_OK; ++rep) {
      // scratch copies: the caller's state and times are left untouched
      if (!hip_ok(hipMemcpyAsync(b->tuneU, U, sizeof(double) * b->state_off[b->n], hipMemcpyDeviceToDevice, st),
                  "autotune copy")) { r = DGX_EDEVICE; break; }
      hipLaunchKernelGGL(k_fill_zero_batch, dim3((b->n + 255) / 256), dim3(256), 0, st, b->tuneT, b->tune_status,
                         b->n);
      if (!make_partition(b, bs)) { r = DGX_EDEVICE; break; }
      if (!upload_args(b, make_args(b, b->tuneU, kTuneSteps, limiter, 0, b->tuneT, b->tune_status), st)) {
        r = DGX_EDEVICE;
        break;
      }
      (void)hipEventRecord(t0, st);
      if (!hip_ok(launch_batch(b, bs, st), "autotune launch")) { r = DGX_EDEVICE; break; }
      (void)hipEventRecord(t1, st);
      if (!hip_ok(hipEventSynchronize(t1), "autotune sync")) { r = DGX_EDEVICE; break; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, t0, t1);
      if (rep > 0) tot = std::min(tot, ms);  // the first run warms the code object and caches
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
    // the flag reset and the kernel, captured on the internal stream
    if (!make_partition(b, b->block) || !upload_args(b, make_args(b, U, nsteps, limiter, limit_initial, T, status), st))
      return DGX_EDEVICE;
    if (!hip_ok(hipStreamBeginCapture(b->cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return DGX_EDEVICE;
    const bool ok = hip_ok(launch_batch(b, b->block, b->cs), "batched launch");
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
