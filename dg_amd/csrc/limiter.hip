// Hierarchical-reconstruction limiter and troubled-cell sensor kernels.
//
// Compiled without FMA contraction: the limiter branches on signs and
// magnitudes (minmod, limiting_kernels.cu:1470-1486) and the sensor flags are
// compared bit for bit, so these kernels reproduce the CPU path's rounding
// exactly for identical inputs.
//
// Global HR (Limiting::HRlimiting, limiting.cc:11-50):
//   A = L2M * U (k_apply), per-slice limiting (k_hrl = hrl1D,
//   limiting_kernels.cu:245-319), U = M2L * Alim (k_apply); in 2D the x sweep
//   is followed by MonoX2MonoY and the y sweep.
// Sensor (SENSOR::sensing, sensor.cc:11-36 + calc_sensors,
//   sensor_kernels.cu:19-163): the reference loops over elements in order and
//   each element skips the test if an earlier element already flagged it,
//   then overwrites its flagged neighbour.  The result depends on that order.
//   Here it is reproduced deterministically in three phases:
//     k_sensor_hit    per element, the first neighbour test that fires
//                     (independent of order);
//     k_sensor_relax  "active" = not flagged by an earlier active element,
//                     iterated to its fixed point (a DAG in element order);
//     k_sensor_final  each element takes the value of its last writer in
//                     element order.
// Sensor-gated HR (Limiting::HRIlimiting, limiting.cc:596-634; hri1D,
//   limiting_kernels.cu:322-427; copy_detected, sensor_kernels.cu:167-200).
#pragma clang fp contract(off)
#include <algorithm>
#include <type_traits>
#include <utility>

#include "dgx_internal.h"
#include "hr.cuh"
#include "hr_tables.h"
#include "physics.cuh"
#include "sensor.cuh"

namespace dgx {

// Out[:,col] = M * In[:,col] for every column (e,f), netlib dgemm order.
// gate: when set and nonzero, the launch does nothing (a step skipped by the
// device-side time control).
template <int Ns>
__global__ __launch_bounds__(256) void k_apply(const double* __restrict__ M, const double* __restrict__ In,
                                               double* __restrict__ Out, long ncol, const int* gate) {
  if (gate && *gate) return;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  double x[Ns], y[Ns];
#pragma unroll
  for (int l = 0; l < Ns; ++l) x[l] = In[col * Ns + l];
#pragma unroll
  for (int r = 0; r < Ns; ++r) y[r] = 0.0;
#pragma unroll
  for (int l = 0; l < Ns; ++l)
#pragma unroll
    for (int r = 0; r < Ns; ++r) y[r] = y[r] + x[l] * M[r + l * Ns];
#pragma unroll
  for (int r = 0; r < Ns; ++r) Out[col * Ns + r] = y[r];
}

// hrl1D: one thread per (element, slice, field); N = polynomial order,
// slices of N+1 coefficients.  ns_slice = N+1, slicenum = 1 (1D) or N+1 (2D).
template <int N>
__global__ __launch_bounds__(256) void k_hrl(int N_E, int NF, int N_N, int slicenum, const int* __restrict__ nb,
                                             int offxy, const double* __restrict__ A, double* __restrict__ Alim) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)N_E * slicenum * NF;
  if (tid >= total) return;
  const int fc = (int)(tid % NF);
  const int slice = (int)((tid / NF) % slicenum);
  const int e = (int)(tid / ((long)NF * slicenum));
  constexpr int Nsl = N + 1;
  const int left = nb[e * N_N + offxy + 0], right = nb[e * N_N + offxy + 1];
  int physical = 0;
  if (left < 0) physical = -left;
  if (right < 0) physical = -right;
  const long idx = ((long)e * NF + fc) * Nsl * slicenum + slice * Nsl;
  if (physical == 4) {
    // unchanged (reference leaves Alim as it was)
  } else if (physical == 2) {
    average_monomial_d<N>(&A[idx], &Alim[idx]);
  } else if (physical == 3) {
    int un_idx = (offxy == 0) ? 1 : (offxy == 2 ? 2 : 0);
    if (fc != un_idx)
      average_monomial_d<N>(&A[idx], &Alim[idx]);
    else
      for (int n = 0; n <= N; n++) Alim[idx + n] = A[idx + n];
  } else {
    const long idxL = ((long)left * NF + fc) * Nsl * slicenum + slice * Nsl;
    const long idxR = ((long)right * NF + fc) * Nsl * slicenum + slice * Nsl;
    double al[Nsl], ac[Nsl], ar[Nsl], lim[Nsl];
#pragma unroll
    for (int n = 0; n < Nsl; ++n) { al[n] = A[idxL + n]; ac[n] = A[idx + n]; ar[n] = A[idxR + n]; lim[n] = 0; }
    limit_monomial_d<N>(al, ac, ar, lim);
#pragma unroll
    for (int n = 0; n < Nsl; ++n) Alim[idx + n] = lim[n];
  }
}

// ------------------------------------------------- fused global HR sweeps
// The same arithmetic as k_apply + k_hrl + k_apply, per column (element,
// field), without the intermediate full-state passes: a sweep kernel forms
// the Taylor coefficients of its own column and of the two neighbour columns
// it needs, limits every slice and applies the next transform.

// y = M x in netlib dgemm order
template <int Ns>
__device__ inline void col_apply(const double* __restrict__ M, const double* x, double* y) {
#pragma unroll
  for (int r = 0; r < Ns; ++r) y[r] = 0.0;
#pragma unroll
  for (int l = 0; l < Ns; ++l)
#pragma unroll
    for (int r = 0; r < Ns; ++r) y[r] = y[r] + x[l] * M[r + l * Ns];
}

// y = M x for one of the constant tables of hr_tables.h (WHICH: 0 Lag2MonoX,
// 1 MonoX2MonoY, 2 MonoY2Lag) in col_apply's order.  The loops run over
// integer sequences so that every entry is a constant expression: the zero
// entries are left out, which keeps col_apply's bits (see skip_zero_terms in
// hr.cuh), and the multiplications by +-1 fold away.
template <int P, int WHICH>
__host__ __device__ constexpr double hr_entry(int i) {
  return WHICH == 0 ? HrTables2D<P>::L2MX[i] : (WHICH == 1 ? HrTables2D<P>::MX2MY[i] : HrTables2D<P>::MY2L[i]);
}

template <int P, int WHICH, int L, int... R>
__device__ inline void col_apply_c_col(const double* x, double* y, std::integer_sequence<int, R...>) {
  constexpr int Ns = HrTables2D<P>::Ns;
  auto term = [&](auto r) {
    constexpr int rr = decltype(r)::value;
    constexpr double m = hr_entry<P, WHICH>(rr + L * Ns);
    if constexpr (m != 0.0) y[rr] = y[rr] + x[L] * m;
  };
  (term(std::integral_constant<int, R>{}), ...);
}

template <int P, int WHICH, int... L>
__device__ inline void col_apply_c_all(const double* x, double* y, std::integer_sequence<int, L...>) {
  constexpr int Ns = HrTables2D<P>::Ns;
  (col_apply_c_col<P, WHICH, L>(x, y, std::make_integer_sequence<int, Ns>{}), ...);
}

template <int P, int WHICH>
__device__ inline void col_apply_c(const double* x, double* y) {
  constexpr int Ns = HrTables2D<P>::Ns;
#pragma unroll
  for (int r = 0; r < Ns; ++r) y[r] = 0.0;
  col_apply_c_all<P, WHICH>(x, y, std::make_integer_sequence<int, Ns>{});
}

// hrl1D for every slice of one column (limiting_kernels.cu:245-319)
template <int N, int SL>
__device__ inline void hrl_column(int physical, int fc, int offxy, const double* AL, const double* AC,
                                  const double* AR, double* Alim) {
  constexpr int Nsl = N + 1;
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) {
    const int o = sl * Nsl;
    if (physical == 2 || (physical == 3 && fc != ((offxy == 0) ? 1 : 2))) {
      average_monomial_d<N>(AC + o, Alim + o);
    } else if (physical == 3 || physical == 4) {  // wall-normal velocity (3) / gravity tag (4): unchanged
#pragma unroll
      for (int n = 0; n < Nsl; ++n) Alim[o + n] = AC[o + n];
    } else {
      double lim[Nsl];
#pragma unroll
      for (int n = 0; n < Nsl; ++n) lim[n] = 0.0;
      limit_monomial_d<N>(AL + o, AC + o, AR + o, lim);
#pragma unroll
      for (int n = 0; n < Nsl; ++n) Alim[o + n] = lim[n];
    }
  }
}

__device__ inline int hr_physical(int left, int right) {
  int physical = 0;
  if (left < 0) physical = -left;
  if (right < 0) physical = -right;
  return physical;
}

// first sweep: Out = Mnext * hrl(Mfirst * U) along offxy (x in 2D; the only
// sweep in 1D, where Mnext = Mono2Lag)
template <int Ns, int N, int SL>
__global__ __launch_bounds__(256) void k_hr_sweep(int N_E, int NF, int N_N, const int* __restrict__ nb, int offxy,
                                                  const double* __restrict__ Mfirst, const double* __restrict__ Mnext,
                                                  const double* __restrict__ In, double* __restrict__ Out,
                                                  const int* gate) {
  if (gate && *gate) return;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (long)N_E * NF) return;
  const int e = (int)(col / NF), fc = (int)(col % NF);
  const int left = nb[e * N_N + offxy], right = nb[e * N_N + offxy + 1];
  const int physical = hr_physical(left, right);
  double x[Ns], AC[Ns], AL[Ns], AR[Ns], Alim[Ns];
#pragma unroll
  for (int l = 0; l < Ns; ++l) x[l] = In[col * Ns + l];
  if (Mfirst) col_apply<Ns>(Mfirst, x, AC); else for (int l = 0; l < Ns; ++l) AC[l] = x[l];
  if (physical == 0) {
    const double* pl = In + ((long)left * NF + fc) * Ns;
    const double* pr = In + ((long)right * NF + fc) * Ns;
#pragma unroll
    for (int l = 0; l < Ns; ++l) x[l] = pl[l];
    if (Mfirst) col_apply<Ns>(Mfirst, x, AL); else for (int l = 0; l < Ns; ++l) AL[l] = x[l];
#pragma unroll
    for (int l = 0; l < Ns; ++l) x[l] = pr[l];
    if (Mfirst) col_apply<Ns>(Mfirst, x, AR); else for (int l = 0; l < Ns; ++l) AR[l] = x[l];
  }
  hrl_column<N, SL>(physical, fc, offxy, AL, AC, AR, Alim);
  col_apply<Ns>(Mnext, Alim, x);
#pragma unroll
  for (int r = 0; r < Ns; ++r) Out[col * Ns + r] = x[r];
}

// k_hr_sweep in 2D for p <= 2 with the transforms of hr_tables.h as
// compile-time constants (the host has found the solver's tables equal to
// them, Solver::hr_const): YS 0 is the x sweep (Lag2MonoX, then MonoX2MonoY),
// YS 1 the y sweep (MonoY2Lag).  The same bits as k_hr_sweep for finite
// states (the skipped zero terms do not spread an Inf or NaN through a column).
template <int P, int YS>
__global__ __launch_bounds__(256) void k_hr_sweep_c(int N_E, int NF, int N_N, const int* __restrict__ nb,
                                                    const double* __restrict__ In, double* __restrict__ Out,
                                                    const int* gate) {
  if (gate && *gate) return;
  constexpr int Ns = HrTables2D<P>::Ns, SL = P + 1, offxy = YS ? 2 : 0;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (long)N_E * NF) return;
  const int e = (int)(col / NF), fc = (int)(col % NF);
  const int left = nb[e * N_N + offxy], right = nb[e * N_N + offxy + 1];
  const int physical = hr_physical(left, right);
  auto first = [&](const double* src, double* A) {
    double x[Ns];
#pragma unroll
    for (int l = 0; l < Ns; ++l) x[l] = src[l];
    if constexpr (YS) {
#pragma unroll
      for (int l = 0; l < Ns; ++l) A[l] = x[l];
    } else {
      col_apply_c<P, 0>(x, A);
    }
  };
  double AC[Ns], AL[Ns], AR[Ns], Alim[Ns], y[Ns];
  first(In + col * Ns, AC);
  if (physical == 0) {
    first(In + ((long)left * NF + fc) * Ns, AL);
    first(In + ((long)right * NF + fc) * Ns, AR);
  }
  hrl_column<P, SL>(physical, fc, offxy, AL, AC, AR, Alim);
  col_apply_c<P, YS ? 2 : 1>(Alim, y);
#pragma unroll
  for (int r = 0; r < Ns; ++r) Out[col * Ns + r] = y[r];
}

// Both sweeps of the 2D global HR limiter in one pass, out of place, on a
// Cartesian grid (element e = ey*nx + ex) for a tile of TXE x TYE elements:
//   1. the nodal values of every element the tile needs, with coalesced loads
//      (an element's NF*Ns values are contiguous), into LDS: the tile, the
//      rows just below and above it, and the element left and right of each
//      of these rows (all found through the neighbour table, so periodic ends
//      and walls need no special case);
//   2. their Taylor-x coefficients (Lag2MonoX), in place, one column per thread;
//      the three transforms are the compile-time tables of hr_tables.h, which
//      the host has found equal to the solver's (Solver::hr2d);
//   3. the x sweep (hrl along x, MonoX2MonoY) of the tile rows and of the two
//      halo rows, from LDS into LDS;
//   4. the y sweep (hrl along y, MonoY2Lag) of the tile rows, staged in LDS
//      and stored row by row (a tile row is TXE*NF*Ns contiguous values).
// Every column of a finite state gets the bits of the two k_hr_sweep
// launches (the products with zero table entries that it leaves out change
// no finite sum; an Inf or NaN is not spread through its column), and the
// intermediate state never reaches HBM.  In and Out must differ:
// neighbouring tiles read In's rows.  At least six waves per SIMD (three
// 512-thread blocks of the 2D p=2 five-field limiter per CU).
template <int Ns, int N, int SL, int NF, int TXE, int TYE, int BS, bool OUT_VEC2>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(6))) void k_hr2d(
    int nx, int N_N, const int* __restrict__ nb, const double* __restrict__ In, double* __restrict__ Out,
    const int* gate) {
  if (gate && *gate) return;
  constexpr int ROWS = TYE + 2, W = TXE + 2, EL = NF * Ns, ROW_OUT = TXE * EL;
  // one x-sweep column per thread, so that its result can wait in registers
  // while the Taylor-x coefficients are still read, then take their place
  static_assert(ROWS * TXE * NF <= BS, "one x-sweep column per thread");
  static_assert(ROW_OUT % 2 == 0, "rows stored as pairs of doubles");
  // nodal values, then Taylor-x [row][col][f]; then Taylor-y [row][i][f]; then the output [j][i][f]
  __shared__ __align__(16) double sA[ROWS * W * EL];
  __shared__ int sE[ROWS * W];  // element of each slot, -1 if absent
  const int tiles_x = nx / TXE;
  const int ex0 = (blockIdx.x % tiles_x) * TXE, ey0 = (blockIdx.x / tiles_x) * TYE;
  // slot elements: row 0 below the tile, rows 1..TYE the tile, row TYE+1
  // above; columns 0 and W-1 the left and right neighbours of the row's end
  // elements.  One thread per slot, so that no thread waits for more than the
  // two dependent table reads of a halo row's end.
  static_assert(ROWS * W <= BS, "one thread per slot");
  if (threadIdx.x < ROWS * W) {
    const int c = threadIdx.x % W, r = threadIdx.x / W;
    const int i = c == 0 ? 0 : (c == W - 1 ? TXE - 1 : c - 1);  // the row element it starts from
    int e;
    if (r == 0) e = nb[(ey0 * nx + ex0 + i) * N_N + 2];
    else if (r == ROWS - 1) e = nb[((ey0 + TYE - 1) * nx + ex0 + i) * N_N + 3];
    else e = (ey0 + r - 1) * nx + ex0 + i;
    if (c == 0 || c == W - 1) e = e >= 0 ? nb[e * N_N + (c == 0 ? 0 : 1)] : -1;
    sE[threadIdx.x] = e;
  }
  __syncthreads();
  // every load of the slots in flight at once (a loop that stores each value
  // before it loads the next waits one memory round trip per value), and the
  // neighbour entries the two sweeps test with them
  constexpr int TOT = ROWS * W * EL, PER = (TOT + BS - 1) / BS;
  const int t2 = threadIdx.x;
  const int f2 = t2 % NF, i2 = (t2 / NF) % TXE, r2 = t2 / (NF * TXE);
  const bool on2 = t2 < ROWS * TXE * NF && sE[r2 * W + i2 + 1] >= 0;
  const bool on3 = t2 < TYE * TXE * NF;
  double v[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int q = t2 + k * BS;
    v[k] = 0.0;
    if (q < TOT) {
      const int slot = q / EL;
      const int e = sE[slot];
      if (e >= 0) v[k] = In[(long)e * EL + (q - slot * EL)];
    }
  }
  int nbx0 = 0, nbx1 = 0, nby0 = 0, nby1 = 0;
  if (on2) {
    const int e = sE[r2 * W + i2 + 1];
    nbx0 = nb[e * N_N + 0];
    nbx1 = nb[e * N_N + 1];
  }
  if (on3) {
    const int e = (ey0 + r2) * nx + ex0 + i2;
    nby0 = nb[e * N_N + 2];
    nby1 = nb[e * N_N + 3];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int q = t2 + k * BS;
    if (q < TOT && sE[q / EL] >= 0) sA[q] = v[k];
  }
  // both sweeps' boundary kinds in one register
  const int phys = hr_physical(nbx0, nbx1) | (hr_physical(nby0, nby1) << 16);
  __syncthreads();
  for (int c = threadIdx.x; c < ROWS * W * NF; c += BS) {
    if (sE[c / NF] < 0) continue;
    double x[Ns], A[Ns];
    double* a = sA + c * Ns;
#pragma unroll
    for (int l = 0; l < Ns; ++l) x[l] = a[l];
    col_apply_c<N, 0>(x, A);
#pragma unroll
    for (int l = 0; l < Ns; ++l) a[l] = A[l];
  }
  __syncthreads();
  // x sweep of every row present (a wall below or above leaves its halo row out:
  // the y sweep does not read it)
  double xs[Ns];
  if (on2) {
    const int slot = r2 * W + i2 + 1;
    const int physical = phys & 0xffff;
    double AC[Ns], AL[Ns], AR[Ns], Alim[Ns];
    const double* ac = sA + (slot * NF + f2) * Ns;
#pragma unroll
    for (int l = 0; l < Ns; ++l) AC[l] = ac[l];
    if (physical == 0) {
      const double* al = sA + ((slot - 1) * NF + f2) * Ns;
      const double* ar = sA + ((slot + 1) * NF + f2) * Ns;
#pragma unroll
      for (int l = 0; l < Ns; ++l) { AL[l] = al[l]; AR[l] = ar[l]; }
    }
    hrl_column<N, SL>(physical, f2, 0, AL, AC, AR, Alim);
    col_apply_c<N, 1>(Alim, xs);
  }
  __syncthreads();
  if (on2) {
    double* b = sA + ((r2 * TXE + i2) * NF + f2) * Ns;
#pragma unroll
    for (int l = 0; l < Ns; ++l) b[l] = xs[l];
  }
  __syncthreads();
  // y sweep of the tile rows (TYE*TXE*NF <= BS columns: one pass)
  if (on3) {
    const int f = f2, i = i2, j = r2;
    const int physical = phys >> 16;
    double AC[Ns], AL[Ns], AR[Ns], Alim[Ns];
    const double* bc = sA + (((j + 1) * TXE + i) * NF + f) * Ns;
#pragma unroll
    for (int l = 0; l < Ns; ++l) AC[l] = bc[l];
    if (physical == 0) {
      const double* bl = sA + ((j * TXE + i) * NF + f) * Ns;
      const double* br = sA + (((j + 2) * TXE + i) * NF + f) * Ns;
#pragma unroll
      for (int l = 0; l < Ns; ++l) { AL[l] = bl[l]; AR[l] = br[l]; }
    }
    hrl_column<N, SL>(physical, f, 2, AL, AC, AR, Alim);
    col_apply_c<N, 2>(Alim, xs);
  }
  __syncthreads();
  if (on3) {
    double* o = sA + t2 * Ns;  // [j][i][f]: the element-major order of the tile rows
#pragma unroll
    for (int l = 0; l < Ns; ++l) o[l] = xs[l];
  }
  __syncthreads();
  // pairs of doubles when Out is 16-byte aligned (every row start then is:
  // nx and ex0 are multiples of TXE), else single doubles.  A template
  // parameter, not a run-time branch: the branch cost two VGPRs, one wave
  // per SIMD and 25 % of the time.
  if constexpr (OUT_VEC2) {
    const double2* so = reinterpret_cast<const double2*>(sA);
    for (int q = threadIdx.x; q < TYE * ROW_OUT / 2; q += BS) {
      const int j = q / (ROW_OUT / 2), r = q - j * (ROW_OUT / 2);
      reinterpret_cast<double2*>(Out + ((long)(ey0 + j) * nx + ex0) * EL)[r] = so[q];
    }
  } else {
    for (int q = threadIdx.x; q < TYE * ROW_OUT; q += BS) {
      const int j = q / ROW_OUT, r = q - j * ROW_OUT;
      Out[((long)(ey0 + j) * nx + ex0) * EL + r] = sA[q];
    }
  }
}

// ------------------------------------------------- m2 and primitive limiting
// Limiting::M2limiting (limiting.cc:123-594, two-fluid GAMNCON) and
// PRlimiting (limiting.cc:102-121), launcher by launcher as the reference
// runs them on per-field arrays x[e*Ns+i]: the per-field HR reuses k_apply
// and k_hrl (one field at a time, as the reference calls hrl1D), and these
// kernels restate the rest.
__global__ __launch_bounds__(256) void k_field_get(long n, int Ns, int NF, int f, const double* __restrict__ U,
                                                   double* __restrict__ x) {  // Lstridedcopy U -> field
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = U[((k / Ns) * NF + f) * Ns + k % Ns];
}
__global__ __launch_bounds__(256) void k_field_put(long n, int Ns, int NF, int f, const double* __restrict__ x,
                                                   double* __restrict__ U) {  // field -> U
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) U[((k / Ns) * NF + f) * Ns + k % Ns] = x[k];
}
// pressure (physics.cu:891-960), GAMNCON
template <int D>
__global__ __launch_bounds__(256) void k_m2_pressure(long n, int Ns, const double* __restrict__ U, double* __restrict__ p) {
  constexpr int NF = 2 + D + 1;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* u = U + (k / Ns) * NF * Ns + k % Ns;
  if (D == 1) {
    const double rho = u[0], rhou = u[Ns], E = u[2 * Ns], gamma = 1.0 + 1.0 / u[3 * Ns];
    p[k] = (gamma - 1) * (E - 0.5 * rhou * rhou / rho);
  } else {
    const double rho = u[0], rhou = u[Ns], rhov = u[2 * Ns], E = u[3 * Ns], gamma = 1.0 + 1.0 / u[4 * Ns];
    p[k] = (gamma - 1) * (E - 0.5 * (rhou * rhou + rhov * rhov) / rho);
  }
}
// kinetic_energy1D / kinetic_energy2D (physics.cu:814-890)
template <int D>
__global__ __launch_bounds__(256) void k_m2_kinetic(long n, const double* __restrict__ rho, const double* __restrict__ rhou,
                                                    const double* __restrict__ rhov, double* __restrict__ K) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  K[k] = D == 1 ? 0.5 * rhou[k] * rhou[k] / rho[k] : 0.5 * (rhou[k] * rhou[k] + rhov[k] * rhov[k]) / rho[k];
}
__device__ inline int binomial_d(int n, int k) {
  int r = 1;
  for (int j = 1; j <= k; j++) r = r * (n - k + j) / j;
  return r;
}
// internal_energy_multifluid (limiting_kernels.cu:160-200): one thread per (element, slice)
__global__ __launch_bounds__(256) void k_m2_internal(int Ns, int N_E, int slicenum, const double* __restrict__ p,
                                                     const double* __restrict__ g, double* __restrict__ rhoe) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * slicenum) return;
  const long o = k * Ns;  // (e*slicenum + slice)*Ns
  for (int i = 0; i < Ns; i++) {
    double I = 0;
    for (int q = 0; q < i + 1; q++) I += (double)binomial_d(i, q) * p[o + i - q] * g[o + q];
    rhoe[o + i] = I;
  }
}
// reconstruct_energy (limiting_kernels.cu:111-158)
__global__ __launch_bounds__(256) void k_m2_reconstruct(int Ns, int N_E, int slicenum, const double* __restrict__ rhoeLim,
                                                        const double* __restrict__ KLim, const double* __restrict__ EMono,
                                                        double* __restrict__ ELim) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * slicenum) return;
  const long o = k * Ns;
  for (int i = 1; i < Ns; i++) ELim[o + i] = rhoeLim[o + i] + KLim[o + i];
  double E0 = EMono[o];
  for (int q = 2; q < Ns; q += 2) E0 -= 1.0 / ((double)ifact(q + 1)) * (ELim[o + q] - EMono[o + q]);
  ELim[o] = E0;
}
// Cons2Prim / Prim2Cons (limiting_kernels.cu:957-1068), GAMNCON: alpha is
// field D+2 (two-fluid 1/(gamma-1); for passive states the reference's
// expressions take rho*phic there), beta field D+3 for stiffened states, else 0
template <int D, int PH = 1>
__global__ __launch_bounds__(256) void k_cons_prim(long n, int Ns, int to_prim, double* __restrict__ U) {
  constexpr int NF = Fields<D, PH>::NF;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double* u = U + (k / Ns) * NF * Ns + k % Ns;
  const double beta = PH == 3 ? u[(D + 3) * Ns] : 0.0;
  if (to_prim) {
    const double rho = u[0], rhou = u[Ns], rhov = D == 2 ? u[2 * Ns] : 0.0, E = u[(D + 1) * Ns],
                 alpha = u[(D + 2) * Ns];
    u[Ns] = rhou / rho;
    if (D == 2) u[2 * Ns] = rhov / rho;
    u[(D + 1) * Ns] = (E - beta - 0.5 * (rhou * rhou + rhov * rhov) / rho) / alpha;
  } else {
    const double rho = u[0], uu = u[Ns], v = D == 2 ? u[2 * Ns] : 0.0, p = u[(D + 1) * Ns], alpha = u[(D + 2) * Ns];
    u[Ns] = rho * uu;
    if (D == 2) u[2 * Ns] = rho * v;
    u[(D + 1) * Ns] = p * alpha + beta + 0.5 * rho * (uu * uu + v * v);
  }
}

// ------------------------------------- reference-named helper launchers
// The kernels behind Lpressure, Lkinetic_energy1D/2D, Linternal_energy_*,
// Lreconstruct_energy, Laverage_cell_p0, LChangeBasis, Lstridedcopy,
// Lformat_sensor, Lp0i (compat.hip), in the reference's operation order and
// without contraction, so their outputs are the reference CPU loops' bits.

// pressure (physics.cu:891-960) for every physics: gamma 1.4 (single fluid,
// passive), 1 + 1/alpha (two-fluid, stiffened), E - beta (stiffened)
template <int D, int PH>
__global__ __launch_bounds__(256) void k_pressure(long n, int Ns, const double* __restrict__ U, double* __restrict__ p) {
  constexpr int NF = Fields<D, PH>::NF;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double* u = U + (k / Ns) * NF * Ns + k % Ns;
  const double rho = u[0], rhou = u[Ns];
  double E = u[(D + 1) * Ns];
  double gamma = kGammaSF;
  if (PH == 1 || PH == 3) gamma = 1.0 + 1.0 / u[(D + 2) * Ns];
  if (PH == 3) E = E - u[(D + 3) * Ns];
  if (D == 1) {
    p[k] = (gamma - 1) * (E - 0.5 * rhou * rhou / rho);
  } else {
    const double rhov = u[2 * Ns];
    p[k] = (gamma - 1) * (E - 0.5 * (rhou * rhou + rhov * rhov) / rho);
  }
}
// internal_energy_stiffened (limiting_kernels.cu:203-243): the two-fluid
// product plus the limited gamma*pinf/(gamma-1)
__global__ __launch_bounds__(256) void k_internal_stiffened(int Ns, int N_E, int slicenum, const double* __restrict__ p,
                                                            const double* __restrict__ g, const double* __restrict__ b,
                                                            double* __restrict__ rhoe) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * slicenum) return;
  const long o = k * Ns;
  for (int i = 0; i < Ns; i++) {
    double I = 0;
    for (int q = 0; q < i + 1; q++) I += (double)binomial_d(i, q) * p[o + i - q] * g[o + q];
    rhoe[o + i] = I + b[o + i];
  }
}
// average_cell_p0 (rk_kernels.cu:61-92): one thread per (element, field)
__global__ __launch_bounds__(256) void k_average_p0(long n, int Ns, double* __restrict__ DU) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double* d = DU + k * Ns;
  double average = 0.0;
  for (int i = 0; i < Ns; i++) average += d[i];
  average = average / Ns;
  for (int i = 0; i < Ns; i++) d[i] = average;
}
// ChangeBasis (limiting_kernels.cu:918-955): Unew[(e,fc),i] = sum_ii T[e][i][ii] U[(e,fc),ii]
__global__ __launch_bounds__(256) void k_change_basis(int size1, int size2, int N_E, int NF, const double* __restrict__ T,
                                                      const double* __restrict__ U, double* __restrict__ Unew) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * NF * size1) return;
  const int i = (int)(k % size1);
  const long efc = k / size1, e = efc / NF;
  double sol = 0.0;
  for (int ii = 0; ii < size2; ii++) sol += T[(e * size1 + i) * size2 + ii] * U[efc * size2 + ii];
  Unew[k] = sol;
}
// stridedcopy (limiting_kernels.cu:51-108): block i of blocklen values
__global__ __launch_bounds__(256) void k_stridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA,
                                                     int offsetB, const double* __restrict__ A, double* __restrict__ B) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)numblocks * blocklen) return;
  const long i = k / blocklen, j = k % blocklen;
  B[offsetB + i * strideB + j] = A[offsetA + i * strideA + j];
}
// format_sensor (printer_sensor.cu:19-43)
__global__ __launch_bounds__(256) void k_format_sensor(int N_s, int N_E, const int* __restrict__ sensor,
                                                       double* __restrict__ out) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < (long)N_E * N_s) out[k] = (double)sensor[k / N_s];
}
// p0i (limiting_kernels.cu:723-768) into Unew for the flagged elements
__global__ __launch_bounds__(256) void k_p0i(int N_E, int NF, int Ns, int NG, double refArea,
                                             const int* __restrict__ sensors, const double* __restrict__ phi_w,
                                             const double* __restrict__ U, double* __restrict__ Unew) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * NF || sensors[k / NF] == 0) return;
  const double* u = U + k * Ns;
  double avg = 0;
  for (int g = 0; g < NG; ++g)
    for (int i = 0; i < Ns; ++i) avg += phi_w[i * NG + g] * u[i];
  avg = avg / refArea;
  for (int i = 0; i < Ns; ++i) Unew[k * Ns + i] = avg;
}

// ---------------------------------------------------------------- sensors
// Uavg = sum_l (1/N_s * U_l), sequential (blasGemm M=1, sensor.cc:29)
template <int Ns>
__global__ __launch_bounds__(256) void k_uavg(const double* __restrict__ U, long ncol, double* __restrict__ Uavg) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  Uavg[col] = sensor_nodal_mean<Ns>(U + col * Ns);
}

struct SensorTh {
  int on[3];
  double th[3];
};

// First test that fires for element e (as if e is tested): tgt = neighbour
// index, -1 for a boundary neighbour (flags only e), -2 for none.
template <int D, int PH>
// With init and act (the solver's sensing), the pass also sets init[e] = 0
// (sensors start at zero, sensor.cc:17-20) and act[e] = 1 (every element
// active before the first relaxation sweep); with count, zeroes the flag
// list's count that k_sensor_final appends to.
__global__ __launch_bounds__(256) void k_sensor_hit(int N_E, int N_N, const int* __restrict__ nb,
                                                    const double* __restrict__ Uavg, SensorTh s,
                                                    int* __restrict__ tgt, int* __restrict__ val,
                                                    int* __restrict__ init = nullptr, int* __restrict__ act = nullptr,
                                                    int* __restrict__ count = nullptr) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (count && e == 0) *count = 0;
  if (e >= N_E) return;
  if (init) {
    init[e] = 0;
    act[e] = 1;
  }
  const SensorCell L = sensor_cell<D, PH>(Uavg, e);
  int t_out = -2, v_out = 0;
  for (int nn = 0; nn < N_N; nn++) {
    int right = nb[e * N_N + nn];
    if (right < 0) { t_out = -1; v_out = 1; break; }
    const SensorCell R = sensor_cell<D, PH>(Uavg, right);
    const SensorRoe m = sensor_roe(L, R);
    if (s.on[0] && sensor_contact(L, R, m, s.th[0])) { t_out = right; v_out = 1; break; }
    if (s.on[1] && sensor_shock(L, R, m, s.th[1])) { t_out = right; v_out = 2; break; }
    if (s.on[2] && sensor_gamma(L, R, s.th[2])) { t_out = right; v_out = 3; break; }
  }
  tgt[e] = t_out;
  val[e] = v_out;
}

// act_in/act_out: 1 if the element runs its tests in the sequential order.
// prev: the previous sweep's change flag when sweeps are queued in batches;
// zero there means the fixed point is reached and this sweep does nothing.
__global__ __launch_bounds__(256) void k_sensor_relax(int N_E, int N_N, const int* __restrict__ nb,
                                                      const int* __restrict__ init, const int* __restrict__ tgt,
                                                      const int* __restrict__ act_in, int* __restrict__ act_out,
                                                      int* __restrict__ changed, const int* prev = nullptr) {
  if (prev && *prev == 0) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= N_E) return;
  int a = init[e] == 0 ? 1 : 0;
  if (a) {
    for (int nn = 0; nn < N_N; ++nn) {
      int o = nb[e * N_N + nn];
      if (o >= 0 && o < e && act_in[o] && tgt[o] == e) { a = 0; break; }
    }
  }
  act_out[e] = a;
  if (a != act_in[e]) atomicOr(changed, 1);
}

__global__ __launch_bounds__(256) void k_fill_i32(int* p, long n, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// last writer in element order wins; untouched elements keep init.  With
// list, the flagged elements are also appended to list[0 .. *count) (one
// atomic per block, as k_flag_list; *count zeroed by the hit pass).
__global__ __launch_bounds__(256) void k_sensor_final(int N_E, int N_N, const int* __restrict__ nb,
                                                      const int* __restrict__ init, const int* __restrict__ tgt,
                                                      const int* __restrict__ val, const int* __restrict__ act,
                                                      int* __restrict__ sensors, int* __restrict__ list = nullptr,
                                                      int* __restrict__ count = nullptr) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  int bval = 0;
  if (x < N_E) {
    int best = -1;
    bval = init[x];
    if (act[x] && tgt[x] != -2) { best = x; bval = val[x]; }
    for (int nn = 0; nn < N_N; ++nn) {
      int o = nb[x * N_N + nn];
      if (o >= 0 && act[o] && tgt[o] == x && o > best) { best = o; bval = val[o]; }
    }
    sensors[x] = bval;
  }
  if (list) {  // uniform over the launch
    __shared__ int sN, sBase;
    __shared__ int sIdx[256];
    if (threadIdx.x == 0) sN = 0;
    __syncthreads();
    const bool f = x < N_E && bval != 0;
    if (f) sIdx[atomicAdd(&sN, 1)] = x;
    __syncthreads();
    if (threadIdx.x == 0) sBase = sN ? atomicAdd(count, sN) : 0;
    __syncthreads();
    if ((int)threadIdx.x < sN) list[sBase + threadIdx.x] = sIdx[threadIdx.x];
  }
}

// set2average (limiting_kernels.cu:1792-1828) on one nodal column of Ns
// values, SL slices of N+1: UC <- M2L * average(L2M * UC); Alim receives the
// averaged monomial coefficients.
template <int Ns, int N, int SL>
__device__ inline void hr_avg_nodal(const double* __restrict__ L2M, const double* __restrict__ M2L, double* UC,
                                    double* Alim) {
  double A[Ns];
#pragma unroll
  for (int m = 0; m < Ns; ++m) {
    double s = 0;
#pragma unroll
    for (int n = 0; n < Ns; ++n) s += L2M[n * Ns + m] * UC[n];
    A[m] = s;
  }
#pragma unroll
  for (int sl = 0; sl < SL; ++sl) average_monomial_d<N>(&A[sl * (N + 1)], &Alim[sl * (N + 1)]);
#pragma unroll
  for (int m = 0; m < Ns; ++m) {
    double s = 0;
#pragma unroll
    for (int n = 0; n < Ns; ++n) s += M2L[n * Ns + m] * Alim[n];
    UC[m] = s;
  }
}

// HR (limiting_kernels.cu:1830-1868) on one nodal column with its left and
// right neighbours' columns: UC <- M2L * limit(L2M * {UL, UC, UR}); Alim
// receives the limited monomial coefficients.
template <int Ns, int N, int SL>
__device__ inline void hr_full_nodal(const double* __restrict__ L2M, const double* __restrict__ M2L, const double* UL,
                                     double* UC, const double* UR, double* Alim) {
  double AL[Ns], AC[Ns], AR[Ns];
#pragma unroll
  for (int m = 0; m < Ns; ++m) {
    double s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int n = 0; n < Ns; ++n) {
      double a = L2M[n * Ns + m];
      s1 += a * UL[n];
      s2 += a * UC[n];
      s3 += a * UR[n];
    }
    AL[m] = s1; AC[m] = s2; AR[m] = s3;
  }
#pragma unroll
  for (int i = 0; i < Ns; ++i) Alim[i] = 0;
#pragma unroll
  for (int sl = 0; sl < SL; ++sl)
    limit_monomial_d<N>(&AL[sl * (N + 1)], &AC[sl * (N + 1)], &AR[sl * (N + 1)], &Alim[sl * (N + 1)]);
#pragma unroll
  for (int m = 0; m < Ns; ++m) {
    double s = 0;
#pragma unroll
    for (int n = 0; n < Ns; ++n) s += M2L[n * Ns + m] * Alim[n];
    UC[m] = s;
  }
}

// hri1D (limiting_kernels.cu:322-427) on one column of a flagged element
// whose nodal values are in UC (in/out), with the boundary rules of the
// direction offxy; UL and UR (the neighbours' columns) are read only when
// both neighbours exist (physical == 0).
template <int D, int P>
__device__ inline void hri_limit(int physical, int fc, int offxy, const double* __restrict__ L2M,
                                 const double* __restrict__ M2L, const double* UL, double* UC, const double* UR) {
  constexpr int Ns = Shape<D, P>::Ns, N = P;
  constexpr int slicenum = (D == 1) ? 1 : P + 1;
  double Alim[Ns];
  if (physical == 4) {
  } else if (physical == 2) {
    hr_avg_nodal<Ns, N, slicenum>(L2M, M2L, UC, Alim);
  } else if (physical == 3) {
    const int un_idx = (offxy == 0) ? 1 : (offxy == 2 ? 2 : 0);
    if (fc != un_idx) hr_avg_nodal<Ns, N, slicenum>(L2M, M2L, UC, Alim);
  } else {
    hr_full_nodal<Ns, N, slicenum>(L2M, M2L, UL, UC, UR, Alim);
  }
}

// The limited column (element e, field fc) of one sweep, from U, into UC.
template <int D, int P>
__device__ inline void hri_sweep_column(int e, int fc, int NF, int N_N, const int* __restrict__ nb, int offxy,
                                        const double* __restrict__ L2M, const double* __restrict__ M2L,
                                        const double* __restrict__ U, double* UC) {
  constexpr int Ns = Shape<D, P>::Ns;
  const int left = nb[e * N_N + offxy + 0], right = nb[e * N_N + offxy + 1];
  int physical = 0;
  if (left < 0) physical = -left;
  if (right < 0) physical = -right;
#pragma unroll
  for (int i = 0; i < Ns; ++i) UC[i] = U[((long)e * NF + fc) * Ns + i];
  double UL[Ns], UR[Ns];
  if (physical == 0) {
#pragma unroll
    for (int i = 0; i < Ns; ++i) {
      UL[i] = U[((long)left * NF + fc) * Ns + i];
      UR[i] = U[((long)right * NF + fc) * Ns + i];
    }
  }
  hri_limit<D, P>(physical, fc, offxy, L2M, M2L, UL, UC, UR);
}

// The same column of a second sweep with each element's values taken from W
// where the element is flagged (the first sweep's result for it) and from U
// elsewhere: the values U holds after the first sweep's copy-back.
template <int D, int P>
__device__ inline void hri_sweep_column_sel(int e, int fc, int NF, int N_N, const int* __restrict__ nb, int offxy,
                                            const double* __restrict__ L2M, const double* __restrict__ M2L,
                                            const int* __restrict__ sensors, const double* U,
                                            const double* __restrict__ W, double* UC) {
  constexpr int Ns = Shape<D, P>::Ns;
  const int left = nb[e * N_N + offxy + 0], right = nb[e * N_N + offxy + 1];
  int physical = 0;
  if (left < 0) physical = -left;
  if (right < 0) physical = -right;
  const double* Se = sensors[e] ? W : U;
#pragma unroll
  for (int i = 0; i < Ns; ++i) UC[i] = Se[((long)e * NF + fc) * Ns + i];
  double UL[Ns], UR[Ns];
  if (physical == 0) {
    const double* Sl = sensors[left] ? W : U;
    const double* Sr = sensors[right] ? W : U;
#pragma unroll
    for (int i = 0; i < Ns; ++i) {
      UL[i] = Sl[((long)left * NF + fc) * Ns + i];
      UR[i] = Sr[((long)right * NF + fc) * Ns + i];
    }
  }
  hri_limit<D, P>(physical, fc, offxy, L2M, M2L, UL, UC, UR);
}

// hri1D (limiting_kernels.cu:322-427) for one column (element, field) of a
// flagged element: HR in Lagrange space through Lag2Mono / Mono2Lag, or the
// boundary rules; Unew receives the column.
template <int D, int P>
__device__ inline void hri_column(int e, int fc, int NF, int N_N, const int* __restrict__ nb, int offxy,
                                  const double* __restrict__ L2M, const double* __restrict__ M2L,
                                  const double* __restrict__ U, double* __restrict__ Unew) {
  constexpr int Ns = Shape<D, P>::Ns;
  double UC[Ns];
  hri_sweep_column<D, P>(e, fc, NF, N_N, nb, offxy, L2M, M2L, U, UC);
#pragma unroll
  for (int i = 0; i < Ns; ++i) Unew[((long)e * NF + fc) * Ns + i] = UC[i];
}

// one thread per (element, field); unflagged elements return at once
template <int D, int P>
__global__ __launch_bounds__(128) void k_hri(int N_E, int NF, int N_N, const int* __restrict__ nb, int offxy,
                                             const double* __restrict__ L2M, const double* __restrict__ M2L,
                                             const int* __restrict__ sensors, const double* __restrict__ U,
                                             double* __restrict__ Unew) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)N_E * NF) return;
  const int e = (int)(tid / NF), fc = (int)(tid % NF);
  if (sensors[e] == 0) return;
  hri_column<D, P>(e, fc, NF, N_N, nb, offxy, L2M, M2L, U, Unew);
}

// The same over a list of flagged elements (k_flag_list): a fixed grid
// strides over list[0 .. count) x fields.
template <int D, int P>
__global__ __launch_bounds__(128) void k_hri_list(int NF, int N_N, const int* __restrict__ nb, int offxy,
                                                  const double* __restrict__ L2M, const double* __restrict__ M2L,
                                                  const int* __restrict__ list, const int* __restrict__ count,
                                                  const double* __restrict__ U, double* __restrict__ Unew) {
  const long n = (long)count[0] * NF;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    hri_column<D, P>(list[t / NF], (int)(t % NF), NF, N_N, nb, offxy, L2M, M2L, U, Unew);
}

// The second sweep over the list without either copy-back: the listed
// (flagged) elements' first-sweep values are in W, so the result goes straight
// to U[e] -- no thread reads U of a flagged element here.  Same bits as the
// two sweeps with their copy-backs.
template <int D, int P>
__global__ __launch_bounds__(128) void k_hri_list_second(int NF, int N_N, const int* __restrict__ nb, int offxy,
                                                         const double* __restrict__ L2M,
                                                         const double* __restrict__ M2L, const int* __restrict__ list,
                                                         const int* __restrict__ count,
                                                         const int* __restrict__ sensors, double* U,
                                                         const double* __restrict__ W) {
  constexpr int Ns = Shape<D, P>::Ns;
  const long n = (long)count[0] * NF;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int e = list[t / NF], fc = (int)(t % NF);
    double UC[Ns];
    hri_sweep_column_sel<D, P>(e, fc, NF, N_N, nb, offxy, L2M, M2L, sensors, U, W, UC);
#pragma unroll
    for (int i = 0; i < Ns; ++i) U[((long)e * NF + fc) * Ns + i] = UC[i];
  }
}

// m2i1D (limiting_kernels.cu:431-722) over the listed elements (two-fluid:
// rho, rho u[, rho v], E, 1/(gamma-1)).  Eight lanes per element: lane c < D+3
// limits one column (c <= D: rho and the momenta; D+1: 1/(gamma-1); D+2: the
// pressure (:1870-1897), computed node by node from the five fields) through
// HR or the boundary rules; then lane 0 rebuilds the energy from the limited
// columns: kinetic_energy (:1899-1922), internal_energy (:1924-1953),
// reconstruct_total_energy (:1955-2001).  A block of 64 lanes takes eight
// elements per pass; the loop bound is uniform over the block.
// With list == nullptr the kernel walks elements 0 .. N_E-1 and limits those
// with sensors[e] != 0 (Lm2i1D's own loop).
template <int D, int P>
__global__ __launch_bounds__(64) void k_m2i_list(int N_N, const int* __restrict__ nb, int offxy,
                                                 const double* __restrict__ L2M, const double* __restrict__ M2L,
                                                 const int* __restrict__ list, const int* __restrict__ count,
                                                 const double* __restrict__ U, double* __restrict__ Unew,
                                                 int N_E = 0, const int* __restrict__ sensors = nullptr) {
  constexpr int Ns = Shape<D, P>::Ns, N = P, N1 = P + 1, NF = D + 3;
  constexpr int SL = (D == 1) ? 1 : N1;
  constexpr int fE = D + 1, fG = D + 2;
  __shared__ double sNod[8][D + 1][Ns];  // limited rho, rho u[, rho v]
  __shared__ double sG[8][Ns], sP[8][Ns];  // limited monomials of 1/(gamma-1) and p
  const int g = threadIdx.x >> 3, c = threadIdx.x & 7;
  const long n = list ? (long)count[0] : (long)N_E;
  auto pressure = [&](long el, int i) -> double {
    const double* u = U + (long)el * NF * Ns;
    const double rho = u[i], ru = u[Ns + i], E = u[fE * Ns + i], ga = u[fG * Ns + i];
    if (D == 1) return (E - 0.5 * ru * ru / rho) / ga;
    const double rv = u[2 * Ns + i];
    return (E - 0.5 * (ru * ru + rv * rv) / rho) / ga;
  };
  for (long base = (long)blockIdx.x * 8; base < n; base += (long)gridDim.x * 8) {
    const bool live = base + g < n && (list || sensors[base + g] != 0);
    const int e = live ? (list ? list[base + g] : (int)(base + g)) : 0;
    int physical = 0;
    if (live) {
      const int left = nb[e * N_N + offxy + 0], right = nb[e * N_N + offxy + 1];
      if (left < 0) physical = -left;
      if (right < 0) physical = -right;
      if (c < NF) {
        const int fld = c <= D ? c : fG;  // column c's field (the pressure column has none)
        double UC[Ns], Alim[Ns] = {};
#pragma unroll
        for (int i = 0; i < Ns; ++i) UC[i] = c == D + 2 ? pressure(e, i) : U[((long)e * NF + fld) * Ns + i];
        if (physical == 4) {
        } else if (physical == 2 || physical == 3) {
          const int un_idx = (offxy == 0) ? 1 : (offxy == 2 ? 2 : 0);  // wall-normal momentum
          if (physical == 2 || c != un_idx) hr_avg_nodal<Ns, N, SL>(L2M, M2L, UC, Alim);
        } else {
          double UL[Ns], UR[Ns];
#pragma unroll
          for (int i = 0; i < Ns; ++i) {
            UL[i] = c == D + 2 ? pressure(left, i) : U[((long)left * NF + fld) * Ns + i];
            UR[i] = c == D + 2 ? pressure(right, i) : U[((long)right * NF + fld) * Ns + i];
          }
          hr_full_nodal<Ns, N, SL>(L2M, M2L, UL, UC, UR, Alim);
        }
        if (c <= D) {
#pragma unroll
          for (int i = 0; i < Ns; ++i) sNod[g][c][i] = UC[i];
        } else if (c == D + 1) {
#pragma unroll
          for (int i = 0; i < Ns; ++i) sG[g][i] = Alim[i];
        } else {
#pragma unroll
          for (int i = 0; i < Ns; ++i) sP[g][i] = Alim[i];
        }
        if (c <= D + 1) {
#pragma unroll
          for (int i = 0; i < Ns; ++i) Unew[((long)e * NF + fld) * Ns + i] = UC[i];
        }
      }
    }
    __syncthreads();
    if (live && c == 0) {
      double EC[Ns];
#pragma unroll
      for (int i = 0; i < Ns; ++i) EC[i] = U[((long)e * NF + fE) * Ns + i];
      if (physical != 4) {
        double tmp[Ns], KLim[Ns], rhoeLim[Ns], EMono[Ns], ELim[Ns];
#pragma unroll
        for (int i = 0; i < Ns; ++i) {
          const double rho = sNod[g][0][i], ru = sNod[g][1][i];
          if (D == 1) {
            tmp[i] = 0.5 * (ru * ru) / rho;
          } else {
            const double rv = sNod[g][D][i];
            tmp[i] = 0.5 * (ru * ru + rv * rv) / rho;
          }
        }
#pragma unroll
        for (int m = 0; m < Ns; ++m) {
          double s1 = 0, s2 = 0;
#pragma unroll
          for (int q = 0; q < Ns; ++q) {
            s1 += L2M[q * Ns + m] * tmp[q];
            s2 += L2M[q * Ns + m] * EC[q];
          }
          KLim[m] = s1;
          EMono[m] = s2;
        }
#pragma unroll
        for (int sl = 0; sl < SL; ++sl)
#pragma unroll
          for (int i = 0; i < N1; ++i) {
            double I = 0;
#pragma unroll
            for (int q = 0; q < i + 1; ++q)
              I += (double)binomial_d(i, q) * sP[g][sl * N1 + i - q] * sG[g][sl * N1 + q];
            rhoeLim[sl * N1 + i] = I;
          }
#pragma unroll
        for (int sl = 0; sl < SL; ++sl) {
          const int i0 = sl * N1;
#pragma unroll
          for (int i = 1; i < N1; ++i) ELim[i0 + i] = rhoeLim[i0 + i] + KLim[i0 + i];
          double E0 = EMono[i0];
#pragma unroll
          for (int q = 2; q < N1; q += 2) E0 -= 1.0 / ((double)ifact(q + 1)) * (ELim[i0 + q] - EMono[i0 + q]);
          ELim[i0] = E0;
        }
#pragma unroll
        for (int m = 0; m < Ns; ++m) {
          double s = 0;
#pragma unroll
          for (int q = 0; q < Ns; ++q) s += M2L[q * Ns + m] * ELim[q];
          EC[m] = s;
        }
      }
#pragma unroll
      for (int i = 0; i < Ns; ++i) Unew[((long)e * NF + fE) * Ns + i] = EC[i];
    }
    __syncthreads();
  }
}

// p0i (limiting_kernels.cu:723-768) over the listed elements: one thread per
// (element, field) sets the column to its cell average
// sum_g sum_i phi_w[i*N_G+g] U_i / refArea.  In place: a thread reads and
// writes only its own column, which equals the reference's Unew +
// copy_detected.
__global__ __launch_bounds__(256) void k_p0i_list(int NF, int Ns, int NG, double refArea,
                                                  const double* __restrict__ phi_w, const int* __restrict__ list,
                                                  const int* __restrict__ count, double* __restrict__ U) {
  const long n = (long)count[0] * NF;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    double* u = U + ((long)list[t / NF] * NF + t % NF) * Ns;
    double avg = 0;
    for (int g = 0; g < NG; ++g)
      for (int i = 0; i < Ns; ++i) avg += phi_w[i * NG + g] * u[i];
    avg = avg / refArea;
    for (int i = 0; i < Ns; ++i) u[i] = avg;
  }
}

// list[0 .. count) = the flagged elements (sensors[e] != 0), in no fixed
// order (each is limited independently); count must be zero on entry.  One
// atomic add per block, on a global counter.
__global__ __launch_bounds__(256) void k_flag_list(int N_E, const int* __restrict__ sensors, int* __restrict__ list,
                                                   int* __restrict__ count) {
  __shared__ int sN, sBase;
  __shared__ int sIdx[256];
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (threadIdx.x == 0) sN = 0;
  __syncthreads();
  const bool f = e < N_E && sensors[e] != 0;
  int slot = -1;
  if (f) slot = atomicAdd(&sN, 1);
  if (f) sIdx[slot] = e;
  __syncthreads();
  if (threadIdx.x == 0) sBase = sN ? atomicAdd(count, sN) : 0;
  __syncthreads();
  if ((int)threadIdx.x < sN) list[sBase + threadIdx.x] = sIdx[threadIdx.x];
}

// U[e] = Uold[e] for the listed elements
__global__ __launch_bounds__(256) void k_copy_list(int per_elem, const int* __restrict__ list,
                                                   const int* __restrict__ count, const double* __restrict__ Uold,
                                                   double* __restrict__ U) {
  const long n = (long)count[0] * per_elem;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long i = (long)list[t / per_elem] * per_elem + t % per_elem;
    U[i] = Uold[i];
  }
}

// U[e] = Uold[e] for the flagged elements (copy_detected, sensor_kernels.cu:167-200);
// a block covers 64 elements and reads each flag once.
__global__ __launch_bounds__(256) void k_copy_detected(int N_E, int per_elem, const int* __restrict__ sensors,
                                                       const double* __restrict__ Uold, double* __restrict__ U) {
  __shared__ int sF[64];
  const long e0 = (long)blockIdx.x * 64;
  const int ne = (int)min(64L, (long)N_E - e0);
  if (threadIdx.x < 64) sF[threadIdx.x] = (int)threadIdx.x < ne ? sensors[e0 + threadIdx.x] : 0;
  __syncthreads();
  const long base = e0 * per_elem;
  for (int k = threadIdx.x; k < ne * per_elem; k += 256)
    if (sF[k / per_elem] != 0) U[base + k] = Uold[base + k];
}

// ------------------------------------------------------------ host side
static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// fn(dim, physics) with both as integral constants (physics 0..3)
template <typename Fn>
static hipError_t by_dim_phys(int D, int physics, Fn&& fn) {
  auto ph = [&](auto d) -> hipError_t {
    switch (physics) {
      case 0: return fn(d, std::integral_constant<int, 0>{});
      case 1: return fn(d, std::integral_constant<int, 1>{});
      case 2: return fn(d, std::integral_constant<int, 2>{});
      default: return fn(d, std::integral_constant<int, 3>{});
    }
  };
  return D == 1 ? ph(std::integral_constant<int, 1>{}) : ph(std::integral_constant<int, 2>{});
}

template <typename Fn>
static hipError_t by_order_ns(int D, int P, Fn&& fn) {
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  if (D == 1) {
    if (P == 1) return fn(I1{}, I1{});
    if (P == 2) return fn(I1{}, I2{});
    return fn(I1{}, I3{});
  }
  if (P == 1) return fn(I2{}, I1{});
  if (P == 2) return fn(I2{}, I2{});
  return fn(I2{}, I3{});
}

// Two launches per application.  1D: A = Lag2Mono U (k_apply), then
// U = Mono2Lag hrl(A) (k_hr_sweep).  2D: work = MonoX2MonoY hrl_x(Lag2MonoX U),
// then U = MonoY2Lag hrl_y(work).
hipError_t launch_limit_hr(Solver& s, double* U, hipStream_t st) {
  const long ncol = (long)s.N_E * s.NF;
  return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value;
    constexpr int Ns = Shape<D, P>::Ns, SL = (D == 1) ? 1 : P + 1;
    const dim3 g(nblk(ncol, 256)), b(256);
    if (D == 1) {
      hipLaunchKernelGGL((k_apply<Ns>), g, b, 0, st, s.t.Lag2Mono, U, s.work, ncol, s.gate);
      hipLaunchKernelGGL((k_hr_sweep<Ns, P, SL>), g, b, 0, st, s.N_E, s.NF, s.N_N, s.t.neighbors, 0, nullptr,
                         s.t.Mono2Lag, s.work, U, s.gate);
    } else if (s.hr_const) {
      if constexpr (D == 2 && P <= 2) {
        hipLaunchKernelGGL((k_hr_sweep_c<P, 0>), g, b, 0, st, s.N_E, s.NF, s.N_N, s.t.neighbors, U, s.work, s.gate);
        hipLaunchKernelGGL((k_hr_sweep_c<P, 1>), g, b, 0, st, s.N_E, s.NF, s.N_N, s.t.neighbors, s.work, U, s.gate);
      }
    } else {
      hipLaunchKernelGGL((k_hr_sweep<Ns, P, SL>), g, b, 0, st, s.N_E, s.NF, s.N_N, s.t.neighbors, 0, s.t.Lag2MonoX,
                         s.t.MonoX2MonoY, U, s.work, s.gate);
      hipLaunchKernelGGL((k_hr_sweep<Ns, P, SL>), g, b, 0, st, s.N_E, s.NF, s.N_N, s.t.neighbors, 2, nullptr,
                         s.t.MonoY2Lag, s.work, U, s.gate);
    }
    return hipGetLastError();
  });
}

// Out = HR(In) with both 2D sweeps in one launch (k_hr2d), p <= 2; needs the
// tile layout and a neighbour table that follows it (Solver::hr2d).
hipError_t launch_limit_hr_oop(Solver& s, const double* In, double* Out, hipStream_t st) {
  if (In == Out) return hipErrorInvalidValue;
  if (s.D == 1) {
    // 1D: one sweep that forms its own and its neighbours' Taylor
    // coefficients (col_apply, the order of k_apply: the same bits as the
    // two launches of launch_limit_hr)
    const long ncol = (long)s.N_E * s.NF;
    return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
      constexpr int D = decltype(d)::value, P = decltype(p)::value;
      if constexpr (D == 1) {
        constexpr int Ns = Shape<D, P>::Ns;
        hipLaunchKernelGGL((k_hr_sweep<Ns, P, 1>), dim3(nblk(ncol, 256)), dim3(256), 0, st, s.N_E, s.NF, s.N_N,
                           s.t.neighbors, 0, s.t.Lag2Mono, s.t.Mono2Lag, In, Out, s.gate);
      }
      return hipGetLastError();
    });
  }
  if (s.D != 2 || !s.hr2d || (s.NF != 4 && s.NF != 5)) return hipErrorInvalidValue;
  // 8 x 8 tiles (512 threads) where the rows allow, else 8 x 4: the taller
  // tile recomputes fewer halo columns (1.56 against 1.88 Taylor-x columns
  // per output column) and ran 167 against 177 us on the 1024 x 512 p = 2
  // two-fluid state
  constexpr int TXE = 8;
  // every element must lie in a tile (the eligibility check, Solver::hr2d, asks the same)
  if (s.cart.nx % TXE != 0 || s.cart.ny % 4 != 0) return hipErrorInvalidValue;
  const int tye = s.cart.ny % 8 == 0 ? 8 : 4;
  const int tiles = (s.cart.nx / TXE) * (s.cart.ny / tye);
  return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value;
    if constexpr (D == 2 && P <= 2) {
      constexpr int Ns = Shape<D, P>::Ns, SL = P + 1;
      auto go = [&](auto nf, auto ty) -> hipError_t {
        constexpr int NF = decltype(nf)::value, TYE = decltype(ty)::value, BS = TYE == 4 ? 256 : 512;
        if ((reinterpret_cast<uintptr_t>(Out) & 15) == 0)
          hipLaunchKernelGGL((k_hr2d<Ns, P, SL, NF, TXE, TYE, BS, true>), dim3(tiles), dim3(BS), 0, st, s.cart.nx,
                             s.N_N, s.t.neighbors, In, Out, s.gate);
        else
          hipLaunchKernelGGL((k_hr2d<Ns, P, SL, NF, TXE, TYE, BS, false>), dim3(tiles), dim3(BS), 0, st, s.cart.nx,
                             s.N_N, s.t.neighbors, In, Out, s.gate);
        return hipGetLastError();
      };
      using I4 = std::integral_constant<int, 4>;
      using I5 = std::integral_constant<int, 5>;
      using I8 = std::integral_constant<int, 8>;
      if (s.NF == 4) return tye == 8 ? go(I4{}, I8{}) : go(I4{}, I4{});
      return tye == 8 ? go(I5{}, I8{}) : go(I5{}, I4{});
    }
    return hipErrorInvalidValue;
  });
}

// Limiting::M2limiting: the reference's launch sequence on per-field arrays
// (s.m2buf: 15 arrays of N_E*N_s doubles, allocated on first use).
hipError_t launch_limit_m2(Solver& s, double* U, hipStream_t st) {
  if (s.cfg.physics != DGX_MULTIFLUID) return hipErrorInvalidValue;
  const long n = (long)s.N_E * s.N_s;
  if (!s.m2buf) {
    hipError_t e = hipMalloc(&s.m2buf, sizeof(double) * 15 * n);
    if (e != hipSuccess) return e;
    s.allocations.push_back(s.m2buf);
  }
  double* b = s.m2buf;
  double *rho = b, *rhou = b + n, *rhov = b + 2 * n, *E = b + 3 * n, *gam = b + 4 * n, *pres = b + 5 * n,
         *K = b + 6 * n, *mono = b + 7 * n, *lim = b + 8 * n, *gLim = b + 9 * n, *pLim = b + 10 * n,
         *KLim = b + 11 * n, *EMono = b + 12 * n, *rhoeLim = b + 13 * n, *ELim = b + 14 * n;
  const dim3 g(nblk(n, 256)), bl(256);
  const int NF = s.NF, Ns = s.N_s, NE = s.N_E, NN = s.N_N, n1 = s.N1;
  const int* nb = s.t.neighbors;
  return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value, Nsc = Shape<D, P>::Ns;
    auto get = [&](int f, double* x) { hipLaunchKernelGGL(k_field_get, g, bl, 0, st, n, Ns, NF, f, U, x); };
    auto put = [&](int f, const double* x) { hipLaunchKernelGGL(k_field_put, g, bl, 0, st, n, Ns, NF, f, x, U); };
    auto gemm = [&](const double* M, const double* X, double* Y) {
      hipLaunchKernelGGL((k_apply<Nsc>), dim3(nblk(NE, 256)), bl, 0, st, M, X, Y, (long)NE, nullptr);
    };
    const int sl = D == 1 ? 1 : n1, hs = D == 1 ? Ns : n1;  // slices and slice length
    auto hrl = [&](int offxy, const double* A, double* Alim) {
      hipLaunchKernelGGL((k_hrl<P>), dim3(nblk((long)NE * sl, 256)), bl, 0, st, NE, 1, NN, sl, nb, offxy, A, Alim);
    };
    auto hr = [&](const double* L2M, const double* M2L, int offxy, double* x, double* Lm, bool back) {
      gemm(L2M, x, mono);
      hrl(offxy, mono, Lm);
      if (back) gemm(M2L, Lm, x);
    };
    auto energy = [&](const double* L2M, const double* M2L) {
      hipLaunchKernelGGL((k_m2_kinetic<D>), g, bl, 0, st, n, rho, rhou, rhov, K);
      gemm(L2M, K, KLim);
      hipLaunchKernelGGL(k_m2_internal, dim3(nblk((long)NE * sl, 256)), bl, 0, st, hs, NE, sl, pLim, gLim, rhoeLim);
      gemm(L2M, E, EMono);
      hipLaunchKernelGGL(k_m2_reconstruct, dim3(nblk((long)NE * sl, 256)), bl, 0, st, hs, NE, sl, rhoeLim, KLim,
                         EMono, ELim);
      gemm(M2L, ELim, E);
    };
    if constexpr (D == 1) {
      const double *L2M = s.t.Lag2Mono, *M2L = s.t.Mono2Lag;
      get(0, rho); hr(L2M, M2L, 0, rho, lim, true);
      get(1, rhou); hr(L2M, M2L, 0, rhou, lim, true);
      get(2, E);
      get(3, gam); hr(L2M, M2L, 0, gam, gLim, true);
      hipLaunchKernelGGL((k_m2_pressure<1>), g, bl, 0, st, n, Ns, U, pres);
      hr(L2M, M2L, 0, pres, pLim, false);
      energy(L2M, M2L);
      put(0, rho); put(1, rhou); put(2, E); put(3, gam);
    } else {
      const double *LX = s.t.Lag2MonoX, *XL = s.t.MonoX2Lag, *LY = s.t.Lag2MonoY, *YL = s.t.MonoY2Lag;
      get(0, rho); hr(LX, XL, 0, rho, lim, true);
      get(1, rhou); hr(LX, XL, 0, rhou, lim, true);
      get(2, rhov); hr(LX, XL, 0, rhov, lim, true);
      get(4, gam); hr(LX, XL, 0, gam, gLim, true);
      hipLaunchKernelGGL((k_m2_pressure<2>), g, bl, 0, st, n, Ns, U, pres);
      hr(LX, XL, 0, pres, pLim, true);
      get(3, E);
      energy(LX, XL);
      hr(LY, YL, 2, rho, lim, true);
      hr(LY, YL, 2, rhou, lim, true);
      hr(LY, YL, 2, rhov, lim, true);
      hr(LY, YL, 2, gam, gLim, true);
      hr(LY, YL, 2, pres, pLim, false);
      energy(LY, YL);
      put(0, rho); put(1, rhou); put(2, rhov); put(3, E); put(4, gam);
    }
    return hipGetLastError();
  });
}

// Limiting::PRIlimiting (limiting.cc:636-662): the sensor on the conserved
// state, then the flagged HR sweeps on the primitive variables.
hipError_t launch_limit_pri(Solver& s, double* U, hipStream_t st) {
  if (s.cfg.physics != DGX_MULTIFLUID) return hipErrorInvalidValue;
  hipError_t e = launch_sensing(s, U, s.sensors, s.Uavg, st);
  if (e != hipSuccess) return e;
  const long n = (long)s.N_E * s.N_s;
  const dim3 g(nblk(n, 256)), bl(256);
  if (s.D == 1) hipLaunchKernelGGL((k_cons_prim<1, 1>), g, bl, 0, st, n, s.N_s, 1, U);
  else hipLaunchKernelGGL((k_cons_prim<2, 1>), g, bl, 0, st, n, s.N_s, 1, U);
  if ((e = launch_limit_flagged(s, U, st)) != hipSuccess) return e;
  if (s.D == 1) hipLaunchKernelGGL((k_cons_prim<1, 1>), g, bl, 0, st, n, s.N_s, 0, U);
  else hipLaunchKernelGGL((k_cons_prim<2, 1>), g, bl, 0, st, n, s.N_s, 0, U);
  return hipGetLastError();
}

// Limiting::PRlimiting: HR of the primitive variables.
hipError_t launch_limit_pr(Solver& s, double* U, hipStream_t st) {
  if (s.cfg.physics != DGX_MULTIFLUID) return hipErrorInvalidValue;
  const long n = (long)s.N_E * s.N_s;
  const dim3 g(nblk(n, 256)), bl(256);
  if (s.D == 1) hipLaunchKernelGGL((k_cons_prim<1, 1>), g, bl, 0, st, n, s.N_s, 1, U);
  else hipLaunchKernelGGL((k_cons_prim<2, 1>), g, bl, 0, st, n, s.N_s, 1, U);
  hipError_t e = launch_limit_hr(s, U, st);
  if (e != hipSuccess) return e;
  if (s.D == 1) hipLaunchKernelGGL((k_cons_prim<1, 1>), g, bl, 0, st, n, s.N_s, 0, U);
  else hipLaunchKernelGGL((k_cons_prim<2, 1>), g, bl, 0, st, n, s.N_s, 0, U);
  return hipGetLastError();
}

// Resolve sequential-order flags from per-element first hits.
// act_set: the hit pass has already set every element active in s.sens_act.
static hipError_t sensor_resolve(Solver& s, const int* nb, const int* init, const int* tgt, const int* val,
                                 int* sensors, hipStream_t st, bool act_set = false, int* list = nullptr,
                                 int* count = nullptr) {
  int* act0 = s.sens_act;
  int* act1 = s.sens_act + s.N_E;
  if (!act_set) hipLaunchKernelGGL(k_fill_i32, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, act0, (long)s.N_E, 1);
  // Jacobi sweeps converge to the sequential fixed point after (longest chain + 1) sweeps.
  if (s.sens_changed) {
    // batches of kRelaxBatch sweeps with one read-back each: a sweep after one
    // that changed nothing returns at once (its input already equals its
    // output, so either buffer holds the fixed point); the batch length is
    // even, so every batch starts from act0
    static_assert(kRelaxBatch % 2 == 0, "batches end on act0");
    for (int it = 0; it <= s.N_E; it += kRelaxBatch) {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kRelaxBatch), 0, st, s.sens_changed, (long)kRelaxBatch, 0);
      for (int i = 0; i < kRelaxBatch; ++i) {
        hipLaunchKernelGGL(k_sensor_relax, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, s.N_N, nb, init, tgt,
                           (i % 2) ? act1 : act0, (i % 2) ? act0 : act1, s.sens_changed + i,
                           i ? s.sens_changed + i - 1 : nullptr);
      }
      int h = 0;
      hipError_t e = hipMemcpyAsync(&h, s.sens_changed + kRelaxBatch - 1, sizeof(int), hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) return e;
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
      if (!h) break;
    }
    hipLaunchKernelGGL(k_sensor_final, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, s.N_N, nb, init, tgt, val,
                       act0, sensors, list, count);
    return hipGetLastError();
  }
  for (int it = 0; it <= s.N_E; ++it) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, s.flag, 1L, 0);
    hipLaunchKernelGGL(k_sensor_relax, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, s.N_N, nb, init, tgt, act0,
                       act1, s.flag);
    int h = 0;
    hipError_t e = hipMemcpyAsync(&h, s.flag, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    std::swap(act0, act1);
    if (!h) break;
  }
  hipLaunchKernelGGL(k_sensor_final, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, s.N_N, nb, init, tgt, val, act0,
                     sensors);
  return hipGetLastError();
}

hipError_t launch_uavg(const Solver& s, const double* U, double* Uavg, hipStream_t st) {
  const long ncol = (long)s.N_E * s.NF;
  return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int Ns = Shape<decltype(d)::value, decltype(p)::value>::Ns;
    hipLaunchKernelGGL((k_uavg<Ns>), dim3(nblk(ncol, 256)), dim3(256), 0, st, U, ncol, Uavg);
    return hipGetLastError();
  });
}

hipError_t launch_sensing(Solver& s, const double* U, int* sensors, double* Uavg, hipStream_t st, int* list,
                          int* count) {
  if (!Uavg) Uavg = s.Uavg;
  long ncol = (long)s.N_E * s.NF;
  hipError_t e = by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int Ns = Shape<decltype(d)::value, decltype(p)::value>::Ns;
    hipLaunchKernelGGL((k_uavg<Ns>), dim3(nblk(ncol, 256)), dim3(256), 0, st, U, ncol, Uavg);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  SensorTh th{};
  for (int k = 0; k < 3; ++k) { th.on[k] = k < s.n_thresh; th.th[k] = s.thresh[k]; }
  int* tgt = s.sens_tgt;
  int* val = s.sens_tgt + s.N_E;
  // sensors start at zero (sensor.cc:17-20): init, set by the hit pass with the first active set
  int* init = s.sens_act + 2 * (long)s.N_E;
  auto hit = [&](auto d, auto ph) -> hipError_t {
    hipLaunchKernelGGL((k_sensor_hit<decltype(d)::value, decltype(ph)::value>), dim3(nblk(s.N_E, 256)), dim3(256), 0,
                       st, s.N_E, s.N_N, s.t.neighbors, Uavg, th, tgt, val, init, s.sens_act,
                       list && s.sens_changed ? count : nullptr);
    return hipGetLastError();
  };
  e = by_dim_phys(s.D, s.cfg.physics, hit);
  if (e != hipSuccess) return e;
  if (list && !s.sens_changed) {  // the per-sweep loop below lists nothing: list afterwards
    hipError_t r = sensor_resolve(s, s.t.neighbors, init, tgt, val, sensors, st, true);
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, count, 1L, 0);
    hipLaunchKernelGGL(k_flag_list, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, sensors, list, count);
    return hipGetLastError();
  }
  return sensor_resolve(s, s.t.neighbors, init, tgt, val, sensors, st, true, list, count);
}

// One sweep of the flagged HR over the compacted list (see launch_limit_flagged).
static hipError_t hri_list(Solver& s, int offxy, const double* L2M, const double* M2L, double* U, hipStream_t st) {
  const int* count = s.flag + 2;
  hipError_t e = by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    hipLaunchKernelGGL((k_hri_list<decltype(d)::value, decltype(p)::value>), dim3(2048), dim3(128), 0, st, s.NF,
                       s.N_N, s.t.neighbors, offxy, L2M, M2L, s.flag_list, count, U, s.work);
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_copy_list, dim3(1024), dim3(256), 0, st, s.NF * s.N_s, s.flag_list, count, s.work, U);
  return hipGetLastError();
}

// HR sweeps on the cells flagged in s.sensors: the flagged elements are
// listed once (k_flag_list), then each sweep and its copy-back run over the
// list only, so the cost follows the number of flagged cells.  The same
// arithmetic per column as k_hri over the whole grid.
hipError_t launch_limit_flagged(Solver& s, double* U, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, st, s.flag + 2, 1L, 0);
  hipLaunchKernelGGL(k_flag_list, dim3(nblk(s.N_E, 256)), dim3(256), 0, st, s.N_E, s.sensors, s.flag_list,
                     s.flag + 2);
  hipError_t e;
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_limit_listed(s, U, st);
}

// 2D: the x sweep into s.work, then the y sweep from s.work (flagged
// elements) and U (the rest) straight into U (k_hri_list_second): no copy-back.
hipError_t launch_limit_listed(Solver& s, double* U, hipStream_t st) {
  if (s.D == 1) return hri_list(s, 0, s.t.Lag2Mono, s.t.Mono2Lag, U, st);
  const int* count = s.flag + 2;
  return by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
    constexpr int D = decltype(d)::value, P = decltype(p)::value;
    hipLaunchKernelGGL((k_hri_list<D, P>), dim3(2048), dim3(128), 0, st, s.NF, s.N_N, s.t.neighbors, 0,
                       s.t.Lag2MonoX, s.t.MonoX2Lag, s.flag_list, count, U, s.work);
    hipLaunchKernelGGL((k_hri_list_second<D, P>), dim3(2048), dim3(128), 0, st, s.NF, s.N_N, s.t.neighbors, 2,
                       s.t.Lag2MonoY, s.t.MonoY2Lag, s.flag_list, count, s.sensors, U, s.work);
    return hipGetLastError();
  });
}

// Limiting::M2Ilimiting (limiting.cc:682-720), two-fluid: sensing, which
// lists the flagged elements once, then m2i1D + copy_detected per direction
// over the list.
hipError_t launch_limit_m2i(Solver& s, double* U, hipStream_t st) {
  if (s.cfg.physics != DGX_MULTIFLUID) return hipErrorInvalidValue;
  hipError_t e = launch_sensing(s, U, s.sensors, s.Uavg, st, s.flag_list, s.flag + 2);  // also lists the flags
  if (e != hipSuccess) return e;
  const int* count = s.flag + 2;
  auto sweep = [&](int offxy, const double* L2M, const double* M2L) -> hipError_t {
    hipError_t r = by_order_ns(s.D, s.P, [&](auto d, auto p) -> hipError_t {
      hipLaunchKernelGGL((k_m2i_list<decltype(d)::value, decltype(p)::value>), dim3(1024), dim3(64), 0, st, s.N_N,
                         s.t.neighbors, offxy, L2M, M2L, s.flag_list, count, U, s.work);
      return hipGetLastError();
    });
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(k_copy_list, dim3(1024), dim3(256), 0, st, s.NF * s.N_s, s.flag_list, count, s.work, U);
    return hipGetLastError();
  };
  if (s.D == 1) return sweep(0, s.t.Lag2Mono, s.t.Mono2Lag);
  if ((e = sweep(0, s.t.Lag2MonoX, s.t.MonoX2Lag)) != hipSuccess) return e;
  return sweep(2, s.t.Lag2MonoY, s.t.MonoY2Lag);
}

// Limiting::P0Ilimiting (limiting.cc:723-744): sensing, then the flagged
// elements set to their cell averages.  refArea is 4 for the quadrilateral
// (main.cc:286); the reference leaves it unset for segments, here 2.
hipError_t launch_limit_p0i(Solver& s, double* U, hipStream_t st) {
  hipError_t e = launch_sensing(s, U, s.sensors, s.Uavg, st, s.flag_list, s.flag + 2);  // also lists the flags
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_p0i_list, dim3(1024), dim3(256), 0, st, s.NF, s.N_s, s.N_G, s.D == 1 ? 2.0 : 4.0, s.t.phi_w,
                     s.flag_list, s.flag + 2, U);
  return hipGetLastError();
}

// the sensing lists the flagged elements itself (k_sensor_final)
hipError_t launch_limit_hri(Solver& s, double* U, hipStream_t st) {
  hipError_t e = launch_sensing(s, U, s.sensors, s.Uavg, st, s.flag_list, s.flag + 2);
  if (e != hipSuccess) return e;
  return launch_limit_listed(s, U, st);
}

// exported for the reference-named launchers (compat.hip)
hipError_t compat_hrl(int N_s, int N_E, int Nfields, int N_N, int slicenum, const int* nb, int offxy, const double* A,
                      double* Alim) {
  long total = (long)N_E * slicenum * Nfields;
  int N = (slicenum == 1) ? N_s - 1 : N_s - 1;
  auto go = [&](auto n) -> hipError_t {
    hipLaunchKernelGGL((k_hrl<decltype(n)::value>), dim3(nblk(total, 256)), dim3(256), 0, 0, N_E, Nfields, N_N,
                       slicenum, nb, offxy, A, Alim);
    return hipGetLastError();
  };
  switch (N) {
    case 1: return go(std::integral_constant<int, 1>{});
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    default: return hipErrorInvalidValue;
  }
}

hipError_t compat_sensors(Solver& tmp, const int* nb, const double* Uavg, int* sensors, const int* on, const double* th_in,
                          int* scratch) {
  // scratch: 5*N_E ints: tgt, val, act0, act1, init
  SensorTh th{};
  for (int k = 0; k < 3; ++k) { th.on[k] = on[k]; th.th[k] = th_in[k]; }
  int* tgt = scratch;
  int* val = scratch + tmp.N_E;
  auto hit = [&](auto d, auto ph) -> hipError_t {
    hipLaunchKernelGGL((k_sensor_hit<decltype(d)::value, decltype(ph)::value>), dim3(nblk(tmp.N_E, 256)), dim3(256),
                       0, 0, tmp.N_E, tmp.N_N, nb, Uavg, th, tgt, val);
    return hipGetLastError();
  };
  hipError_t e = by_dim_phys(tmp.D, tmp.cfg.physics, hit);
  if (e != hipSuccess) return e;
  tmp.sens_act = scratch + 2 * tmp.N_E;
  int* init = scratch + 4 * tmp.N_E;
  hipError_t ec = hipMemcpyAsync(init, sensors, sizeof(int) * tmp.N_E, hipMemcpyDeviceToDevice, 0);
  if (ec != hipSuccess) return ec;
  return sensor_resolve(tmp, nb, init, tgt, val, sensors, 0);
}

hipError_t compat_hri(int N_s, int N_E, int NF, int N_N, const int* nb, int N_s1D, int slicenum, int offxy, int D,
                      const double* L2M, const double* M2L, const int* sensors, const double* U, double* Unew) {
  long n = (long)N_E * NF;
  int P = N_s1D - 1;
  (void)slicenum; (void)N_s;
  return by_order_ns(D, P, [&](auto d, auto p) -> hipError_t {
    hipLaunchKernelGGL((k_hri<decltype(d)::value, decltype(p)::value>), dim3(nblk(n, 128)), dim3(128), 0, 0, N_E, NF,
                       N_N, nb, offxy, L2M, M2L, sensors, U, Unew);
    return hipGetLastError();
  });
}

hipError_t compat_uavg(int Ns, const double* U, long ncol, double* Uavg) {
  auto go = [&](auto n) -> hipError_t {
    hipLaunchKernelGGL((k_uavg<decltype(n)::value>), dim3(nblk(ncol, 256)), dim3(256), 0, 0, U, ncol, Uavg);
    return hipGetLastError();
  };
  switch (Ns) {
    case 2: return go(std::integral_constant<int, 2>{});
    case 3: return go(std::integral_constant<int, 3>{});
    case 4: return go(std::integral_constant<int, 4>{});
    case 9: return go(std::integral_constant<int, 9>{});
    case 16: return go(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
  }
}

hipError_t compat_copy_detected(int N_s, int N_E, int NF, const int* sensors, const double* Uold, double* U) {
  int per = NF * N_s;
  hipLaunchKernelGGL(k_copy_detected, dim3(nblk(N_E, 64)), dim3(256), 0, 0, N_E, per, sensors, Uold, U);
  return hipGetLastError();
}

// ---------------------------------------- reference-named helper launchers
// Host halves of the launchers of compat.hip whose arithmetic must round like
// the reference's CPU loops (this file is compiled without contraction).
hipError_t compat_pressure(const dgx_config& c, int N_s, int N_E, const double* U, double* p) {
  const long n = (long)N_E * N_s;
  return by_dim_phys(c.dim, c.physics, [&](auto d, auto ph) -> hipError_t {
    hipLaunchKernelGGL((k_pressure<decltype(d)::value, decltype(ph)::value>), dim3(nblk(n, 256)), dim3(256), 0, 0, n,
                       N_s, U, p);
    return hipGetLastError();
  });
}
hipError_t compat_kinetic(int D, int N_s, int N_E, const double* rho, const double* rhou, const double* rhov, double* K) {
  const long n = (long)N_E * N_s;
  if (D == 1) hipLaunchKernelGGL((k_m2_kinetic<1>), dim3(nblk(n, 256)), dim3(256), 0, 0, n, rho, rhou, rhov, K);
  else hipLaunchKernelGGL((k_m2_kinetic<2>), dim3(nblk(n, 256)), dim3(256), 0, 0, n, rho, rhou, rhov, K);
  return hipGetLastError();
}
hipError_t compat_internal_energy(int N_s, int N_E, int slicenum, const double* p, const double* g, const double* b,
                                  double* rhoe) {
  const unsigned nb = nblk((long)N_E * slicenum, 256);
  if (b) hipLaunchKernelGGL(k_internal_stiffened, dim3(nb), dim3(256), 0, 0, N_s, N_E, slicenum, p, g, b, rhoe);
  else hipLaunchKernelGGL(k_m2_internal, dim3(nb), dim3(256), 0, 0, N_s, N_E, slicenum, p, g, rhoe);
  return hipGetLastError();
}
hipError_t compat_reconstruct_energy(int N_s, int N_E, int slicenum, const double* rhoeLim, const double* KLim,
                                     const double* EMono, double* ELim) {
  hipLaunchKernelGGL(k_m2_reconstruct, dim3(nblk((long)N_E * slicenum, 256)), dim3(256), 0, 0, N_s, N_E, slicenum,
                     rhoeLim, KLim, EMono, ELim);
  return hipGetLastError();
}
hipError_t compat_average_p0(int N_s, int N_E, int NF, double* DU) {
  const long n = (long)N_E * NF;
  hipLaunchKernelGGL(k_average_p0, dim3(nblk(n, 256)), dim3(256), 0, 0, n, N_s, DU);
  return hipGetLastError();
}
hipError_t compat_change_basis(int size1, int size2, int N_E, int NF, const double* T, const double* U, double* Unew) {
  hipLaunchKernelGGL(k_change_basis, dim3(nblk((long)N_E * NF * size1, 256)), dim3(256), 0, 0, size1, size2, N_E, NF, T,
                     U, Unew);
  return hipGetLastError();
}
hipError_t compat_cons_prim(const dgx_config& c, int N_s, int N_E, double* U, int to_prim) {
  if (c.physics == DGX_SINGLEFLUID) return hipErrorInvalidValue;
  const long n = (long)N_E * N_s;
  return by_dim_phys(c.dim, c.physics, [&](auto d, auto ph) -> hipError_t {
    hipLaunchKernelGGL((k_cons_prim<decltype(d)::value, decltype(ph)::value>), dim3(nblk(n, 256)), dim3(256), 0, 0, n,
                       N_s, to_prim, U);
    return hipGetLastError();
  });
}
hipError_t compat_stridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB,
                              const double* A, double* B) {
  const long n = (long)numblocks * blocklen;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_stridedcopy, dim3(nblk(n, 256)), dim3(256), 0, 0, numblocks, blocklen, strideA, strideB,
                     offsetA, offsetB, A, B);
  return hipGetLastError();
}
hipError_t compat_format_sensor(int N_s, int N_E, const int* sensor, double* out) {
  hipLaunchKernelGGL(k_format_sensor, dim3(nblk((long)N_E * N_s, 256)), dim3(256), 0, 0, N_s, N_E, sensor, out);
  return hipGetLastError();
}
hipError_t compat_p0i(int N_s, int N_E, int NF, int N_G, double refArea, const int* sensors, const double* phi_w,
                      const double* U, double* Unew) {
  hipLaunchKernelGGL(k_p0i, dim3(nblk((long)N_E * NF, 256)), dim3(256), 0, 0, N_E, NF, N_s, N_G, refArea, sensors,
                     phi_w, U, Unew);
  return hipGetLastError();
}
// m2i1D (two-fluid): the flagged elements of U limited into Unew
hipError_t compat_m2i(int D, int N_s1D, int N_E, int N_N, const int* nb, int offxy, const double* L2M,
                      const double* M2L, const int* sensors, const double* U, double* Unew) {
  return by_order_ns(D, N_s1D - 1, [&](auto d, auto p) -> hipError_t {
    const unsigned grid = (unsigned)std::min<long>(1024, nblk(N_E, 8));
    hipLaunchKernelGGL((k_m2i_list<decltype(d)::value, decltype(p)::value>), dim3(grid), dim3(64), 0, 0, N_N, nb,
                       offxy, L2M, M2L, nullptr, nullptr, U, Unew, N_E, sensors);
    return hipGetLastError();
  });
}

}  // namespace dgx
