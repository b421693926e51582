// Reference-named launchers (drop-in for the reference's extern "C" L*
// functions and its BLAS macro layer).  Device pointers, reference layouts,
// null stream, synchronous return.  Physics configuration comes from
// dgx_set_config() because the reference fixes it at compile time.
#include <mutex>
#include <string>
#include <type_traits>

#include "dgx_internal.h"
#include "physics.cuh"

namespace dgx {

hipError_t compat_hrl(int N_s, int N_E, int Nfields, int N_N, int slicenum, const int* nb, int offxy, const double* A,
                      double* Alim);
hipError_t compat_sensors(Solver& tmp, const int* nb, const double* Uavg, int* sensors, const int* on,
                          const double* th, int* scratch);
hipError_t compat_hri(int N_s, int N_E, int NF, int N_N, const int* nb, int N_s1D, int slicenum, int offxy, int D,
                      const double* L2M, const double* M2L, const int* sensors, const double* U, double* Unew);
hipError_t compat_copy_detected(int N_s, int N_E, int NF, const int* sensors, const double* Uold, double* U);
hipError_t compat_pressure(const dgx_config& c, int N_s, int N_E, const double* U, double* p);
hipError_t compat_kinetic(int D, int N_s, int N_E, const double* rho, const double* rhou, const double* rhov, double* K);
hipError_t compat_internal_energy(int N_s, int N_E, int slicenum, const double* p, const double* g, const double* b,
                                  double* rhoe);
hipError_t compat_reconstruct_energy(int N_s, int N_E, int slicenum, const double* rhoeLim, const double* KLim,
                                     const double* EMono, double* ELim);
hipError_t compat_average_p0(int N_s, int N_E, int NF, double* DU);
hipError_t compat_change_basis(int size1, int size2, int N_E, int NF, const double* T, const double* U, double* Unew);
hipError_t compat_cons_prim(const dgx_config& c, int N_s, int N_E, double* U, int to_prim);
hipError_t compat_stridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB,
                              const double* A, double* B);
hipError_t compat_format_sensor(int N_s, int N_E, const int* sensor, double* out);
hipError_t compat_p0i(int N_s, int N_E, int NF, int N_G, double refArea, const int* sensors, const double* phi_w,
                      const double* U, double* Unew);
hipError_t compat_m2i(int D, int N_s1D, int N_E, int N_N, const int* nb, int offxy, const double* L2M,
                      const double* M2L, const int* sensors, const double* U, double* Unew);

static dgx_config g_cfg = {1, 2, DGX_MULTIFLUID, DGX_RUSANOV};

// Fixed device scratch of dgx_calc_sensors_noalloc: five ints per element
// (hit targets, hit values, two activity vectors, initial flags) for up to
// kSensorsNoallocMaxElements elements, then the sweep flag.  Each GPU the
// library runs on holds its own copy (40 MiB) until the process ends.  Host
// code never uses the array's name as an address: noalloc_scratch() looks up
// each device's copy once, with that device current.
constexpr long kSensorsNoallocMaxElements = 1L << 21;
static __device__ __align__(256) int scratch[5 * kSensorsNoallocMaxElements + 64];

static std::mutex g_scratch_mu;
static int* g_scratch_dev[64];  // the array's address on each device, once looked up

static bool noalloc_scratch(int** out) {
  int dev = -1;
  if (!hip_ok(hipGetDevice(&dev), "dgx_calc_sensors_noalloc: hipGetDevice")) return false;
  if (dev < 0 || dev >= 64) {
    set_error("dgx_calc_sensors_noalloc: device ordinal out of range");
    return false;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (!hip_ok(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "dgx_calc_sensors_noalloc: hipGetSymbolAddress"))
      return false;
    g_scratch_dev[dev] = static_cast<int*>(p);
  }
  *out = g_scratch_dev[dev];
  return true;
}
const dgx_config& current_config() { return g_cfg; }

static inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

static void finish(const char* what) {
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) hip_ok(e, what);
}

template <typename Fn>
static hipError_t dispatch_phys(Fn&& fn) {
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  const dgx_config& c = g_cfg;
  auto fl = [&](auto d, auto ph) { return c.flux == 0 ? fn(d, ph, I0{}) : fn(d, ph, I1{}); };
  using I3 = std::integral_constant<int, 3>;
  auto ph = [&](auto d) {
    switch (c.physics) {
      case 0: return fl(d, I0{});
      case 1: return fl(d, I1{});
      case 2: return fl(d, I2{});
      default: return fl(d, I3{});
    }
  };
  return c.dim == 1 ? ph(I1{}) : ph(I2{});
}

// ---------------------------------------------------------------- kernels
__global__ void kc_map_to_face(long n, const int* map, const double* U, double* UF) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) UF[k] = U[map[k]];
}

// Q[e,f,i] = sum over invmap entries with node i, in invmap order
__global__ void kc_map_to_element(int N_s, int N_E, int NF, int MsNN, const int* invmap, double* Q, const double* Qt) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * NF * N_s) return;
  int i = (int)(k % N_s);
  long col = k / N_s;
  double sol = 0;
  for (int m = 0; m < MsNN; ++m) {
    const int* p = invmap + (col * MsNN + m) * 2;
    if (p[0] == i) sol += Qt[p[1]];
  }
  Q[k] = sol;
}

template <int D>
__global__ void kc_redistribute_sf(int N_G, int N_E, int NF, double* sJ, double* fJ, const double* s, const double* f,
                                   const double* J, const double* invJac) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * N_G * NF) return;
  int fc = (int)(k % NF);
  int g = (int)((k / NF) % N_G);
  int e = (int)(k / ((long)NF * N_G));
  double j = J[e];
  sJ[((long)e * NF + fc) * N_G + g] = s[((long)e * NF + fc) * N_G + g] * j;
  for (int al = 0; al < D; al++) {
    double sol = 0.0;
    for (int a = 0; a < D; a++)
      sol += invJac[(((long)e * N_G + g) * D + al) * D + a] * f[(((long)e * NF + fc) * N_G + g) * D + a] * j;
    fJ[(((long)e * NF + fc) * N_G + g) * D + al] = sol;
  }
}

// k = ((t*NF+f)*2+d)*M_G+g
__global__ void kc_redistribute_q2(int M_G, int M_T, int NF, double* qJ, const double* q, const double* JF) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)M_T * NF * 2 * M_G) return;
  int d = (int)((k / M_G) % 2);
  long t = k / ((long)M_G * 2 * NF);
  qJ[k] = q[k] * JF[t * 2 + d];
}

__global__ void kc_add_sfq(long n, double* A, const double* S, const double* F, const double* Q) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) A[k] = S[k] + F[k] + Q[k];
}

template <int D, int PH>
__global__ void kc_evaluate_sf(int N_G, int N_E, double* s, double* f, const double* Ug, const double* dUg,
                               const double* invJac) {
  constexpr int NF = Fields<D, PH>::NF, NNC = Fields<D, PH>::NNC, NFX = NF - NNC;
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * N_G) return;
  int e = (int)(k / N_G), g = (int)(k % N_G);
  double u[NF], dua[(NNC ? NNC : 1) * D], ij[D * D], fx[NF * D], sa[NNC ? NNC : 1];
  for (int fc = 0; fc < NF; ++fc) u[fc] = Ug[((long)e * NF + fc) * N_G + g];
  for (int m = 0; m < NNC; ++m)
    for (int al = 0; al < D; ++al) dua[m * D + al] = dUg[(((long)e * NF + (NFX + m)) * N_G + g) * D + al];
  for (int q = 0; q < D * D; ++q) ij[q] = invJac[((long)e * N_G + g) * D * D + q];
  volume_flux<D, PH>(u, dua, ij, fx, sa);
  for (int fc = 0; fc < NF; ++fc) {
    s[((long)e * NF + fc) * N_G + g] = fc >= NFX ? sa[fc - NFX] : 0.0;
    for (int a = 0; a < D; ++a) f[(((long)e * NF + fc) * N_G + g) * D + a] = fx[fc * D + a];
  }
}

template <int D, int PH, int FL>
__global__ void kc_evaluate_q(int M_G, int M_T, double* q, const double* UgF, const double* normals) {
  constexpr int NF = Fields<D, PH>::NF;
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)M_T * M_G) return;
  int t = (int)(k / M_G), g = (int)(k % M_G);
  double uL[NF], uR[NF], F[NF], nc[NF];
  for (int fc = 0; fc < NF; ++fc) {
    uL[fc] = UgF[(((long)t * NF + fc) * 2 + 0) * M_G + g];
    uR[fc] = UgF[(((long)t * NF + fc) * 2 + 1) * M_G + g];
  }
  riemann<D, PH, FL>(uL, uR, normals[t * D], D == 2 ? normals[t * D + 1] : 0.0, F, nc);
  for (int fc = 0; fc < NF; ++fc) {
    q[(((long)t * NF + fc) * 2 + 0) * M_G + g] = -F[fc] + nc[fc];
    q[(((long)t * NF + fc) * 2 + 1) * M_G + g] = F[fc] + nc[fc];
  }
}

template <int D>
__global__ void kc_rflctive(int M_s, int M_B, int NF, const int* boundaryMap, const double* normals, int start,
                            double* UF) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)M_B * M_s) return;
  int t = boundaryMap[start + k / M_s], j = (int)(k % M_s);
  if (D == 1) {
    UF[(((long)t * NF + 1) * 2 + 1) * M_s + j] = -UF[(((long)t * NF + 1) * 2 + 1) * M_s + j];
  } else {
    double nx = normals[t * D + 0], ny = normals[t * D + 1];
    double nx2 = nx * nx, ny2 = ny * ny, m2nxny = -2 * nx * ny, inv = 1.0 / (nx2 + ny2);
    double vxL = UF[(((long)t * NF + 1) * 2 + 0) * M_s + j];
    double vyL = UF[(((long)t * NF + 2) * 2 + 0) * M_s + j];
    UF[(((long)t * NF + 1) * 2 + 1) * M_s + j] = inv * ((-nx2 + ny2) * vxL + m2nxny * vyL);
    UF[(((long)t * NF + 2) * 2 + 1) * M_s + j] = inv * (m2nxny * vxL + (nx2 - ny2) * vyL);
  }
}

// solve (rk_kernels.cu:22-56): DU = Dt * (Minv_e f), the sum in node order, no contraction
__global__ void kc_solve(int N_s, int N_E, int NF, double Dt, const double* Minv, const double* f, double* DU) {
#pragma clang fp contract(off)
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * NF * N_s) return;
  int i = (int)(k % N_s);
  long col = k / N_s;
  int e = (int)(col / NF);
  double sol = 0.0;
  for (int ii = 0; ii < N_s; ii++) sol += Minv[((long)e * N_s + ii) * N_s + i] * f[col * N_s + ii];
  DU[k] = Dt * sol;
}

template <int D, int PH>
__global__ void kc_find_upa(int N_s, int N_E, const double* U, double* UPA) {
  constexpr int NF = Fields<D, PH>::NF;
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * N_s) return;
  int e = (int)(k / N_s), i = (int)(k % N_s);
  double u[NF];
  for (int fc = 0; fc < NF; ++fc) u[fc] = U[((long)e * NF + fc) * N_s + i];
  UPA[k] = upa<D, PH>(u);
}

// BLAS layer: column-major, one thread per C entry, each entry in the
// reference BLAS's order and rounding (no contraction): NN C(i,j) starts from
// 0, C or beta*C and adds (alpha*B(l,j))*A(i,l) for l = 1..K in turn; TN forms
// the dot product first, then alpha*temp (+ beta*C).
__global__ void kc_dgemm(int tA, int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb,
                         double beta, double* C, int ldc) {
#pragma clang fp contract(off)
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)M * N) return;
  int i = (int)(k % M), j = (int)(k / M);
  double* c = C + i + (long)j * ldc;
  if (!tA) {
    double s = beta == 0.0 ? 0.0 : (beta == 1.0 ? *c : beta * *c);
    for (int l = 0; l < K; ++l) s = s + (alpha * B[l + (long)j * ldb]) * A[i + (long)l * lda];
    *c = s;
  } else {
    double s = 0.0;
    for (int l = 0; l < K; ++l) s = s + A[l + (long)i * lda] * B[l + (long)j * ldb];
    *c = (beta == 0.0) ? alpha * s : alpha * s + beta * *c;
  }
}

// y = y + a*x (reference DAXPY; a zero a leaves y as it is)
__global__ void kc_daxpy(int n, double a, const double* x, int incx, double* y, int incy) {
#pragma clang fp contract(off)
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && a != 0.0) y[k * incy] = y[k * incy] + a * x[k * incx];
}
__global__ void kc_dcopy(int n, const double* x, int incx, double* y, int incy) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) y[k * incy] = x[k * incx];
}
// BLAS idamax: the first index of the largest |x|, scanning from x[0] and
// moving on only to a strictly larger value, so a NaN after x[0] is passed
// over and a NaN x[0] is the answer.  The non-negative doubles order like
// their bit patterns, so the maximum over non-NaN |x| is an integer atomicMax;
// the second kernel takes the lowest index holding it.  A NaN x[0] is caught
// on the host side of the launcher.
__global__ void kc_idamax(int n, const double* x, int incx, unsigned long long* best_val, int* best_idx) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = fabs(x[k * incx]);
  if (v != v) return;
  unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  unsigned long long old = atomicMax(best_val, bits);
  (void)old;
}
__global__ void kc_idamax_index(int n, const double* x, int incx, const unsigned long long* best_val, int* best_idx) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = fabs(x[k * incx]);
  if ((unsigned long long)__double_as_longlong(v) == *best_val) atomicMin(best_idx, (int)k);
}

}  // namespace dgx

using namespace dgx;

extern "C" {

dgx_status dgx_set_config(const dgx_config* cfg) {
  if (!cfg || (cfg->dim != 1 && cfg->dim != 2) || cfg->order < 1 || cfg->order > 3 || cfg->physics < 0 ||
      cfg->physics > 3 || cfg->flux < 0 || cfg->flux > 1) {
    set_error("dgx_set_config: invalid configuration");
    return DGX_EINVAL;
  }
  g_cfg = *cfg;
  return DGX_OK;
}

void LmapToFace(int M_s, int M_T, int N_s, int* map, double* U, double* UF) {
  (void)N_s;
  long n = (long)M_T * dgx_nfields(&g_cfg) * 2 * M_s;
  hipLaunchKernelGGL(kc_map_to_face, dim3(nb256(n)), dim3(256), 0, 0, n, map, U, UF);
  finish("LmapToFace");
}

void LmapToElement(int N_s, int N_E, int M_s, int N_N, int* invmap, double* Q, double* q) {
  int NF = dgx_nfields(&g_cfg);
  hipLaunchKernelGGL(kc_map_to_element, dim3(nb256((long)N_E * NF * N_s)), dim3(256), 0, 0, N_s, N_E, NF, M_s * N_N,
                     invmap, Q, q);
  finish("LmapToElement");
}

void Lredistribute_sf(int N_G, int N_E, double* sJ, double* fJ, double* s, double* f, double* J, double* invJac) {
  int NF = dgx_nfields(&g_cfg);
  long n = (long)N_E * N_G * NF;
  if (g_cfg.dim == 1)
    hipLaunchKernelGGL(kc_redistribute_sf<1>, dim3(nb256(n)), dim3(256), 0, 0, N_G, N_E, NF, sJ, fJ, s, f, J, invJac);
  else
    hipLaunchKernelGGL(kc_redistribute_sf<2>, dim3(nb256(n)), dim3(256), 0, 0, N_G, N_E, NF, sJ, fJ, s, f, J, invJac);
  finish("Lredistribute_sf");
}

void Lredistribute_q(int M_G, int M_T, double* qJ, double* q, double* JF) {
  int NF = dgx_nfields(&g_cfg);
  long n = (long)M_T * NF * 2 * M_G;
  hipLaunchKernelGGL(kc_redistribute_q2, dim3(nb256(n)), dim3(256), 0, 0, M_G, M_T, NF, qJ, q, JF);
  finish("Lredistribute_q");
}

void LaddSFQ(int N_s, int N_E, double* A, double* S, double* F, double* Q) {
  long n = (long)N_E * dgx_nfields(&g_cfg) * N_s;
  hipLaunchKernelGGL(kc_add_sfq, dim3(nb256(n)), dim3(256), 0, 0, n, A, S, F, Q);
  finish("LaddSFQ");
}

void Levaluate_sf(int N_G, int N_E, double* s, double* f, double* Ug, double* dUg, double* invJac) {
  long n = (long)N_E * N_G;
  dispatch_phys([&](auto d, auto ph, auto fl) -> hipError_t {
    (void)fl;
    hipLaunchKernelGGL((kc_evaluate_sf<decltype(d)::value, decltype(ph)::value>), dim3(nb256(n)), dim3(256), 0, 0, N_G,
                       N_E, s, f, Ug, dUg, invJac);
    return hipSuccess;
  });
  finish("Levaluate_sf");
}

void Levaluate_q(int M_G, int M_T, double* q, double* UgF, double* normals) {
  long n = (long)M_T * M_G;
  dispatch_phys([&](auto d, auto ph, auto fl) -> hipError_t {
    hipLaunchKernelGGL((kc_evaluate_q<decltype(d)::value, decltype(ph)::value, decltype(fl)::value>), dim3(nb256(n)),
                       dim3(256), 0, 0, M_G, M_T, q, UgF, normals);
    return hipSuccess;
  });
  finish("Levaluate_q");
}

void LrflctiveBoundary(int M_s, int M_B, int* boundaryMap, double* normals, int start, double* UF) {
  if (M_B <= 0) return;
  int NF = dgx_nfields(&g_cfg);
  long n = (long)M_B * M_s;
  if (g_cfg.dim == 1)
    hipLaunchKernelGGL(kc_rflctive<1>, dim3(nb256(n)), dim3(256), 0, 0, M_s, M_B, NF, boundaryMap, normals, start, UF);
  else
    hipLaunchKernelGGL(kc_rflctive<2>, dim3(nb256(n)), dim3(256), 0, 0, M_s, M_B, NF, boundaryMap, normals, start, UF);
  finish("LrflctiveBoundary");
}

void Lsolver(int N_s, int N_E, double Dt, double* Minv, double* f, double* DU) {
  int NF = dgx_nfields(&g_cfg);
  hipLaunchKernelGGL(kc_solve, dim3(nb256((long)N_E * NF * N_s)), dim3(256), 0, 0, N_s, N_E, NF, Dt, Minv, f, DU);
  finish("Lsolver");
}

void LfindUPA(const int N_s, const int N_E, double* U, double* UPA) {
  long n = (long)N_E * N_s;
  dispatch_phys([&](auto d, auto ph, auto fl) -> hipError_t {
    (void)fl;
    hipLaunchKernelGGL((kc_find_upa<decltype(d)::value, decltype(ph)::value>), dim3(nb256(n)), dim3(256), 0, 0, N_s,
                       N_E, U, UPA);
    return hipSuccess;
  });
  finish("LfindUPA");
}

void Lhrl1D(int N_s, int N_E, int Nfields, int N_N, int slicenum, int* neighbors, int offxy, double* A, double* Alim) {
  hipError_t e = compat_hrl(N_s, N_E, Nfields, N_N, slicenum, neighbors, offxy, A, Alim);
  if (e != hipSuccess) { hip_ok(e, "Lhrl1D"); return; }
  finish("Lhrl1D");
}

void Lhri1D(int N_s, int N_E, int N_N, int* neighbors, int N_s1D, int slicenum, int offxy, double* Lag2Mono,
            double* Mono2Lag, int* sensors, double* U, double* Unew) {
  hipError_t e = compat_hri(N_s, N_E, dgx_nfields(&g_cfg), N_N, neighbors, N_s1D, slicenum, offxy, g_cfg.dim, Lag2Mono,
                            Mono2Lag, sensors, U, Unew);
  if (e != hipSuccess) { hip_ok(e, "Lhri1D"); return; }
  finish("Lhri1D");
}

void Lcalc_sensors(int N_E, int N_N, dgx_bool sensor1, double thresh1, dgx_bool sensor2, double thresh2, dgx_bool sensor3,
                   double thresh3, int* neighbors, double* Uavg, int* sensors) {
  Solver tmp;
  tmp.cfg = g_cfg;
  tmp.D = g_cfg.dim;
  tmp.N_E = N_E;
  tmp.N_N = N_N;
  int* buf = nullptr;
  int* flag = nullptr;
  if (!hip_ok(hipMalloc(&buf, sizeof(int) * 5 * (size_t)N_E), "Lcalc_sensors") ||
      !hip_ok(hipMalloc(&flag, sizeof(int)), "Lcalc_sensors")) {
    hipFree(buf);
    return;
  }
  tmp.flag = flag;
  int on[3] = {sensor1, sensor2, sensor3};
  double th[3] = {thresh1, thresh2, thresh3};
  hipError_t e = compat_sensors(tmp, neighbors, Uavg, sensors, on, th, buf);
  if (e != hipSuccess) hip_ok(e, "Lcalc_sensors");
  finish("Lcalc_sensors");
  hipFree(buf);
  hipFree(flag);
}

dgx_status dgx_calc_sensors_noalloc(int N_E, int N_N, dgx_bool sensor1, double thresh1, dgx_bool sensor2,
                                    double thresh2, dgx_bool sensor3, double thresh3, int* neighbors, double* Uavg,
                                    int* sensors) {
  if (N_E < 0 || N_N < 0 || (N_E > 0 && (!neighbors || !Uavg || !sensors))) {
    set_error("dgx_calc_sensors_noalloc: bad arguments");
    return DGX_EINVAL;
  }
  if (N_E > kSensorsNoallocMaxElements) {
    set_error("dgx_calc_sensors_noalloc: " + std::to_string(N_E) + " elements exceed the fixed scratch of " +
              std::to_string(kSensorsNoallocMaxElements) + " elements (use Lcalc_sensors)");
    return DGX_EINVAL;
  }
  if (N_E == 0) return DGX_OK;
  int* base = nullptr;
  if (!noalloc_scratch(&base)) return DGX_EDEVICE;
  Solver tmp;
  tmp.cfg = g_cfg;
  tmp.D = g_cfg.dim;
  tmp.N_E = N_E;
  tmp.N_N = N_N;
  // [0, 5*N_E): hit targets and values, two activity vectors, initial flags;
  // the sweep flag after the largest such span, so its offset never moves
  tmp.flag = base + 5 * (size_t)kSensorsNoallocMaxElements;
  int on[3] = {sensor1, sensor2, sensor3};
  double th[3] = {thresh1, thresh2, thresh3};
  hipError_t e = compat_sensors(tmp, neighbors, Uavg, sensors, on, th, base);
  if (e != hipSuccess) {
    hip_ok(e, "dgx_calc_sensors_noalloc");
    return DGX_EDEVICE;
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return hip_ok(e, "dgx_calc_sensors_noalloc") ? DGX_OK : DGX_EDEVICE;
}

void Lcopy_detected(int N_s, int N_E, int* sensors, double* Uold, double* U) {
  hipError_t e = compat_copy_detected(N_s, N_E, dgx_nfields(&g_cfg), sensors, Uold, U);
  if (e != hipSuccess) { hip_ok(e, "Lcopy_detected"); return; }
  finish("Lcopy_detected");
}

// ----------------------------------------------- helper launchers (m2, pr, p0, output)
static void run(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    hip_ok(e, what);
    return;
  }
  finish(what);
}

void Lstridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB, double* A,
                  double* B) {
  run(compat_stridedcopy(numblocks, blocklen, strideA, strideB, offsetA, offsetB, A, B), "Lstridedcopy");
}

void Lreconstruct_energy(int N_s, int N_E, int slicenum, double* rhoeLim, double* KLim, double* EMono, double* ELim) {
  run(compat_reconstruct_energy(N_s, N_E, slicenum, rhoeLim, KLim, EMono, ELim), "Lreconstruct_energy");
}

void Linternal_energy_multifluid(int N_s, int N_E, int slicenum, double* p, double* g, double* rhoe) {
  run(compat_internal_energy(N_s, N_E, slicenum, p, g, nullptr, rhoe), "Linternal_energy_multifluid");
}

void Linternal_energy_stiffened(int N_s, int N_E, int slicenum, double* p, double* g, double* b, double* rhoe) {
  run(compat_internal_energy(N_s, N_E, slicenum, p, g, b, rhoe), "Linternal_energy_stiffened");
}

void Lm2i1D(int N_s, int N_E, int N_N, int* neighbors, int N_s1D, int slicenum, int offxy, double* Lag2Mono,
            double* Mono2Lag, int* sensors, double* U, double* Unew) {
  (void)N_s; (void)slicenum;
  if (g_cfg.physics != DGX_MULTIFLUID || N_s1D < 2 || N_s1D > 4) {
    set_error("Lm2i1D: two-fluid states of order 1..3 only");
    return;
  }
  run(compat_m2i(g_cfg.dim, N_s1D, N_E, N_N, neighbors, offxy, Lag2Mono, Mono2Lag, sensors, U, Unew), "Lm2i1D");
}

void Lp0i(int N_s, int N_E, int N_G, double refArea, int* sensors, double* phi_w, double* U, double* Unew) {
  run(compat_p0i(N_s, N_E, dgx_nfields(&g_cfg), N_G, refArea, sensors, phi_w, U, Unew), "Lp0i");
}

void LChangeBasis(int size1, int size2, int N_E, double* Transform, double* U, double* Unew) {
  run(compat_change_basis(size1, size2, N_E, dgx_nfields(&g_cfg), Transform, U, Unew), "LChangeBasis");
}

void LPrim2Cons(int N_s, int N_E, double* U) {
  if (g_cfg.physics == DGX_SINGLEFLUID) {
    set_error("LPrim2Cons: single-fluid states have no alpha field");
    return;
  }
  run(compat_cons_prim(g_cfg, N_s, N_E, U, 0), "LPrim2Cons");
}

void LCons2Prim(int N_s, int N_E, double* U) {
  if (g_cfg.physics == DGX_SINGLEFLUID) {
    set_error("LCons2Prim: single-fluid states have no alpha field");
    return;
  }
  run(compat_cons_prim(g_cfg, N_s, N_E, U, 1), "LCons2Prim");
}

void Lpressure(int N_s, int N_E, double* U, double* p) { run(compat_pressure(g_cfg, N_s, N_E, U, p), "Lpressure"); }

void Lkinetic_energy1D(int N_s, int N_E, double* rho, double* rhou, double* K) {
  run(compat_kinetic(1, N_s, N_E, rho, rhou, nullptr, K), "Lkinetic_energy1D");
}

void Lkinetic_energy2D(int N_s, int N_E, double* rho, double* rhou, double* rhov, double* K) {
  run(compat_kinetic(2, N_s, N_E, rho, rhou, rhov, K), "Lkinetic_energy2D");
}

void Laverage_cell_p0(const int N_s, const int N_E, double* DU) {
  run(compat_average_p0(N_s, N_E, dgx_nfields(&g_cfg), DU), "Laverage_cell_p0");
}

void Lformater(int N_s, int N_E, double* U, double* output, dgx_bool inverse) {
  Solver tmp;
  tmp.cfg = g_cfg;
  tmp.D = g_cfg.dim;
  tmp.N_s = N_s;
  tmp.N_E = N_E;
  run(launch_format(tmp, U, output, inverse ? 1 : 0, 0), "Lformater");
}

void Lformat_sensor(int N_s, int N_E, int* sensor, double* output) {
  run(compat_format_sensor(N_s, N_E, sensor, output), "Lformat_sensor");
}

void dgx_blasDgemm(char transa, char transb, int M, int N, int K, double alpha, const double* A, int lda,
                   const double* B, int ldb, double beta, double* C, int ldc) {
  (void)transb;
  int tA = (transa == 'T' || transa == 't');
  hipLaunchKernelGGL(kc_dgemm, dim3(nb256((long)M * N)), dim3(256), 0, 0, tA, M, N, K, alpha, A, lda, B, ldb, beta, C,
                     ldc);
  finish("dgx_blasDgemm");
}

void dgx_blasDaxpy(int n, double alpha, const double* x, int incx, double* y, int incy) {
  hipLaunchKernelGGL(kc_daxpy, dim3(nb256(n)), dim3(256), 0, 0, n, alpha, x, incx, y, incy);
  finish("dgx_blasDaxpy");
}

void dgx_blasDcopy(int n, const double* x, int incx, double* y, int incy) {
  hipLaunchKernelGGL(kc_dcopy, dim3(nb256(n)), dim3(256), 0, 0, n, x, incx, y, incy);
  finish("dgx_blasDcopy");
}

int dgx_blasIdamax(int n, const double* x, int incx) {
  if (n <= 0) return 0;
  unsigned long long* bv = nullptr;
  int* bi = nullptr;
  if (!hip_ok(hipMalloc(&bv, sizeof(unsigned long long)), "dgx_blasIdamax") ||
      !hip_ok(hipMalloc(&bi, sizeof(int)), "dgx_blasIdamax")) {
    hipFree(bv);
    return 0;
  }
  unsigned long long z = 0;
  int big = n;
  hipMemcpy(bv, &z, sizeof(z), hipMemcpyHostToDevice);
  hipMemcpy(bi, &big, sizeof(big), hipMemcpyHostToDevice);
  hipLaunchKernelGGL(kc_idamax, dim3(nb256(n)), dim3(256), 0, 0, n, x, incx, bv, bi);
  hipLaunchKernelGGL(kc_idamax_index, dim3(nb256(n)), dim3(256), 0, 0, n, x, incx, bv, bi);
  int idx = 0;
  double x0 = 0;
  hipMemcpy(&idx, bi, sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(&x0, x, sizeof(double), hipMemcpyDeviceToHost);
  finish("dgx_blasIdamax");
  hipFree(bv);
  hipFree(bi);
  if (x0 != x0 || idx >= n) return 1;  // a NaN first entry, or every entry NaN after it
  return idx + 1;
}

}  // extern "C"
