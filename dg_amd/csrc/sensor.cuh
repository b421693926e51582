// The arithmetic of the troubled-cell sensor (calc_sensors,
// sensor_kernels.cu:19-163), shared by the sensor kernels (limiter.hip) and
// the learned indicator's shock guard (mlp.hip) so that both take every
// decision on the same bits.  No FMA contraction, as in the CPU path.
#pragma once

#include "physics.cuh"

namespace dgx {

// Uavg = sum_l (1/N_s * U_l) of one nodal column, in the reference's order
// (blasGemm with M = 1 and a column of ones, sensor.cc:29).
template <int Ns>
__device__ inline double sensor_nodal_mean(const double* __restrict__ u) {
#pragma clang fp contract(off)
  const double w = 1.0 / (double)Ns;
  double c = 0.0;
#pragma unroll
  for (int l = 0; l < Ns; ++l) c = c + (w * u[l]) * 1.0;
  return c;
}

// The quantities of one cell from its averages (sensor_kernels.cu:45-75):
// passive states use the single-fluid gamma, stiffened states read beta after
// alpha (sensor_kernels.cu:57-64).
struct SensorCell {
  double rho, vx, vy, alpha, gamma, p, i, a, H;
};

template <int D, int PH>
__device__ inline SensorCell sensor_cell(const double* __restrict__ Uavg, int e) {
#pragma clang fp contract(off)
  constexpr int NF = Fields<D, PH>::NF;
  SensorCell c;
  double beta = 0;
  c.vy = 0;
  int fc = 0;
  c.rho = Uavg[e * NF + fc]; fc++;
  c.vx = Uavg[e * NF + fc] / c.rho; fc++;
  if (D == 2) { c.vy = Uavg[e * NF + fc] / c.rho; fc++; }
  const double Et = Uavg[e * NF + fc]; fc++;
  c.alpha = (PH == 0 || PH == 2) ? 1 / (kGammaSF - 1) : Uavg[e * NF + fc];
  if (PH == 3) beta = Uavg[e * NF + fc + 1];
  c.gamma = 1.0 + 1.0 / c.alpha;
  const double pinf = (1 - 1.0 / c.gamma) * beta;
  c.p = (c.gamma - 1) * (Et - beta - 0.5 * c.rho * (c.vx * c.vx + c.vy * c.vy));
  c.i = c.p * c.alpha;
  c.a = sqrt((c.gamma * (c.p + pinf)) / c.rho);
  c.H = (Et + c.p) / c.rho;
  return c;
}

// Roe averages of a pair, L = the tested element, R = its neighbour
// (sensor_kernels.cu:104-112).
struct SensorRoe {
  double vx, vy, alpha, gamma, a, i;
};

__device__ inline SensorRoe sensor_roe(const SensorCell& L, const SensorCell& R) {
#pragma clang fp contract(off)
  SensorRoe r;
  const double RT = sqrt(R.rho / L.rho);
  r.vx = (L.vx + RT * R.vx) / (1 + RT);
  r.vy = (L.vy + RT * R.vy) / (1 + RT);
  r.alpha = (L.alpha + RT * R.alpha) / (1 + RT);
  r.gamma = 1 + 1.0 / r.alpha;
  const double H = (L.H + RT * R.H) / (1 + RT);
  r.a = sqrt((r.gamma - 1) * (H - 0.5 * (r.vx * r.vx + r.vy * r.vy)));
  r.i = (L.i + RT * R.i) / (1 + RT);
  return r;
}

// Test 1, contact (sensor_kernels.cu:113-124).
__device__ inline bool sensor_contact(const SensorCell& L, const SensorCell& R, const SensorRoe& m, double th) {
#pragma clang fp contract(off)
  const double drho = R.rho - L.rho;
  const double dp = (m.gamma - 1) * (m.gamma - 1) * (m.alpha * (R.i - L.i) - m.i * (R.alpha - L.alpha));
  const double dV2 = drho - dp / (m.a * m.a);
  const double G = fabs(dV2) / (L.rho + R.rho);
  const double PHI = 2 * G / ((1 + G) * (1 + G));
  return PHI > th;
}

// The pressure-jump measure of test 2, W = 2 psi / (1 + psi)^2 with
// psi = |pR - pL| / (pL + pR) (sensor_kernels.cu:131-132).  It does not depend
// on which cell is tested, so W > th is a necessary condition for the shock
// test in either order.
__device__ inline double sensor_shock_w(double pL, double pR) {
#pragma clang fp contract(off)
  const double PSI = fabs(pR - pL) / (pL + pR);
  return 2 * PSI / ((1 + PSI) * (1 + PSI));
}

// Test 2, shock: converging characteristics and a pressure jump
// (sensor_kernels.cu:125-136).
__device__ inline bool sensor_shock(const SensorCell& L, const SensorCell& R, const SensorRoe& m, double th) {
#pragma clang fp contract(off)
  if (((L.vx - L.a > m.vx - m.a) && (m.vx - m.a > R.vx - R.a)) ||
      ((L.vy - L.a > m.vy - m.a) && (m.vy - m.a > R.vy - R.a)))
    return sensor_shock_w(L.p, R.p) > th;
  return false;
}

// The shock guard of a pair (e, n): the shock test fires with e as the tested
// cell or with n as the tested cell.  pe, pn: the cells' sensor pressures
// (SensorCell::p), which decide the necessary W > th first.
template <int D, int PH>
__device__ inline bool sensor_shock_pair(const double* __restrict__ Uavg, int e, int n, double pe, double pn,
                                         double th) {
  if (!(sensor_shock_w(pe, pn) > th)) return false;
  const SensorCell C = sensor_cell<D, PH>(Uavg, e), R = sensor_cell<D, PH>(Uavg, n);
  return sensor_shock(C, R, sensor_roe(C, R), th) || sensor_shock(R, C, sensor_roe(R, C), th);
}

// The sensor's averages of a 2D p=2 two-fluid element (nodal values
// u[f*9 + i]) into ua[0..4], and its sensor pressure (SensorCell::p) into *p:
// what the shock guard reads, written by the producers of the indicator's
// cell means.
__device__ inline void sensor_averages_2d(const double* u, double* ua, double* p) {
  double v[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) v[f] = ua[f] = sensor_nodal_mean<9>(u + f * 9);
  *p = sensor_cell<2, 1>(v, 0).p;
}

// Test 3, material (gamma) interface (sensor_kernels.cu:137-145).
__device__ inline bool sensor_gamma(const SensorCell& L, const SensorCell& R, double th) {
#pragma clang fp contract(off)
  const double G = fabs(R.gamma - L.gamma) / (L.gamma + R.gamma);
  const double PHI = 2 * G / ((1 + G) * (1 + G));
  return PHI > th;
}

}  // namespace dgx
