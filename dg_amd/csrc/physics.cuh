// Pointwise Euler physics for the HIP kernels: volume fluxes
// (physics.cu:27-363) and the Rusanov/Roe Riemann solvers
// (face_fluxes/{oned,twod}_{singlefluid,multifluid}_fluxes.h), templated on
// dimension D, physics PHYS (0 single fluid gamma=1.4, 1 two-fluid gamma
// model with alpha = 1/(gamma-1) advected non-conservatively, GAMNCON) and
// flux FLUX (0 Rusanov, 1 Roe).  Expressions keep the reference's operand
// order so results differ from the FP64 CPU path only by FMA contraction.
#pragma once
#include <hip/hip_runtime.h>

namespace dgx {

template <int D, int P>
struct Shape {
  static constexpr int Ns = (D == 1) ? (P + 1) : (P + 1) * (P + 1);
  static constexpr int NG = (D == 1) ? (P + 1) : (P == 1 ? 4 : (P + 2) * (P + 2));
  static constexpr int Ms = (D == 1) ? 1 : (P + 1);
  static constexpr int MG = (D == 1) ? 1 : (P + 1);
  static constexpr int NN = 2 * D;
  static constexpr int N1 = P + 1;
};

// Physics codes: 0 single fluid (gamma = 1.4); 1 two-fluid, alpha =
// 1/(gamma-1) advected non-conservatively (GAMNCON); 2 passive scalars,
// rho*phic conservative and phinc non-conservative (single-fluid gamma);
// 3 stiffened gas, alpha and beta = gamma*pinf/(gamma-1) both
// non-conservative.  The non-conservative fields are the last NNC fields.
template <int D, int PHYS>
struct Fields {
  static constexpr int NX = PHYS == 0 ? 0 : (PHYS == 1 ? 1 : 2);  // fields after rho, momentum, E
  static constexpr int NF = 2 + D + NX;
  static constexpr int NNC = PHYS == 0 ? 0 : (PHYS == 3 ? 2 : 1);  // non-conservative fields (no volume flux)
};

constexpr double kGammaSF = 1.4;  // constants::GLOBAL_GAMMA

__host__ __device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// 1/x for the positive, normal operands of the flux kernels (densities,
// 1 + sqrt ratios, sound speeds, 1/(gamma-1)): the hardware reciprocal
// refined by two Newton steps, within an ulp or two of the correctly
// rounded quotient, without the scaling and special-case fix-up of IEEE
// division.
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// ---------------------------------------------------------------------------
// Riemann solvers.  uL/uR: conserved face states; n: unit normal (side 0
// outward).  Outputs F (flux) and nc (non-conservative term), both NF long.
// ---------------------------------------------------------------------------
template <int D, int PHYS, int FLUX>
__device__ inline void riemann(const double* uL, const double* uR, double nx, double ny, double* F, double* nc);

// 1D single fluid
template <>
__device__ inline void riemann<1, 0, 0>(const double* uL, const double* uR, double nx, double, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0], vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], EtL = uL[2], EtR = uR[2];
  double gamma = kGammaSF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double aL = sqrt((gamma * pL) / rhoL), aR = sqrt((gamma * pR) / rhoR);
  double maxvap = dmax(fabs(vxL) + aL, fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
  nc[0] = nc[1] = nc[2] = 0;
}
template <>
__device__ inline void riemann<1, 0, 1>(const double* uL, const double* uR, double nx, double, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0], vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], EtL = uL[2], EtR = uR[2];
  double gamma = kGammaSF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double RT = sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
  double H = (HL + RT * HR) / (1 + RT);
  double a = sqrt((gamma - 1) * (H - 0.5 * v * v));
  double Dp = pR - pL;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double ws0 = fabs(v - a), ws1 = fabs(v), ws2 = fabs(v + a);
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx - 0.5 * (ws0 * dV0 + ws1 * dV1 + ws2 * dV2);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx -
         0.5 * (ws0 * dV0 * (v - a) + ws1 * dV1 * v + ws2 * dV2 * (v + a));
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx -
         0.5 * (ws0 * dV0 * (H - v * a) + ws1 * dV1 * (0.5 * v * v) + ws2 * dV2 * (H + v * a));
  nc[0] = nc[1] = nc[2] = 0;
}

// 1D two fluid (GAMNCON)
template <>
__device__ inline void riemann<1, 1, 0>(const double* uL, const double* uR, double nx, double, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0], vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], EtL = uL[2], EtR = uR[2];
  double alphaL = uL[3], alphaR = uR[3];
  double gammaL = 1.0 + 1.0 / alphaL, gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double aL = sqrt((gammaL * pL) / rhoL), aR = sqrt((gammaR * pR) / rhoR);
  double maxvap = dmax(fabs(vxL) + aL, fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
  F[3] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[0] = nc[1] = nc[2] = 0;
  nc[3] = -0.5 * 0.5 * (vxL + vxR) * (alphaR - alphaL) * nx;
}
template <>
__device__ inline void riemann<1, 1, 1>(const double* uL, const double* uR, double nx, double, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0], vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], EtL = uL[2], EtR = uR[2];
  double alphaL = uL[3], alphaR = uR[3];
  double gammaL = 1.0 + 1.0 / alphaL, gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double RT = sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
  double H = (HL + RT * HR) / (1 + RT);
  double alpha = (alphaL + RT * alphaR) / (1 + RT);
  double gamma = 1 + 1.0 / alpha;
  double a = sqrt((gamma - 1) * (H - 0.5 * v * v));
  double iL = pL * alphaL, iR = pR * alphaR;
  double sL = sqrt(rhoL), sR = sqrt(rhoR);
  double i = (sL * iL + sR * iR) / (sL + sR);
  double Dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double p = (gamma - 1) * i;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double dV3 = alphaR - alphaL;
  double ws0 = fabs(v - a), ws1 = fabs(v), ws2 = fabs(v + a), ws3 = fabs(v);
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx - 0.5 * (ws0 * dV0 + ws1 * dV1 + ws2 * dV2 + ws3 * dV3 * 0.0);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx -
         0.5 * (ws0 * dV0 * (v - a) + ws1 * dV1 * v + ws2 * dV2 * (v + a) + ws3 * dV3 * 0.0);
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx -
         0.5 * (ws0 * dV0 * (H - v * a) + ws1 * dV1 * (0.5 * v * v) + ws2 * dV2 * (H + v * a) + ws3 * dV3 * p);
  F[3] = 0.0 + -0.5 * (ws0 * dV0 * 0.0 + ws1 * dV1 * 0.0 + ws2 * dV2 * 0.0 + ws3 * dV3 * 1.0);
  nc[0] = nc[1] = nc[2] = 0;
  nc[3] = -0.5 * v * (alphaR - alphaL) * nx;
}

// 2D single fluid
template <>
__device__ inline void riemann<2, 0, 0>(const double* uL, const double* uR, double nx, double ny, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0];
  double vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], vyL = uL[2] / uL[0], vyR = uR[2] / uR[0];
  double EtL = uL[3], EtR = uR[3];
  double vnL = vxL * nx + vyL * ny, vnR = vxR * nx + vyR * ny;
  double gamma = kGammaSF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = sqrt((gamma * pL) / rhoL), aR = sqrt((gamma * pR) / rhoR);
  double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
  double maxvap = dmax(fabs(vnL) + aL, fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
  nc[0] = nc[1] = nc[2] = nc[3] = 0;
}
template <>
__device__ inline void riemann<2, 0, 1>(const double* uL, const double* uR, double nx, double ny, double* F, double* nc) {
  // divisions by a common divisor as products with its reciprocal (FP64
  // division is a long sequence here; a few ulps against the reference)
  double rhoL = uL[0], rhoR = uR[0];
  const double irL = rcp_nr(rhoL), irR = rcp_nr(rhoR);
  double vxL = uL[1] * irL, vxR = uR[1] * irR, vyL = uL[2] * irL, vyR = uR[2] * irR;
  double EtL = uL[3], EtR = uR[3];
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny, vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty, vtR = vxR * tx + vyR * ty;
  double gamma = kGammaSF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) * irL, HR = (EtR + pR) * irR;
  double RT = sqrt(rhoR * irL);
  const double i1RT = rcp_nr(1 + RT);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) * i1RT;
  double vy = (vyL + RT * vyR) * i1RT;
  double vn = vx * nx + vy * ny, vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) * i1RT;
  double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  const double ia = rcp_nr(a), ia2 = ia * ia;
  double drho = rhoR - rhoL, dp = pR - pL, dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) * (0.5 * ia2);
  double dV1 = rho * dvt * ia;
  double dV2 = drho - dp * ia2;
  double dV3 = (dp + rho * a * dvn) * (0.5 * ia2);
  double w0 = fabs(vn - a) * dV0, w1 = fabs(vn) * dV1, w2 = fabs(vn) * dV2, w3 = fabs(vn + a) * dV3;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) - 0.5 * (w0 * 1.0 + w1 * 0.0 + w2 * 1.0 + w3 * 1.0);
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (w0 * (vx - a * nx) + w1 * (a * tx) + w2 * vx + w3 * (vx + a * nx));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (w0 * (vy - a * ny) + w1 * (a * ty) + w2 * vy + w3 * (vy + a * ny));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) -
         0.5 * (w0 * (H - vn * a) + w1 * (vt * a) + w2 * (0.5 * (vx * vx + vy * vy)) + w3 * (H + vn * a));
  nc[0] = nc[1] = nc[2] = nc[3] = 0;
}

// 2D two fluid (GAMNCON)
template <>
__device__ inline void riemann<2, 1, 0>(const double* uL, const double* uR, double nx, double ny, double* F, double* nc) {
  double rhoL = uL[0], rhoR = uR[0];
  double vxL = uL[1] / uL[0], vxR = uR[1] / uR[0], vyL = uL[2] / uL[0], vyR = uR[2] / uR[0];
  double EtL = uL[3], EtR = uR[3], alphaL = uL[4], alphaR = uR[4];
  double vnL = vxL * nx + vyL * ny, vnR = vxR * nx + vyR * ny;
  double gammaL = 1.0 + 1.0 / alphaL, gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = sqrt((gammaL * pL) / rhoL), aR = sqrt((gammaR * pR) / rhoR);
  double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
  double maxvap = dmax(fabs(vnL) + aL, fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
  F[4] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[0] = nc[1] = nc[2] = nc[3] = 0;
  nc[4] = -0.5 * 0.5 * (vnL + vnR) * (alphaR - alphaL);
}
template <>
__device__ inline void riemann<2, 1, 1>(const double* uL, const double* uR, double nx, double ny, double* F, double* nc) {
  // divisions by a common divisor as products with its reciprocal (FP64
  // division is a long sequence here; a few ulps against the reference)
  double rhoL = uL[0], rhoR = uR[0];
  const double irL = rcp_nr(rhoL), irR = rcp_nr(rhoR);
  double vxL = uL[1] * irL, vxR = uR[1] * irR, vyL = uL[2] * irL, vyR = uR[2] * irR;
  double EtL = uL[3], EtR = uR[3], alphaL = uL[4], alphaR = uR[4];
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny, vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty, vtR = vxR * tx + vyR * ty;
  double gammaL = 1.0 + rcp_nr(alphaL), gammaR = 1.0 + rcp_nr(alphaR);
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) * irL, HR = (EtR + pR) * irR;
  double iL = pL * alphaL, iR = pR * alphaR;
  double RT = sqrt(rhoR * irL);
  const double i1RT = rcp_nr(1 + RT);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) * i1RT;
  double vy = (vyL + RT * vyR) * i1RT;
  double vn = vx * nx + vy * ny, vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) * i1RT;
  double alpha = (alphaL + RT * alphaR) * i1RT;
  double gamma = 1 + rcp_nr(alpha);
  double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  const double ia = rcp_nr(a), ia2 = ia * ia;
  double i = (iL + RT * iR) * i1RT;
  double p = (gamma - 1) * i;
  double drho = rhoR - rhoL;
  double dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) * (0.5 * ia2);
  double dV1 = rho * dvt * ia;
  double dV2 = drho - dp * ia2;
  double dV3 = (dp + rho * a * dvn) * (0.5 * ia2);
  double dV4 = alphaR - alphaL;
  double w0 = fabs(vn - a) * dV0, w1 = fabs(vn) * dV1, w2 = fabs(vn) * dV2, w3 = fabs(vn + a) * dV3,
         w4 = fabs(vn) * dV4;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) - 0.5 * (w0 * 1.0 + w1 * 0.0 + w2 * 1.0 + w3 * 1.0 + w4 * 0.0);
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (w0 * (vx - a * nx) + w1 * (a * tx) + w2 * vx + w3 * (vx + a * nx) + w4 * 0.0);
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (w0 * (vy - a * ny) + w1 * (a * ty) + w2 * vy + w3 * (vy + a * ny) + w4 * 0.0);
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) -
         0.5 * (w0 * (H - vn * a) + w1 * (vt * a) + w2 * (0.5 * (vx * vx + vy * vy)) + w3 * (H + vn * a) + w4 * p);
  F[4] = 0.0 + -0.5 * (w0 * 0.0 + w1 * 0.0 + w2 * 0.0 + w3 * 0.0 + w4 * 1.0);
  nc[0] = nc[1] = nc[2] = nc[3] = 0;
  nc[4] = -0.5 * vn * (alphaR - alphaL);
}

// Passive scalars (PHYS 2): oned_passive_fluxes.h:26-65 (Rusanov), 175-266
// (Roe); twod_passive_fluxes.h:27-75, 197-323.  The Euler part is the
// single-fluid solver's; rho*phic is carried with the Rusanov spread or the
// Roe density waves, phinc by its non-conservative term.
template <int D, int FL>
__device__ inline void riemann_passive(const double* uL, const double* uR, double nx, double ny, double* F,
                                       double* nc) {
  constexpr int NF = Fields<D, 2>::NF;
  const double rhoL = uL[0], rhoR = uR[0];
  const double phicL = uL[D + 2] / rhoL, phicR = uR[D + 2] / rhoR, phincL = uL[D + 3], phincR = uR[D + 3];
  riemann<D, 0, FL>(uL, uR, nx, ny, F, nc);
#pragma unroll
  for (int k = D + 2; k < NF; ++k) nc[k] = 0;
  const double vxL = uL[1] / rhoL, vxR = uR[1] / rhoR;
  const double vyL = D == 2 ? uL[2] / rhoL : 0.0, vyR = D == 2 ? uR[2] / rhoR : 0.0;
  const double EtL = uL[D + 1], EtR = uR[D + 1];
  const double gamma = kGammaSF;
  const double vnL = D == 2 ? vxL * nx + vyL * ny : vxL, vnR = D == 2 ? vxR * nx + vyR * ny : vxR;
  const double pL = D == 2 ? (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL))
                           : (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  const double pR = D == 2 ? (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR))
                           : (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  if constexpr (FL == 0) {
    const double aL = sqrt((gamma * pL) / rhoL), aR = sqrt((gamma * pR) / rhoR);
    const double maxvap = dmax(fabs(vnL) + aL, fabs(vnR) + aR);
    if constexpr (D == 1) {
      F[3] = 0.5 * ((rhoL * vxL * phicL + rhoR * vxR * phicR) * nx - maxvap * (rhoR * phicR - rhoL * phicL));
      F[4] = 0.5 * (-maxvap * (phincR - phincL));
      nc[4] = 0.5 * 0.5 * (vxL + vxR) * (phincL - phincR) * nx;
    } else {
      F[4] = 0.5 * (rhoL * vnL * phicL + rhoR * vnR * phicR - maxvap * (rhoR * phicR - rhoL * phicL));
      F[5] = 0.5 * (-maxvap * (phincR - phincL));
      nc[5] = 0.5 * 0.5 * (vnL + vnR) * (phincL - phincR);
    }
  } else {
    const double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
    const double RT = sqrt(rhoR / rhoL);
    const double rho = RT * rhoL;
    const double vx = (vxL + RT * vxR) / (1 + RT), vy = (vyL + RT * vyR) / (1 + RT);
    const double H = (HL + RT * HR) / (1 + RT);
    if constexpr (D == 1) {
      const double a = sqrt((gamma - 1) * (H - 0.5 * vx * vx));
      const double Dp = pR - pL;
      const double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
      const double dV1 = (rhoR - rhoL) - Dp / (a * a);
      const double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
      const double w = fabs(vx - a) * dV0 * 1.0 + fabs(vx) * dV1 * 1.0 + fabs(vx + a) * dV2 * 1.0;
      F[3] = 0.5 * (rhoL * vxL * phicL + rhoR * vxR * phicR) * nx - 0.5 * w;
      F[4] = 0.0 - 0.5 * w;
      nc[4] = -0.5 * vx * (phincR - phincL) * nx;
    } else {
      const double tx = -ny, ty = nx;
      const double vn = vx * nx + vy * ny;
      const double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
      const double dp = pR - pL, dvn = vnR - vnL, dvt = (vxR * tx + vyR * ty) - (vxL * tx + vyL * ty);
      const double dV0 = (dp - rho * a * dvn) / (2 * a * a);
      const double dV1 = rho * dvt / a;
      const double dV2 = (rhoR - rhoL) - dp / (a * a);
      const double dV3 = (dp + rho * a * dvn) / (2 * a * a);
      const double w = fabs(vn - a) * dV0 * 1.0 + fabs(vn) * dV1 * 0.0 + fabs(vn) * dV2 * 1.0 + fabs(vn + a) * dV3 * 1.0;
      F[4] = 0.5 * (rhoL * vnL * phicL + rhoR * vnR * phicR) - 0.5 * w;
      F[5] = -0.5 * w;
      nc[5] = 0.5 * vn * (phincL - phincR);
    }
  }
}

// Stiffened gas (PHYS 3): oned_stiffened_fluxes.h:29-88 (Rusanov), 142-295
// (Roe); twod_stiffened_fluxes.h:30-104, 153-362.
template <int D, int FL>
__device__ inline void riemann_stiffened(const double* uL, const double* uR, double nx, double ny, double* F,
                                         double* nc) {
  constexpr int NF = Fields<D, 3>::NF;
  const double rhoL = uL[0], rhoR = uR[0];
  const double vxL = uL[1] / rhoL, vxR = uR[1] / rhoR;
  const double vyL = D == 2 ? uL[2] / rhoL : 0.0, vyR = D == 2 ? uR[2] / rhoR : 0.0;
  const double EtL = uL[D + 1], EtR = uR[D + 1];
  const double alphaL = uL[D + 2], alphaR = uR[D + 2], betaL = uL[D + 3], betaR = uR[D + 3];
  const double gammaL = 1.0 + 1.0 / alphaL, gammaR = 1.0 + 1.0 / alphaR;
  const double pL = D == 2 ? (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * (vxL * vxL + vyL * vyL))
                           : (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * vxL * vxL);
  const double pR = D == 2 ? (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * (vxR * vxR + vyR * vyR))
                           : (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * vxR * vxR);
#pragma unroll
  for (int k = 0; k < NF; ++k) nc[k] = 0;
  const double vnL = D == 2 ? vxL * nx + vyL * ny : vxL, vnR = D == 2 ? vxR * nx + vyR * ny : vxR;
  if constexpr (FL == 0) {
    const double pinfL = (1 - 1.0 / gammaL) * betaL, pinfR = (1 - 1.0 / gammaR) * betaR;
    const double aL = sqrt((gammaL * (pL + pinfL)) / rhoL), aR = sqrt((gammaR * (pR + pinfR)) / rhoR);
    const double maxvap = dmax(fabs(vnL) + aL, fabs(vnR) + aR);
    if constexpr (D == 1) {
      F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
      F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
      F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
      F[3] = 0.5 * (-maxvap * (alphaR - alphaL));
      nc[3] = -0.5 * 0.5 * (vxL + vxR) * (alphaR - alphaL) * nx;
      F[4] = 0.5 * (-maxvap * (betaR - betaL));
      nc[4] = -0.5 * 0.5 * (vxL + vxR) * (betaR - betaL) * nx;
    } else {
      const double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
      F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
      F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
      F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
      F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
      F[4] = 0.5 * (-maxvap * (alphaR - alphaL));
      nc[4] = -0.5 * 0.5 * (vnL + vnR) * (alphaR - alphaL);
      F[5] = 0.5 * (-maxvap * (betaR - betaL));
      nc[5] = -0.5 * 0.5 * (vnL + vnR) * (betaR - betaL);
    }
  } else {
    const double HL = (EtL + pL) / rhoL, HR = (EtR + pR) / rhoR;
    const double iL = pL * alphaL, iR = pR * alphaR;
    const double RT = sqrt(rhoR / rhoL);
    const double rho = RT * rhoL;
    const double vx = (vxL + RT * vxR) / (1 + RT), vy = (vyL + RT * vyR) / (1 + RT);
    const double H = (HL + RT * HR) / (1 + RT);
    const double alpha = (alphaL + RT * alphaR) / (1 + RT);
    const double gamma = 1 + 1.0 / alpha;
    if constexpr (D == 1) {
      const double v = vx;
      const double a = sqrt((gamma - 1) * (H - 0.5 * v * v));
      const double i = (sqrt(rhoL) * iL + sqrt(rhoR) * iR) / (sqrt(rhoL) + sqrt(rhoR));
      const double Dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
      const double p = (gamma - 1) * i;
      const double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
      const double dV1 = (rhoR - rhoL) - Dp / (a * a);
      const double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
      const double dV3 = alphaR - alphaL, dV4 = betaR - betaL;
      const double ws0 = fabs(v - a), ws1 = fabs(v), ws2 = fabs(v + a), ws3 = fabs(v), ws4 = fabs(v);
      F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx -
             0.5 * (ws0 * dV0 * 1.0 + ws1 * dV1 * 1.0 + ws2 * dV2 * 1.0 + ws3 * dV3 * 0.0 + ws4 * dV4 * 0.0);
      F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx -
             0.5 * (ws0 * dV0 * (v - a) + ws1 * dV1 * v + ws2 * dV2 * (v + a) + ws3 * dV3 * 0.0 + ws4 * dV4 * 0.0);
      F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx -
             0.5 * (ws0 * dV0 * (H - v * a) + ws1 * dV1 * (0.5 * v * v) + ws2 * dV2 * (H + v * a) + ws3 * dV3 * p +
                    ws4 * dV4 * 1.0);
      nc[3] = -0.5 * v * (alphaR - alphaL) * nx;
      F[3] = 0.0 + -0.5 * (ws0 * dV0 * 0.0 + ws1 * dV1 * 0.0 + ws2 * dV2 * 0.0 + ws3 * dV3 * 1.0 + ws4 * dV4 * 0.0);
      nc[4] = -0.5 * v * (betaR - betaL) * nx;
      F[4] = 0.0 + -0.5 * (ws0 * dV0 * 0.0 + ws1 * dV1 * 0.0 + ws2 * dV2 * 0.0 + ws3 * dV3 * 0.0 + ws4 * dV4 * 1.0);
    } else {
      const double tx = -ny, ty = nx;
      const double vn = vx * nx + vy * ny, vt = vx * tx + vy * ty;
      const double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
      const double i = (iL + RT * iR) / (1 + RT);
      const double p = (gamma - 1) * i;
      const double dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
      const double dvn = vnR - vnL, dvt = (vxR * tx + vyR * ty) - (vxL * tx + vyL * ty);
      const double dV0 = (dp - rho * a * dvn) / (2 * a * a);
      const double dV1 = rho * dvt / a;
      const double dV2 = (rhoR - rhoL) - dp / (a * a);
      const double dV3 = (dp + rho * a * dvn) / (2 * a * a);
      const double dV4 = alphaR - alphaL, dV5 = betaR - betaL;
      const double w0 = fabs(vn - a) * dV0, w1 = fabs(vn) * dV1, w2 = fabs(vn) * dV2, w3 = fabs(vn + a) * dV3;
      const double w4 = fabs(vn) * dV4, w5 = fabs(vn) * dV5;
      F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) -
             0.5 * (w0 * 1.0 + w1 * 0.0 + w2 * 1.0 + w3 * 1.0 + w4 * 0.0 + w5 * 0.0);
      F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
             0.5 * (w0 * (vx - a * nx) + w1 * (a * tx) + w2 * vx + w3 * (vx + a * nx) + w4 * 0.0 + w5 * 0.0);
      F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
             0.5 * (w0 * (vy - a * ny) + w1 * (a * ty) + w2 * vy + w3 * (vy + a * ny) + w4 * 0.0 + w5 * 0.0);
      F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) -
             0.5 * (w0 * (H - vn * a) + w1 * (vt * a) + w2 * (0.5 * (vx * vx + vy * vy)) + w3 * (H + vn * a) + w4 * p +
                    w5 * 1.0);
      nc[4] = -0.5 * vn * (alphaR - alphaL);
      F[4] = 0.0 + -0.5 * (w0 * 0.0 + w1 * 0.0 + w2 * 0.0 + w3 * 0.0 + w4 * 1.0 + w5 * 0.0);
      nc[5] = -0.5 * vn * (betaR - betaL);
      F[5] = 0.0 + -0.5 * (w0 * 0.0 + w1 * 0.0 + w2 * 0.0 + w3 * 0.0 + w4 * 0.0 + w5 * 1.0);
    }
  }
}

template <> __device__ inline void riemann<1, 2, 0>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_passive<1, 0>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<1, 2, 1>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_passive<1, 1>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<2, 2, 0>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_passive<2, 0>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<2, 2, 1>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_passive<2, 1>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<1, 3, 0>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_stiffened<1, 0>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<1, 3, 1>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_stiffened<1, 1>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<2, 3, 0>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_stiffened<2, 0>(a, b, nx, ny, F, nc); }
template <> __device__ inline void riemann<2, 3, 1>(const double* a, const double* b, double nx, double ny, double* F, double* nc) { riemann_stiffened<2, 1>(a, b, nx, ny, F, nc); }

// ---------------------------------------------------------------------------
// Volume flux f[k*D+a] and the non-conservative sources snc[m] of the last
// NNC fields at one point (physics.cu:27-363 with gravity zero).  dua:
// reference-space gradients of those fields (NNC x D values); ij: inverse
// Jacobian [alpha*D+a].
// ---------------------------------------------------------------------------
template <int D, int PHYS>
__device__ inline void volume_flux(const double* u, const double* dua, const double* ij, double* f, double* snc) {
  if constexpr (PHYS >= 2) {
    // physics.cu:57-78, 144-170 (1D); 183-220, 302-352 (2D)
    constexpr int NNC = Fields<D, PHYS>::NNC;
    const double rho = u[0];
    const double vx = u[1] / rho, vy = D == 2 ? u[2] / rho : 0.0, Et = u[D + 1];
    double gamma, beta = 0.0;
    if constexpr (PHYS == 2) {
      gamma = kGammaSF;
    } else {
      gamma = 1 + 1.0 / u[D + 2];
      beta = u[D + 3];
    }
    const double p = D == 1 ? (gamma - 1) * (Et - beta - 0.5 * rho * vx * vx)
                            : (gamma - 1) * (Et - beta - 0.5 * rho * (vx * vx + vy * vy));
    if constexpr (D == 1) {
      f[0] = rho * vx;
      f[1] = rho * vx * vx + p;
      f[2] = (Et + p) * vx;
      if constexpr (PHYS == 2) {
        f[3] = rho * vx * (u[3] / rho);
        f[4] = 0;
      } else {
        f[3] = 0;
        f[4] = 0;
      }
#pragma unroll
      for (int m = 0; m < NNC; ++m) snc[m] = -vx * dua[m] * ij[0];
    } else {
      f[0 * 2 + 0] = rho * vx;
      f[1 * 2 + 0] = rho * vx * vx + p;
      f[2 * 2 + 0] = rho * vx * vy;
      f[3 * 2 + 0] = (Et + p) * vx;
      f[0 * 2 + 1] = rho * vy;
      f[1 * 2 + 1] = rho * vx * vy;
      f[2 * 2 + 1] = rho * vy * vy + p;
      f[3 * 2 + 1] = (Et + p) * vy;
      if constexpr (PHYS == 2) {
        const double phic = u[4] / rho;
        f[4 * 2 + 0] = rho * vx * phic;
        f[4 * 2 + 1] = rho * vy * phic;
      } else {
        f[4 * 2 + 0] = 0;
        f[4 * 2 + 1] = 0;
      }
      f[5 * 2 + 0] = 0;
      f[5 * 2 + 1] = 0;
#pragma unroll
      for (int m = 0; m < NNC; ++m) {
        double s = 0;
        for (int al = 0; al < 2; al++)
          s += -vx * dua[m * 2 + al] * ij[0 * 2 + al] - vy * dua[m * 2 + al] * ij[1 * 2 + al];
        snc[m] = s;
      }
    }
    return;
  }
  double& salpha = snc[0];
  if constexpr (D == 1) {
    double rho = u[0], vx = u[1] / rho, Et = u[2];
    double gamma = (PHYS == 0) ? kGammaSF : 1 + 1.0 / u[3];
    double p = (gamma - 1) * (Et - 0.5 * rho * vx * vx);
    f[0] = rho * vx;
    f[1] = rho * vx * vx + p;
    f[2] = (Et + p) * vx;
    if constexpr (PHYS == 1) {
      f[3] = 0;
      salpha = -vx * dua[0] * ij[0];
    } else {
      salpha = 0;
    }
  } else {
    double rho = u[0];
    const double ir = rcp_nr(rho);
    double vx = u[1] * ir, vy = u[2] * ir, Et = u[3];
    double gamma = (PHYS == 0) ? kGammaSF : 1 + rcp_nr(u[4]);
    double p = (gamma - 1) * (Et - 0.5 * rho * (vx * vx + vy * vy));
    f[0 * 2 + 0] = rho * vx;
    f[1 * 2 + 0] = rho * vx * vx + p;
    f[2 * 2 + 0] = rho * vx * vy;
    f[3 * 2 + 0] = (Et + p) * vx;
    f[0 * 2 + 1] = rho * vy;
    f[1 * 2 + 1] = rho * vx * vy;
    f[2 * 2 + 1] = rho * vy * vy + p;
    f[3 * 2 + 1] = (Et + p) * vy;
    if constexpr (PHYS == 1) {
      f[4 * 2 + 0] = 0;
      f[4 * 2 + 1] = 0;
      double s = 0;
      // index use as in physics.cu:283-288
      for (int al = 0; al < 2; al++) s += -vx * dua[al] * ij[0 * 2 + al] - vy * dua[al] * ij[1 * 2 + al];
      salpha = s;
    } else {
      salpha = 0;
    }
  }
}

// |u|+a (1D) or max(|u|+a, |v|+a) (2D) at one node (upa.h:20-130).
template <int D, int PHYS>
__device__ inline double upa(const double* u) {
  if constexpr (PHYS == 3) {
    const double rho = u[0], vx = u[1] / u[0], vy = D == 2 ? u[2] / u[0] : 0.0, E = u[D + 1];
    const double alpha = u[D + 2], beta = u[D + 3];
    const double gamma = 1.0 + 1.0 / alpha;
    const double pinf = (1 - 1.0 / gamma) * beta;
    if constexpr (D == 1) {
      const double p = (gamma - 1) * (E - beta - 0.5 * rho * vx * vx);
      return fabs(vx) + sqrt(gamma * (p + pinf) / rho);
    } else {
      const double p = (gamma - 1) * (E - beta - 0.5 * rho * (vx * vx + vy * vy));
      const double a = sqrt(gamma * (p + pinf) / rho);
      return dmax(fabs(vx) + a, fabs(vy) + a);
    }
  }
  if constexpr (D == 1) {
    double rho = u[0], vx = u[1] / u[0], E = u[2];
    double gamma = (PHYS != 1) ? kGammaSF : 1.0 + 1.0 / u[3];
    double p = (gamma - 1) * (E - 0.5 * rho * vx * vx);
    return fabs(vx) + sqrt(gamma * p / rho);
  } else {
    double rho = u[0], vx = u[1] / u[0], vy = u[2] / u[0], E = u[3];
    double gamma = (PHYS != 1) ? kGammaSF : 1.0 + 1.0 / u[4];
    double p = (gamma - 1) * (E - 0.5 * rho * (vx * vx + vy * vy));
    double a = sqrt(gamma * p / rho);
    return dmax(fabs(vx) + a, fabs(vy) + a);
  }
}

// Element-local node index of position j of face closure cl (Gmsh closures,
// polynomialBasis.cc:911-934).  1D: closure c is node c.
template <int D, int P>
__host__ __device__ constexpr int closure_node(int cl, int j) {
  if (D == 1) return cl;
  if (cl < 4) {
    if (j == 0) return cl;
    if (j == 1) return (cl + 1) % 4;
    return 4 + (P - 1) * cl + (j - 2);
  }
  int k = cl - 4;
  if (j == 0) return (k + 1) % 4;
  if (j == 1) return k;
  return 4 + (P - 1) * (k + 1) - (j - 2) - 1;
}

// position in reversed closure (4+k) of the node at position j of closure k
template <int P>
__host__ __device__ constexpr int reversed_pos(int j) {
  return j == 0 ? 1 : (j == 1 ? 0 : P + 2 - j);
}

// Cell means of rho, p and alpha = 1/(gamma-1) of a 2D p=2 two-fluid element
// (nodal values u[f*9 + i]): the tensor-Legendre mode 0 of rho, the momenta,
// E and alpha (w0 = the first row of the Lagrange-to-Legendre transform),
// then p = (gamma-1)(E - (m_x^2 + m_y^2)/(2 rho)).  One function for the
// learned indicator's own pass (k_mlp_means) and the stage kernel's
// epilogue, so that both give the same bits (checked by the tests).
// Split in two for callers that form the five field means on separate lanes.
__device__ inline double cell_mode0(const double* uf, const double* w0) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) s += w0[i] * uf[i];
  return s;
}
__device__ inline void rpa_of_means(const double* m, double* avg3) {
  const double gam = 1.0 + 1.0 / m[4];
  avg3[0] = m[0];
  avg3[1] = (gam - 1.0) * (m[3] - 0.5 * (m[1] * m[1] + m[2] * m[2]) / m[0]);
  avg3[2] = m[4];
}
__device__ inline void cell_means_rpa(const double* u, const double* w0, double* avg3) {
  double m[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) m[f] = cell_mode0(u + f * 9, w0);
  rpa_of_means(m, avg3);
}

}  // namespace dgx
