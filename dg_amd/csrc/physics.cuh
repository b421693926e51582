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

template <int D, int PHYS>
struct Fields {
  static constexpr int NF = 2 + D + PHYS;  // rho, momentum(D), E, [alpha]
};

constexpr double kGammaSF = 1.4;  // constants::GLOBAL_GAMMA

__host__ __device__ inline double dmax(double a, double b) { return a > b ? a : b; }

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
  const double irL = 1.0 / rhoL, irR = 1.0 / rhoR;
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
  const double i1RT = 1.0 / (1 + RT);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) * i1RT;
  double vy = (vyL + RT * vyR) * i1RT;
  double vn = vx * nx + vy * ny, vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) * i1RT;
  double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  const double ia = 1.0 / a, ia2 = ia * ia;
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
  const double irL = 1.0 / rhoL, irR = 1.0 / rhoR;
  double vxL = uL[1] * irL, vxR = uR[1] * irR, vyL = uL[2] * irL, vyR = uR[2] * irR;
  double EtL = uL[3], EtR = uR[3], alphaL = uL[4], alphaR = uR[4];
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny, vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty, vtR = vxR * tx + vyR * ty;
  double gammaL = 1.0 + 1.0 / alphaL, gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) * irL, HR = (EtR + pR) * irR;
  double iL = pL * alphaL, iR = pR * alphaR;
  double RT = sqrt(rhoR * irL);
  const double i1RT = 1.0 / (1 + RT);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) * i1RT;
  double vy = (vyL + RT * vyR) * i1RT;
  double vn = vx * nx + vy * ny, vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) * i1RT;
  double alpha = (alphaL + RT * alphaR) * i1RT;
  double gamma = 1 + 1.0 / alpha;
  double a = sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  const double ia = 1.0 / a, ia2 = ia * ia;
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

// ---------------------------------------------------------------------------
// Volume flux f[k*D+a] and the non-conservative alpha source at one point
// (physics.cu:96-142, 222-300 with gravity zero).  dua: reference-space
// gradient of alpha (D values); ij: inverse Jacobian [alpha*D+a].
// ---------------------------------------------------------------------------
template <int D, int PHYS>
__device__ inline void volume_flux(const double* u, const double* dua, const double* ij, double* f, double& salpha) {
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
    const double ir = 1.0 / rho;
    double vx = u[1] * ir, vy = u[2] * ir, Et = u[3];
    double gamma = (PHYS == 0) ? kGammaSF : 1 + 1.0 / u[4];
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

// |u|+a (1D) or max(|u|+a, |v|+a) (2D) at one node (upa.h:41-83).
template <int D, int PHYS>
__device__ inline double upa(const double* u) {
  if constexpr (D == 1) {
    double rho = u[0], vx = u[1] / u[0], E = u[2];
    double gamma = (PHYS == 0) ? kGammaSF : 1.0 + 1.0 / u[3];
    double p = (gamma - 1) * (E - 0.5 * rho * vx * vx);
    return fabs(vx) + sqrt(gamma * p / rho);
  } else {
    double rho = u[0], vx = u[1] / u[0], vy = u[2] / u[0], E = u[3];
    double gamma = (PHYS == 0) ? kGammaSF : 1.0 + 1.0 / u[4];
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

}  // namespace dgx
