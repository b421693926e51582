// CPU oracle: FP64 restatement of the reference's CPU path for the DG
// right-hand side, RK4, the hierarchical-reconstruction limiter and the
// troubled-cell sensor.  TEST INFRASTRUCTURE ONLY (see dg_oracle.h).
//
// Arithmetic follows the reference expression by expression so that, built
// with the same compiler flags, it reproduces the reference's launchers bit
// for bit; the dense products follow netlib dgemm's loop order.
#include "dg_oracle.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

inline double dmax(double a, double b) { return a > b ? a : b; }
inline int signum(double v) { return v > 0 ? 1 : (v < 0 ? -1 : 0); }

// ---------------------------------------------------------------------------
// Riemann solvers (face_fluxes/*.h).  F and nc have N_F entries, zeroed by
// the caller (physics.cu:390-392).
// ---------------------------------------------------------------------------
const double GAMMA_SF = 1.4;  // constants::GLOBAL_GAMMA (constants.h:16)

// oned_singlefluid_fluxes.h:27-59
void rus_1d_sf(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double nx, double* F) {
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double aL = std::sqrt((gamma * pL) / rhoL);
  double aR = std::sqrt((gamma * pR) / rhoR);
  double maxvap = dmax(std::fabs(vxL) + aL, std::fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
}

// oned_singlefluid_fluxes.h:106-173
void roe_1d_sf(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double nx, double* F) {
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double H = (HL + RT * HR) / (1 + RT);
  double a = std::sqrt((gamma - 1) * (H - 0.5 * v * v));
  double Dp = pR - pL;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double ws0 = std::fabs(v - a), ws1 = std::fabs(v), ws2 = std::fabs(v + a);
  double R01 = v - a, R02 = H - v * a;
  double R11 = v, R12 = 0.5 * v * v;
  double R21 = v + a, R22 = H + v * a;
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx - 0.5 * (ws0 * dV0 * 1.0 + ws1 * dV1 * 1.0 + ws2 * dV2 * 1.0);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - 0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21);
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - 0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22);
}

// twod_singlefluid_fluxes.h:31-72
void rus_2d_sf(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double nx, double ny, double* F) {
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = std::sqrt((gamma * pL) / rhoL);
  double aR = std::sqrt((gamma * pR) / rhoR);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double maxvap = dmax(std::fabs(vnL) + aL, std::fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
}

// twod_singlefluid_fluxes.h:120-222
void roe_2d_sf(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double nx, double ny, double* F) {
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty;
  double vtR = vxR * tx + vyR * ty;
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) / (1 + RT);
  double vy = (vyL + RT * vyR) / (1 + RT);
  double vn = vx * nx + vy * ny;
  double vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) / (1 + RT);
  double a = std::sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  double drho = rhoR - rhoL, dp = pR - pL, dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) / (2 * a * a);
  double dV1 = rho * dvt / a;
  double dV2 = drho - dp / (a * a);
  double dV3 = (dp + rho * a * dvn) / (2 * a * a);
  double ws0 = std::fabs(vn - a), ws1 = std::fabs(vn), ws2 = std::fabs(vn), ws3 = std::fabs(vn + a);
  double R00 = 1, R01 = vx - a * nx, R02 = vy - a * ny, R03 = H - vn * a;
  double R10 = 0, R11 = a * tx, R12 = a * ty, R13 = vt * a;
  double R20 = 1, R21 = vx, R22 = vy, R23 = 0.5 * (vx * vx + vy * vy);
  double R30 = 1, R31 = vx + a * nx, R32 = vy + a * ny, R33 = H + vn * a;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) - 0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30);
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31);
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32);
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) - 0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33);
}

// ---- passive scalars: rho, momenta, E, rho*phic (conservative), phinc
// (non-conservative), single-fluid gamma

// oned_passive_fluxes.h:26-65
void rus_1d_pa(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double phicL, double phicR,
               double phincL, double phincR, double nx, double* F, double* nc) {
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double aL = std::sqrt((gamma * pL) / rhoL);
  double aR = std::sqrt((gamma * pR) / rhoR);
  double maxvap = dmax(std::fabs(vxL) + aL, std::fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
  F[3] = 0.5 * ((rhoL * vxL * phicL + rhoR * vxR * phicR) * nx - maxvap * (rhoR * phicR - rhoL * phicL));
  F[4] = 0.5 * (-maxvap * (phincR - phincL));
  nc[4] = 0.5 * 0.5 * (vxL + vxR) * (phincL - phincR) * nx;
}

// oned_passive_fluxes.h:175-266
void roe_1d_pa(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double phicL, double phicR,
               double phincL, double phincR, double nx, double* F, double* nc) {
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double H = (HL + RT * HR) / (1 + RT);
  double a = std::sqrt((gamma - 1) * (H - 0.5 * v * v));
  double Dp = pR - pL;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double ws0 = std::fabs(v - a), ws1 = std::fabs(v), ws2 = std::fabs(v + a);
  double R00 = 1, R01 = v - a, R02 = H - v * a;
  double R10 = 1, R11 = v, R12 = 0.5 * v * v;
  double R20 = 1, R21 = v + a, R22 = H + v * a;
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx - 0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - 0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21);
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - 0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22);
  F[3] = 0.5 * (rhoL * vxL * phicL + rhoR * vxR * phicR) * nx - 0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20);
  F[4] = 0.0 - 0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20);
  nc[4] = -0.5 * v * (phincR - phincL) * nx;
}

// twod_passive_fluxes.h:27-75
void rus_2d_pa(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double phicL, double phicR, double phincL, double phincR, double nx, double ny, double* F, double* nc) {
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = std::sqrt((gamma * pL) / rhoL);
  double aR = std::sqrt((gamma * pR) / rhoR);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double maxvap = dmax(std::fabs(vnL) + aL, std::fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
  F[4] = 0.5 * (rhoL * vnL * phicL + rhoR * vnR * phicR - maxvap * (rhoR * phicR - rhoL * phicL));
  F[5] = 0.5 * (-maxvap * (phincR - phincL));
  nc[5] = 0.5 * 0.5 * (vnL + vnR) * (phincL - phincR);
}

// twod_passive_fluxes.h:197-323
void roe_2d_pa(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double phicL, double phicR, double phincL, double phincR, double nx, double ny, double* F, double* nc) {
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty;
  double vtR = vxR * tx + vyR * ty;
  double gamma = GAMMA_SF;
  double pL = (gamma - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gamma - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) / (1 + RT);
  double vy = (vyL + RT * vyR) / (1 + RT);
  double vn = vx * nx + vy * ny;
  double vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) / (1 + RT);
  double a = std::sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  double drho = rhoR - rhoL, dp = pR - pL, dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) / (2 * a * a);
  double dV1 = rho * dvt / a;
  double dV2 = drho - dp / (a * a);
  double dV3 = (dp + rho * a * dvn) / (2 * a * a);
  double ws0 = std::fabs(vn - a), ws1 = std::fabs(vn), ws2 = std::fabs(vn), ws3 = std::fabs(vn + a);
  double R00 = 1, R01 = vx - a * nx, R02 = vy - a * ny, R03 = H - vn * a;
  double R10 = 0, R11 = a * tx, R12 = a * ty, R13 = vt * a;
  double R20 = 1, R21 = vx, R22 = vy, R23 = 0.5 * (vx * vx + vy * vy);
  double R30 = 1, R31 = vx + a * nx, R32 = vy + a * ny, R33 = H + vn * a;
  double w0 = ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) - 0.5 * w0;
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31);
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32);
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) - 0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33);
  F[4] = 0.5 * (rhoL * vnL * phicL + rhoR * vnR * phicR) - 0.5 * w0;
  F[5] = -0.5 * w0;
  nc[5] = 0.5 * vn * (phincL - phincR);
}

// ---- stiffened gas: rho, momenta, E, alpha = 1/(gamma-1), beta =
// gamma*pinf/(gamma-1), both advected non-conservatively

// oned_stiffened_fluxes.h:29-88 (no mass fractions)
void rus_1d_st(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double alphaL, double alphaR,
               double betaL, double betaR, double nx, double* F, double* nc) {
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pinfL = (1 - 1.0 / gammaL) * betaL;
  double pinfR = (1 - 1.0 / gammaR) * betaR;
  double pL = (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * vxR * vxR);
  double aL = std::sqrt((gammaL * (pL + pinfL)) / rhoL);
  double aR = std::sqrt((gammaR * (pR + pinfR)) / rhoR);
  double maxvap = dmax(std::fabs(vxL) + aL, std::fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
  F[3] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[3] = -0.5 * 0.5 * (vxL + vxR) * (alphaR - alphaL) * nx;
  F[4] = 0.5 * (-maxvap * (betaR - betaL));
  nc[4] = -0.5 * 0.5 * (vxL + vxR) * (betaR - betaL) * nx;
}

// oned_stiffened_fluxes.h:142-295 (no mass fractions); F and nc zeroed by the caller
void roe_1d_st(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double alphaL, double alphaR,
               double betaL, double betaR, double nx, double* F, double* nc) {
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * vxR * vxR);
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double H = (HL + RT * HR) / (1 + RT);
  double alpha = (alphaL + RT * alphaR) / (1 + RT);
  double gamma = 1 + 1.0 / alpha;
  double a = std::sqrt((gamma - 1) * (H - 0.5 * v * v));
  double iL = pL * alphaL;
  double iR = pR * alphaR;
  double i = (std::sqrt(rhoL) * iL + std::sqrt(rhoR) * iR) / (std::sqrt(rhoL) + std::sqrt(rhoR));
  double Dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double p = (gamma - 1) * i;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double dV3 = alphaR - alphaL;
  double dV4 = betaR - betaL;
  double ws0 = std::fabs(v - a), ws1 = std::fabs(v), ws2 = std::fabs(v + a), ws3 = std::fabs(v), ws4 = std::fabs(v);
  double R00 = 1, R01 = v - a, R02 = H - v * a, R03 = 0, R04 = 0;
  double R10 = 1, R11 = v, R12 = 0.5 * v * v, R13 = 0, R14 = 0;
  double R20 = 1, R21 = v + a, R22 = H + v * a, R23 = 0, R24 = 0;
  double R30 = 0, R31 = 0, R32 = p, R33 = 1, R34 = 0;
  double R40 = 0, R41 = 0, R42 = 1, R43 = 0, R44 = 1;
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx -
         0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30 + ws4 * dV4 * R40);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31 + ws4 * dV4 * R41);
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx -
         0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32 + ws4 * dV4 * R42);
  nc[3] = -0.5 * v * (alphaR - alphaL) * nx;
  F[3] += -0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33 + ws4 * dV4 * R43);
  nc[4] = -0.5 * v * (betaR - betaL) * nx;
  F[4] += -0.5 * (ws0 * dV0 * R04 + ws1 * dV1 * R14 + ws2 * dV2 * R24 + ws3 * dV3 * R34 + ws4 * dV4 * R44);
}

// twod_stiffened_fluxes.h:30-104 (no mass fractions)
void rus_2d_st(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double alphaL, double alphaR, double betaL, double betaR, double nx, double ny, double* F, double* nc) {
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pinfL = (1 - 1.0 / gammaL) * betaL;
  double pinfR = (1 - 1.0 / gammaR) * betaR;
  double pL = (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = std::sqrt((gammaL * (pL + pinfL)) / rhoL);
  double aR = std::sqrt((gammaR * (pR + pinfR)) / rhoR);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double maxvap = dmax(std::fabs(vnL) + aL, std::fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
  F[4] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[4] = -0.5 * 0.5 * (vnL + vnR) * (alphaR - alphaL);
  F[5] = 0.5 * (-maxvap * (betaR - betaL));
  nc[5] = -0.5 * 0.5 * (vnL + vnR) * (betaR - betaL);
}

// twod_stiffened_fluxes.h:153-362 (no mass fractions); F and nc zeroed by the caller
void roe_2d_st(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double alphaL, double alphaR, double betaL, double betaR, double nx, double ny, double* F, double* nc) {
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty;
  double vtR = vxR * tx + vyR * ty;
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double iL = pL * alphaL;
  double iR = pR * alphaR;
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) / (1 + RT);
  double vy = (vyL + RT * vyR) / (1 + RT);
  double vn = vx * nx + vy * ny;
  double vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) / (1 + RT);
  double alpha = (alphaL + RT * alphaR) / (1 + RT);
  double gamma = 1 + 1.0 / alpha;
  double a = std::sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  double i = (iL + RT * iR) / (1 + RT);
  double p = (gamma - 1) * i;
  double drho = rhoR - rhoL;
  double dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) / (2 * a * a);
  double dV1 = rho * dvt / a;
  double dV2 = drho - dp / (a * a);
  double dV3 = (dp + rho * a * dvn) / (2 * a * a);
  double dV4 = alphaR - alphaL;
  double dV5 = betaR - betaL;
  double ws0 = std::fabs(vn - a), ws1 = std::fabs(vn), ws2 = std::fabs(vn), ws3 = std::fabs(vn + a);
  double ws4 = std::fabs(vn), ws5 = std::fabs(vn);
  double R00 = 1, R01 = vx - a * nx, R02 = vy - a * ny, R03 = H - vn * a, R04 = 0, R05 = 0;
  double R10 = 0, R11 = a * tx, R12 = a * ty, R13 = vt * a, R14 = 0, R15 = 0;
  double R20 = 1, R21 = vx, R22 = vy, R23 = 0.5 * (vx * vx + vy * vy), R24 = 0, R25 = 0;
  double R30 = 1, R31 = vx + a * nx, R32 = vy + a * ny, R33 = H + vn * a, R34 = 0, R35 = 0;
  double R40 = 0, R41 = 0, R42 = 0, R43 = p, R44 = 1, R45 = 0;
  double R50 = 0, R51 = 0, R52 = 0, R53 = 1, R54 = 0, R55 = 1;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) -
         0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30 + ws4 * dV4 * R40 + ws5 * dV5 * R50);
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31 + ws4 * dV4 * R41 + ws5 * dV5 * R51);
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32 + ws4 * dV4 * R42 + ws5 * dV5 * R52);
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) -
         0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33 + ws4 * dV4 * R43 + ws5 * dV5 * R53);
  nc[4] = -0.5 * vn * (alphaR - alphaL);
  F[4] += -0.5 * (ws0 * dV0 * R04 + ws1 * dV1 * R14 + ws2 * dV2 * R24 + ws3 * dV3 * R34 + ws4 * dV4 * R44 + ws5 * dV5 * R54);
  nc[5] = -0.5 * vn * (betaR - betaL);
  F[5] += -0.5 * (ws0 * dV0 * R05 + ws1 * dV1 * R15 + ws2 * dV2 * R25 + ws3 * dV3 * R35 + ws4 * dV4 * R45 + ws5 * dV5 * R55);
}

// oned_multifluid_fluxes.h:30-94 (GAMNCON, no mass fractions)
void rus_1d_mf(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double alphaL,
               double alphaR, double nx, double* F, double* nc) {
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double aL = std::sqrt((gammaL * pL) / rhoL);
  double aR = std::sqrt((gammaR * pR) / rhoR);
  double maxvap = dmax(std::fabs(vxL) + aL, std::fabs(vxR) + aR);
  F[0] = 0.5 * ((rhoL * vxL + rhoR * vxR) * nx - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * (((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - maxvap * (EtR - EtL));
  F[3] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[3] = -0.5 * 0.5 * (vxL + vxR) * (alphaR - alphaL) * nx;
}

// oned_multifluid_fluxes.h:237-382 (GAMNCON, no mass fractions)
void roe_1d_mf(double rhoL, double rhoR, double vxL, double vxR, double EtL, double EtR, double alphaL,
               double alphaR, double nx, double* F, double* nc) {
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * vxL * vxL);
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * vxR * vxR);
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double v = (vxL + RT * vxR) / (1 + RT);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double H = (HL + RT * HR) / (1 + RT);
  double alpha = (alphaL + RT * alphaR) / (1 + RT);
  double gamma = 1 + 1.0 / alpha;
  double a = std::sqrt((gamma - 1) * (H - 0.5 * v * v));
  double iL = pL * alphaL, iR = pR * alphaR;
  double i = (std::sqrt(rhoL) * iL + std::sqrt(rhoR) * iR) / (std::sqrt(rhoL) + std::sqrt(rhoR));
  double Dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double p = (gamma - 1) * i;
  double dV0 = (Dp - rho * a * (vxR - vxL)) / (2 * a * a);
  double dV1 = (rhoR - rhoL) - Dp / (a * a);
  double dV2 = (Dp + rho * a * (vxR - vxL)) / (2 * a * a);
  double dV3 = alphaR - alphaL;
  double ws0 = std::fabs(v - a), ws1 = std::fabs(v), ws2 = std::fabs(v + a), ws3 = std::fabs(v);
  double R00 = 1, R01 = v - a, R02 = H - v * a, R03 = 0;
  double R10 = 1, R11 = v, R12 = 0.5 * v * v, R13 = 0;
  double R20 = 1, R21 = v + a, R22 = H + v * a, R23 = 0;
  double R30 = 0, R31 = 0, R32 = p, R33 = 1;
  F[0] = 0.5 * (rhoL * vxL + rhoR * vxR) * nx - 0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30);
  F[1] = 0.5 * (rhoL * vxL * vxL + pL + (rhoR * vxR * vxR + pR)) * nx -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31);
  F[2] = 0.5 * ((EtL + pL) * vxL + (EtR + pR) * vxR) * nx - 0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32);
  F[3] = 0;
  nc[3] = -0.5 * v * (alphaR - alphaL) * nx;
  F[3] += -0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33);
}

// twod_multifluid_fluxes.h:30-97 (GAMNCON, no mass fractions)
void rus_2d_mf(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double alphaL, double alphaR, double nx, double ny, double* F, double* nc) {
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double aL = std::sqrt((gammaL * pL) / rhoL);
  double aR = std::sqrt((gammaR * pR) / rhoR);
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double maxvap = dmax(std::fabs(vnL) + aL, std::fabs(vnR) + aR);
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR - maxvap * (rhoR - rhoL));
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx) - maxvap * (rhoR * vxR - rhoL * vxL));
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny) - maxvap * (rhoR * vyR - rhoL * vyL));
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR - maxvap * (EtR - EtL));
  F[4] = 0.5 * (-maxvap * (alphaR - alphaL));
  nc[4] = -0.5 * 0.5 * (vnL + vnR) * (alphaR - alphaL);
}

// twod_multifluid_fluxes.h:151-342 (GAMNCON, no mass fractions)
void roe_2d_mf(double rhoL, double rhoR, double vxL, double vxR, double vyL, double vyR, double EtL, double EtR,
               double alphaL, double alphaR, double nx, double ny, double* F, double* nc) {
  double tx = -ny, ty = nx;
  double vnL = vxL * nx + vyL * ny;
  double vnR = vxR * nx + vyR * ny;
  double vtL = vxL * tx + vyL * ty;
  double vtR = vxR * tx + vyR * ty;
  double gammaL = 1.0 + 1.0 / alphaL;
  double gammaR = 1.0 + 1.0 / alphaR;
  double pL = (gammaL - 1) * (EtL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
  double pR = (gammaR - 1) * (EtR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
  double HL = (EtL + pL) / rhoL;
  double HR = (EtR + pR) / rhoR;
  double iL = pL * alphaL, iR = pR * alphaR;
  double RT = std::sqrt(rhoR / rhoL);
  double rho = RT * rhoL;
  double vx = (vxL + RT * vxR) / (1 + RT);
  double vy = (vyL + RT * vyR) / (1 + RT);
  double vn = vx * nx + vy * ny;
  double vt = vx * tx + vy * ty;
  double H = (HL + RT * HR) / (1 + RT);
  double alpha = (alphaL + RT * alphaR) / (1 + RT);
  double gamma = 1 + 1.0 / alpha;
  double a = std::sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
  double i = (iL + RT * iR) / (1 + RT);
  double p = (gamma - 1) * i;
  double drho = rhoR - rhoL;
  double dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
  double dvn = vnR - vnL, dvt = vtR - vtL;
  double dV0 = (dp - rho * a * dvn) / (2 * a * a);
  double dV1 = rho * dvt / a;
  double dV2 = drho - dp / (a * a);
  double dV3 = (dp + rho * a * dvn) / (2 * a * a);
  double dV4 = alphaR - alphaL;
  double ws0 = std::fabs(vn - a), ws1 = std::fabs(vn), ws2 = std::fabs(vn), ws3 = std::fabs(vn + a), ws4 = std::fabs(vn);
  double R00 = 1, R01 = vx - a * nx, R02 = vy - a * ny, R03 = H - vn * a, R04 = 0;
  double R10 = 0, R11 = a * tx, R12 = a * ty, R13 = vt * a, R14 = 0;
  double R20 = 1, R21 = vx, R22 = vy, R23 = 0.5 * (vx * vx + vy * vy), R24 = 0;
  double R30 = 1, R31 = vx + a * nx, R32 = vy + a * ny, R33 = H + vn * a, R34 = 0;
  double R40 = 0, R41 = 0, R42 = 0, R43 = p, R44 = 1;
  F[0] = 0.5 * (rhoL * vnL + rhoR * vnR) -
         0.5 * (ws0 * dV0 * R00 + ws1 * dV1 * R10 + ws2 * dV2 * R20 + ws3 * dV3 * R30 + ws4 * dV4 * R40);
  F[1] = 0.5 * ((rhoL * vnL * vxL + pL * nx) + (rhoR * vnR * vxR + pR * nx)) -
         0.5 * (ws0 * dV0 * R01 + ws1 * dV1 * R11 + ws2 * dV2 * R21 + ws3 * dV3 * R31 + ws4 * dV4 * R41);
  F[2] = 0.5 * ((rhoL * vnL * vyL + pL * ny) + (rhoR * vnR * vyR + pR * ny)) -
         0.5 * (ws0 * dV0 * R02 + ws1 * dV1 * R12 + ws2 * dV2 * R22 + ws3 * dV3 * R32 + ws4 * dV4 * R42);
  F[3] = 0.5 * (rhoL * vnL * HL + rhoR * vnR * HR) -
         0.5 * (ws0 * dV0 * R03 + ws1 * dV1 * R13 + ws2 * dV2 * R23 + ws3 * dV3 * R33 + ws4 * dV4 * R43);
  F[4] = 0;
  nc[4] = -0.5 * vn * (alphaR - alphaL);
  F[4] += -0.5 * (ws0 * dV0 * R04 + ws1 * dV1 * R14 + ws2 * dV2 * R24 + ws3 * dV3 * R34 + ws4 * dV4 * R44);
}

// ---------------------------------------------------------------------------
// HR limiter device functions (limiting_kernels.cu:1470-1868)
// ---------------------------------------------------------------------------
double minmod2(double a, double b) {  // minmod(a,b), :1470-1486
  int signa = signum(a);
  if (signa != signum(b)) return 0;
  double fa = std::fabs(a), fb = std::fabs(b);
  if (fa < fb) return signa * fa;
  return signa * fb;
}
int lim_factorial(int n) {
  static const int f[13] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880, 3628800, 39916800, 479001600};
  return (n >= 0 && n <= 12) ? f[n] : 1;
}
double integ_deriv(int k, int n) {  // :1748-1760
  int num = n - k + 1;
  if (num % 2) return 2.0 / (double)lim_factorial(num);
  return 0.0;
}
double integ_deriv_bounds(int k, int n, double a, double b) {  // :1762-1775
  int num = n - k + 1;
  return (std::pow(b, num) - std::pow(a, num)) / (double)lim_factorial(num);
}
void limit_monomial(int N, const double* AL, const double* AC, const double* AR, double* Alim) {  // :1534-1583
  double avgLC = 0;
  for (int m = N; m > 0; m--) {
    double avgdUL = 0, avgdUC = 0, avgdUR = 0, avgRL = 0, avgRC = 0, avgRR = 0;
    for (int n = m - 1; n <= N; n++) {
      double integral = integ_deriv(m - 1, n);
      avgdUL += AL[n] * integral;
      avgdUC += AC[n] * integral;
      avgdUR += AR[n] * integral;
      if (n >= m + 1) {
        double alim = Alim[n];
        avgRL += alim * integ_deriv_bounds(m - 1, n, -3, -1);
        avgRC += alim * integral;
        avgRR += alim * integ_deriv_bounds(m - 1, n, 1, 3);
      }
    }
    double avgLL = 0.5 * (avgdUL - avgRL);
    avgLC = 0.5 * (avgdUC - avgRC);
    double avgLR = 0.5 * (avgdUR - avgRR);
    double c1 = 0.5 * (avgLC - avgLL);
    double c2 = 0.5 * (avgLR - avgLC);
    Alim[m] = minmod2(c1, c2);
  }
  Alim[0] = avgLC;
}
void average_monomial(int N, const double* A, double* Alim) {  // :1777-1790
  double avg = 0;
  for (int n = 0; n <= N; n++) {
    avg += A[n] * integ_deriv(0, n);
    Alim[n] = 0;
  }
  Alim[0] = 0.5 * avg;
}
void gemv_cm(int M, int N, const double* A, const double* B, double* C) {  // :1699-1716
  for (int m = 0; m < M; m++) {
    double sum = 0;
    for (int n = 0; n < N; n++) sum += A[n * M + m] * B[n];
    C[m] = sum;
  }
}
void gemv3_cm(int M, int N, const double* A, const double* B1, double* C1, const double* B2, double* C2,
              const double* B3, double* C3) {  // :1718-1746
  for (int m = 0; m < M; m++) {
    double s1 = 0, s2 = 0, s3 = 0;
    for (int n = 0; n < N; n++) {
      double a = A[n * M + m];
      s1 += a * B1[n];
      s2 += a * B2[n];
      s3 += a * B3[n];
    }
    C1[m] = s1; C2[m] = s2; C3[m] = s3;
  }
}
// mono (optional) receives the limited monomial coefficients
void set2average(int N_s, int N, int N_s1D, int slicenum, const double* L2M, const double* M2L, double* tmp,
                 double* U, double* mono = nullptr) {  // :1792-1828
  double* A = tmp;
  double* Alim = tmp + N_s;
  gemv_cm(N_s, N_s, L2M, U, A);
  for (int s = 0; s < slicenum; s++) average_monomial(N, &A[s * N_s1D], &Alim[s * N_s1D]);
  gemv_cm(N_s, N_s, M2L, Alim, U);
  if (mono) for (int i = 0; i < N_s; i++) mono[i] = Alim[i];
}
void HRslices(int N_s, int N, int N_s1D, int slicenum, const double* L2M, const double* M2L, double* tmp,
              const double* UL, double* UC, const double* UR, double* mono = nullptr) {  // :1830-1868
  double* AL = tmp;
  double* AC = tmp + N_s;
  double* AR = tmp + 2 * N_s;
  double* Alim = tmp + 3 * N_s;
  gemv3_cm(N_s, N_s, L2M, UL, AL, UC, AC, UR, AR);
  for (int s = 0; s < slicenum; s++) limit_monomial(N, &AL[s * N_s1D], &AC[s * N_s1D], &AR[s * N_s1D], &Alim[s * N_s1D]);
  gemv_cm(N_s, N_s, M2L, Alim, UC);
  if (mono) for (int i = 0; i < N_s; i++) mono[i] = Alim[i];
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// netlib dgemm (column-major)
// ---------------------------------------------------------------------------
void or_dgemm(char ta, char tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B,
              int ldb, double beta, double* C, int ldc) {
  bool nota = (ta == 'N' || ta == 'n');
  // only the forms the reference uses: NN and TN
  for (int j = 0; j < N; ++j) {
    if (nota) {
      if (beta == 0.0)
        for (int i = 0; i < M; ++i) C[i + j * ldc] = 0.0;
      else if (beta != 1.0)
        for (int i = 0; i < M; ++i) C[i + j * ldc] = beta * C[i + j * ldc];
      for (int l = 0; l < K; ++l) {
        double temp = alpha * B[l + j * ldb];
        for (int i = 0; i < M; ++i) C[i + j * ldc] = C[i + j * ldc] + temp * A[i + l * lda];
      }
    } else {
      for (int i = 0; i < M; ++i) {
        double temp = 0.0;
        for (int l = 0; l < K; ++l) temp = temp + A[l + i * lda] * B[l + j * ldb];
        if (beta == 0.0)
          C[i + j * ldc] = alpha * temp;
        else
          C[i + j * ldc] = alpha * temp + beta * C[i + j * ldc];
      }
    }
  }
  (void)tb;
}

// kernels.cu:41-75
void or_mapToFace(const or_cfg* c, const int* map, const double* U, double* UF) {
  const int M_s = c->M_s, N_F = c->N_F;
  for (int t = 0; t < c->M_T; t++)
    for (int j = 0; j < M_s; j++)
      for (int fc = 0; fc < N_F; fc++)
        for (int d = 0; d < 2; d++) {
          int face = ((t * N_F + fc) * 2 + d) * M_s + j;
          UF[face] = U[map[face]];
        }
}

// boundaries.cu:20-67
void or_rflctive(const or_cfg* c, int M_B, const int* boundaryMap, const double* normals, int start, double* UF) {
  const int M_s = c->M_s, N_F = c->N_F, D = c->D;
  for (int k = 0; k < M_B; k++) {
    int t = boundaryMap[start + k];
    for (int j = 0; j < M_s; j++) {
      if (D == 1) {
        UF[((t * N_F + 1) * 2 + 1) * M_s + j] = -UF[((t * N_F + 1) * 2 + 1) * M_s + j];
      } else {
        double nx = normals[t * D + 0], ny = normals[t * D + 1];
        double nx2 = nx * nx, ny2 = ny * ny;
        double m2nxny = -2 * nx * ny;
        double invnx2ny2 = 1.0 / (nx2 + ny2);
        double vxL = UF[((t * N_F + 1) * 2 + 0) * M_s + j];
        double vyL = UF[((t * N_F + 2) * 2 + 0) * M_s + j];
        double vxR = invnx2ny2 * ((-nx2 + ny2) * vxL + m2nxny * vyL);
        double vyR = invnx2ny2 * (m2nxny * vxL + (nx2 - ny2) * vyL);
        UF[((t * N_F + 1) * 2 + 1) * M_s + j] = vxR;
        UF[((t * N_F + 2) * 2 + 1) * M_s + j] = vyR;
      }
    }
  }
}

// physics.cu:27-363 (SINGLEFLUID, MULTIFLUID/GAMNCON; N_Y = 0)
void or_evaluate_sf(const or_cfg* c, double* s, double* f, const double* Ug, const double* dUg, const double* invJac) {
  const int N_G = c->N_G, N_F = c->N_F, D = c->D;
  const double GX = c->gx, GY = c->gy;
  for (int e = 0; e < c->N_E; e++)
    for (int g = 0; g < N_G; g++) {
#define UG(k) Ug[(e * N_F + (k)) * N_G + g]
#define S(k) s[(e * N_F + (k)) * N_G + g]
#define FX(k, a) f[((e * N_F + (k)) * N_G + g) * D + (a)]
      if (D == 1) {
        double rho = UG(0);
        double u = UG(1) / rho;
        double Et = UG(2);
        if (c->physics == 0) {
          double gamma = GAMMA_SF;
          double p = (gamma - 1) * (Et - 0.5 * rho * u * u);
          S(0) = 0; S(1) = 0; S(2) = 0;
          FX(0, 0) = rho * u;
          FX(1, 0) = rho * u * u + p;
          FX(2, 0) = (Et + p) * u;
        } else if (c->physics == 1) {
          double gamma = 1 + 1.0 / UG(3);
          double p = (gamma - 1) * (Et - 0.5 * rho * u * u);
          S(0) = 0;
          S(1) = rho * GX;
          S(2) = rho * u * GX;
          FX(0, 0) = rho * u;
          FX(1, 0) = rho * u * u + p;
          FX(2, 0) = (Et + p) * u;
          S(3) = -u * dUg[(e * N_F + 3) * N_G + g] * invJac[e * N_G + g];
          FX(3, 0) = 0;
        } else if (c->physics == 2) {  // physics.cu:57-78
          double gamma = GAMMA_SF;
          double p = (gamma - 1) * (Et - 0.5 * rho * u * u);
          S(0) = 0; S(1) = 0; S(2) = 0; S(3) = 0;
          S(4) = -u * dUg[(e * N_F + 4) * N_G + g] * invJac[e * N_G + g];
          FX(0, 0) = rho * u;
          FX(1, 0) = rho * u * u + p;
          FX(2, 0) = (Et + p) * u;
          FX(3, 0) = rho * u * (UG(3) / rho);
          FX(4, 0) = 0;
        } else {  // physics.cu:144-170
          double gamma = 1 + 1.0 / UG(3);
          double beta = UG(4);
          double p = (gamma - 1) * (Et - beta - 0.5 * rho * u * u);
          S(0) = 0;
          S(1) = rho * GX;
          S(2) = rho * u * GX;
          S(3) = -u * dUg[(e * N_F + 3) * N_G + g] * invJac[e * N_G + g];
          S(4) = -u * dUg[(e * N_F + 4) * N_G + g] * invJac[e * N_G + g];
          FX(0, 0) = rho * u;
          FX(1, 0) = rho * u * u + p;
          FX(2, 0) = (Et + p) * u;
          FX(3, 0) = 0;
          FX(4, 0) = 0;
        }
      } else {
        double rho = UG(0);
        double u = UG(1) / rho;
        double v = UG(2) / rho;
        double Et = UG(3);
        if (c->physics == 0) {
          double gamma = GAMMA_SF;
          double p = (gamma - 1) * (Et - 0.5 * rho * (u * u + v * v));
          S(0) = 0; S(1) = 0; S(2) = 0; S(3) = 0;
          FX(0, 0) = rho * u; FX(1, 0) = rho * u * u + p; FX(2, 0) = rho * u * v; FX(3, 0) = (Et + p) * u;
          FX(0, 1) = rho * v; FX(1, 1) = rho * u * v; FX(2, 1) = rho * v * v + p; FX(3, 1) = (Et + p) * v;
        } else if (c->physics == 2) {  // physics.cu:183-220
          double phic = UG(4) / rho;
          double vdotgradphinc = 0;
          for (int al = 0; al < D; al++)
            vdotgradphinc += -u * dUg[((e * N_F + 5) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 0) * D + al] -
                             v * dUg[((e * N_F + 5) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 1) * D + al];
          double gamma = GAMMA_SF;
          double p = (gamma - 1) * (Et - 0.5 * rho * (u * u + v * v));
          S(0) = 0; S(1) = 0; S(2) = 0; S(3) = 0; S(4) = 0;
          S(5) = vdotgradphinc;
          FX(0, 0) = rho * u; FX(1, 0) = rho * u * u + p; FX(2, 0) = rho * u * v; FX(3, 0) = (Et + p) * u;
          FX(4, 0) = rho * u * phic; FX(5, 0) = 0;
          FX(0, 1) = rho * v; FX(1, 1) = rho * u * v; FX(2, 1) = rho * v * v + p; FX(3, 1) = (Et + p) * v;
          FX(4, 1) = rho * v * phic; FX(5, 1) = 0;
        } else if (c->physics == 3) {  // physics.cu:302-352
          double gamma = 1 + 1.0 / UG(4);
          double beta = UG(5);
          double p = (gamma - 1) * (Et - beta - 0.5 * rho * (u * u + v * v));
          double vdotgradalpha = 0, vdotgradbeta = 0;
          for (int al = 0; al < D; al++) {
            vdotgradalpha += -u * dUg[((e * N_F + 4) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 0) * D + al] -
                             v * dUg[((e * N_F + 4) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 1) * D + al];
            vdotgradbeta += -u * dUg[((e * N_F + 5) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 0) * D + al] -
                            v * dUg[((e * N_F + 5) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 1) * D + al];
          }
          S(0) = 0;
          S(1) = rho * GX;
          S(2) = rho * GY;
          S(3) = rho * (u * GX + v * GY);
          S(4) = vdotgradalpha;
          S(5) = vdotgradbeta;
          FX(0, 0) = rho * u; FX(1, 0) = rho * u * u + p; FX(2, 0) = rho * u * v; FX(3, 0) = (Et + p) * u;
          FX(4, 0) = 0; FX(5, 0) = 0;
          FX(0, 1) = rho * v; FX(1, 1) = rho * u * v; FX(2, 1) = rho * v * v + p; FX(3, 1) = (Et + p) * v;
          FX(4, 1) = 0; FX(5, 1) = 0;
        } else {
          double gamma = 1 + 1.0 / UG(4);
          double p = (gamma - 1) * (Et - 0.5 * rho * (u * u + v * v));
          S(0) = 0;
          S(1) = rho * GX;
          S(2) = rho * GY;
          S(3) = rho * (u * GX + v * GY);
          FX(0, 0) = rho * u; FX(1, 0) = rho * u * u + p; FX(2, 0) = rho * u * v; FX(3, 0) = (Et + p) * u;
          FX(0, 1) = rho * v; FX(1, 1) = rho * u * v; FX(2, 1) = rho * v * v + p; FX(3, 1) = (Et + p) * v;
          double vdotgradalpha = 0;
          for (int al = 0; al < D; al++) {
            // same index use as physics.cu:283-288
            vdotgradalpha += -u * dUg[((e * N_F + 4) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 0) * D + al] -
                             v * dUg[((e * N_F + 4) * N_G + g) * D + al] * invJac[((e * N_G + g) * D + 1) * D + al];
          }
          S(4) = vdotgradalpha;
          FX(4, 0) = 0;
          FX(4, 1) = 0;
        }
      }
#undef UG
#undef S
#undef FX
    }
}

// physics.cu:367-811
void or_evaluate_q(const or_cfg* c, double* q, const double* UgF, const double* normals) {
  const int M_G = c->M_G, N_F = c->N_F, D = c->D;
  std::vector<double> buf(2 * N_F);
  for (int t = 0; t < c->M_T; t++)
    for (int g = 0; g < M_G; g++) {
      double* F = buf.data();
      double* nc = buf.data() + N_F;
      for (int k = 0; k < 2 * N_F; k++) buf[k] = 0;
#define UF(k, d) UgF[((t * N_F + (k)) * 2 + (d)) * M_G + g]
      if (D == 1) {
        double nx = normals[t * D + 0];
        if (c->physics == 0) {
          if (c->flux == 0)
            rus_1d_sf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), nx, F);
          else
            roe_1d_sf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), nx, F);
        } else if (c->physics == 2) {
          auto fn = c->flux == 0 ? rus_1d_pa : roe_1d_pa;
          fn(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), UF(3, 0) / UF(0, 0),
             UF(3, 1) / UF(0, 1), UF(4, 0), UF(4, 1), nx, F, nc);
        } else if (c->physics == 3) {
          auto fn = c->flux == 0 ? rus_1d_st : roe_1d_st;
          fn(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), UF(3, 0), UF(3, 1),
             UF(4, 0), UF(4, 1), nx, F, nc);
        } else {
          if (c->flux == 0)
            rus_1d_mf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), UF(3, 0),
                      UF(3, 1), nx, F, nc);
          else
            roe_1d_mf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0), UF(2, 1), UF(3, 0),
                      UF(3, 1), nx, F, nc);
        }
      } else {
        double nx = normals[t * D + 0], ny = normals[t * D + 1];
        if (c->physics == 0) {
          if (c->flux == 0)
            rus_2d_sf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0),
                      UF(2, 1) / UF(0, 1), UF(3, 0), UF(3, 1), nx, ny, F);
          else
            roe_2d_sf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0),
                      UF(2, 1) / UF(0, 1), UF(3, 0), UF(3, 1), nx, ny, F);
        } else if (c->physics == 2) {
          auto fn = c->flux == 0 ? rus_2d_pa : roe_2d_pa;
          fn(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0), UF(2, 1) / UF(0, 1),
             UF(3, 0), UF(3, 1), UF(4, 0) / UF(0, 0), UF(4, 1) / UF(0, 1), UF(5, 0), UF(5, 1), nx, ny, F, nc);
        } else if (c->physics == 3) {
          auto fn = c->flux == 0 ? rus_2d_st : roe_2d_st;
          fn(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0), UF(2, 1) / UF(0, 1),
             UF(3, 0), UF(3, 1), UF(4, 0), UF(4, 1), UF(5, 0), UF(5, 1), nx, ny, F, nc);
        } else {
          if (c->flux == 0)
            rus_2d_mf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0),
                      UF(2, 1) / UF(0, 1), UF(3, 0), UF(3, 1), UF(4, 0), UF(4, 1), nx, ny, F, nc);
          else
            roe_2d_mf(UF(0, 0), UF(0, 1), UF(1, 0) / UF(0, 0), UF(1, 1) / UF(0, 1), UF(2, 0) / UF(0, 0),
                      UF(2, 1) / UF(0, 1), UF(3, 0), UF(3, 1), UF(4, 0), UF(4, 1), nx, ny, F, nc);
        }
      }
#undef UF
      for (int fc = 0; fc < N_F; fc++) {
        q[((t * N_F + fc) * 2 + 0) * M_G + g] = -F[fc] + nc[fc];
        q[((t * N_F + fc) * 2 + 1) * M_G + g] = F[fc] + nc[fc];
      }
    }
}

// kernels.cu:154-196
void or_redistribute_sf(const or_cfg* c, double* sJ, double* fJ, const double* s, const double* f, const double* J,
                        const double* invJac) {
  const int N_G = c->N_G, N_F = c->N_F, D = c->D;
  for (int e = 0; e < c->N_E; e++)
    for (int g = 0; g < N_G; g++)
      for (int fc = 0; fc < N_F; fc++) {
        double j = J[e];
        double sol = 0.0;
        sJ[(e * N_F + fc) * N_G + g] = s[(e * N_F + fc) * N_G + g] * j;
        for (int al = 0; al < D; al++) {
          for (int a = 0; a < D; a++) sol += invJac[((e * N_G + g) * D + al) * D + a] * f[((e * N_F + fc) * N_G + g) * D + a] * j;
          fJ[((e * N_F + fc) * N_G + g) * D + al] = sol;
          sol = 0;
        }
      }
}

// kernels.cu:199-228
void or_redistribute_q(const or_cfg* c, double* qJ, const double* q, const double* JF) {
  const int M_G = c->M_G, N_F = c->N_F;
  for (int t = 0; t < c->M_T; t++)
    for (int g = 0; g < M_G; g++)
      for (int fc = 0; fc < N_F; fc++) {
        qJ[((t * N_F + fc) * 2 + 0) * M_G + g] = q[((t * N_F + fc) * 2 + 0) * M_G + g] * JF[t * 2 + 0];
        qJ[((t * N_F + fc) * 2 + 1) * M_G + g] = q[((t * N_F + fc) * 2 + 1) * M_G + g] * JF[t * 2 + 1];
      }
}

// kernels.cu:78-150
void or_mapToElement(const or_cfg* c, const int* invmap, double* Q, const double* Qtcj) {
  const int N_s = c->N_s, N_F = c->N_F, M_s = c->M_s, N_N = c->N_N;
  std::vector<double> sol(N_s);
  for (int e = 0; e < c->N_E; e++)
    for (int fc = 0; fc < N_F; fc++) {
      for (int i = 0; i < N_s; i++) sol[i] = 0;
      for (int i = 0; i < M_s * N_N; i++) {
        int solidx = invmap[((e * N_F + fc) * M_s * N_N + i) * 2 + 0];
        int qidx = invmap[((e * N_F + fc) * M_s * N_N + i) * 2 + 1];
        sol[solidx] += Qtcj[qidx];
      }
      for (int i = 0; i < N_s; i++) Q[(e * N_F + fc) * N_s + i] = sol[i];
    }
}

// kernels.cu:231-261
void or_addSFQ(const or_cfg* c, double* A, const double* S, const double* F, const double* Q) {
  const int n = c->N_E * c->N_F * c->N_s;
  for (int k = 0; k < n; k++) A[k] = S[k] + F[k] + Q[k];
}

// rk_kernels.cu:22-56
void or_solve(const or_cfg* c, double Dt, const double* Minv, const double* f, double* DU) {
  const int N_s = c->N_s, N_F = c->N_F;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++)
      for (int fc = 0; fc < N_F; fc++) {
        double sol = 0.0;
        for (int ii = 0; ii < N_s; ii++) sol += Minv[(e * N_s + ii) * N_s + i] * f[(e * N_F + fc) * N_s + ii];
        DU[(e * N_F + fc) * N_s + i] = Dt * sol;
      }
}

// rk_kernels.cu:94-173 with upa.h:41-83
void or_findUPA(const or_cfg* c, const double* U, double* UPA) {
  const int N_s = c->N_s, N_F = c->N_F;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
#define UU(k) U[(e * N_F + (k)) * N_s + i]
      double r;
      if (c->physics == 3) {  // oned_/twod_stiffened_upa (upa.h:98-130)
        if (c->D == 1) {
          double rho = UU(0), u = UU(1) / UU(0), E = UU(2), alpha = UU(3), beta = UU(4);
          double gamma = 1.0 + 1.0 / alpha;
          double pinf = (1 - 1.0 / gamma) * beta;
          double p = (gamma - 1) * (E - beta - 0.5 * rho * u * u);
          r = std::fabs(u) + std::sqrt(gamma * (p + pinf) / rho);
        } else {
          double rho = UU(0), u = UU(1) / UU(0), v = UU(2) / UU(0), E = UU(3), alpha = UU(4), beta = UU(5);
          double gamma = 1.0 + 1.0 / alpha;
          double pinf = (1 - 1.0 / gamma) * beta;
          double p = (gamma - 1) * (E - beta - 0.5 * rho * (u * u + v * v));
          double a = std::sqrt(gamma * (p + pinf) / rho);
          r = dmax(std::fabs(u) + a, std::fabs(v) + a);
        }
      } else if (c->D == 1) {  // passive: the single-fluid form (upa.h:20-40)
        double rho = UU(0), u = UU(1) / UU(0), E = UU(2);
        double gamma = c->physics != 1 ? GAMMA_SF : 1.0 + 1.0 / UU(3);
        double p = (gamma - 1) * (E - 0.5 * rho * u * u);
        r = std::fabs(u) + std::sqrt(gamma * p / rho);
      } else {
        double rho = UU(0), u = UU(1) / UU(0), v = UU(2) / UU(0), E = UU(3);
        double gamma = c->physics != 1 ? GAMMA_SF : 1.0 + 1.0 / UU(4);
        double p = (gamma - 1) * (E - 0.5 * rho * (u * u + v * v));
        double a = std::sqrt(gamma * p / rho);
        r = dmax(std::fabs(u) + a, std::fabs(v) + a);
      }
#undef UU
      UPA[e * N_s + i] = r;
    }
}

// dg_solver.h:292-364
void or_rhs(const or_cfg* c, const or_tables* t, const double* U, double* f_rk) {
  const int N_s = c->N_s, N_G = c->N_G, M_s = c->M_s, M_G = c->M_G, N_F = c->N_F, D = c->D;
  const int N_E = c->N_E, M_T = c->M_T;
  std::vector<double> UF((size_t)2 * N_F * M_s * M_T), Ug((size_t)N_F * N_G * N_E),
      dUg((size_t)D * N_G * N_E * N_F), UgF((size_t)2 * N_F * M_G * M_T), s((size_t)N_G * N_E * N_F),
      sJ((size_t)N_G * N_E * N_F), S((size_t)N_s * N_E * N_F), f((size_t)D * N_F * N_G * N_E),
      fJ((size_t)D * N_G * N_E * N_F), F((size_t)N_s * N_E * N_F), q((size_t)M_G * M_T * N_F * 2),
      qJ((size_t)M_G * M_T * N_F * 2), Qtcj((size_t)M_s * M_T * N_F * 2), Q((size_t)N_s * N_E * N_F);
  or_mapToFace(c, t->map, U, UF.data());
  or_rflctive(c, t->rflctiveIdx, t->boundaryMap, t->normals, 0, UF.data());
  or_dgemm('N', 'N', N_G, N_E * N_F, N_s, 1, t->phi, N_G, U, N_s, 0.0, Ug.data(), N_G);
  or_dgemm('N', 'N', N_G * D, N_E * N_F, N_s, 1, t->dphi, N_G * D, U, N_s, 0.0, dUg.data(), N_G * D);
  or_dgemm('N', 'N', M_G, M_T * N_F * 2, M_s, 1, t->psi, M_G, UF.data(), M_s, 0.0, UgF.data(), M_G);
  or_evaluate_sf(c, s.data(), f.data(), Ug.data(), dUg.data(), t->invJac);
  or_evaluate_q(c, q.data(), UgF.data(), t->normals);
  or_redistribute_sf(c, sJ.data(), fJ.data(), s.data(), f.data(), t->J, t->invJac);
  or_dgemm('T', 'N', N_s, N_E * N_F, N_G, 1, t->phi_w, N_G, sJ.data(), N_G, 0.0, S.data(), N_s);
  or_dgemm('T', 'N', N_s, N_E * N_F, N_G * D, 1, t->dphi_w, N_G * D, fJ.data(), N_G * D, 0.0, F.data(), N_s);
  or_redistribute_q(c, qJ.data(), q.data(), t->JF);
  or_dgemm('T', 'N', M_s, M_T * N_F * 2, M_G, 1, t->psi_w, M_G, qJ.data(), M_G, 0.0, Qtcj.data(), M_s);
  or_mapToElement(c, t->invmap, Q.data(), Qtcj.data());
  or_addSFQ(c, f_rk, S.data(), F.data(), Q.data());
}

// rk.cc:250-272 (blasIamax: first index of the largest |UPA|)
double or_dt_from_cfl(const or_cfg* c, const double* U, double CFL) {
  std::vector<double> UPA((size_t)c->N_E * c->N_s);
  or_findUPA(c, U, UPA.data());
  int idx = 0;
  double best = std::fabs(UPA[0]);
  for (size_t k = 1; k < UPA.size(); k++)
    if (std::fabs(UPA[k]) > best) { best = std::fabs(UPA[k]); idx = (int)k; }
  return CFL / UPA[idx];
}

// limiting_kernels.cu:245-319
void or_hrl1D(int N_s, int N_E, int Nfields, int N_N, int slicenum, const int* neighbors, int offxy, const double* A,
              double* Alim) {
  for (int e = 0; e < N_E; e++)
    for (int slice = 0; slice < slicenum; slice++)
      for (int fc = 0; fc < Nfields; fc++) {
        int N = N_s - 1;
        int left = neighbors[e * N_N + offxy + 0];
        int right = neighbors[e * N_N + offxy + 1];
        int physical = 0;
        if (left < 0) physical = -left;
        if (right < 0) physical = -right;
        int idx = (e * Nfields + fc) * N_s * slicenum + slice * N_s;
        if (physical == 4) {
        } else if (physical == 2) {
          average_monomial(N, &A[idx], &Alim[idx]);
        } else if (physical == 3) {
          int un_idx = 0;
          if (offxy == 0) un_idx = 1;
          else if (offxy == 2) un_idx = 2;
          if (fc != un_idx)
            average_monomial(N, &A[idx], &Alim[idx]);
          else
            for (int n = 0; n <= N; n++) Alim[idx + n] = A[idx + n];
        } else {
          int idxL = (left * Nfields + fc) * N_s * slicenum + slice * N_s;
          int idxR = (right * Nfields + fc) * N_s * slicenum + slice * N_s;
          limit_monomial(N, &A[idxL], &A[idx], &A[idxR], &Alim[idx]);
        }
      }
}

// limiting.cc:11-50 (1D and 2D Cartesian)
void or_limit_hr(const or_cfg* c, const or_tables* t, double* U) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F, N_N = c->N_N;
  std::vector<double> A((size_t)N_s * N_E * N_F), Alim((size_t)N_s * N_E * N_F, 0.0);
  if (c->D == 1) {
    or_dgemm('N', 'N', N_s, N_E * N_F, N_s, 1, t->Lag2Mono, N_s, U, N_s, 0.0, A.data(), N_s);
    or_hrl1D(N_s, N_E, N_F, N_N, 1, t->neighbors, 0, A.data(), Alim.data());
    or_dgemm('N', 'N', N_s, N_E * N_F, N_s, 1, t->Mono2Lag, N_s, Alim.data(), N_s, 0.0, U, N_s);
  } else {
    const int n1 = c->N_s1D;
    or_dgemm('N', 'N', N_s, N_E * N_F, N_s, 1, t->Lag2MonoX, N_s, U, N_s, 0.0, A.data(), N_s);
    or_hrl1D(n1, N_E, N_F, N_N, n1, t->neighbors, 0, A.data(), Alim.data());
    or_dgemm('N', 'N', N_s, N_E * N_F, N_s, 1, t->MonoX2MonoY, N_s, Alim.data(), N_s, 0.0, A.data(), N_s);
    or_hrl1D(n1, N_E, N_F, N_N, n1, t->neighbors, 2, A.data(), Alim.data());
    or_dgemm('N', 'N', N_s, N_E * N_F, N_s, 1, t->MonoY2Lag, N_s, Alim.data(), N_s, 0.0, U, N_s);
  }
}

// sensor_kernels.cu:19-163, sequential element order
void or_calc_sensors(const or_cfg* c, const int* on, const double* thresh, const int* neighbors, const double* Uavg,
                     int* sensors) {
  const int N_F = c->N_F, N_N = c->N_N, D = c->D;
  for (int e = 0; e < c->N_E; e++) {
    if (sensors[e] != 0) continue;
    double vyL = 0, vyR = 0, betaL = 0, betaR = 0, pinfL = 0, pinfR = 0;
    int fc = 0;
    double rhoL = Uavg[e * N_F + fc]; fc++;
    double vxL = Uavg[e * N_F + fc] / rhoL; fc++;
    if (D == 2) { vyL = Uavg[e * N_F + fc] / rhoL; fc++; }
    double EtL = Uavg[e * N_F + fc]; fc++;
    double alphaL;
    if (c->physics == 0 || c->physics == 2) alphaL = 1 / (GAMMA_SF - 1);
    else { alphaL = Uavg[e * N_F + fc]; fc++; }
    if (c->physics == 3) { betaL = Uavg[e * N_F + fc]; fc++; }
    double gammaL = 1.0 + 1.0 / alphaL;
    pinfL = (1 - 1.0 / gammaL) * betaL;
    double pL = (gammaL - 1) * (EtL - betaL - 0.5 * rhoL * (vxL * vxL + vyL * vyL));
    double iL = pL * alphaL;
    double aL = std::sqrt((gammaL * (pL + pinfL)) / rhoL);
    double HL = (EtL + pL) / rhoL;
    bool done = false;
    for (int nn = 0; (nn < N_N) && (!done); nn++) {
      int right = neighbors[e * N_N + nn];
      if (right >= 0) {
        fc = 0;
        double rhoR = Uavg[right * N_F + fc]; fc++;
        double vxR = Uavg[right * N_F + fc] / rhoR; fc++;
        if (D == 2) { vyR = Uavg[right * N_F + fc] / rhoR; fc++; }
        double EtR = Uavg[right * N_F + fc]; fc++;
        double alphaR;
        if (c->physics == 0 || c->physics == 2) alphaR = 1 / (GAMMA_SF - 1);
        else { alphaR = Uavg[right * N_F + fc]; fc++; }
        if (c->physics == 3) { betaR = Uavg[right * N_F + fc]; fc++; }
        double gammaR = 1.0 + 1.0 / alphaR;
        pinfR = (1 - 1.0 / gammaR) * betaR;
        double pR = (gammaR - 1) * (EtR - betaR - 0.5 * rhoR * (vxR * vxR + vyR * vyR));
        double iR = pR * alphaR;
        double aR = std::sqrt((gammaR * (pR + pinfR)) / rhoR);
        double HR = (EtR + pR) / rhoR;
        double RT = std::sqrt(rhoR / rhoL);
        double vx = (vxL + RT * vxR) / (1 + RT);
        double vy = (vyL + RT * vyR) / (1 + RT);
        double alpha = (alphaL + RT * alphaR) / (1 + RT);
        double gamma = 1 + 1.0 / alpha;
        double H = (HL + RT * HR) / (1 + RT);
        double a = std::sqrt((gamma - 1) * (H - 0.5 * (vx * vx + vy * vy)));
        double i = (iL + RT * iR) / (1 + RT);
        if (on[0]) {
          double drho = rhoR - rhoL;
          double dp = (gamma - 1) * (gamma - 1) * (alpha * (iR - iL) - i * (alphaR - alphaL));
          double dV2 = drho - dp / (a * a);
          double G = std::fabs(dV2) / (rhoL + rhoR);
          double PHI = 2 * G / ((1 + G) * (1 + G));
          if (PHI > thresh[0]) { sensors[e] = 1; sensors[right] = 1; done = true; }
        }
        if (on[1] && !done) {
          if (((vxL - aL > vx - a) && (vx - a > vxR - aR)) || ((vyL - aL > vy - a) && (vy - a > vyR - aR))) {
            double PSI = std::fabs(pR - pL) / (pL + pR);
            double W = 2 * PSI / ((1 + PSI) * (1 + PSI));
            if (W > thresh[1]) { sensors[e] = 2; sensors[right] = 2; done = true; }
          }
        }
        if (on[2] && !done) {
          double G = std::fabs(gammaR - gammaL) / (gammaL + gammaR);
          double PHI = 2 * G / ((1 + G) * (1 + G));
          if (PHI > thresh[2]) { sensors[e] = 3; sensors[right] = 3; done = true; }
        }
      } else {
        sensors[e] = 1;
        done = true;
      }
    }
  }
}

// sensor.cc:11-36
void or_sensing(const or_cfg* c, const int* neighbors, const double* U, const int* on, const double* thresh,
                int* sensors, double* Uavg) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  for (int e = 0; e < N_E; e++) sensors[e] = 0;
  std::vector<double> ones(N_s, 1.0);
  or_dgemm('N', 'N', 1, N_E * N_F, N_s, 1.0 / (double)N_s, ones.data(), 1, U, N_s, 0.0, Uavg, 1);
  or_calc_sensors(c, on, thresh, neighbors, Uavg, sensors);
}

// limiting_kernels.cu:322-427
void or_hri1D(const or_cfg* c, const int* neighbors, int N_s1D, int slicenum, int offxy, const double* L2M,
              const double* M2L, const int* sensors, const double* U, double* Unew) {
  const int N_s = c->N_s, N_F = c->N_F, N_N = c->N_N;
  int N = N_s1D - 1;
  std::vector<double> share(7 * N_s);
  for (int e = 0; e < c->N_E; e++) {
    if (sensors[e] == 0) continue;
    for (int fc = 0; fc < N_F; fc++) {
      double* UL = &share[0];
      double* UC = &share[N_s];
      double* UR = &share[2 * N_s];
      double* tmp = &share[3 * N_s];
      int left = neighbors[e * N_N + offxy + 0];
      int right = neighbors[e * N_N + offxy + 1];
      int physical = 0;
      if (left < 0) physical = -left;
      if (right < 0) physical = -right;
      for (int i = 0; i < N_s; i++) UC[i] = U[(e * N_F + fc) * N_s + i];
      if (physical == 4) {
      } else if (physical == 2) {
        set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp, UC);
      } else if (physical == 3) {
        int un_idx = 0;
        if (offxy == 0) un_idx = 1;
        else if (offxy == 2) un_idx = 2;
        if (fc != un_idx) set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp, UC);
      } else {
        for (int i = 0; i < N_s; i++) UL[i] = U[(left * N_F + fc) * N_s + i];
        for (int i = 0; i < N_s; i++) UR[i] = U[(right * N_F + fc) * N_s + i];
        HRslices(N_s, N, N_s1D, slicenum, L2M, M2L, tmp, UL, UC, UR);
      }
      for (int i = 0; i < N_s; i++) Unew[(e * N_F + fc) * N_s + i] = UC[i];
    }
  }
}

// sensor_kernels.cu:167-200
void or_copy_detected(const or_cfg* c, const int* sensors, const double* Uold, double* U) {
  const int N_s = c->N_s, N_F = c->N_F;
  for (int e = 0; e < c->N_E; e++) {
    if (sensors[e] == 0) continue;
    for (int i = 0; i < N_s; i++)
      for (int fc = 0; fc < N_F; fc++) U[(e * N_F + fc) * N_s + i] = Uold[(e * N_F + fc) * N_s + i];
  }
}

// The HR sweeps of limiting.cc:596-634 on given flags (sensor values), without
// the sensor: used with the flags of the learned indicator.
void or_limit_flagged(const or_cfg* c, const or_tables* t, const int* flags, double* U) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  std::vector<double> Utmp((size_t)N_s * N_E * N_F, 0.0);
  if (c->D == 1) {
    or_hri1D(c, t->neighbors, N_s, 1, 0, t->Lag2Mono, t->Mono2Lag, flags, U, Utmp.data());
    or_copy_detected(c, flags, Utmp.data(), U);
  } else {
    const int n1 = c->N_s1D;
    or_hri1D(c, t->neighbors, n1, n1, 0, t->Lag2MonoX, t->MonoX2Lag, flags, U, Utmp.data());
    or_copy_detected(c, flags, Utmp.data(), U);
    or_hri1D(c, t->neighbors, n1, n1, 2, t->Lag2MonoY, t->MonoY2Lag, flags, U, Utmp.data());
    or_copy_detected(c, flags, Utmp.data(), U);
  }
}

// limiting.cc:596-634
void or_limit_hri(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  std::vector<double> Uavg((size_t)N_E * N_F), Utmp((size_t)N_s * N_E * N_F, 0.0);
  or_sensing(c, t->neighbors, U, on, thresh, sensors, Uavg.data());
  if (c->D == 1) {
    or_hri1D(c, t->neighbors, N_s, 1, 0, t->Lag2Mono, t->Mono2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
  } else {
    const int n1 = c->N_s1D;
    or_hri1D(c, t->neighbors, n1, n1, 0, t->Lag2MonoX, t->MonoX2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
    or_hri1D(c, t->neighbors, n1, n1, 2, t->Lag2MonoY, t->MonoY2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
  }
}

// rk.cc:132-247 (no particles, single rank).  Output times: out_times[0..n_out-1] are the reference's
// _output_time_array[1..] (rk.h:58-83; the initial time 0 is entry 0 there), the last one the final time;
// out_steps[k] receives the step count after which output k happened (CFL mode; the printing itself is left
// to the caller).
int or_rk4_out(const or_cfg* c, const or_tables* t, double* U, int mode, double dt_or_cfl, int nsteps,
               const double* out_times, int n_out, int limiter, const int* on, const double* thresh, double* t_out,
               double* dt_hist, int max_hist, int* out_steps) {
  const double beta[4] = {0.0, 0.5, 0.5, 1.0};
  const double gam[4] = {1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0};
  const size_t n = (size_t)c->N_s * c->N_E * c->N_F;
  std::vector<double> Us(n), Ustar(n), DU(n, 0.0), f(n);
  std::vector<int> sens(c->N_E, 0);
  auto limit = [&](double* X) {
    if (limiter == 1) or_limit_hr(c, t, X);
    else if (limiter == 2) or_limit_m2(c, t, X);
    else if (limiter == 3) or_limit_hri(c, t, on, thresh, X, sens.data());
    else if (limiter == 4) or_limit_m2i(c, t, on, thresh, X, sens.data());
    else if (limiter == 5) or_limit_pr(c, t, X);
    else if (limiter == 6) or_limit_pri(c, t, on, thresh, X, sens.data());
    else if (limiter == 7) or_limit_p0i(c, t, on, thresh, X, sens.data());
  };
  double T = 0;
  int steps = 0;
  bool done = false;
  const double tf = n_out > 0 ? out_times[n_out - 1] : 0.0;
  int count = 0;  // next output: out_times[count] (rk.cc:127-128)
  if (mode == 1) limit(U);  // CFL runs limit the initial condition first (rk.cc:100-114)
  while (!done) {
    double Dt;
    bool output = false;
    if (mode == 0) {
      if (steps >= nsteps) break;
      Dt = dt_or_cfl;
    } else {
      if (nsteps > 0 && steps >= nsteps) break;  // optional step cap (not in the reference)
      Dt = or_dt_from_cfl(c, U, dt_or_cfl);
      if (Dt < 1e-14 || Dt != Dt) { if (t_out) *t_out = T; return -1; }
      const double Tout = out_times[count < n_out ? count : n_out - 1];
      if (Dt > (tf - T)) { Dt = tf - T; output = true; done = true; }  // rk.cc:149
      else if (Dt > (Tout - T)) { Dt = Tout - T; output = true; }      // rk.cc:150
    }
    if (dt_hist && steps < max_hist) dt_hist[steps] = Dt;
    std::memcpy(Us.data(), U, n * sizeof(double));
    for (int k = 0; k < 4; k++) {
      std::memcpy(Ustar.data(), Us.data(), n * sizeof(double));
      if (k > 0)
        for (size_t m = 0; m < n; m++) Ustar[m] = Ustar[m] + beta[k] * DU[m];  // daxpy
      if (k > 0) limit(Ustar.data());
      or_rhs(c, t, Ustar.data(), f.data());
      or_solve(c, Dt, t->Minv, f.data(), DU.data());
      for (size_t m = 0; m < n; m++) U[m] = U[m] + gam[k] * DU[m];  // daxpy
    }
    limit(U);
    T = T + Dt;
    steps++;
    if (output) {  // rk.cc:215-233
      if (out_steps && count < n_out) out_steps[count] = steps;
      count++;
    }
  }
  if (t_out) *t_out = T;
  return steps;
}

int or_rk4(const or_cfg* c, const or_tables* t, double* U, int mode, double dt_or_cfl, int nsteps, double tf,
           int limiter, const int* on, const double* thresh, double* t_out, double* dt_hist, int max_hist) {
  return or_rk4_out(c, t, U, mode, dt_or_cfl, nsteps, &tf, 1, limiter, on, thresh, t_out, dt_hist, max_hist, nullptr);
}

// ---- launcher-level restatements of the reference's helper kernels (CPU loop order)
// stridedcopy (limiting_kernels.cu:51-108)
void or_stridedcopy(int numblocks, int blocklen, int strideA, int strideB, int offsetA, int offsetB, const double* A,
                    double* B) {
  long indA = offsetA, indB = offsetB;
  for (int i = 0; i < numblocks; i++) {
    for (int j = 0; j < blocklen; j++) B[indB + j] = A[indA + j];
    indA += strideA;
    indB += strideB;
  }
}
// reconstruct_energy (limiting_kernels.cu:111-158)
void or_reconstruct_energy(int N_s, int N_E, int slicenum, const double* rhoeLim, const double* KLim,
                           const double* EMono, double* ELim) {
  for (int e = 0; e < N_E; e++)
    for (int sl = 0; sl < slicenum; sl++) {
      const size_t o = ((size_t)e * slicenum + sl) * N_s;
      for (int i = 1; i < N_s; i++) ELim[o + i] = rhoeLim[o + i] + KLim[o + i];
      double E0 = EMono[o];
      for (int k = 2; k < N_s; k += 2) E0 -= 1.0 / ((double)lim_factorial(k + 1)) * (ELim[o + k] - EMono[o + k]);
      ELim[o] = E0;
    }
}
static int binom(int n, int k) {
  int r = 1;
  for (int j = 1; j <= k; j++) r = r * (n - k + j) / j;
  return r;
}
// internal_energy_multifluid / internal_energy_stiffened (limiting_kernels.cu:160-243); b may be NULL
void or_internal_energy(int N_s, int N_E, int slicenum, const double* p, const double* g, const double* b,
                        double* rhoe) {
  for (int e = 0; e < N_E; e++)
    for (int sl = 0; sl < slicenum; sl++)
      for (int i = 0; i < N_s; i++) {
        const size_t o = ((size_t)e * slicenum + sl) * N_s;
        double I = 0;
        for (int k = 0; k < i + 1; k++) I += (double)binom(i, k) * p[o + i - k] * g[o + k];
        rhoe[o + i] = b ? I + b[o + i] : I;
      }
}
// pressure (physics.cu:891-960), every physics (GAMNCON; stiffened E - beta)
void or_pressure(const or_cfg* c, const double* U, double* p) {
  const int N_s = c->N_s, N_F = c->N_F, D = c->D, ph = c->physics;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
      auto u = [&](int f) { return U[((size_t)e * N_F + f) * N_s + i]; };
      const double rho = u(0), rhou = u(1), rhov = D == 2 ? u(2) : 0.0;
      double E = u(D + 1);
      double gamma = GAMMA_SF;
      if (ph == 1 || ph == 3) gamma = 1.0 + 1.0 / u(D + 2);
      if (ph == 3) E = E - u(D + 3);
      p[(size_t)e * N_s + i] = D == 1 ? (gamma - 1) * (E - 0.5 * rhou * rhou / rho)
                                      : (gamma - 1) * (E - 0.5 * (rhou * rhou + rhov * rhov) / rho);
    }
}
// kinetic_energy1D / kinetic_energy2D (physics.cu:814-890); rhov NULL in 1D
void or_kinetic_energy(int N_s, int N_E, const double* rho, const double* rhou, const double* rhov, double* K) {
  for (size_t k = 0; k < (size_t)N_s * N_E; k++)
    K[k] = rhov ? 0.5 * (rhou[k] * rhou[k] + rhov[k] * rhov[k]) / rho[k] : 0.5 * rhou[k] * rhou[k] / rho[k];
}
// average_cell_p0 (rk_kernels.cu:61-92)
void or_average_cell_p0(const or_cfg* c, double* DU) {
  const int N_s = c->N_s, N_F = c->N_F;
  for (int e = 0; e < c->N_E; e++)
    for (int fc = 0; fc < N_F; fc++) {
      double average = 0.0;
      for (int i = 0; i < N_s; i++) average += DU[((size_t)e * N_F + fc) * N_s + i];
      average = average / N_s;
      for (int i = 0; i < N_s; i++) DU[((size_t)e * N_F + fc) * N_s + i] = average;
    }
}
// ChangeBasis (limiting_kernels.cu:918-955): per-element transform
void or_change_basis(const or_cfg* c, int size1, int size2, const double* T, const double* U, double* Unew) {
  const int N_F = c->N_F;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < size1; i++)
      for (int fc = 0; fc < N_F; fc++) {
        double sol = 0.0;
        for (int ii = 0; ii < size2; ii++)
          sol += T[((size_t)e * size1 + i) * size2 + ii] * U[((size_t)e * N_F + fc) * size2 + ii];
        Unew[((size_t)e * N_F + fc) * size1 + i] = sol;
      }
}
// Cons2Prim / Prim2Cons (limiting_kernels.cu:957-1068): alpha is field D+2 (two-fluid 1/(gamma-1), passive rho
// phic), beta field D+3 for stiffened states, 0 otherwise; GAMNCON
void or_cons_prim(const or_cfg* c, double* U, int to_prim) {
  const int N_s = c->N_s, N_F = c->N_F, D = c->D, ph = c->physics;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
      auto u = [&](int f) -> double& { return U[((size_t)e * N_F + f) * N_s + i]; };
      const double beta = ph == 3 ? u(D + 3) : 0.0;
      if (to_prim) {
        const double rho = u(0), rhou = u(1), rhov = D == 2 ? u(2) : 0.0, E = u(D + 1), alpha = u(D + 2);
        u(0) = rho;
        u(1) = rhou / rho;
        if (D == 2) u(2) = rhov / rho;
        u(D + 1) = (E - beta - 0.5 * (rhou * rhou + rhov * rhov) / rho) / alpha;
      } else {
        const double rho = u(0), uu = u(1), v = D == 2 ? u(2) : 0.0, p = u(D + 1), alpha = u(D + 2);
        u(0) = rho;
        u(1) = rho * uu;
        if (D == 2) u(2) = rho * v;
        u(D + 1) = p * alpha + beta + 0.5 * rho * (uu * uu + v * v);
      }
    }
}
// format_sensor (printer_sensor.cu:19-43)
void or_format_sensor(int N_s, int N_E, const int* sensor, double* output) {
  for (int e = 0; e < N_E; e++)
    for (int i = 0; i < N_s; i++) output[(size_t)e * N_s + i] = sensor[e];
}

// DG_SOLVER::conservation (dg_solver.h:367-389): UgC = phi*U (netlib gemm), then
// I[fc] += UgC*J_e*w_g over elements and points, one running sum per field.
void or_conservation(const or_cfg* c, const double* phi, const double* J, const double* weight, const double* U,
                     double* I) {
  const int N_G = c->N_G, N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  std::vector<double> UgC((size_t)N_G * N_E * N_F);
  or_dgemm('N', 'N', N_G, N_E * N_F, N_s, 1, phi, N_G, U, N_s, 0.0, UgC.data(), N_G);
  for (int fc = 0; fc < N_F; fc++) I[fc] = 0.0;
  for (int fc = 0; fc < N_F; fc++)
    for (int e = 0; e < N_E; e++)
      for (int g = 0; g < N_G; g++) I[fc] += UgC[((size_t)e * N_F + fc) * N_G + g] * J[e] * weight[g];
}

// PRINTER::print formatter (printer_twod_multifluid.cu:20-110,
// printer_twod_singlefluid.cu:20-90, printer_oned_*.cu): node-wise
// conserved <-> printed primitive fields, GAMNCON (gamma = 1 + 1/alpha).
void or_format(const or_cfg* c, double* U, double* out, int inverse) {
  const int NF = c->N_F, Ns = c->N_s, D = c->D, ph = c->physics;
  const int x = D + 1;  // first field after the velocities
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < Ns; i++) {
      auto u = [&](int f) -> double& { return U[(e * NF + f) * Ns + i]; };
      auto o = [&](int f) -> double& { return out[(e * NF + f) * Ns + i]; };
      if (!inverse) {
        const double rho = u(0), ux = u(1) / rho;
        const double uy = D == 2 ? u(2) / rho : 0.0;
        const double et = u(x);
        o(0) = rho;
        o(1) = ux;
        if (D == 2) o(2) = uy;
        const double ke = D == 2 ? 0.5 * (ux * ux + uy * uy) * rho : 0.5 * ux * ux * rho;
        if (ph == 3) {  // printer_{oned,twod}_stiffened.cu
          const double gamma = 1 + 1.0 / u(x + 1);
          const double pinf = (1 - 1.0 / gamma) * u(x + 2);
          o(x) = gamma;
          o(x + 1) = pinf;
          o(x + 2) = (gamma - 1) * (et - gamma * pinf / (gamma - 1) - ke);
        } else {
          const double gamma = ph == 1 ? 1 + 1.0 / u(x + 1) : GAMMA_SF;
          if (ph == 1) o(x) = gamma;
          o(ph == 1 ? x + 1 : x) = (gamma - 1) * (et - ke);
          if (ph == 2) {  // printer_{oned,twod}_passive.cu
            o(x + 1) = u(x + 1) / rho;
            o(x + 2) = u(x + 2);
          }
        }
      } else {
        const double rho = o(0), ux = o(1), uy = D == 2 ? o(2) : 0.0;
        const double ke = D == 2 ? 0.5 * rho * (ux * ux + uy * uy) : 0.5 * rho * ux * ux;
        u(0) = rho;
        u(1) = rho * ux;
        if (D == 2) u(2) = rho * uy;
        if (ph == 3) {
          const double gamma = o(x), pinf = o(x + 1), p = o(x + 2);
          u(x) = p / (gamma - 1) + gamma * pinf / (gamma - 1) + ke;
          u(x + 1) = 1.0 / (gamma - 1);
          u(x + 2) = gamma * pinf / (gamma - 1);
        } else {
          const double gamma = ph == 1 ? o(x) : GAMMA_SF;
          const double p = o(ph == 1 ? x + 1 : x);
          u(x) = p / (gamma - 1) + ke;
          if (ph == 1) u(x + 1) = 1.0 / (gamma - 1);
          if (ph == 2) {
            const double phic = o(x + 1), phinc = o(x + 2);
            u(x + 1) = rho * phic;
            u(x + 2) = phinc;
          }
        }
      }
    }
}

// ---- m2 limiting pieces (reference launcher semantics)
static void m2_field(int N_s, int N_E, int N_F, int f, const double* U, double* x) {  // Lstridedcopy
  for (int e = 0; e < N_E; e++)
    for (int i = 0; i < N_s; i++) x[(size_t)e * N_s + i] = U[((size_t)e * N_F + f) * N_s + i];
}
static void m2_store(int N_s, int N_E, int N_F, int f, const double* x, double* U) {
  for (int e = 0; e < N_E; e++)
    for (int i = 0; i < N_s; i++) U[((size_t)e * N_F + f) * N_s + i] = x[(size_t)e * N_s + i];
}
// pressure (physics.cu:891-960), GAMNCON
static void m2_pressure(const or_cfg* c, const double* U, double* p) {
  const int N_s = c->N_s, N_F = c->N_F, D = c->D;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
      auto u = [&](int f) { return U[((size_t)e * N_F + f) * N_s + i]; };
      if (D == 1) {
        double rho = u(0), rhou = u(1), E = u(2), gamma = 1.0 + 1.0 / u(3);
        p[(size_t)e * N_s + i] = (gamma - 1) * (E - 0.5 * rhou * rhou / rho);
      } else {
        double rho = u(0), rhou = u(1), rhov = u(2), E = u(3), gamma = 1.0 + 1.0 / u(4);
        p[(size_t)e * N_s + i] = (gamma - 1) * (E - 0.5 * (rhou * rhou + rhov * rhov) / rho);
      }
    }
}
// kinetic_energy1D / kinetic_energy2D (physics.cu:814-890)
static void m2_kinetic(int D, size_t n, const double* rho, const double* rhou, const double* rhov, double* K) {
  for (size_t k = 0; k < n; k++)
    K[k] = D == 1 ? 0.5 * rhou[k] * rhou[k] / rho[k] : 0.5 * (rhou[k] * rhou[k] + rhov[k] * rhov[k]) / rho[k];
}
static int binomial(int n, int k) {
  int r = 1;
  for (int j = 1; j <= k; j++) r = r * (n - k + j) / j;
  return r;
}
// internal_energy_multifluid (limiting_kernels.cu:160-200): Taylor product of p and 1/(gamma-1)
static void m2_internal(int N_s, int N_E, int slicenum, const double* p, const double* g, double* rhoe) {
  for (int e = 0; e < N_E; e++)
    for (int sl = 0; sl < slicenum; sl++)
      for (int i = 0; i < N_s; i++) {
        double I = 0;
        const size_t o = (size_t)e * N_s * slicenum + sl * N_s;
        for (int k = 0; k < i + 1; k++) I += (double)binomial(i, k) * p[o + i - k] * g[o + k];
        rhoe[o + i] = I;
      }
}
// reconstruct_energy (limiting_kernels.cu:111-158)
static void m2_reconstruct(int N_s, int N_E, int slicenum, const double* rhoeLim, const double* KLim,
                           const double* EMono, double* ELim) {
  for (int e = 0; e < N_E; e++)
    for (int sl = 0; sl < slicenum; sl++) {
      const size_t o = (size_t)e * N_s * slicenum + sl * N_s;
      for (int i = 1; i < N_s; i++) ELim[o + i] = rhoeLim[o + i] + KLim[o + i];
      double E0 = EMono[o];
      for (int k = 2; k < N_s; k += 2) E0 -= 1.0 / ((double)lim_factorial(k + 1)) * (ELim[o + k] - EMono[o + k]);
      ELim[o] = E0;
    }
}

void or_limit_m2(const or_cfg* c, const or_tables* t, double* U) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F, N_N = c->N_N, D = c->D;
  const size_t n = (size_t)N_s * N_E;
  std::vector<double> rho(n), rhou(n), rhov(n), E(n), gam(n), pres(n), K(n), mono(n), lim(n, 0.0), gLim(n, 0.0),
      pLim(n, 0.0), KLim(n), EMono(n), rhoeLim(n), ELim(n, 0.0);
  auto gemm = [&](const double* M, const double* X, double* Y) {
    or_dgemm('N', 'N', N_s, N_E, N_s, 1, M, N_s, X, N_s, 0.0, Y, N_s);
  };
  if (D == 1) {
    auto hr = [&](double* X, double* Lim, bool back) {
      gemm(t->Lag2Mono, X, mono.data());
      or_hrl1D(N_s, N_E, 1, N_N, 1, t->neighbors, 0, mono.data(), Lim);
      if (back) gemm(t->Mono2Lag, Lim, X);
    };
    m2_field(N_s, N_E, N_F, 0, U, rho.data());
    hr(rho.data(), lim.data(), true);
    m2_field(N_s, N_E, N_F, 1, U, rhou.data());
    hr(rhou.data(), lim.data(), true);
    m2_field(N_s, N_E, N_F, 2, U, E.data());
    gemm(t->Lag2Mono, E.data(), EMono.data());
    m2_field(N_s, N_E, N_F, 3, U, gam.data());
    hr(gam.data(), gLim.data(), true);
    m2_pressure(c, U, pres.data());
    hr(pres.data(), pLim.data(), false);
    m2_kinetic(1, n, rho.data(), rhou.data(), nullptr, K.data());
    gemm(t->Lag2Mono, K.data(), KLim.data());
    m2_internal(N_s, N_E, 1, pLim.data(), gLim.data(), rhoeLim.data());
    m2_reconstruct(N_s, N_E, 1, rhoeLim.data(), KLim.data(), EMono.data(), ELim.data());
    gemm(t->Mono2Lag, ELim.data(), E.data());
    m2_store(N_s, N_E, N_F, 0, rho.data(), U);
    m2_store(N_s, N_E, N_F, 1, rhou.data(), U);
    m2_store(N_s, N_E, N_F, 2, E.data(), U);
    m2_store(N_s, N_E, N_F, 3, gam.data(), U);
    return;
  }
  const int n1 = c->N_s1D;
  auto hr = [&](const double* L2M, const double* M2L, int offxy, double* X, double* Lim, bool back) {
    gemm(L2M, X, mono.data());
    or_hrl1D(n1, N_E, 1, N_N, n1, t->neighbors, offxy, mono.data(), Lim);
    if (back) gemm(M2L, Lim, X);
  };
  // x
  m2_field(N_s, N_E, N_F, 0, U, rho.data());
  hr(t->Lag2MonoX, t->MonoX2Lag, 0, rho.data(), lim.data(), true);
  m2_field(N_s, N_E, N_F, 1, U, rhou.data());
  hr(t->Lag2MonoX, t->MonoX2Lag, 0, rhou.data(), lim.data(), true);
  m2_field(N_s, N_E, N_F, 2, U, rhov.data());
  hr(t->Lag2MonoX, t->MonoX2Lag, 0, rhov.data(), lim.data(), true);
  m2_field(N_s, N_E, N_F, 4, U, gam.data());
  hr(t->Lag2MonoX, t->MonoX2Lag, 0, gam.data(), gLim.data(), true);
  m2_pressure(c, U, pres.data());
  hr(t->Lag2MonoX, t->MonoX2Lag, 0, pres.data(), pLim.data(), true);
  m2_kinetic(2, n, rho.data(), rhou.data(), rhov.data(), K.data());
  gemm(t->Lag2MonoX, K.data(), KLim.data());
  m2_internal(n1, N_E, n1, pLim.data(), gLim.data(), rhoeLim.data());
  m2_field(N_s, N_E, N_F, 3, U, E.data());
  gemm(t->Lag2MonoX, E.data(), EMono.data());
  m2_reconstruct(n1, N_E, n1, rhoeLim.data(), KLim.data(), EMono.data(), ELim.data());
  gemm(t->MonoX2Lag, ELim.data(), E.data());
  // y
  hr(t->Lag2MonoY, t->MonoY2Lag, 2, rho.data(), lim.data(), true);
  hr(t->Lag2MonoY, t->MonoY2Lag, 2, rhou.data(), lim.data(), true);
  hr(t->Lag2MonoY, t->MonoY2Lag, 2, rhov.data(), lim.data(), true);
  hr(t->Lag2MonoY, t->MonoY2Lag, 2, gam.data(), gLim.data(), true);
  hr(t->Lag2MonoY, t->MonoY2Lag, 2, pres.data(), pLim.data(), false);
  m2_kinetic(2, n, rho.data(), rhou.data(), rhov.data(), K.data());
  gemm(t->Lag2MonoY, K.data(), KLim.data());
  m2_internal(n1, N_E, n1, pLim.data(), gLim.data(), rhoeLim.data());
  gemm(t->Lag2MonoY, E.data(), EMono.data());
  m2_reconstruct(n1, N_E, n1, rhoeLim.data(), KLim.data(), EMono.data(), ELim.data());
  gemm(t->MonoY2Lag, ELim.data(), E.data());
  m2_store(N_s, N_E, N_F, 0, rho.data(), U);
  m2_store(N_s, N_E, N_F, 1, rhou.data(), U);
  m2_store(N_s, N_E, N_F, 2, rhov.data(), U);
  m2_store(N_s, N_E, N_F, 3, E.data(), U);
  m2_store(N_s, N_E, N_F, 4, gam.data(), U);
}

// Cons2Prim / Prim2Cons (limiting_kernels.cu:957-1068), two-fluid GAMNCON
// (beta = 0, rhov = v = 0 in 1D, in the reference's expressions)
static void pr_cons2prim(const or_cfg* c, double* U) {
  const int N_s = c->N_s, N_F = c->N_F, D = c->D;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
      auto u = [&](int f) -> double& { return U[((size_t)e * N_F + f) * N_s + i]; };
      const double beta = 0;
      double rho = u(0), rhou = u(1), rhov = D == 2 ? u(2) : 0.0, E = u(D + 1), alpha = u(D + 2);
      u(0) = rho;
      u(1) = rhou / rho;
      if (D == 2) u(2) = rhov / rho;
      u(D + 1) = (E - beta - 0.5 * (rhou * rhou + rhov * rhov) / rho) / alpha;
    }
}
static void pr_prim2cons(const or_cfg* c, double* U) {
  const int N_s = c->N_s, N_F = c->N_F, D = c->D;
  for (int e = 0; e < c->N_E; e++)
    for (int i = 0; i < N_s; i++) {
      auto u = [&](int f) -> double& { return U[((size_t)e * N_F + f) * N_s + i]; };
      const double beta = 0;
      double rho = u(0), uu = u(1), v = D == 2 ? u(2) : 0.0, p = u(D + 1), alpha = u(D + 2);
      u(0) = rho;
      u(1) = rho * uu;
      if (D == 2) u(2) = rho * v;
      u(D + 1) = p * alpha + beta + 0.5 * rho * (uu * uu + v * v);
    }
}

void or_limit_pr(const or_cfg* c, const or_tables* t, double* U) {
  pr_cons2prim(c, U);
  or_limit_hr(c, t, U);
  pr_prim2cons(c, U);
}

// ---- m2i: per-element m2 limiting on the flagged cells (two-fluid, GAMNCON)

// pressure (limiting_kernels.cu:1870-1897), MULTIFLUID
static void m2i_pressure(int D, int N_s, const double* rho, const double* rhou, const double* rhov, const double* E,
                         const double* g, double* p) {
  for (int i = 0; i < N_s; i++)
    p[i] = D == 1 ? (E[i] - 0.5 * rhou[i] * rhou[i] / rho[i]) / g[i]
                  : (E[i] - 0.5 * (rhou[i] * rhou[i] + rhov[i] * rhov[i]) / rho[i]) / g[i];
}

// m2i1D (limiting_kernels.cu:431-722) with kinetic_energy, internal_energy and
// reconstruct_total_energy (:1899-2001): each flagged element is limited in
// one direction from its own and its two neighbours' nodal values; rho, the
// momenta, 1/(gamma-1) and the pressure go through HR (or the boundary
// rules), and the energy is rebuilt from the limited pressure and
// 1/(gamma-1) (internal energy, Taylor product) plus the kinetic energy of the
// limited fields, with the zeroth coefficient corrected for conservation.
void or_m2i1D(const or_cfg* c, const int* neighbors, int N_s1D, int slicenum, int offxy, const double* L2M,
              const double* M2L, const int* sensors, const double* U, double* Unew) {
  const int N_s = c->N_s, N_F = c->N_F, N_N = c->N_N, D = c->D;
  const int N = N_s1D - 1;
  const int fE = D + 1, fG = D + 2;
  std::vector<double> L(N_F * N_s), C(N_F * N_s), R(N_F * N_s), pL(N_s), pC(N_s), pR(N_s), gLim(N_s), pLim(N_s),
      KLim(N_s), rhoeLim(N_s), tmp(4 * N_s);
  std::vector<double> zero(N_s, 0.0);
  auto fld = [&](std::vector<double>& X, int f) { return &X[f * N_s]; };
  for (int e = 0; e < c->N_E; e++) {
    if (sensors[e] == 0) continue;
    const int left = neighbors[e * N_N + offxy + 0], right = neighbors[e * N_N + offxy + 1];
    int physical = 0;
    if (left < 0) physical = -left;
    if (right < 0) physical = -right;
    for (int f = 0; f < N_F; f++)
      for (int i = 0; i < N_s; i++) C[f * N_s + i] = U[((size_t)e * N_F + f) * N_s + i];
    const double* vC = D == 2 ? fld(C, 2) : zero.data();
    m2i_pressure(D, N_s, fld(C, 0), fld(C, 1), vC, fld(C, fE), fld(C, fG), pC.data());
    bool energy = true;
    if (physical == 4) {
      energy = false;
    } else if (physical == 2 || physical == 3) {
      set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(C, 0));
      if (physical == 2) {
        for (int f = 1; f <= D; f++) set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(C, f));
      } else if (D == 2) {
        // the wall-normal momentum keeps its values
        if (offxy == 0) set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(C, 2));
        else if (offxy == 2) set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(C, 1));
      }
      set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(C, fG), gLim.data());
      set2average(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), pC.data(), pLim.data());
    } else {
      for (int f = 0; f < N_F; f++)
        for (int i = 0; i < N_s; i++) {
          L[f * N_s + i] = U[((size_t)left * N_F + f) * N_s + i];
          R[f * N_s + i] = U[((size_t)right * N_F + f) * N_s + i];
        }
      m2i_pressure(D, N_s, fld(L, 0), fld(L, 1), D == 2 ? fld(L, 2) : zero.data(), fld(L, fE), fld(L, fG), pL.data());
      m2i_pressure(D, N_s, fld(R, 0), fld(R, 1), D == 2 ? fld(R, 2) : zero.data(), fld(R, fE), fld(R, fG), pR.data());
      for (int f = 0; f <= D; f++) HRslices(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(L, f), fld(C, f), fld(R, f));
      HRslices(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), fld(L, fG), fld(C, fG), fld(R, fG), gLim.data());
      HRslices(N_s, N, N_s1D, slicenum, L2M, M2L, tmp.data(), pL.data(), pC.data(), pR.data(), pLim.data());
    }
    if (energy) {
      // kinetic_energy (:1899-1922): nodal K of the limited fields -> monomial
      const double* rho = fld(C, 0);
      const double* ru = fld(C, 1);
      const double* rv = D == 2 ? fld(C, 2) : zero.data();
      for (int i = 0; i < N_s; i++)
        tmp[i] = D == 1 ? 0.5 * (ru[i] * ru[i]) / rho[i] : 0.5 * (ru[i] * ru[i] + rv[i] * rv[i]) / rho[i];
      gemv_cm(N_s, N_s, L2M, tmp.data(), KLim.data());
      // internal_energy (:1924-1953)
      for (int sl = 0; sl < slicenum; sl++)
        for (int i = 0; i < N_s1D; i++) {
          double I = 0;
          for (int k = 0; k < i + 1; k++)
            I += (double)binomial(i, k) * pLim[sl * N_s1D + i - k] * gLim[sl * N_s1D + k];
          rhoeLim[sl * N_s1D + i] = I;
        }
      // reconstruct_total_energy (:1955-2001)
      double* EMono = tmp.data();
      double* ELim = tmp.data() + N_s;
      double* EC = fld(C, fE);
      gemv_cm(N_s, N_s, L2M, EC, EMono);
      for (int sl = 0; sl < slicenum; sl++) {
        const int i0 = sl * N_s1D;
        for (int i = 1; i < N_s1D; i++) ELim[i0 + i] = rhoeLim[i0 + i] + KLim[i0 + i];
        double E0 = EMono[i0];
        for (int k = 2; k < N_s1D; k += 2)
          E0 -= 1.0 / ((double)lim_factorial(k + 1)) * (ELim[i0 + k] - EMono[i0 + k]);
        ELim[i0] = E0;
      }
      gemv_cm(N_s, N_s, M2L, ELim, EC);
    }
    for (int f = 0; f < N_F; f++)
      for (int i = 0; i < N_s; i++) Unew[((size_t)e * N_F + f) * N_s + i] = C[f * N_s + i];
  }
}

// Limiting::M2Ilimiting (limiting.cc:682-720): sensing, then m2i1D +
// copy_detected in x and in y (2D Cartesian) or once (1D)
void or_limit_m2i(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  std::vector<double> Uavg((size_t)N_E * N_F), Utmp((size_t)N_s * N_E * N_F, 0.0);
  or_sensing(c, t->neighbors, U, on, thresh, sensors, Uavg.data());
  if (c->D == 1) {
    or_m2i1D(c, t->neighbors, N_s, 1, 0, t->Lag2Mono, t->Mono2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
  } else {
    const int n1 = c->N_s1D;
    or_m2i1D(c, t->neighbors, n1, n1, 0, t->Lag2MonoX, t->MonoX2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
    or_m2i1D(c, t->neighbors, n1, n1, 2, t->Lag2MonoY, t->MonoY2Lag, sensors, U, Utmp.data());
    or_copy_detected(c, sensors, Utmp.data(), U);
  }
}

// p0i (limiting_kernels.cu:723-768): each flagged element's fields set to
// their cell average, sum_g sum_i phi_w[i][g] U_i / refArea
void or_p0i(const or_cfg* c, const double* phi_w, double refArea, const int* sensors, const double* U,
            double* Unew) {
  const int N_s = c->N_s, N_F = c->N_F, N_G = c->N_G;
  for (int e = 0; e < c->N_E; e++)
    for (int fc = 0; fc < N_F; fc++) {
      if (sensors[e] == 0) continue;
      double Uavg = 0;
      for (int g = 0; g < N_G; g++)
        for (int i = 0; i < N_s; i++) Uavg += phi_w[i * N_G + g] * U[((size_t)e * N_F + fc) * N_s + i];
      Uavg = Uavg / refArea;
      for (int i = 0; i < N_s; i++) Unew[((size_t)e * N_F + fc) * N_s + i] = Uavg;
    }
}

// reference element measure: 4 for the quadrilateral [-1,1]^2 (main.cc:286);
// the reference leaves it unset for segments (main.cc:284), here 2 = |[-1,1]|
double or_ref_area(const or_cfg* c) { return c->D == 1 ? 2.0 : 4.0; }

// Limiting::P0Ilimiting (limiting.cc:723-744): sensing, p0i, copy_detected
void or_limit_p0i(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors) {
  const int N_s = c->N_s, N_E = c->N_E, N_F = c->N_F;
  std::vector<double> Uavg((size_t)N_E * N_F), Utmp((size_t)N_s * N_E * N_F, 0.0);
  or_sensing(c, t->neighbors, U, on, thresh, sensors, Uavg.data());
  or_p0i(c, t->phi_w, or_ref_area(c), sensors, U, Utmp.data());
  or_copy_detected(c, sensors, Utmp.data(), U);
}

// Limiting::PRIlimiting (limiting.cc:636-662): sensing on the conserved state,
// then the hri sweeps on the primitive variables
void or_limit_pri(const or_cfg* c, const or_tables* t, const int* on, const double* thresh, double* U,
                  int* sensors) {
  std::vector<double> Uavg((size_t)c->N_E * c->N_F);
  or_sensing(c, t->neighbors, U, on, thresh, sensors, Uavg.data());
  pr_cons2prim(c, U);
  or_limit_flagged(c, t, sensors, U);
  pr_prim2cons(c, U);
}

}  // extern "C"
