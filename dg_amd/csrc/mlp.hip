// Learned troubled-cell indicator: a 49 -> 64 -> 64 -> 1 MLP on per-element
// features of the 2D p=2 two-fluid state, evaluated ahead of the limiter and
// trained on the flags of the reference sensor (calc_sensors,
// sensor_kernels.cu:19-163, computed by limiter.hip).  The indicator has no
// counterpart in the reference.
//
// Features of element e (49, padded with zeros to 64):
//   45: the tensor-Legendre modal coefficients m[f][a*3+b] of the 5 fields,
//       divided by a per-field scale (|mean| for rho, E and 1/(gamma-1);
//       sqrt(|mean rho| |mean E|) for the momenta), clamped to [-8, 8];
//    4: for each neighbour (L, R, D, U) the largest relative jump of the cell
//       means of rho, p and 1/(gamma-1): |a_n - a_e| / (a_n + a_e); 0 at a
//       domain boundary.
// Network: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3,
// the cell is flagged when z > threshold.  Weights and activations are BF16
// (round to nearest even), accumulation FP32: the two 64-wide layers run on
// MFMA (v_mfma_f32_16x16x16_bf16, one wave per 16 elements, weights held in
// registers, the layer-1 output fragments feed layer 2 without data
// movement).  Training: weighted binary cross-entropy, gradients by the
// backward kernels below (FP32, deterministic fixed-order reductions), Adam
// on FP32 master weights.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "dgx_internal.h"
#include "physics.cuh"
#include "sensor.cuh"

#ifndef DGX_IND_BLOCKS
#define DGX_IND_BLOCKS 768  // blocks of the flagging pass: three per CU resident at once, each wave looping over groups of 16 elements (768: 124 us, 4096: 130 us per pass on the 1024x512 state)
#endif

namespace dgx {

constexpr int kIn = 49, kInP = 64, kH = 64;
// parameter vector: W1 [64][64] (inputs >= 49 stay zero), b1, W2 [64][64], b2, w3, b3
constexpr int kW1 = 0, kB1 = kW1 + kH * kInP, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kNParam = kB3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4m __attribute__((ext_vector_type(4)));

// float -> BF16, round to nearest even.  Device code uses gfx950's
// conversion instruction (v_cvt_pk_bf16_f32, one instruction for the
// bit arithmetic below): the same bits for every value that is not a NaN
// (a NaN stays a NaN; features and activations never are one).
__host__ __device__ inline uint16_t f2bf(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
#endif
}
__host__ __device__ inline float bf2f(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// ------------------------------------------------------------- features
// cell means of rho, p and alpha = 1/(gamma-1) (cell_means_rpa), and with
// uavg the sensor's cell averages and pressure (sensor_averages_2d); a block
// stages 64 elements with coalesced loads, one lane per element reduces.  The
// stage kernel emits the same values from its output tile when the RK driver
// limits with the indicator (Solver::means_of), and this pass is skipped then.
__global__ __launch_bounds__(256) void k_mlp_means(const double* __restrict__ U, const double* __restrict__ w0, int N_E,
                                                   double* __restrict__ avg, double* __restrict__ uavg,
                                                   double* __restrict__ psens) {
  constexpr int NF = 5, Ns = 9, NU = NF * Ns;
  __shared__ double sU[64 * NU];
  const int e0 = blockIdx.x * 64, ne = min(64, N_E - e0);
  constexpr int R = (64 * NU + 255) / 256;
  double r[R];  // all loads in flight before the first LDS store
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int k = threadIdx.x + 256 * i;
    r[i] = k < ne * NU ? U[(size_t)e0 * NU + k] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int k = threadIdx.x + 256 * i;
    if (k < 64 * NU) sU[k] = r[i];
  }
  __syncthreads();
  if ((int)threadIdx.x >= ne) return;
  const int e = e0 + threadIdx.x;
  double a3[3];
  cell_means_rpa(sU + threadIdx.x * NU, w0, a3);
  avg[e * 3 + 0] = a3[0];
  avg[e * 3 + 1] = a3[1];
  avg[e * 3 + 2] = a3[2];
  if (uavg) sensor_averages_2d(sU + threadIdx.x * NU, uavg + (size_t)e * NF, psens + e);
}

__device__ inline double rel_jump(double a, double b) {
  const double s = fabs(a) + fabs(b);
  return s > 0.0 ? fabs(a - b) / s : 0.0;
}

// Modal coefficients of the 16 elements a wave has staged in LDS (element c
// at u + c*45): the tensor-Legendre transform m[f][mode] = sum_i
// Vinv[mode][i] u[f][i] of each field as v_mfma_f64_16x16x4f64 tiles (rows =
// modes padded to 16, K = nodes padded to 12, columns = elements; result
// register v of lane l holds mode l/16 + 4v of element l%16).  The modes
// replace the nodal values in place; the five field scales of element c
// (|mean| for rho, E and alpha, sqrt(|mean rho| |mean E|) for the momenta,
// the mean being mode 0) go to inv[c*5 + f] as reciprocals (0 for a zero
// scale).  Wave-local: no block barrier.
__device__ inline void wave_modes(double* u, const double* Vinv, double* inv, int lane) {
  const int c = lane & 15, q = lane >> 4;
  double a[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int k = 4 * s + q;
    a[s] = (c < 9 && k < 9) ? Vinv[c * 9 + k] : 0.0;
  }
  f64x4m d[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    f64x4m acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int k = 4 * s + q;
      const double b = k < 9 ? u[c * 45 + f * 9 + k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], b, acc, 0, 0, 0);
    }
    d[f] = acc;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();  // every lane has read its nodal values
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int f = 0; f < 5; ++f)
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const int mode = q + 4 * v;
      if (mode < 9) u[c * 45 + f * 9 + mode] = d[f][v];
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();  // the modes are in place
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // the scales from the modes 0 in place, lane group q taking field q (and
  // q = 0 field 4 too): two divisions per lane instead of five on one group
  {
    const double* m = u + c * 45;
    const double sc = (q == 1 || q == 2) ? sqrt(fabs(m[0]) * fabs(m[27])) : fabs(m[q * 9]);
    inv[c * 5 + q] = sc > 0.0 ? 1.0 / sc : 0.0;
    if (q == 0) {
      const double s4 = fabs(m[36]);
      inv[c * 5 + 4] = s4 > 0.0 ? 1.0 / s4 : 0.0;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The four neighbour features of every element (L, R, D, U): the largest
// relative jump of the cell means of rho, p and alpha (avg, k_mlp_means),
// as the float the network input takes; -1 marks a domain boundary (the
// feature is 0 there and the cell is flagged by rule).
// With uavg (the sensor's cell averages [e*5 + f]) and psens (their sensor
// pressures) the same pass writes the shock guard: guard[e] = 1 when the
// sensor's shock test (sensor_shock, the same arithmetic as calc_sensors)
// fires for a pair of e and a neighbour in either order.  Every cell the
// sensor marks as a shock is in such a pair, so the guard holds all of them,
// whatever the network says.  The pressure jump W > th, necessary for the
// test, is checked first from psens.
// eight waves per SIMD (62 registers): the pass waits on dependent gathers
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_mlp_jumps(const double* __restrict__ avg, const int* __restrict__ nb, int N,
                                                   float4* __restrict__ jumps, int* __restrict__ count_reset,
                                                   const double* __restrict__ uavg, const double* __restrict__ psens,
                                                   double th_shock, int* __restrict__ guard) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  // the flagged-element count that k_mlp_indicator appends to starts at zero
  if (count_reset && e == 0) *count_reset = 0;
  if (e >= N) return;
  int o[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) o[n] = nb[e * 4 + n];
  if (uavg) {
    const double pe = psens[e];
    int g = 0;
    for (int n = 0; n < 4; ++n)
      if (o[n] >= 0 && sensor_shock_pair<2, 1>(uavg, e, o[n], pe, psens[o[n]], th_shock)) g = 1;
    guard[e] = g;
  }
  const double r0 = avg[e * 3 + 0], p0 = avg[e * 3 + 1], a0 = avg[e * 3 + 2];
  double an[4][3];
#pragma unroll
  for (int n = 0; n < 4; ++n)
#pragma unroll
    for (int c = 0; c < 3; ++c) an[n][c] = o[n] >= 0 ? avg[o[n] * 3 + c] : 0.0;
  float jv[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    double j = rel_jump(an[n][0], r0);
    j = fmax(j, rel_jump(an[n][1], p0));
    j = fmax(j, rel_jump(an[n][2], a0));
    jv[n] = o[n] < 0 ? -1.0f : (float)j;
  }
  jumps[e] = make_float4(jv[0], jv[1], jv[2], jv[3]);
}

__device__ inline bool at_wall(float4 j) { return j.x < 0.f || j.y < 0.f || j.z < 0.f || j.w < 0.f; }

// Feature k of an element from its modes m and scales inv (wave_modes) and
// its neighbour jumps (k_mlp_jumps): scaled modes clamped to [-8, 8] (NaN to
// 0), then the four jumps (0 toward a boundary), then zeros.
__device__ inline float feature_of(int k, const double* m, const double* inv, float4 jmp) {
  if (k < 45) {
    const double v = m[k] * inv[k / 9];
    return (float)(v > 8.0 ? 8.0 : (v < -8.0 ? -8.0 : (v == v ? v : 0.0)));
  }
  const float j = k == 45 ? jmp.x : k == 46 ? jmp.y : k == 47 ? jmp.z : k == 48 ? jmp.w : 0.0f;
  return j < 0.0f ? 0.0f : j;
}

// The 16 features k = 16s + 4q + j (s, j = 0..3) of one element as the
// B fragments of the first layer: the values of feature_of, without a branch
// on the lane group q.  k < 32 are modes; k = 32 + 4q + j are modes up to 44
// and the jumps L, R, D at 45..47 (q = 3, j = 1..3); k = 48 is the jump U
// (q = 0, j = 0); the rest are zero.
__device__ inline double clamp_feature(double v) { return v > 8.0 ? 8.0 : (v < -8.0 ? -8.0 : (v == v ? v : 0.0)); }
// (float)clamp_feature(v), clamped after the conversion: the conversion is
// monotonic and +-8 are floats, so the same value for every v (infinities to
// +-8, NaN to +0, -0 kept), in single-precision instructions.
__device__ inline float clamp_feature_f(double v) {
  const float x = (float)v;
  return x == x ? fminf(fmaxf(x, -8.0f), 8.0f) : 0.0f;
}
__device__ inline float jump_feature(float j) { return j < 0.0f ? 0.0f : j; }
__device__ inline void element_features(int q, const double* m, const double* inv, float4 jmp, bf16x4 (&bx)[4]) {
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = 16 * s + 4 * q + j;
      bx[s][j] = (short)f2bf(clamp_feature_f(m[k] * inv[k / 9]));
    }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = min(32 + 4 * q + j, 44);
    const float mf = clamp_feature_f(m[k] * inv[k / 9]);
    if (j == 0) {
      bx[2][j] = (short)f2bf(mf);
    } else {
      const float jv = jump_feature(j == 1 ? jmp.x : (j == 2 ? jmp.y : jmp.z));
      bx[2][j] = (short)f2bf(q == 3 ? jv : mf);
    }
  }
  bx[3][0] = (short)f2bf(q == 0 ? jump_feature(jmp.w) : 0.0f);
  bx[3][1] = bx[3][2] = bx[3][3] = 0;
}

// Staging of 16 elements for a wave: the loads of elements e0 .. e0+ne-1
// (ne <= 16, absent ones zero) go to registers (issue), then to the wave's
// LDS rows (commit), so that a group's loads can be in flight while the
// previous group computes.
constexpr int kStageR = (16 * 45 + 63) / 64;
__device__ inline void wave_issue(const double* __restrict__ U, long e0, int ne, double (&r)[kStageR], int lane) {
#pragma unroll
  for (int i = 0; i < kStageR; ++i) {
    const int k = lane + 64 * i;
    r[i] = k < ne * 45 ? U[(size_t)e0 * 45 + k] : 0.0;
  }
}
__device__ inline void wave_commit(const double (&r)[kStageR], double* u, int lane) {
#pragma unroll
  for (int i = 0; i < kStageR; ++i) {
    const int k = lane + 64 * i;
    if (k < 16 * 45) u[k] = r[i];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Feature rows X[e][64] for training and inspection: one wave per 16
// elements, lane (column c, group q) writes the features k = 16s + 4q + j of
// element c, exactly the values the fused indicator feeds its network.
__global__ __launch_bounds__(256) void k_mlp_features(const double* __restrict__ U, const double* __restrict__ Vinv,
                                                      const float4* __restrict__ jumps, int N, float* __restrict__ X) {
  __shared__ double sU[4][16 * 45];
  __shared__ double sV[81];
  __shared__ double sInv[4][16 * 5];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, col = lane & 15, q = lane >> 4;
  for (int k = threadIdx.x; k < 81; k += 256) sV[k] = Vinv[k];
  __syncthreads();
  const int nwaves = gridDim.x * 4;
  for (int g = blockIdx.x * 4 + wv; g * 16 < N; g += nwaves) {
    const int e0 = g * 16, ne = min(16, N - e0);
    const float4 jmp = col < ne ? jumps[e0 + col] : make_float4(0.f, 0.f, 0.f, 0.f);
    double r[kStageR];
    wave_issue(U, e0, ne, r, lane);
    __builtin_amdgcn_wave_barrier();  // the previous group's LDS reads are done
    wave_commit(r, sU[wv], lane);
    wave_modes(sU[wv], sV, sInv[wv], lane);
    if (col < ne) {
      const int e = e0 + col;
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = 16 * s + 4 * q + j;
          X[(size_t)e * kInP + k] = feature_of(k, &sU[wv][col * 45], &sInv[wv][col * 5], jmp);
        }
    }
  }
}

// --------------------------------------------------------------- forward
// One wave per 16 elements (persistent over element groups).  MFMA operand
// layouts (16x16x16 bf16): A lane l = row l%16, k 4*(l/16)+0..3; B lane l =
// column l%16, k 4*(l/16)+0..3; D lane l = rows 4*(l/16)+0..3, column l%16.
// With rows = hidden units and columns = elements, the D fragment of M-tile
// t is the B fragment of k-step t of the next layer.
__global__ __launch_bounds__(256) void k_mlp_fwd(const uint16_t* __restrict__ w1b, const uint16_t* __restrict__ w2b,
                                                 const uint16_t* __restrict__ w3b, const float* __restrict__ P,
                                                 const float* __restrict__ X, const int* __restrict__ idx, int N,
                                                 float* __restrict__ z, uint16_t* __restrict__ h1s,
                                                 uint16_t* __restrict__ h2s) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, q = lane >> 4;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  bf16x4 a1[4][4], a2[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a1[t][s][j] = (short)w1b[(16 * t + col) * kInP + 16 * s + 4 * q + j];
        a2[t][s][j] = (short)w2b[(16 * t + col) * kH + 16 * s + 4 * q + j];
      }
  float b1v[4][4], b2v[4][4], w3v[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int u = 16 * t + 4 * q + v;
      b1v[t][v] = P[kB1 + u];
      b2v[t][v] = P[kB2 + u];
      w3v[t][v] = bf2f(w3b[u]);
    }
  const float b3 = P[kB3];
  for (int g = wave; g * 16 < N; g += nwaves) {
    const int n = g * 16 + col;
    const bool valid = n < N;
    const int row = valid ? (idx ? idx[n] : n) : 0;
    bf16x4 bx[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        bx[s][j] = valid ? (short)f2bf(X[(size_t)row * kInP + 16 * s + 4 * q + j]) : (short)0;
    bf16x4 bh1[4], bh2[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1[t][s], bx[s], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float h = fmaxf(acc[v] + b1v[t][v], 0.0f);
        bh1[t][v] = (short)f2bf(h);
      }
    }
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[t][s], bh1[s], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const uint16_t hb = f2bf(fmaxf(acc[v] + b2v[t][v], 0.0f));
        bh2[t][v] = (short)hb;
        part += w3v[t][v] * bf2f(hb);
      }
    }
    // sum the four lanes that hold the same element (q = 0..3), fixed order
    const float p1 = __shfl_xor(part, 16);
    const float s01 = (q & 1) ? p1 + part : part + p1;
    const float p2 = __shfl_xor(s01, 32);
    const float tot = (q & 2) ? p2 + s01 : s01 + p2;
    if (valid) {
      if (q == 0) z[n] = tot + b3;
      if (h1s) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            h1s[(size_t)n * kH + 16 * t + 4 * q + v] = (uint16_t)bh1[t][v];
            h2s[(size_t)n * kH + 16 * t + 4 * q + v] = (uint16_t)bh2[t][v];
          }
      }
    }
  }
}

// ------------------------------------------------- fused indicator pass
// Features and network in one pass (the solver's limiter 8): each wave
// stages 16 consecutive elements, forms their modes on the FP64 matrix
// cores (wave_modes) and lane (column c, group q) picks up exactly the 16
// features of element c that its MFMA B fragments hold (k = 16s + 4q + j);
// the neighbour jumps come from k_mlp_jumps.  Same features, rounding and
// network as k_mlp_features + k_mlp_fwd.  Waves work
// independently (no block barrier after the table load).
__global__ __launch_bounds__(256) void k_mlp_indicator(const double* __restrict__ U, const double* __restrict__ Vinv,
                                                       const float4* __restrict__ jumps, int N,
                                                       const uint16_t* __restrict__ w1b,
                                                       const uint16_t* __restrict__ w2b,
                                                       const uint16_t* __restrict__ w3b, const float* __restrict__ P,
                                                       float threshold, const int* __restrict__ guard,
                                                       float* __restrict__ z, int* __restrict__ flags,
                                                       int* __restrict__ list, int* __restrict__ count) {
  __shared__ double sU[4][16 * 45];
  __shared__ double sV[81];
  __shared__ double sInv[4][16 * 5];
  // the layers' A fragments per lane ([t*4 + s][lane], 8 bytes: conflict-free
  // reads) and the biases and output weights: LDS, not registers, so that
  // three waves fit on a SIMD
  __shared__ bf16x4 sA1[16][64], sA2[16][64];
  __shared__ float sB1[kH], sB2[kH], sW3[kH];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, q = lane >> 4;
  for (int k = threadIdx.x; k < 81; k += 256) sV[k] = Vinv[k];
  if (wv == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        bf16x4 f1, f2;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          f1[j] = (short)w1b[(16 * t + col) * kInP + 16 * s + 4 * q + j];
          f2[j] = (short)w2b[(16 * t + col) * kH + 16 * s + 4 * q + j];
        }
        sA1[t * 4 + s][lane] = f1;
        sA2[t * 4 + s][lane] = f2;
      }
  }
  for (int k = threadIdx.x; k < kH; k += 256) {
    sB1[k] = P[kB1 + k];
    sB2[k] = P[kB2 + k];
    sW3[k] = bf2f(w3b[k]);
  }
  const float b3 = P[kB3];
  __syncthreads();
  const int nwaves = gridDim.x * 4;
  // software pipeline: group g + nwaves is loading while group g computes
  int g = blockIdx.x * 4 + wv;
  double r[kStageR];
  float4 jn = make_float4(0.f, 0.f, 0.f, 0.f);
  if (g * 16 < N) {
    const int ne0 = min(16, N - g * 16);
    wave_issue(U, (long)g * 16, ne0, r, lane);
    if (col < ne0) jn = jumps[g * 16 + col];
  }
  for (; g * 16 < N; g += nwaves) {
    const int e0 = g * 16, ne = min(16, N - e0);
    const int e = e0 + col;
    const bool valid = col < ne;
    const float4 jmp = jn;
    __builtin_amdgcn_wave_barrier();  // the previous group's LDS reads are done
    wave_commit(r, sU[wv], lane);
    const int gn = g + nwaves;
    if (gn * 16 < N) {
      const int nen = min(16, N - gn * 16);
      wave_issue(U, (long)gn * 16, nen, r, lane);
      jn = col < nen ? jumps[gn * 16 + col] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wave_modes(sU[wv], sV, sInv[wv], lane);
    bf16x4 bx[4];
    if (valid) {
      element_features(q, &sU[wv][col * 45], &sInv[wv][col * 5], jmp, bx);
    } else {
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) bx[s][j] = 0;
    }
    bf16x4 bh1[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(sA1[t * 4 + s][lane], bx[s], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v) bh1[t][v] = (short)f2bf(fmaxf(acc[v] + sB1[16 * t + 4 * q + v], 0.0f));
    }
    float part = 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s)
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(sA2[t * 4 + s][lane], bh1[s], acc, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 4; ++v)
        part += sW3[16 * t + 4 * q + v] * bf2f(f2bf(fmaxf(acc[v] + sB2[16 * t + 4 * q + v], 0.0f)));
    }
    const float p1 = __shfl_xor(part, 16);
    const float s01 = (q & 1) ? p1 + part : part + p1;
    const float p2 = __shfl_xor(s01, 32);
    const float tot = (q & 2) ? p2 + s01 : s01 + p2;
    int fv = 0;
    if (valid && q == 0) {
      const float zz = tot + b3;
      if (z) z[e] = zz;
      // cells next to a domain boundary are flagged by rule, as by the
      // sensor (a boundary neighbour flags the cell, sensor_kernels.cu:19-163)
      fv = (at_wall(jmp) || zz > threshold || (guard && guard[e])) ? 1 : 0;
      flags[e] = fv;
    }
    if (list) {
      // append the group's flagged elements to list[0 .. *count): one atomic per group with flags
      const unsigned long long bal = __ballot(fv != 0);
      if (bal) {
        int base = 0;
        if (lane == 0) base = atomicAdd(count, (int)__popcll(bal));
        base = __shfl(base, 0);
        if (fv) list[base + (int)__popcll(bal & ((1ull << lane) - 1ull))] = e;
      }
    }
  }
}

// -------------------------------------------------------------- backward
// 32 samples per block: dz and the layer-2 deltas per sample, then layer-1
// deltas per (sample, unit), then one lane per parameter summing its
// gradient over the block's samples in sample order.  Partial sums per block
// are reduced by k_mlp_reduce in block order.
constexpr int kSB = 32;

__global__ __launch_bounds__(256) void k_mlp_bwd(const uint16_t* __restrict__ w2b, const uint16_t* __restrict__ w3b,
                                                 const float* __restrict__ X, const int* __restrict__ idx,
                                                 const float* __restrict__ y, int N, float pos_weight, float inv_n,
                                                 const float* __restrict__ z, const uint16_t* __restrict__ h1s,
                                                 const uint16_t* __restrict__ h2s, float* __restrict__ partial) {
  __shared__ float sX[kSB][kInP], sH1[kSB][kH], sH2[kSB][kH], sD1[kSB][kH], sD2[kSB][kH];
  __shared__ float sDz[kSB], sLoss[kSB];
  const int tid = threadIdx.x;
  const int n0 = blockIdx.x * kSB;
  for (int k = tid; k < kSB * kH; k += 256) {
    const int s = k / kH, c = k % kH, n = n0 + s;
    const bool v = n < N;
    const int row = v ? (idx ? idx[n] : n) : 0;
    sX[s][c] = v ? bf2f(f2bf(X[(size_t)row * kInP + c])) : 0.0f;
    sH1[s][c] = v ? bf2f(h1s[(size_t)n * kH + c]) : 0.0f;
    sH2[s][c] = v ? bf2f(h2s[(size_t)n * kH + c]) : 0.0f;
  }
  __syncthreads();
  if (tid < kSB) {
    const int n = n0 + tid;
    float dz = 0.0f, loss = 0.0f;
    if (n < N) {
      const float zz = z[n], yy = y[idx ? idx[n] : n];
      const float sg = 1.0f / (1.0f + expf(-zz));
      dz = (pos_weight * yy * (sg - 1.0f) + (1.0f - yy) * sg) * inv_n;
      // softplus(t) = log(1 + e^t), stable
      const float sp_neg = fmaxf(-zz, 0.0f) + log1pf(expf(-fabsf(zz)));
      const float sp_pos = fmaxf(zz, 0.0f) + log1pf(expf(-fabsf(zz)));
      loss = (pos_weight * yy * sp_neg + (1.0f - yy) * sp_pos) * inv_n;
    }
    sDz[tid] = dz;
    sLoss[tid] = loss;
  }
  __syncthreads();
  for (int k = tid; k < kSB * kH; k += 256) {
    const int s = k / kH, u = k % kH;
    sD2[s][u] = sH2[s][u] > 0.0f ? sDz[s] * bf2f(w3b[u]) : 0.0f;
  }
  __syncthreads();
  for (int k = tid; k < kSB * kH; k += 256) {
    const int s = k / kH, v = k % kH;
    float acc = 0.0f;
    for (int u = 0; u < kH; ++u) acc += bf2f(w2b[u * kH + v]) * sD2[s][u];
    sD1[s][v] = sH1[s][v] > 0.0f ? acc : 0.0f;
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * (kNParam + 1);
  for (int p = tid; p < kNParam + 1; p += 256) {
    float g = 0.0f;
    if (p < kB1) {  // W1[v][k]
      const int v = p / kInP, k = p % kInP;
      for (int s = 0; s < kSB; ++s) g += sD1[s][v] * sX[s][k];
    } else if (p < kW2) {
      const int v = p - kB1;
      for (int s = 0; s < kSB; ++s) g += sD1[s][v];
    } else if (p < kB2) {  // W2[u][v]
      const int u = (p - kW2) / kH, v = (p - kW2) % kH;
      for (int s = 0; s < kSB; ++s) g += sD2[s][u] * sH1[s][v];
    } else if (p < kW3) {
      const int u = p - kB2;
      for (int s = 0; s < kSB; ++s) g += sD2[s][u];
    } else if (p < kB3) {
      const int u = p - kW3;
      for (int s = 0; s < kSB; ++s) g += sDz[s] * sH2[s][u];
    } else if (p == kB3) {
      for (int s = 0; s < kSB; ++s) g += sDz[s];
    } else {
      for (int s = 0; s < kSB; ++s) g += sLoss[s];
    }
    out[p] = g;
  }
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, int nblocks,
                                                    float* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNParam + 1) return;
  float g = 0.0f;
  for (int b = 0; b < nblocks; ++b) g += partial[(size_t)b * (kNParam + 1) + p];
  grad[p] = g;
}

struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                                                  const float* __restrict__ grad, AdamArgs a, uint16_t* __restrict__ w1b,
                                                  uint16_t* __restrict__ w2b, uint16_t* __restrict__ w3b,
                                                  float* __restrict__ loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0 && loss_acc) *loss_acc += grad[kNParam];
  if (p >= kNParam) return;
  const float g = grad[p];
  const float mm = a.b1 * m[p] + (1.0f - a.b1) * g;
  const float vv = a.b2 * v[p] + (1.0f - a.b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float w = P[p] - a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
  P[p] = w;
  if (p < kB1) w1b[p - kW1] = f2bf(w);
  else if (p >= kW2 && p < kB2) w2b[p - kW2] = f2bf(w);
  else if (p >= kW3 && p < kB3) w3b[p - kW3] = f2bf(w);
}

__global__ void k_mlp_cast(const float* __restrict__ P, uint16_t* __restrict__ w1b, uint16_t* __restrict__ w2b,
                           uint16_t* __restrict__ w3b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kB1) w1b[p] = f2bf(P[kW1 + p]);
  if (p < kH * kH) w2b[p] = f2bf(P[kW2 + p]);
  if (p < kH) w3b[p] = f2bf(P[kW3 + p]);
}

__global__ void k_mlp_flags(const float* __restrict__ z, int N, float threshold, int* __restrict__ flags) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) flags[n] = z[n] > threshold ? 1 : 0;
}

__global__ void k_fill_f32(float* p, int n, float v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

static inline unsigned nblocks_for(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace dgx

using namespace dgx;

struct dgx_mlp {
  float* P = nullptr;  // FP32 master parameters
  float *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr;
  // training scratch, grown on demand
  float* z = nullptr;
  uint16_t *h1 = nullptr, *h2 = nullptr;
  float* partial = nullptr;
  int* idx = nullptr;
  float* loss = nullptr;
  long cap = 0, idx_cap = 0;
  int step = 0;
  std::vector<void*> allocs;
};

namespace {

bool mlp_alloc(dgx_mlp* m, void** p, size_t bytes) {
  if (!hip_ok(hipMalloc(p, bytes ? bytes : 4), "hipMalloc")) return false;
  m->allocs.push_back(*p);
  return true;
}

void mlp_release(dgx_mlp* m) {
  for (void* p : m->allocs) (void)hipFree(p);
  m->allocs.clear();
}

dgx_status mfail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

// scratch for n samples of activations (forward with store) and backward partials
bool ensure_scratch(dgx_mlp* m, long n) {
  if (n <= m->cap) return true;
  const long c = std::max(n, 2 * m->cap);
  void *z, *h1, *h2, *part;
  if (!mlp_alloc(m, &z, sizeof(float) * c) || !mlp_alloc(m, &h1, 2 * kH * c) || !mlp_alloc(m, &h2, 2 * kH * c) ||
      !mlp_alloc(m, &part, sizeof(float) * (kNParam + 1) * ((c + kSB - 1) / kSB)))
    return false;
  m->z = (float*)z;
  m->h1 = (uint16_t*)h1;
  m->h2 = (uint16_t*)h2;
  m->partial = (float*)part;
  m->cap = c;
  return true;
}

unsigned fwd_grid(long n) {
  const long waves = (n + 15) / 16;
  return (unsigned)std::max<long>(1, std::min<long>((waves + 3) / 4, 2048));
}

hipError_t forward(dgx_mlp* m, const float* X, const int* idx, long n, float* z, uint16_t* h1, uint16_t* h2,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_fwd, dim3(fwd_grid(n)), dim3(256), 0, st, m->w1b, m->w2b, m->w3b, m->P, X, idx, (int)n, z,
                     h1, h2);
  return hipGetLastError();
}

// mean gradient of the weighted BCE over samples idx[0..n) (or rows 0..n)
hipError_t gradient(dgx_mlp* m, const float* X, const int* idx, const float* y, long n, float pos_weight,
                    hipStream_t st) {
  // the stored activations and the per-block partials are sized for m->cap samples
  if (n <= 0 || n > m->cap || !m->z || !m->h1 || !m->h2 || !m->partial) return hipErrorInvalidValue;
  hipError_t e = forward(m, X, idx, n, m->z, m->h1, m->h2, st);
  if (e != hipSuccess) return e;
  const unsigned nb = nblocks_for(n, kSB);
  hipLaunchKernelGGL(k_mlp_bwd, dim3(nb), dim3(256), 0, st, m->w2b, m->w3b, X, idx, y, (int)n, pos_weight,
                     1.0f / (float)n, m->z, m->h1, m->h2, m->partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(nblocks_for(kNParam + 1, 256)), dim3(256), 0, st, m->partial, (int)nb,
                     m->grad);
  return hipGetLastError();
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

namespace dgx {

// Indicator state of a solver: features, logits and the flags that gate the
// sensor-driven HR sweeps (limiter 8).
hipError_t mlp_indicator_flags(Solver& s, const double* U, hipStream_t st, bool with_list) {
  dgx_mlp* m = s.mlp;
  const int N = s.N_E;
  int* count = with_list ? s.flag + 2 : nullptr;
  const bool guard = s.shock_guard_on();
  // else the stage kernel that wrote U left the cell means (and the sensor's
  // averages and pressures) in s.mlp_avg, s.Uavg and s.mlp_psens
  if (s.means_of != U)
    hipLaunchKernelGGL(k_mlp_means, dim3(nblocks_for(N, 64)), dim3(256), 0, st, U, s.t.mlp_vinv, N, s.mlp_avg,
                       guard ? s.Uavg : nullptr, s.mlp_psens);
  hipLaunchKernelGGL(k_mlp_jumps, dim3(nblocks_for(N, 256)), dim3(256), 0, st, s.mlp_avg, s.t.neighbors, N,
                     reinterpret_cast<float4*>(s.mlp_jumps), count, guard ? s.Uavg : nullptr, s.mlp_psens,
                     s.thresh[1], s.mlp_guard);
  const unsigned nb = std::max(1u, std::min(nblocks_for(N, 64), (unsigned)DGX_IND_BLOCKS));
  hipLaunchKernelGGL(k_mlp_indicator, dim3(nb), dim3(256), 0, st, U, s.t.mlp_vinv,
                     reinterpret_cast<const float4*>(s.mlp_jumps), N, m->w1b, m->w2b, m->w3b, m->P, s.mlp_threshold,
                     guard ? s.mlp_guard : nullptr, s.mlp_z, s.sensors, with_list ? s.flag_list : nullptr, count);
  return hipGetLastError();
}

hipError_t mlp_features(Solver& s, const double* U, float* X, hipStream_t st) {
  const int N = s.N_E;
  hipLaunchKernelGGL(k_mlp_means, dim3(nblocks_for(N, 64)), dim3(256), 0, st, U, s.t.mlp_vinv, N, s.mlp_avg, nullptr,
                     nullptr);
  hipLaunchKernelGGL(k_mlp_jumps, dim3(nblocks_for(N, 256)), dim3(256), 0, st, s.mlp_avg, s.t.neighbors, N,
                     reinterpret_cast<float4*>(s.mlp_jumps), nullptr, nullptr, nullptr, 0.0, nullptr);
  const unsigned nbf = std::max(1u, std::min(nblocks_for(N, 64), 4096u));
  hipLaunchKernelGGL(k_mlp_features, dim3(nbf), dim3(256), 0, st, U, s.t.mlp_vinv,
                     reinterpret_cast<const float4*>(s.mlp_jumps), N, X);
  return hipGetLastError();
}

}  // namespace dgx

extern "C" {

int dgx_mlp_param_count(void) { return kNParam; }

dgx_status dgx_mlp_create(unsigned long long seed, dgx_mlp** out) {
  if (!out) return mfail(DGX_EINVAL, "dgx_mlp_create: null argument");
  dgx_mlp* m = new dgx_mlp;
  void *P, *mm, *vv, *g, *w1, *w2, *w3, *loss;
  bool ok = mlp_alloc(m, &P, sizeof(float) * kNParam) && mlp_alloc(m, &mm, sizeof(float) * kNParam) &&
            mlp_alloc(m, &vv, sizeof(float) * kNParam) && mlp_alloc(m, &g, sizeof(float) * (kNParam + 1)) &&
            mlp_alloc(m, &w1, 2 * kH * kInP) && mlp_alloc(m, &w2, 2 * kH * kH) && mlp_alloc(m, &w3, 2 * kH) &&
            mlp_alloc(m, &loss, sizeof(float) * 4);
  if (!ok) {
    mlp_release(m);
    delete m;
    return DGX_EDEVICE;
  }
  m->P = (float*)P; m->m = (float*)mm; m->v = (float*)vv; m->grad = (float*)g;
  m->w1b = (uint16_t*)w1; m->w2b = (uint16_t*)w2; m->w3b = (uint16_t*)w3; m->loss = (float*)loss;
  // He-uniform initialisation from a seeded splitmix64 stream; padded inputs stay zero
  std::vector<float> h(kNParam, 0.0f);
  uint64_t s = seed;
  auto uni = [&]() { return (float)((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2.0f - 1.0f; };
  const float a1 = std::sqrt(6.0f / kIn), a2 = std::sqrt(6.0f / kH);
  for (int u = 0; u < kH; ++u)
    for (int k = 0; k < kIn; ++k) h[kW1 + u * kInP + k] = a1 * uni();
  for (int u = 0; u < kH * kH; ++u) h[kW2 + u] = a2 * uni();
  for (int u = 0; u < kH; ++u) h[kW3 + u] = a2 * uni();
  *out = m;
  const dgx_status r = dgx_mlp_set_params(m, h.data());
  if (r != DGX_OK) {
    dgx_mlp_free(m);
    *out = nullptr;
  }
  return r;
}

void dgx_mlp_free(dgx_mlp* m) {
  if (!m) return;
  mlp_release(m);
  delete m;
}

dgx_status dgx_mlp_set_params(dgx_mlp* m, const float* params) {
  if (!m || !params) return mfail(DGX_EINVAL, "dgx_mlp_set_params: null argument");
  if (!hip_ok(hipMemcpy(m->P, params, sizeof(float) * kNParam, hipMemcpyHostToDevice), "params upload") ||
      !hip_ok(hipMemset(m->m, 0, sizeof(float) * kNParam), "hipMemset") ||
      !hip_ok(hipMemset(m->v, 0, sizeof(float) * kNParam), "hipMemset"))
    return DGX_EDEVICE;
  m->step = 0;
  hipLaunchKernelGGL(k_mlp_cast, dim3(nblocks_for(kH * kInP, 256)), dim3(256), 0, 0, m->P, m->w1b, m->w2b, m->w3b);
  if (!hip_ok(hipGetLastError(), "cast") || !hip_ok(hipDeviceSynchronize(), "cast")) return DGX_EDEVICE;
  return DGX_OK;
}

dgx_status dgx_mlp_get_params(const dgx_mlp* m, float* params) {
  if (!m || !params) return mfail(DGX_EINVAL, "dgx_mlp_get_params: null argument");
  return hip_ok(hipMemcpy(params, m->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost), "params download")
             ? DGX_OK
             : DGX_EDEVICE;
}

dgx_status dgx_mlp_forward(dgx_mlp* m, const float* X, long n, float* z, void* stream) {
  if (!m || !X || !z || n < 0) return mfail(DGX_EINVAL, "dgx_mlp_forward: bad arguments");
  if (n == 0) return DGX_OK;
  return hip_ok(forward(m, X, nullptr, n, z, nullptr, nullptr, reinterpret_cast<hipStream_t>(stream)), "mlp forward")
             ? DGX_OK
             : DGX_EDEVICE;
}

dgx_status dgx_mlp_gradient(dgx_mlp* m, const float* X, const float* y, long n, float pos_weight, float* grad,
                            void* stream) {
  if (!m || !X || !y || !grad || n <= 0) return mfail(DGX_EINVAL, "dgx_mlp_gradient: bad arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!ensure_scratch(m, n)) return DGX_EDEVICE;
  if (!hip_ok(gradient(m, X, nullptr, y, n, pos_weight, st), "mlp gradient")) return DGX_EDEVICE;
  return hip_ok(hipMemcpyAsync(grad, m->grad, sizeof(float) * (kNParam + 1), hipMemcpyDeviceToDevice, st),
                "gradient copy")
             ? DGX_OK
             : DGX_EDEVICE;
}

dgx_status dgx_mlp_train(dgx_mlp* m, const float* X, const float* y, long n, int epochs, int batch, float lr,
                         float pos_weight, unsigned long long seed, float* epoch_loss, void* stream) {
  if (!m || !X || !y || n <= 0 || epochs < 0 || batch < kSB || !epoch_loss)
    return mfail(DGX_EINVAL, "dgx_mlp_train: bad arguments");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  if (!ensure_scratch(m, batch)) return DGX_EDEVICE;
  if (n > m->idx_cap) {
    void* p;
    if (!mlp_alloc(m, &p, sizeof(int) * n)) return DGX_EDEVICE;
    m->idx = (int*)p;
    m->idx_cap = n;
  }
  std::vector<int> perm(n);
  uint64_t s = seed;
  AdamArgs a{lr, 0.9f, 0.999f, 1e-8f, 0.f, 0.f};
  for (int ep = 0; ep < epochs; ++ep) {
    for (long k = 0; k < n; ++k) perm[k] = (int)k;
    for (long k = n - 1; k > 0; --k) std::swap(perm[k], perm[splitmix(s) % (uint64_t)(k + 1)]);
    if (!hip_ok(hipMemcpyAsync(m->idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, st), "idx upload"))
      return DGX_EDEVICE;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st, m->loss, 1, 0.0f);
    for (long b0 = 0; b0 < n; b0 += batch) {
      const long nb = std::min<long>(batch, n - b0);
      if (nb > m->cap || b0 + nb > m->idx_cap)  // the minibatch's rows of idx and its activations exist
        return mfail(DGX_EINVAL, "dgx_mlp_train: minibatch beyond the training scratch");
      if (!hip_ok(gradient(m, X, m->idx + b0, y, nb, pos_weight, st), "mlp gradient")) return DGX_EDEVICE;
      m->step += 1;
      a.c1 = 1.0f - std::pow(a.b1, (float)m->step);
      a.c2 = 1.0f - std::pow(a.b2, (float)m->step);
      hipLaunchKernelGGL(k_mlp_adam, dim3(nblocks_for(kNParam, 256)), dim3(256), 0, st, m->P, m->m, m->v, m->grad, a,
                         m->w1b, m->w2b, m->w3b, m->loss);
    }
    float l = 0.0f;
    if (!hip_ok(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, st), "loss readback") ||
        !hip_ok(hipStreamSynchronize(st), "loss readback"))
      return DGX_EDEVICE;
    epoch_loss[ep] = l * (float)batch / (float)n;  // mean of the per-batch means, weighted by batch size
  }
  return DGX_OK;
}

}  // extern "C"
