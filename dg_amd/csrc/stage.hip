// Fused RK stage for Cartesian meshes: one kernel evaluates the whole right-
// hand side of a tile of elements and applies Dt*Minv and the RK4 update.
//
// A 256-thread block owns a TX x TY tile (TX in 1D).  It stages the tile's
// state and the face traces of its halo in LDS, evaluates every face that
// touches the tile (faces on tile edges are evaluated by both neighbouring
// tiles with identical inputs and code), integrates them against psi_w,
// then runs the element phases of volume.cuh and writes the stage update.
// Face sides, closures, normals and boundary treatment follow the reference
// interface construction (simpleMesh.cc:770-989: side 0 = first element in
// element order, its outward normal; boundary faces contribute to side 0
// only, dg_functions.cc:225-233), so the result equals the general
// face/element kernels.  Tiles are mapped to XCDs in contiguous bands so
// that halo rows are shared through each XCD's L2.
#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "dgx_internal.h"
#include "physics.cuh"
#include "sensor.cuh"
#include "volume.cuh"

namespace dgx {

struct StageArgs {
  const double* __restrict__ phi;
  const double* __restrict__ dphi;
  const double* __restrict__ phi_w;
  const double* __restrict__ dphi_w;
  const double* __restrict__ psi;
  const double* __restrict__ psi_w;
  const double* __restrict__ Minv;    // [e*minv_stride] N_s x N_s
  const double* __restrict__ J;       // [e*geo_stride]
  const double* __restrict__ invJac;  // [(e*geo_stride)*D*D + alpha*D + a]
  const double* __restrict__ gx;      // faces normal to x: [(ey*(nx+1) + fx)*4 + {n_x, n_y, JF0, JF1}]
  const double* __restrict__ gy;      // faces normal to y: [(fy*nx + ex)*4 + {n_x, n_y, JF0, JF1}]
  double J0;                          // Jacobian of the element whose Minv is stored when minv_stride == 0
  int geo_stride, minv_stride;
  int nx, ny;
  int bc[4];  // left, right, bottom, top
  int tiles_x, tiles_y, xcd_bands;
  int stage;
  const double* Uin;
  const double* Un;
  double* Uacc;
  double* Uout;
  const double* __restrict__ dt;
  double* fout;
  // 2D p=2 two-fluid RK stages: cell means of rho, p and alpha of the output
  // tile for the learned indicator ([e*3 + k], cell_means_rpa with w0), or null
  double* means;
  const double* __restrict__ w0;
  double* uavg;   // with means: also the sensor's cell averages [e*5 + f] (sensor_nodal_mean), or null
  double* psens;  // with uavg: their sensor pressures [e]
  const int* gate;  // set and nonzero: the launch does nothing (a step skipped by the device-side time control)
};

template <int D, int TX, int TY>
struct Tile {
  static constexpr int NT = TX * TY;
  static constexpr int NFX = (D == 1) ? TX + 1 : (TX + 1) * TY;  // faces normal to x
  static constexpr int NFY = (D == 1) ? 0 : TX * (TY + 1);       // faces normal to y
  static constexpr int NFACE = NFX + NFY;
  static constexpr int NH = (D == 1) ? 2 : 2 * TY + 2 * TX;      // halo trace slots: L, R, B, T
};

// Where each side of a tile face reads its trace and where it contributes.
struct FaceDesc {
  int src[2];     // tile slot (>= 0) or halo slot (-1 - h)
  int cl[2];      // closure id used by the side (tile sources)
  int dst[2];     // tile slot receiving this side's contribution, or -1
  int edge[2];    // local edge of the receiving element
  int tag;        // 0 interior, else boundary physical tag
  int geo;        // index of the face in gx (x-faces) or gy (y-faces)
  bool xf;        // face normal to x
};

template <int D, int TX, int TY>
__device__ inline FaceDesc face_desc(int fi, int ex0, int ey0, const StageArgs& a) {
  FaceDesc d;
  d.dst[0] = d.dst[1] = -1;
  d.edge[0] = d.edge[1] = 0;
  if constexpr (D == 1) {
    const int exl = ex0 + fi - 1, exr = ex0 + fi;
    d.xf = true;
    d.geo = ex0 + fi;
    if (exl < 0) {  // left domain boundary: side 0 = element 0, closure 0
      d.tag = a.bc[0];
      d.src[0] = 0; d.cl[0] = 0;
      if (d.tag == BC_PERIODIC) { d.src[1] = -1 - 0; d.cl[1] = 1; } else { d.src[1] = 0; d.cl[1] = 0; }
      d.dst[0] = 0; d.edge[0] = 0;
    } else if (exr >= a.nx) {  // right domain boundary: side 0 = last element, closure 1
      d.tag = a.bc[1];
      d.src[0] = TX - 1; d.cl[0] = 1;
      if (d.tag == BC_PERIODIC) { d.src[1] = -1 - 1; d.cl[1] = 0; } else { d.src[1] = TX - 1; d.cl[1] = 1; }
      d.dst[0] = TX - 1; d.edge[0] = 1;
    } else {
      d.tag = 0;
      d.src[0] = (fi == 0) ? -1 - 0 : fi - 1; d.cl[0] = 1;
      d.src[1] = (fi == TX) ? -1 - 1 : fi; d.cl[1] = 0;
      if (fi > 0) { d.dst[0] = fi - 1; d.edge[0] = 1; }
      if (fi < TX) { d.dst[1] = fi; d.edge[1] = 0; }
    }
  } else {
    if (fi < Tile<D, TX, TY>::NFX) {
      const int ty = fi / (TX + 1), fx = fi % (TX + 1);
      const int exl = ex0 + fx - 1, exr = ex0 + fx;
      d.xf = true;
      d.geo = (ey0 + ty) * (a.nx + 1) + ex0 + fx;
      if (exl < 0) {
        d.tag = a.bc[0];
        d.src[0] = ty * TX; d.cl[0] = 3;
        if (d.tag == BC_PERIODIC) { d.src[1] = -1 - ty; d.cl[1] = 5; } else { d.src[1] = ty * TX; d.cl[1] = 3; }
        d.dst[0] = ty * TX; d.edge[0] = 3;
      } else if (exr >= a.nx) {
        d.tag = a.bc[1];
        d.src[0] = ty * TX + TX - 1; d.cl[0] = 1;
        if (d.tag == BC_PERIODIC) { d.src[1] = -1 - (TY + ty); d.cl[1] = 7; } else { d.src[1] = ty * TX + TX - 1; d.cl[1] = 1; }
        d.dst[0] = ty * TX + TX - 1; d.edge[0] = 1;
      } else {
        d.tag = 0;
        d.src[0] = (fx == 0) ? -1 - ty : ty * TX + fx - 1; d.cl[0] = 1;
        d.src[1] = (fx == TX) ? -1 - (TY + ty) : ty * TX + fx; d.cl[1] = 7;
        if (fx > 0) { d.dst[0] = ty * TX + fx - 1; d.edge[0] = 1; }
        if (fx < TX) { d.dst[1] = ty * TX + fx; d.edge[1] = 3; }
      }
    } else {
      const int fj = fi - Tile<D, TX, TY>::NFX;
      const int fy = fj / TX, tx = fj % TX;
      const int eyl = ey0 + fy - 1, eyu = ey0 + fy;
      d.xf = false;
      d.geo = (ey0 + fy) * a.nx + ex0 + tx;
      if (eyl < 0) {
        d.tag = a.bc[2];
        d.src[0] = tx; d.cl[0] = 0;
        if (d.tag == BC_PERIODIC) { d.src[1] = -1 - (2 * TY + tx); d.cl[1] = 6; } else { d.src[1] = tx; d.cl[1] = 0; }
        d.dst[0] = tx; d.edge[0] = 0;
      } else if (eyu >= a.ny) {
        d.tag = a.bc[3];
        d.src[0] = (TY - 1) * TX + tx; d.cl[0] = 2;
        if (d.tag == BC_PERIODIC) { d.src[1] = -1 - (2 * TY + TX + tx); d.cl[1] = 4; }
        else { d.src[1] = (TY - 1) * TX + tx; d.cl[1] = 2; }
        d.dst[0] = (TY - 1) * TX + tx; d.edge[0] = 2;
      } else {
        d.tag = 0;
        d.src[0] = (fy == 0) ? -1 - (2 * TY + tx) : (fy - 1) * TX + tx; d.cl[0] = 2;
        d.src[1] = (fy == TY) ? -1 - (2 * TY + TX + tx) : fy * TX + tx; d.cl[1] = 4;
        if (fy > 0) { d.dst[0] = (fy - 1) * TX + tx; d.edge[0] = 2; }
        if (fy < TY) { d.dst[1] = fy * TX + tx; d.edge[1] = 0; }
      }
    }
  }
  return d;
}

// Global element and closure of halo slot h, or -1 when the slot is unused.
template <int D, int TX, int TY>
__device__ inline int halo_source(int h, int ex0, int ey0, const StageArgs& a, int& cl) {
  if constexpr (D == 1) {
    if (h == 0) {
      cl = 1;
      if (ex0 > 0) return ex0 - 1;
      return a.bc[0] == BC_PERIODIC ? a.nx - 1 : -1;
    }
    cl = 0;
    if (ex0 + TX < a.nx) return ex0 + TX;
    return a.bc[1] == BC_PERIODIC ? 0 : -1;
  } else {
    const int nx = a.nx, ny = a.ny;
    if (h < TY) {  // left of row h
      const int row = ey0 + h;
      if (ex0 > 0) { cl = 1; return row * nx + ex0 - 1; }
      cl = 5;
      return a.bc[0] == BC_PERIODIC ? row * nx + nx - 1 : -1;
    }
    if (h < 2 * TY) {  // right of row h - TY
      const int row = ey0 + h - TY;
      cl = 7;
      if (ex0 + TX < nx) return row * nx + ex0 + TX;
      return a.bc[1] == BC_PERIODIC ? row * nx : -1;
    }
    if (h < 2 * TY + TX) {  // below column
      const int col = ex0 + h - 2 * TY;
      if (ey0 > 0) { cl = 2; return (ey0 - 1) * nx + col; }
      cl = 6;
      return a.bc[2] == BC_PERIODIC ? (ny - 1) * nx + col : -1;
    }
    const int col = ex0 + h - 2 * TY - TX;  // above column
    cl = 4;
    if (ey0 + TY < ny) return (ey0 + TY) * nx + col;
    return a.bc[3] == BC_PERIODIC ? col : -1;
  }
}

#ifndef DGX_STAGE_PAD
#define DGX_STAGE_PAD 1  // pad per-element LDS strides against bank conflicts
#endif
#ifndef DGX_STAGE_MFMA64
#define DGX_STAGE_MFMA64 1  // 2D p=2 FP64 element phases on the matrix cores (0: VALU batches)
#endif
#ifndef DGX_STAGE_PERSIST
#define DGX_STAGE_PERSIST 0  // 1: persistent blocks, next tile loaded during the current one
#endif

// Per-element LDS stride for n doubles: with padding, the smallest stride
// >= n whose element starts fall on distinct, non-overlapping 16-byte bank
// groups for the four elements a wave reads at once (2*stride mod 32 dwords
// in [4, 28]).
constexpr bool bank_ok(int n) { return (2 * n) % 32 >= 4 && (2 * n) % 32 <= 28; }
constexpr int lds_stride(int n) { return !DGX_STAGE_PAD || bank_ok(n) ? n : lds_stride(n + 1); }

// Tile origin of the i-th tile a block processes.  Blocks are dispatched
// round-robin over the 8 XCDs; with XCD bands the tiles of each XCD form a
// contiguous band of the grid so that halo rows are shared in its L2.
__device__ inline int tile_of(const StageArgs& a, int i) {
  const int ntiles = a.tiles_x * a.tiles_y;
  if (a.xcd_bands) {
    const int per = ntiles >> 3, nbx = gridDim.x >> 3;
    const int x = blockIdx.x & 7, j = (blockIdx.x >> 3) + i * nbx;
    return j < per ? x * per + j : -1;
  }
  const int t = blockIdx.x + i * gridDim.x;
  return t < ntiles ? t : -1;
}

// VEC: the tile state, the RK operands and the stage output move as 16-byte
// pairs (tile rows are contiguous in HBM and in LDS; every state pointer is
// 16-byte aligned, which the host checks before it picks this instance).
template <int D, int P, int PH, int FL, int MODE, int TX, int TY, int PREC, bool VEC = false>
__global__ __launch_bounds__(256) void k_stage(StageArgs a) {
  if (a.gate && *a.gate) return;
  using V = Vol<D, P, PH>;
  using T = Tile<D, TX, TY>;
  constexpr int NF = V::NF, Ns = V::Ns, NG = V::NG, Ms = V::Ms, MG = Shape<D, P>::MG;
  constexpr int NU = V::NU, NQE = V::NQE, NQ = V::NQ, TPE = V::TPE, EPB = V::EPB;
  constexpr int NT = T::NT, NFACE = T::NFACE, NH = T::NH, HT = NF * Ms;
  constexpr int SU = lds_stride(NU), SQ = lds_stride(NQE), SW = lds_stride(NG * NQ);
  // element phases on the matrix cores (FP64, or FP32 in the mixed-precision mode) or as VALU batches
  // PREC 0: FP64; 1: FP32 contractions; 2: split-BF16 contractions with FP32 accumulation
  constexpr bool MFW = D == 2 && P == 2 && NT % 16 == 0 && PH <= 1 && (PREC >= 1 || DGX_STAGE_MFMA64);
  using TW = typename std::conditional<PREC >= 1, float, double>::type;
  constexpr int NNC = V::NNC;    // non-conservative fields: the last NNC
  constexpr int NWP = NF + NNC;  // per face point: F, and the non-conservative terms
  constexpr int NW_FACE = NFACE * MG * NWP, NW_VOL = MFW ? 0 : EPB * SW;
  constexpr int NW_MF = MFW ? 4 * vol_wave_lds<TW, PH>() : 0;
  constexpr int NW0 = NW_FACE > NW_VOL ? NW_FACE : NW_VOL;
  constexpr int NW = NW0 > NW_MF ? NW0 : NW_MF;
  constexpr int NGEO = 1 + D * D;
  static_assert(PREC == 0 || MFW, "mixed precision: 2D p=2 tiles of a multiple of 16 elements");
  __shared__ int sNodeQ[18];
  // per-thread counts of the loaded quantities (state entries in groups of VW)
  static_assert(!VEC || (D == 2 && SU == NU && (TX * NU) % 2 == 0), "paired state rows");
  constexpr int VW = VEC ? 2 : 1;
  constexpr int R_UV = (NT * NU / VW + 255) / 256, R_U = R_UV * VW;
  constexpr int R_H = (NH * HT + 255) / 256;
  constexpr int R_G = (NT * NGEO + 255) / 256, R_F = (NFACE * 4 + 255) / 256;

  __shared__ __align__(16) double sU[NT * SU];
  __shared__ double sH[NH * HT];
  __shared__ double sQ[NT * SQ];
  __shared__ double sW[NW];  // face-point fluxes, then per-point volume values
  __shared__ double sPhi[V::NTAB_PHI];
  __shared__ double sDphi[PH ? V::NTAB_D : 1];
  __shared__ double sDw[MFW ? 1 : V::NTAB_D];  // the matrix-core path keeps its operand in registers
  __shared__ double sPw[PH ? V::NTAB_PHI : 1];
  __shared__ double sMinv[Ns * Ns];
  __shared__ double sPsi[MG * Ms], sPsiw[MG * Ms];
  __shared__ double sGeo[NT * NGEO];  // per element: J, invJac
  __shared__ double sDte[NT];         // per element: Dt (J_0/J_e with the shared Minv)
  __shared__ double sFG[NFACE * 4];   // per tile face: n_x, n_y, JF side 0, JF side 1
  __shared__ int sFD[NFACE];          // per tile face: receiving slot, edge and closure of each side (packed)
  __shared__ int sFS[NFACE];          // per tile face: source slot of each side (+128, 8 bits each), boundary tag << 16
  static_assert(NT <= 128 && NH <= 128, "face sources fit 8 bits");
  const VolTables<D, P, PH> tb{sPhi, sDphi, sDw, sPw, sMinv};

  const int tid = threadIdx.x;
  const bool shared_minv = a.minv_stride == 0;
  // reference tables and the first tile's inputs are fetched together (one wait)
  static_assert(MG * Ms <= 256, "face tables fit one pass");
  VolTableFetch<D, P, PH, !MFW> tf;
  tf.fetch(a.phi, a.dphi, a.phi_w, a.dphi_w, a.Minv, MODE == 1 && shared_minv, tid);
  double psi_r = 0.0, psiw_r = 0.0;
  if (tid < MG * Ms) { psi_r = a.psi[tid]; psiw_r = a.psi_w[tid]; }
  if constexpr (MFW) node_q_table<NF>(sNodeQ, tid);
  const double dt = (MODE == 1) ? a.dt[0] : 0.0;

  // registers holding the next tile's inputs while the current one computes
  double pu[R_U], ph[R_H], pg[R_G], pf[R_F];
  auto issue = [&](int tile) {
    const int ex0 = (tile % a.tiles_x) * TX, ey0 = (tile / a.tiles_x) * TY;
#pragma unroll
    for (int r = 0; r < R_UV; ++r) {
      const int k = VW * (tid + r * 256);
      if (k < NT * NU) {
        const int ty = k / (TX * NU);
        const double* src = a.Uin + ((size_t)(ey0 + ty) * a.nx + ex0) * NU + (k - ty * TX * NU);
        if constexpr (VEC) {
          const double2 v = *reinterpret_cast<const double2*>(src);
          pu[2 * r] = v.x;
          pu[2 * r + 1] = v.y;
        } else {
          pu[r] = *src;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R_H; ++r) {
      const int k = tid + r * 256;
      if (k < NH * HT) {
        const int h = k / HT, q = k % HT, f = q / Ms, j = q % Ms;
        int cl = 0;
        const int e = halo_source<D, TX, TY>(h, ex0, ey0, a, cl);
        ph[r] = (e >= 0) ? a.Uin[((size_t)e * NF + f) * Ns + closure_node<D, P>(cl, j)] : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < R_G; ++r) {
      const int k = tid + r * 256;
      if (k < NT * NGEO) {
        const int el = k / NGEO, q = k % NGEO;
        const int eg = ((ey0 + el / TX) * a.nx + ex0 + el % TX) * a.geo_stride;
        pg[r] = q == 0 ? a.J[eg] : a.invJac[eg * D * D + q - 1];
      }
    }
#pragma unroll
    for (int r = 0; r < R_F; ++r) {
      const int k = tid + r * 256;
      if (k < NFACE * 4) {
        const int fi = k / 4, q = k % 4;
        const double* src;
        if (D == 1 || fi < T::NFX) {
          const int ty = (D == 1) ? 0 : fi / (TX + 1), fx = (D == 1) ? fi : fi % (TX + 1);
          src = a.gx + ((size_t)(ey0 + ty) * (a.nx + 1) + ex0 + fx) * 4;
        } else {
          const int fj = fi - T::NFX, fy = fj / TX, tx = fj % TX;
          src = a.gy + ((size_t)(ey0 + fy) * a.nx + ex0 + tx) * 4;
        }
        pf[r] = src[q];
      }
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int r = 0; r < R_UV; ++r) {
      const int k = VW * (tid + r * 256);
      if (k < NT * NU) {
        if constexpr (VEC)
          *reinterpret_cast<double2*>(sU + k) = make_double2(pu[2 * r], pu[2 * r + 1]);  // SU == NU
        else
          sU[(k / NU) * SU + k % NU] = pu[r];
      }
    }
#pragma unroll
    for (int r = 0; r < R_H; ++r) {
      const int k = tid + r * 256;
      if (k < NH * HT) sH[k] = ph[r];
    }
#pragma unroll
    for (int r = 0; r < R_G; ++r) {
      const int k = tid + r * 256;
      if (k < NT * NGEO) sGeo[k] = pg[r];
    }
#pragma unroll
    for (int r = 0; r < R_F; ++r) {
      const int k = tid + r * 256;
      if (k < NFACE * 4) sFG[k] = pf[r];
    }
  };

  // the descriptor of every tile face, once per face (one lane each, while
  // the tile's loads are in flight); the face and integration phases read it
  auto descs = [&](int tile) {
    const int ex0 = (tile % a.tiles_x) * TX, ey0 = (tile / a.tiles_x) * TY;
    for (int fi = tid; fi < NFACE; fi += 256) {
      const FaceDesc fd = face_desc<D, TX, TY>(fi, ex0, ey0, a);
      sFD[fi] = ((fd.dst[0] + 1) | (fd.edge[0] << 8) | (fd.cl[0] << 10)) |
                (((fd.dst[1] + 1) | (fd.edge[1] << 8) | (fd.cl[1] << 10)) << 16);
      sFS[fi] = (fd.src[0] + 128) | ((fd.src[1] + 128) << 8) | (fd.tag << 16);
    }
  };
  int tile = tile_of(a, 0);
  if (tile >= 0) {
    issue(tile);
    descs(tile);
  }
  tf.store(tb, MODE == 1 && shared_minv, tid);
  if (tid < MG * Ms) { sPsi[tid] = psi_r; sPsiw[tid] = psiw_r; }
  for (int it = 0; tile >= 0; ++it) {
    const int ex0 = (tile % a.tiles_x) * TX, ey0 = (tile / a.tiles_x) * TY;
    auto gidx = [&](int k) -> size_t {
      const int ty = k / (TX * NU);
      return ((size_t)(ey0 + ty) * a.nx + ex0) * NU + (k - ty * TX * NU);
    };
    commit();
    if (it > 0) descs(tile);  // the previous tile's descriptors were read up to its last barrier
    __syncthreads();
    // the output's Dt factor once per element (read after several barriers)
    if (MODE == 1 && tid < NT) sDte[tid] = shared_minv ? dt * (a.J0 / sGeo[tid * NGEO]) : dt;
    // RK operands of this thread's output entries, then the next tile's
    // inputs: all in flight during the compute below
    double pre_un[R_U], pre_acc[R_U];
#pragma unroll
    for (int r = 0; r < R_UV; ++r) {
#pragma unroll
      for (int v = 0; v < VW; ++v) pre_un[r * VW + v] = pre_acc[r * VW + v] = 0.0;
      const int k = VW * (tid + r * 256);
      if (MODE == 1 && k < NT * NU) {
        const size_t i = gidx(k);
        if constexpr (VEC) {
          if (a.stage < 3) {
            const double2 v = *reinterpret_cast<const double2*>(a.Un + i);
            pre_un[2 * r] = v.x;
            pre_un[2 * r + 1] = v.y;
          }
          if (a.stage > 0) {
            const double2 v = *reinterpret_cast<const double2*>(a.Uacc + i);
            pre_acc[2 * r] = v.x;
            pre_acc[2 * r + 1] = v.y;
          }
        } else {
          if (a.stage < 3) pre_un[r] = a.Un[i];
          if (a.stage > 0) pre_acc[r] = a.Uacc[i];
        }
      }
    }
    const int next = DGX_STAGE_PERSIST ? tile_of(a, it + 1) : -1;

    // ---- faces: one lane per (face, Gauss point)
    for (int itf = tid; itf < NFACE * MG; itf += 256) {
      const int fi = itf / MG, g = itf % MG;
      // descriptor: per side (dst + 1) | edge << 8 | cl << 10 in sFD, source slot + 128 in sFS
      const int fdw = sFD[fi], fsw = sFS[fi];
      const int fsrc[2] = {(fsw & 255) - 128, ((fsw >> 8) & 255) - 128};
      const int fcl[2] = {(fdw >> 10) & 7, (fdw >> 26) & 7};
      const int ftag = fsw >> 16;
      const double* fg = sFG + fi * 4;
      const double nxv = fg[0], nyv = (D == 2) ? fg[1] : 0.0, jf0 = fg[2], jf1 = fg[3];
      double tr[2][NF][Ms];
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int s = fsrc[d];
#pragma unroll
        for (int j = 0; j < Ms; ++j) {
          if (s >= 0) {
            const int node = closure_node<D, P>(fcl[d], j);
#pragma unroll
            for (int f = 0; f < NF; ++f) tr[d][f][j] = sU[s * SU + f * Ns + node];
          } else {
#pragma unroll
            for (int f = 0; f < NF; ++f) tr[d][f][j] = sH[(-1 - s) * HT + f * Ms + j];
          }
        }
      }
      if (ftag == BC_REFLECTIVE) {  // boundaries.cu:20-67 on the nodal face values
#pragma unroll
        for (int j = 0; j < Ms; ++j) {
          if constexpr (D == 1) {
            tr[1][1][j] = -tr[1][1][j];
          } else {
            const double nx2 = nxv * nxv, ny2 = nyv * nyv, m2nxny = -2 * nxv * nyv;
            const double inv = 1.0 / (nx2 + ny2);
            const double vxL = tr[0][1][j], vyL = tr[0][2][j];
            tr[1][1][j] = inv * ((-nx2 + ny2) * vxL + m2nxny * vyL);
            tr[1][2][j] = inv * (m2nxny * vxL + (nx2 - ny2) * vyL);
          }
        }
      }
      double ul[NF], ur[NF], F[NF], nc[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double sl = 0.0, sr = 0.0;
#pragma unroll
        for (int j = 0; j < Ms; ++j) {
          sl += tr[0][f][j] * sPsi[g + j * MG];
          sr += tr[1][f][j] * sPsi[g + j * MG];
        }
        ul[f] = sl;
        ur[f] = sr;
      }
      riemann<D, PH, FL>(ul, ur, nxv, nyv, F, nc);
      // the non-conservative terms are zero except for the last NNC fields
      // (alpha; phinc; alpha and beta): keep F and those; sides and face
      // Jacobians at integration
      double* w = sW + (fi * MG + g) * NWP;
#pragma unroll
      for (int f = 0; f < NF; ++f) w[f] = F[f];
#pragma unroll
      for (int m = 0; m < NNC; ++m) w[NF + m] = nc[NF - NNC + m];
    }
    __syncthreads();
    // the next tile's inputs (persistent blocks): in flight from here to the
    // next commit, past the face phase, whose registers peak
    if (next >= 0) issue(next);
    // ---- face integration against psi_w into the receiving elements' slots
    for (int itf = tid; itf < NFACE * 2 * Ms; itf += 256) {
      const int fi = itf / (2 * Ms), d = (itf / Ms) % 2, j = itf % Ms;
      const int side = (sFD[fi] >> (16 * d)) & 0xffff;
      const int dst = (side & 0xff) - 1;
      if (dst < 0) continue;
      const int cl = side >> 10, edge = (side >> 8) & 3;
      const int jf = (D == 2 && cl >= 4) ? reversed_pos<P>(j) : j;
      double* q = sQ + dst * SQ + edge * NF * Ms;
      const double jfd = sFG[fi * 4 + 2 + d];
      // the side's sign: fma(-1, F, nc) is -F + nc rounded once, the same
      // bits as the sum, without computing both sums and selecting per lane
      const double sgn = d ? 1.0 : -1.0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        double s = 0.0;
#pragma unroll
        for (int g = 0; g < MG; ++g) {
          const double* w = sW + (fi * MG + g) * NWP;
          const double ncf = f >= NF - NNC ? w[NF + f - (NF - NNC)] : 0.0;
          // q_L = (-F + nc) JF_0, q_R = (F + nc) JF_1 (physics.cu:800-804, kernels.cu:199-228)
          const double qv = __builtin_fma(sgn, w[f], ncf) * jfd;
          s += sPsiw[g + j * MG] * qv;
        }
        q[f * Ms + jf] = s;
      }
    }
    __syncthreads();

    // ---- element phases
    if constexpr (MFW) {
      // with the shared Minv the wave also applies Dt * Minv on the matrix cores
      const bool wave_minv = MODE == 1 && shared_minv;
      vol_mfma_wave<TW, PH, NT, SU, SQ, NGEO, TW, PREC == 2>(tb, a.dphi_w, sU, sGeo, sQ, sW, sNodeQ, tid,
                                                             wave_minv ? tb.minv : nullptr, sDte);
      __syncthreads();
    } else {
    for (int b = 0; b < NT; b += EPB) {
      {
        const int el = b + tid / TPE, g = tid % TPE;
        if (el < NT && g < NG)
          vol_phase1<D, P, PH>(tb, g, sU + el * SU, sGeo[el * NGEO], sGeo + el * NGEO + 1,
                               sW + (el - b) * SW + g * NQ);
      }
      __syncthreads();
      {
        const int el = b + tid / TPE, i = tid % TPE;
        double rhs[NF];
        const bool act = el < NT && i < Ns;
        if (act) vol_phase2<D, P, PH>(tb, i, sW + (el - b) * SW, sQ + el * SQ, rhs);
        __syncthreads();
        if (act)
#pragma unroll
          for (int f = 0; f < NF; ++f) sU[el * SU + f * Ns + i] = rhs[f];
      }
      __syncthreads();
    }
    }

    // ---- stage output (coalesced along the tile rows, VW entries per step)
#pragma unroll
    for (int r = 0; r < R_UV; ++r) {
      const int k0 = VW * (tid + r * 256);
      if (k0 >= NT * NU) break;
      const size_t gi = gidx(k0);
      double ov[VW], av[VW];
#pragma unroll
      for (int v = 0; v < VW; ++v) {
        const int k = k0 + v, el = k / NU, rr = k % NU;
        if constexpr (MODE == 0) {
          ov[v] = sU[el * SU + rr];
        } else if (MFW && shared_minv) {
          rk_values(a.stage, sU[el * SU + rr], pre_un[r * VW + v], pre_acc[r * VW + v], av[v], ov[v]);
          if constexpr (PH == 1) {
            if (a.means) sU[el * SU + rr] = ov[v];  // this thread's own entry: kept for the cell means below
          }
        } else {
          // shared Minv: Minv_e = Minv_0 * J_0 / J_e (affine elements); else per element
          const double dte = sDte[el];
          const double* M = shared_minv ? tb.minv
                                        : a.Minv + ((size_t)(ey0 + el / TX) * a.nx + ex0 + el % TX) * Ns * Ns;
          const int f = rr / Ns, i = rr % Ns;
          const double* R = sU + el * SU + f * Ns;
          double sol = 0.0;
#pragma unroll
          for (int ii = 0; ii < Ns; ++ii) sol += M[ii * Ns + i] * R[ii];
          rk_values(a.stage, dte * sol, pre_un[r * VW + v], pre_acc[r * VW + v], av[v], ov[v]);
        }
      }
      double* out = MODE == 0 ? a.fout + gi : a.Uout + gi;
      if constexpr (VEC) {
        *reinterpret_cast<double2*>(out) = make_double2(ov[0], ov[1]);
        if (MODE == 1 && a.stage < 3) *reinterpret_cast<double2*>(a.Uacc + gi) = make_double2(av[0], av[1]);
      } else {
        out[0] = ov[0];
        if (MODE == 1 && a.stage < 3) a.Uacc[gi] = av[0];
      }
    }
    if constexpr (MODE == 1 && MFW && PH == 1) {
      if (a.means && shared_minv) {
        // cell means (cell_means_rpa) and sensor averages (sensor_averages_2d)
        // of the output tile: one lane per (element, field) forms the field's
        // two means, then one lane per element the derived values (the same
        // functions and order as the one-lane forms, so the same bits)
        static_assert(NW >= 2 * NT * NF, "means staging fits the flux buffer");
        __syncthreads();  // the output tile is in sU, sW is free
        double* sM = sW;            // [el*NF + f] mode-0 means
        double* sA = sW + NT * NF;  // [el*NF + f] sensor nodal means
        for (int k = tid; k < NT * NF; k += 256) {
          const int el = k / NF, f = k % NF;
          const double* uf = sU + el * SU + f * Ns;
          sM[k] = cell_mode0(uf, a.w0);
          if (a.uavg) {
            const double v = sensor_nodal_mean<Ns>(uf);
            sA[k] = v;
            a.uavg[((size_t)(ey0 + el / TX) * a.nx + ex0 + el % TX) * NF + f] = v;
          }
        }
        __syncthreads();
        if (tid < NT) {
          double a3[3];
          rpa_of_means(sM + tid * NF, a3);
          const size_t eg = (size_t)(ey0 + tid / TX) * a.nx + ex0 + tid % TX;
          double* m = a.means + eg * 3;
          m[0] = a3[0];
          m[1] = a3[1];
          m[2] = a3[2];
          if (a.uavg) a.psens[eg] = sensor_cell<2, 1>(sA + tid * NF, 0).p;
        }
      }
    }
    if (next >= 0) __syncthreads();  // LDS is refilled for the next tile
    tile = next;
  }
}

// ------------------------------------------------------------------- host
template <int D>
struct TileShape;
// 1D: 128 elements per tile, or 8 where 128 does not divide the mesh
template <>
struct TileShape<1> { static constexpr int TX = 128, TX_SMALL = 8, TY = 1; };
#ifndef DGX_TILE2_TX
#define DGX_TILE2_TX 8
#endif
#ifndef DGX_TILE2_TY
#define DGX_TILE2_TY 4
#endif
template <>
struct TileShape<2> { static constexpr int TX = DGX_TILE2_TX, TY = DGX_TILE2_TY; };

bool stage_tiles(int D, int nx, int ny, int* tx, int* ty) {
  *tx = D == 1 ? (nx % TileShape<1>::TX == 0 ? TileShape<1>::TX : TileShape<1>::TX_SMALL) : TileShape<2>::TX;
  *ty = D == 1 ? TileShape<1>::TY : TileShape<2>::TY;
  return nx % *tx == 0 && (D == 1 || ny % *ty == 0);
}

template <typename Fn>
static hipError_t dispatch_stage(const dgx_config& c, Fn&& fn) {
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  auto by_flux = [&](auto d, auto p, auto ph) -> hipError_t {
    return c.flux == 0 ? fn(d, p, ph, I0{}) : fn(d, p, ph, I1{});
  };
  auto by_phys = [&](auto d, auto p) -> hipError_t {
    switch (c.physics) {
      case 0: return by_flux(d, p, I0{});
      case 1: return by_flux(d, p, I1{});
      case 2: return by_flux(d, p, I2{});
      default: return by_flux(d, p, I3{});
    }
  };
  auto by_order = [&](auto d) -> hipError_t {
    switch (c.order) {
      case 1: return by_phys(d, I1{});
      case 2: return by_phys(d, I2{});
      default: return by_phys(d, I3{});
    }
  };
  return c.dim == 1 ? by_order(I1{}) : by_order(I2{});
}

hipError_t launch_stage(const Solver& s, int mode, int stage, const double* Uin, const double* Un, double* Uacc,
                        double* Uout, const double* dt_dev, double* f_out, hipStream_t st, double* means,
                        double* uavg, double* psens, bool* means_done) {
  if (means_done) *means_done = false;
  const Cart& c = s.cart;
  StageArgs a{};
  a.phi = s.t.phi; a.dphi = s.t.dphi; a.phi_w = s.t.phi_w; a.dphi_w = s.t.dphi_w;
  a.psi = s.t.psi; a.psi_w = s.t.psi_w; a.Minv = s.t.Minv;
  a.J = s.t.J; a.invJac = s.t.invJac; a.gx = s.t.face_geo_x; a.gy = s.t.face_geo_y;
  a.J0 = s.t.J0; a.geo_stride = s.t.geo_stride; a.minv_stride = s.t.minv_stride;
  a.nx = c.nx; a.ny = c.ny;
  for (int k = 0; k < 4; ++k) a.bc[k] = c.bc[k];
  a.tiles_x = c.nx / c.tx;
  a.tiles_y = c.ny / c.ty;
  const int ntiles = a.tiles_x * a.tiles_y;
  a.stage = stage; a.Uin = Uin; a.Un = Un; a.Uacc = Uacc; a.Uout = Uout; a.dt = dt_dev; a.fout = f_out;
  a.gate = s.gate;
  a.w0 = s.t.mlp_vinv;  // first row: the mode-0 weights
  return dispatch_stage(s.cfg, [&](auto d, auto p, auto ph, auto fl) -> hipError_t {
    constexpr int D = decltype(d)::value;
    auto go = [&](auto txc) -> hipError_t {
      constexpr int P = decltype(p)::value, PH = decltype(ph)::value, FL = decltype(fl)::value;
      constexpr int TX = decltype(txc)::value, TY = TileShape<D>::TY;
      auto kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 0> : k_stage<D, P, PH, FL, 1, TX, TY, 0>;
      if constexpr (D == 2 && P == 2 && TX * TY == 32 && PH <= 1) {
        // 16-byte pairs when every state pointer allows them
        auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
        const bool vec = al(Uin) && (mode == 0 ? al(f_out) : al(Un) && al(Uacc) && al(Uout));
        if (vec) {
          kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 0, true> : k_stage<D, P, PH, FL, 1, TX, TY, 0, true>;
          if (s.precision == 1)
            kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 1, true> : k_stage<D, P, PH, FL, 1, TX, TY, 1, true>;
          if (s.precision == 2)
            kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 2, true> : k_stage<D, P, PH, FL, 1, TX, TY, 2, true>;
        } else {
          if (s.precision == 1)
            kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 1> : k_stage<D, P, PH, FL, 1, TX, TY, 1>;
          if (s.precision == 2)
            kern = mode == 0 ? k_stage<D, P, PH, FL, 0, TX, TY, 2> : k_stage<D, P, PH, FL, 1, TX, TY, 2>;
        }
      }
      // persistent grid: as many blocks as fit on the device at once
      int nb = ntiles;
      if (DGX_STAGE_PERSIST) {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) == hipSuccess && per_cu > 0)
          nb = std::min(ntiles, per_cu * cus);
      }
      a.xcd_bands = (ntiles % 8 == 0 && nb % 8 == 0) ? 1 : 0;
      // the epilogue's cell means: RK stages of 2D p=2 two-fluid tiles on the
      // matrix-core element phases with the shared Minv (k_stage's MFW path)
      constexpr bool mfw = D == 2 && P == 2 && PH == 1 && (TX * TY) % 16 == 0;
      a.means = nullptr;
      a.uavg = nullptr;
      if (mfw && mode == 1 && means && s.t.mlp_vinv && s.t.minv_stride == 0 &&
          (s.precision >= 1 || DGX_STAGE_MFMA64)) {
        a.means = means;
        a.uavg = uavg;
        a.psens = psens;
        if (means_done) *means_done = true;
      }
      hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, st, a);
      return hipGetLastError();
    };
    if constexpr (D == 1) {
      if (c.tx == TileShape<1>::TX_SMALL) return go(std::integral_constant<int, TileShape<1>::TX_SMALL>{});
    }
    return go(std::integral_constant<int, TileShape<D>::TX>{});
  });
}

}  // namespace dgx
