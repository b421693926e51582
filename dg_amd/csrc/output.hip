// Solution output and restart: the PRINTER formatter on the GPU and the
// reference's .pos files (Gmsh ElementNodeData views, one file per printed
// field).
//
// Formatter (Lformater, printer_{oned,twod}_{singlefluid,multifluid}.cu:20-110,
// GAMNCON): node by node, conserved -> printed primitive fields
//   1D single fluid  rho, ux, p          1D two-fluid  rho, ux, G, p
//   2D single fluid  rho, ux, uy, p      2D two-fluid  rho, ux, uy, G, p
// (G = gamma = 1 + 1/alpha), and the inverse for restart.  Compiled without
// FMA contraction: the same operations in the same order as the reference,
// so the printed values are its values bit for bit.
//
// .pos files (simpleMesh::writeSolution / readSolution, simpleMesh.cc:322-459):
// "<prefix><fname><step %010d>.pos" per field, header
//   $MeshFormat / 2.1 0 8 / $EndMeshFormat / $ElementNodeData / 1 / "<Name>" /
//   1 / <time> / 4 / <step> / 1 / <N_E> / <partition 0>
// then one line per element "<element id> <N_s> v_0 ... v_{N_s-1} " with 20
// significant digits, and $EndElementNodeData.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <type_traits>

#include "dgx_internal.h"
#include "physics.cuh"

namespace dgx {

// Lformater (printer_{oned,twod}_{singlefluid,multifluid,passive,stiffened}.cu:20-110)
// node by node.  Printed fields: rho, ux[, uy], then p (single fluid);
// G, p (two-fluid); p, phic, phinc (passive); G, Pinf, p (stiffened).
template <int D, int PH>
__global__ __launch_bounds__(256) void k_format(int N_s, int N_E, double* __restrict__ U, double* __restrict__ out,
                                                int inverse) {
  constexpr int NF = Fields<D, PH>::NF;
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long)N_E * N_s) return;
  const long e = k / N_s, i = k % N_s;
  double* u = U + e * NF * N_s + i;
  double* o = out + e * NF * N_s + i;
  constexpr int x = D + 1;  // first field after the velocities
  if (!inverse) {
    const double rho = u[0];
    const double ux = u[1 * N_s] / rho;
    const double uy = D == 2 ? u[2 * N_s] / rho : 0.0;
    const double et = u[x * N_s];
    const double gamma = (PH == 1 || PH == 3) ? 1 + 1.0 / u[(x + 1) * N_s] : kGammaSF;
    o[0] = rho;
    o[1 * N_s] = ux;
    if (D == 2) o[2 * N_s] = uy;
    if constexpr (PH == 3) {
      const double pinf = (1 - 1.0 / gamma) * u[(x + 2) * N_s];
      o[x * N_s] = gamma;
      o[(x + 1) * N_s] = pinf;
      if (D == 2)
        o[(x + 2) * N_s] = (gamma - 1) * (et - gamma * pinf / (gamma - 1) - 0.5 * (ux * ux + uy * uy) * rho);
      else
        o[(x + 2) * N_s] = (gamma - 1) * (et - gamma * pinf / (gamma - 1) - 0.5 * ux * ux * rho);
    } else {
      constexpr int ip = PH == 1 ? x + 1 : x;  // printed pressure
      if (PH == 1) o[x * N_s] = gamma;
      if (D == 2)
        o[ip * N_s] = (gamma - 1) * (et - 0.5 * (ux * ux + uy * uy) * rho);
      else
        o[ip * N_s] = (gamma - 1) * (et - 0.5 * ux * ux * rho);
      if constexpr (PH == 2) {
        o[(x + 1) * N_s] = u[(x + 1) * N_s] / rho;
        o[(x + 2) * N_s] = u[(x + 2) * N_s];
      }
    }
  } else {
    const double rho = o[0], ux = o[1 * N_s];
    const double uy = D == 2 ? o[2 * N_s] : 0.0;
    const double ke = D == 2 ? 0.5 * rho * (ux * ux + uy * uy) : 0.5 * rho * ux * ux;
    u[0] = rho;
    u[1 * N_s] = rho * ux;
    if (D == 2) u[2 * N_s] = rho * uy;
    if constexpr (PH == 3) {
      const double gamma = o[x * N_s], pinf = o[(x + 1) * N_s], p = o[(x + 2) * N_s];
      u[x * N_s] = p / (gamma - 1) + gamma * pinf / (gamma - 1) + ke;
      u[(x + 1) * N_s] = 1.0 / (gamma - 1);
      u[(x + 2) * N_s] = gamma * pinf / (gamma - 1);
    } else {
      const double gamma = PH == 1 ? o[x * N_s] : kGammaSF;
      const double p = o[(PH == 1 ? x + 1 : x) * N_s];
      u[x * N_s] = p / (gamma - 1) + ke;
      if (PH == 1) u[(x + 1) * N_s] = 1.0 / (gamma - 1);
      if constexpr (PH == 2) {
        const double phic = o[(x + 1) * N_s], phinc = o[(x + 2) * N_s];
        u[(x + 1) * N_s] = rho * phic;
        u[(x + 2) * N_s] = phinc;
      }
    }
  }
}

hipError_t launch_format(const Solver& s, double* U, double* out, int inverse, hipStream_t st) {
  const long n = (long)s.N_E * s.N_s;
  const dim3 g((unsigned)((n + 255) / 256)), b(256);
  auto go = [&](auto d) {
    constexpr int D = decltype(d)::value;
    switch (s.cfg.physics) {
      case 0: hipLaunchKernelGGL((k_format<D, 0>), g, b, 0, st, s.N_s, s.N_E, U, out, inverse); break;
      case 1: hipLaunchKernelGGL((k_format<D, 1>), g, b, 0, st, s.N_s, s.N_E, U, out, inverse); break;
      case 2: hipLaunchKernelGGL((k_format<D, 2>), g, b, 0, st, s.N_s, s.N_E, U, out, inverse); break;
      default: hipLaunchKernelGGL((k_format<D, 3>), g, b, 0, st, s.N_s, s.N_E, U, out, inverse); break;
    }
  };
  if (s.D == 1) go(std::integral_constant<int, 1>{});
  else go(std::integral_constant<int, 2>{});
  return hipGetLastError();
}

// DG_SOLVER::conservation (dg_solver.h:367-389): I[fc] = sum over elements e
// and Gauss points g of Ug[(e,fc),g]*J_e*w_g, with Ug = phi*U as a netlib
// gemm (each point's sum over nodes in node order, no contraction).  The
// reference adds the N_E*N_G terms one after another into one double; here
// every term has the reference's bits and the terms are summed as
// double-double values (TwoSum, then the low parts) in a fixed tree: each
// thread its elements in order, the block in LDS, the blocks in one final
// block.  The result is the exact sum of the terms to within about one
// rounding, deterministic for a given N_E.
struct DD {
  double hi, lo;
};
__device__ inline DD dd_add(DD a, double x) {
  const double s = a.hi + x, bp = s - a.hi;
  const double err = (a.hi - (s - bp)) + (x - bp);
  return DD{s, a.lo + err};
}
__device__ inline DD dd_add(DD a, DD b) {
  DD r = dd_add(a, b.hi);
  r.lo = r.lo + b.lo;
  return r;
}

constexpr int kConsThreads = 256;
constexpr int kConsMaxBlocks = 1024;
constexpr int kConsMaxNF = 6;

template <int NF>
__global__ __launch_bounds__(kConsThreads) void k_conservation(int N_s, int N_G, int N_E,
                                                               const double* __restrict__ U,
                                                               const double* __restrict__ phi,
                                                               const double* __restrict__ J, int geo_stride,
                                                               const double* __restrict__ weight,
                                                               double* __restrict__ partial) {
  __shared__ double sPhi[25 * 16], sW[25];
  __shared__ double sHi[kConsThreads], sLo[kConsThreads];
  for (int k = threadIdx.x; k < N_G * N_s; k += blockDim.x) sPhi[k] = phi[k];
  for (int k = threadIdx.x; k < N_G; k += blockDim.x) sW[k] = weight[k];
  __syncthreads();
  DD acc[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) acc[f] = DD{0.0, 0.0};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N_E; e += gridDim.x * blockDim.x) {
    const double Je = J[(long)e * geo_stride];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double* u = U + ((long)e * NF + f) * N_s;
      for (int g = 0; g < N_G; ++g) {
        double ug = 0.0;
        for (int l = 0; l < N_s; ++l) ug = ug + u[l] * sPhi[g + l * N_G];
        acc[f] = dd_add(acc[f], ug * Je * sW[g]);
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    sHi[threadIdx.x] = acc[f].hi;
    sLo[threadIdx.x] = acc[f].lo;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        const DD r = dd_add(DD{sHi[threadIdx.x], sLo[threadIdx.x]}, DD{sHi[threadIdx.x + w], sLo[threadIdx.x + w]});
        sHi[threadIdx.x] = r.hi;
        sLo[threadIdx.x] = r.lo;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      partial[((long)blockIdx.x * NF + f) * 2] = sHi[0];
      partial[((long)blockIdx.x * NF + f) * 2 + 1] = sLo[0];
    }
    __syncthreads();
  }
}

template <int NF>
__global__ __launch_bounds__(kConsThreads) void k_conservation_final(int nblocks, const double* __restrict__ partial,
                                                                     double* __restrict__ totals) {
  __shared__ double sHi[kConsThreads], sLo[kConsThreads];
  for (int f = 0; f < NF; ++f) {
    DD acc{0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
      acc = dd_add(acc, DD{partial[((long)b * NF + f) * 2], partial[((long)b * NF + f) * 2 + 1]});
    sHi[threadIdx.x] = acc.hi;
    sLo[threadIdx.x] = acc.lo;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        const DD r = dd_add(DD{sHi[threadIdx.x], sLo[threadIdx.x]}, DD{sHi[threadIdx.x + w], sLo[threadIdx.x + w]});
        sHi[threadIdx.x] = r.hi;
        sLo[threadIdx.x] = r.lo;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) totals[f] = sHi[0] + sLo[0];
    __syncthreads();
  }
}

int conservation_blocks(int N_E) {
  const int nb = (N_E + kConsThreads - 1) / kConsThreads;
  return nb < kConsMaxBlocks ? nb : kConsMaxBlocks;
}
size_t conservation_scratch_doubles() { return (size_t)kConsMaxBlocks * kConsMaxNF * 2; }

hipError_t launch_conservation(const Solver& s, const double* U, double* totals, double* scratch, hipStream_t st) {
  if (s.N_G > 25 || s.N_s > 16 || s.NF > kConsMaxNF) return hipErrorInvalidValue;
  const int nb = conservation_blocks(s.N_E);
  auto go = [&](auto nf) {
    constexpr int NF = decltype(nf)::value;
    hipLaunchKernelGGL((k_conservation<NF>), dim3(nb), dim3(kConsThreads), 0, st, s.N_s, s.N_G, s.N_E, U, s.t.phi,
                       s.t.J, s.t.geo_stride, s.t.weight, scratch);
    hipLaunchKernelGGL((k_conservation_final<NF>), dim3(1), dim3(kConsThreads), 0, st, nb, scratch, totals);
  };
  switch (s.NF) {
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 5: go(std::integral_constant<int, 5>{}); break;
    default: go(std::integral_constant<int, 6>{}); break;
  }
  return hipGetLastError();
}

// PRINTER::set_names (printer.cc:11-80): printed field names and file names
void pos_field_names(int D, int physics, std::vector<std::string>& names, std::vector<std::string>& fnames) {
  names = {"Rho", "Ux"};
  fnames = {"rho", "ux"};
  if (D == 2) { names.push_back("Uy"); fnames.push_back("uy"); }
  if (physics == DGX_MULTIFLUID || physics == DGX_STIFFENED) { names.push_back("G"); fnames.push_back("g"); }
  if (physics == DGX_STIFFENED) { names.push_back("Pinf"); fnames.push_back("pinf"); }
  names.push_back("P");
  fnames.push_back("p");
  if (physics == DGX_PASSIVE) {
    names.push_back("PhiC"); fnames.push_back("phic");
    names.push_back("PhiNC"); fnames.push_back("phinc");
  }
}

static std::string pos_path(const std::string& prefix, const std::string& fname, int step) {
  char st[24];
  std::snprintf(st, sizeof st, "%010i", step);
  return prefix + fname + st + ".pos";
}

std::string pos_write(const Tables& t, int physics, const double* out, const std::string& prefix, int step,
                      double time) {
  std::vector<std::string> names, fnames;
  pos_field_names(t.D, physics, names, fnames);
  const int NF = (int)names.size(), Ns = t.N_s;
  for (int fc = 0; fc < NF; ++fc) {
    const std::string path = pos_path(prefix, fnames[fc], step);
    std::ofstream o(path);
    if (!o) return "cannot write " + path;
    o.precision(20);
    o << "$MeshFormat\n2.1 0 8\n$EndMeshFormat\n";
    o << "$ElementNodeData\n";
    o << "1\n\"" << names[fc] << "\"\n";
    o << "1\n" << time << "\n";
    o << "4\n" << step << "\n1\n" << t.N_E << "\n" << 0 << "\n";
    for (int e = 0; e < t.N_E; ++e) {
      o << t.elem_id[e] << " " << Ns << " ";
      for (int i = 0; i < Ns; ++i) o << out[((size_t)e * NF + fc) * Ns + i] << " ";
      o << "\n";
    }
    o << "$EndElementNodeData\n";
    if (!o) return "write failed: " + path;
  }
  return "";
}

std::string pos_read(const Tables& t, int physics, const std::string& prefix, int step, double* time,
                     double* out) {
  std::vector<std::string> names, fnames;
  pos_field_names(t.D, physics, names, fnames);
  const int NF = (int)names.size(), Ns = t.N_s;
  for (int fc = 0; fc < NF; ++fc) {
    const std::string path = pos_path(prefix, fnames[fc], step);
    std::ifstream in(path);
    if (!in) return "no file named " + path;
    std::string line;
    for (int k = 0; k < 7; ++k) std::getline(in, line);  // format block, view header, name
    double tm;
    in >> tm;
    std::getline(in, line);
    for (int k = 0; k < 3; ++k) std::getline(in, line);  // 4, step, 1
    long fsize, part;
    in >> fsize;
    std::getline(in, line);
    in >> part;
    std::getline(in, line);
    if (!in) return path + ": bad header";
    if (fsize != t.N_E) return path + ": different number of elements than the mesh";
    if (part != 0) return path + ": different partition";
    for (int e = 0; e < t.N_E; ++e) {
      long id, nn;
      if (!(in >> id >> nn)) return path + ": missing element data";
      if (id != t.elem_id[e] || nn != Ns) return path + ": element id or number of nodes differs from the mesh";
      for (int i = 0; i < Ns; ++i)
        if (!(in >> out[((size_t)e * NF + fc) * Ns + i])) return path + ": bad value";
      std::getline(in, line);
    }
    if (time) *time = tm;
  }
  return "";
}

}  // namespace dgx
