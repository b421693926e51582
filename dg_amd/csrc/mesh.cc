// Structured-mesh and DG setup tables.  See mesh.h for the conventions and
// the reference functions each part restates.
#include "mesh.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace dgx {

// ---------------------------------------------------------------------------
// 1D building blocks
// ---------------------------------------------------------------------------
double lagrange1d(const std::vector<double>& nodes, int i, double x) {
  double v = 1.0;
  for (size_t m = 0; m < nodes.size(); ++m)
    if ((int)m != i) v *= (x - nodes[m]) / (nodes[i] - nodes[m]);
  return v;
}

double lagrange1d_deriv(const std::vector<double>& nodes, int i, double x) {
  // d/dx prod_{m!=i} (x-x_m)/(x_i-x_m) = sum_k 1/(x_i-x_k) prod_{m!=i,k} ...
  double s = 0.0;
  for (size_t k = 0; k < nodes.size(); ++k) {
    if ((int)k == i) continue;
    double t = 1.0 / (nodes[i] - nodes[k]);
    for (size_t m = 0; m < nodes.size(); ++m)
      if ((int)m != i && m != k) t *= (x - nodes[m]) / (nodes[i] - nodes[m]);
    s += t;
  }
  return s;
}

void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int k = 0; k < n; ++k) {
    // Chebyshev-like initial guess for the k-th root (descending), refined by Newton.
    double z = std::cos(M_PI * (k + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int m = 2; m <= n; ++m) {
        double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    // recompute derivative at the converged root
    double p0 = 1.0, p1 = z;
    for (int m = 2; m <= n; ++m) {
      double p2 = ((2.0 * m - 1.0) * z * p1 - (m - 1.0) * p0) / m;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - k] = z;
    w[n - 1 - k] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;  // exact centre
  for (int k = 0; k < n / 2; ++k) {  // exact symmetry
    double a = 0.5 * (x[n - 1 - k] - x[k]);
    x[k] = -a;
    x[n - 1 - k] = a;
    double ww = 0.5 * (w[k] + w[n - 1 - k]);
    w[k] = w[n - 1 - k] = ww;
  }
}

// Gmsh line nodes: -1, 1, then interior ascending (polynomialBasis.cc:20-31).
static std::vector<double> line_nodes(int p) {
  std::vector<double> v(p + 1);
  v[0] = -1.0;
  v[1] = 1.0;
  double dd = 2.0 / p;
  for (int i = 2; i < p + 1; ++i) v[i] = -1.0 + dd * (i - 1);
  return v;
}

// Gmsh quad nodes (polynomialBasis.cc:713-751), appended as (x,y) pairs.
static void quad_nodes(int p, double scale, std::vector<double>& out) {
  if (p == 0) { out.push_back(0.0); out.push_back(0.0); return; }
  const double c[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
  size_t base = out.size();
  for (int k = 0; k < 4; ++k) { out.push_back(c[k][0] * scale); out.push_back(c[k][1] * scale); }
  if (p > 1) {
    const int edges[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    for (int ie = 0; ie < 4; ++ie) {
      int p0 = edges[ie][0], p1 = edges[ie][1];
      for (int i = 1; i < p; ++i) {
        double x0 = out[base + 2 * p0], y0 = out[base + 2 * p0 + 1];
        double x1 = out[base + 2 * p1], y1 = out[base + 2 * p1 + 1];
        out.push_back(x0 + i * (x1 - x0) / p);
        out.push_back(y0 + i * (y1 - y0) / p);
      }
    }
    // interior nodes: recursive quad of order p-2 (order 0 = the centre)
    quad_nodes(p - 2, scale * (1.0 - 2.0 / p), out);
  }
}

// Dense column-major LU inverse with partial pivoting (small N).
static bool invert(int n, const std::vector<double>& A, std::vector<double>& Ainv) {
  std::vector<double> a(A);
  Ainv.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) Ainv[i + i * n] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[r + c * n]) > std::fabs(a[piv + c * n])) piv = r;
    if (a[piv + c * n] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[c + k * n], a[piv + k * n]);
        std::swap(Ainv[c + k * n], Ainv[piv + k * n]);
      }
    double d = a[c + c * n];
    for (int k = 0; k < n; ++k) { a[c + k * n] /= d; Ainv[c + k * n] /= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double f = a[r + c * n];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r + k * n] -= f * a[c + k * n];
        Ainv[r + k * n] -= f * Ainv[c + k * n];
      }
    }
  }
  return true;
}

// C = A*B, column-major, sequential sum over the inner index (netlib dgemm order).
static void matmul(int M, int N, int K, const std::vector<double>& A,
                   const std::vector<double>& B, std::vector<double>& C) {
  C.assign((size_t)M * N, 0.0);
  for (int j = 0; j < N; ++j)
    for (int l = 0; l < K; ++l) {
      double t = B[l + j * K];
      for (int i = 0; i < M; ++i) C[i + j * M] = C[i + j * M] + t * A[i + l * M];
    }
}

// Lagrange (Gmsh line order) -> Taylor coefficients about 0:
// T(j,i) = d^j phi_i / dx^j (0), exact polynomial arithmetic in double.
static std::vector<double> lagrange_to_taylor(const std::vector<double>& nodes) {
  int n = (int)nodes.size();
  std::vector<double> T((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    // coefficients of prod_{m!=i} (x - x_m) / (x_i - x_m), ascending powers
    std::vector<double> c(1, 1.0);
    for (int m = 0; m < n; ++m) {
      if (m == i) continue;
      double d = nodes[i] - nodes[m];
      std::vector<double> nc(c.size() + 1, 0.0);
      for (size_t k = 0; k < c.size(); ++k) {
        nc[k + 1] += c[k] / d;
        nc[k] += -nodes[m] * c[k] / d;
      }
      c.swap(nc);
    }
    double fact = 1.0;
    for (int j = 0; j < n; ++j) {
      if (j > 0) fact *= j;
      T[j + i * n] = c[j] * fact;  // j-th derivative at 0 = j! * c_j
    }
  }
  return T;
}

// Taylor -> Lagrange: Tinv(i,j) = x_i^j / j!.
static std::vector<double> taylor_to_lagrange(const std::vector<double>& nodes) {
  int n = (int)nodes.size();
  std::vector<double> T((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double pw = 1.0, fact = 1.0;
    for (int j = 0; j < n; ++j) {
      if (j > 0) { pw *= nodes[i]; fact *= j; }
      T[i + j * n] = pw / fact;
    }
  }
  return T;
}

// ---------------------------------------------------------------------------
// Gmsh 2.2 ASCII meshes (simpleMesh::load, simpleMesh.cc:141-204)
// ---------------------------------------------------------------------------
// Gmsh element types: lines of 2, 3, 4 nodes; quadrilaterals of 4, 9, 16; point.
int gmsh_element_type(int D, int p) {
  static const int line[4] = {0, 1, 8, 26}, quad[4] = {0, 3, 10, 36};
  return D == 1 ? line[p] : quad[p];
}
static int gmsh_boundary_type(int D, int p) { return D == 1 ? 15 : gmsh_element_type(1, p); }

bool gmsh_dim_order(const GmshMesh& g, int& D, int& p) {
  for (int dim = 2; dim >= 1; --dim)
    for (int q = 1; q <= 3; ++q) {
      const int ty = gmsh_element_type(dim, q);
      for (const auto& el : g.elems)
        if (el.type == ty) {
          D = dim;
          p = q;
          return true;
        }
    }
  return false;
}

std::string read_gmsh(const std::string& path, GmshMesh& g) {
  std::ifstream in(path);
  if (!in) return "cannot open " + path;
  g = GmshMesh();
  std::string line;
  bool nodes = false, elems = false;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "$MeshFormat") {
      std::getline(in, line);
      std::istringstream f(line);
      double version = 0;
      int ascii = -1;
      f >> version >> ascii;
      if (version < 2.0 || version >= 3.0 || ascii != 0) return "only Gmsh 2.x ASCII meshes are supported";
    } else if (line == "$Nodes") {
      long n = 0;
      if (!(in >> n) || n < 0) return "bad $Nodes count";
      g.xyz.resize((size_t)n * 3);
      for (long i = 0; i < n; ++i) {
        long id;
        if (!(in >> id >> g.xyz[i * 3] >> g.xyz[i * 3 + 1] >> g.xyz[i * 3 + 2])) return "bad node line";
        if (!g.node_index.emplace(id, i).second) return "duplicate node id";
      }
      nodes = true;
    } else if (line == "$Elements") {
      long n = 0;
      if (!(in >> n) || n < 0) return "bad $Elements count";
      std::getline(in, line);
      for (long i = 0; i < n; ++i) {
        if (!std::getline(in, line)) return "missing element lines";
        std::istringstream f(line);
        GmshMesh::Elem el;
        int ntags;
        if (!(f >> el.id >> el.type >> ntags) || ntags < 0) return "bad element line";
        for (int j = 0; j < ntags; ++j) {
          int tag;
          if (!(f >> tag)) return "bad element tags";
          if (j == 0) el.tag = tag;  // physical tag (simpleMesh.cc:184)
          if (j == 3 && tag != 1) return "partitioned meshes (partition tag > 1) are not supported";
        }
        long node;
        while (f >> node) el.nodes.push_back(node);
        g.elems.push_back(std::move(el));
      }
      elems = true;
    }
  }
  if (!nodes || !elems) return "no $Nodes or $Elements section";
  return "";
}

// pairPeriodic (simpleMesh.cc:21-91): boundary faces that are images of one
// another across the domain.  1D: any two distinct points; 2D: a vertical
// face pairs with one of the same y extent, a horizontal with the same x.
static bool pair_periodic(const GmshMesh& g, int D, const std::vector<int64_t>& n1, const std::vector<int64_t>& n2) {
  const double eps = 1e-8;
  auto xy = [&](int64_t gid, int a) { return g.xyz[(size_t)g.node_index.at(gid + 1) * 3 + a]; };
  if (D == 1) return std::fabs(xy(n1[0], 0) - xy(n2[0], 0)) >= eps;
  double b1[4] = {xy(n1[0], 0), xy(n1[0], 0), xy(n1[0], 1), xy(n1[0], 1)};  // xs, xe, ys, ye
  double b2[4] = {xy(n2[0], 0), xy(n2[0], 0), xy(n2[0], 1), xy(n2[0], 1)};
  for (size_t j = 1; j < n1.size(); ++j) {
    b1[0] = std::min(b1[0], xy(n1[j], 0)); b1[1] = std::max(b1[1], xy(n1[j], 0));
    b1[2] = std::min(b1[2], xy(n1[j], 1)); b1[3] = std::max(b1[3], xy(n1[j], 1));
    b2[0] = std::min(b2[0], xy(n2[j], 0)); b2[1] = std::max(b2[1], xy(n2[j], 0));
    b2[2] = std::min(b2[2], xy(n2[j], 1)); b2[3] = std::max(b2[3], xy(n2[j], 1));
  }
  bool same = true;
  for (int k = 0; k < 4; ++k) same = same && std::fabs(b1[k] - b2[k]) < eps;
  if (same) return false;
  if (std::fabs(b1[2] - b1[3]) > eps) return std::fabs(b1[2] - b2[2]) < eps && std::fabs(b1[3] - b2[3]) < eps;
  if (std::fabs(b1[0] - b1[1]) > eps) return std::fabs(b1[0] - b2[0]) < eps && std::fabs(b1[1] - b2[1]) < eps;
  return false;
}

// BuildInterfaces (simpleMesh.cc:770-989) for one partition: faces keyed by
// their sorted node ids; a boundary element gives its face the physical tag
// (2, 3, 4: the element is its own neighbour through the same closure);
// the second element found uses closure j + nsides; periodic faces (tag 1)
// are paired by geometry.  Faces come out in key order.
std::string gmsh_interfaces(const GmshMesh& g, int D, int p, int N_E, int N_s, int M_s, int nsides,
                            const std::vector<int64_t>& gid, Tables& t) {
  struct Face { int tag = 0, e0 = -1, c0 = -1, e1 = -1, c1 = -1; };
  std::map<std::vector<int64_t>, Face> faces;
  const int bt = gmsh_boundary_type(D, p);
  for (const auto& el : g.elems)
    if (el.type == bt) {
      std::vector<int64_t> key;
      for (long n : el.nodes) key.push_back(n - 1);
      std::sort(key.begin(), key.end());
      faces[key] = Face{el.tag};
    }
  const int ncl = (D == 1) ? 2 : 4;
  std::vector<int64_t> key(M_s);
  for (int e = 0; e < N_E; ++e)
    for (int j = 0; j < ncl; ++j) {
      for (int k = 0; k < M_s; ++k) key[k] = gid[(size_t)e * N_s + t.closures[j * M_s + k]];
      std::sort(key.begin(), key.end());
      Face& f = faces[key];
      if (f.e0 < 0) {
        f.e0 = e;
        f.c0 = j;
        if (f.tag == 2 || f.tag == 3 || f.tag == 4) {
          f.e1 = e;
          f.c1 = j;
        }
      } else if (f.e0 != e) {
        if (f.e1 < 0) {
          f.e1 = e;
          f.c1 = j + nsides;
        } else if (f.e1 != e) {
          return "a face shared by more than two elements";
        }
      }
    }
  for (auto& a : faces)
    if (a.second.tag == 1 && a.second.e1 < 0)
      for (auto& b : faces)
        if (b.second.tag == 1 && b.second.e1 < 0 && pair_periodic(g, D, a.first, b.first)) {
          a.second.e1 = b.second.e0;
          a.second.c1 = b.second.c0 + nsides;
          b.second.e1 = a.second.e0;
          b.second.c1 = a.second.c0 + nsides;
        }
  for (const auto& a : faces) {
    const Face& f = a.second;
    if (f.e0 < 0) return "a boundary element that touches no element";
    if (f.e1 < 0) return "a face with one element and no boundary tag (or an unpaired periodic face)";
    if (f.tag < 0 || f.tag > 4) return "boundary physical tags must be 1..4";
    t.face_elem.push_back(f.e0);
    t.face_elem.push_back(f.e1);
    t.face_closure.push_back(f.c0);
    t.face_closure.push_back(f.c1);
    t.face_tag.push_back(f.tag);
  }
  return "";
}

std::string write_gmsh(const Tables& t, const std::string& path) {
  const int D = t.D, N_s = t.N_s, M_s = t.M_s;
  // file node ids: the rank of each node's lattice id, so that the sorted
  // face keys (and hence the face order) are the same when read back; a
  // node shared by several elements is the same lattice point
  std::vector<std::pair<double, double>> pts;
  std::map<std::pair<long long, long long>, long> id_of;  // exact coordinates (bits) -> id
  auto bits = [](double v) {
    long long b;
    std::memcpy(&b, &v, 8);
    return b;
  };
  std::vector<long> node_id((size_t)t.N_E * N_s);
  std::vector<std::pair<std::pair<double, double>, size_t>> order;  // ((y, x), slot): lattice order is y-major
  for (int e = 0; e < t.N_E; ++e)
    for (int i = 0; i < N_s; ++i) {
      const double x = t.xyz[((size_t)e * N_s + i) * D], y = D == 2 ? t.xyz[((size_t)e * N_s + i) * D + 1] : 0.0;
      order.push_back({{y, x}, (size_t)e * N_s + i});
    }
  std::stable_sort(order.begin(), order.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  long next = 0;
  for (const auto& o : order) {
    const auto key = std::make_pair(bits(o.first.first), bits(o.first.second));
    auto it = id_of.find(key);
    if (it == id_of.end()) {
      it = id_of.emplace(key, ++next).first;
      pts.push_back({o.first.second, o.first.first});
    }
    node_id[o.second] = it->second;
  }
  std::ofstream out(path);
  if (!out) return "cannot write " + path;
  out << "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n" << pts.size() << "\n";
  char buf[128];
  for (size_t i = 0; i < pts.size(); ++i) {
    std::snprintf(buf, sizeof buf, "%zu %.17g %.17g 0\n", i + 1, pts[i].first, pts[i].second);
    out << buf;
  }
  out << "$EndNodes\n$Elements\n";
  long nb = 0;
  for (int f = 0; f < t.M_T; ++f) nb += t.face_tag[f] != 0;
  out << nb + t.N_E << "\n";
  // boundary elements first, numbered 1..nb unless the element ids (kept for
  // .pos files) take those numbers, then after the largest element id
  long id = 0;
  if (!t.elem_id.empty() && *std::min_element(t.elem_id.begin(), t.elem_id.end()) <= nb)
    id = *std::max_element(t.elem_id.begin(), t.elem_id.end());
  const int bt = gmsh_boundary_type(D, t.order), vt = gmsh_element_type(D, t.order);
  for (int f = 0; f < t.M_T; ++f) {
    if (!t.face_tag[f]) continue;
    const int e = t.face_elem[2 * f], cl = t.face_closure[2 * f];
    out << ++id << " " << bt << " 2 " << t.face_tag[f] << " " << t.face_tag[f];
    for (int k = 0; k < M_s; ++k) out << " " << node_id[(size_t)e * N_s + t.closures[cl * M_s + k]];
    out << "\n";
  }
  for (int e = 0; e < t.N_E; ++e) {
    out << (t.elem_id.empty() ? ++id : (long)t.elem_id[e]) << " " << vt << " 2 88 1";
    for (int i = 0; i < N_s; ++i) out << " " << node_id[(size_t)e * N_s + i];
    out << "\n";
  }
  out << "$EndElements\n";
  return out ? "" : "write failed: " + path;
}

// ---------------------------------------------------------------------------
// Table builder
// ---------------------------------------------------------------------------
// Structured mesh (gm == nullptr) or a Gmsh mesh: the two differ only in
// where the element nodes come from and how boundary faces find their tag
// and periodic partner; every table after that is built by the same code.
static std::string build_tables_impl(const MeshSpec& s, const GmshMesh* gm, Tables& t) {
  const int D = s.dim, p = s.order;
  if (D != 1 && D != 2) return "dim must be 1 or 2";
  if (p < 1 || p > 3) return "order must be in 1..3";
  if (!gm) {
  if (s.nx < 1 || (D == 2 && s.ny < 1)) return "mesh needs at least one element per direction";
  for (int k = 0; k < 2 * D; ++k)
    if (s.bc[k] < 1 || s.bc[k] > 3) return "boundary tags must be 1 (periodic), 2 (farfield) or 3 (wall)";
  if (D == 1 && (s.bc[0] == BC_PERIODIC) != (s.bc[1] == BC_PERIODIC)) return "periodic tags must pair";
  if (D == 2 && ((s.bc[0] == BC_PERIODIC) != (s.bc[1] == BC_PERIODIC) ||
                 (s.bc[2] == BC_PERIODIC) != (s.bc[3] == BC_PERIODIC)))
    return "periodic tags must pair";
  if (D == 2 && ((s.bc[0] == BC_PERIODIC && s.nx < 2) || (s.bc[2] == BC_PERIODIC && s.ny < 2)))
    return "periodic direction needs at least two elements";
  }

  t = Tables();
  t.D = D;
  t.order = p;
  t.N_F = s.nfields;
  t.nx = s.nx;
  t.ny = (D == 2) ? s.ny : 1;
  for (int k = 0; k < 4; ++k) t.bc[k] = (D == 2 || k < 2) ? s.bc[k] : 0;
  t.N_s1D = p + 1;
  const std::vector<double> ln = line_nodes(p);

  // --- reference element nodes
  if (D == 1) {
    t.ref_nodes = ln;
  } else {
    quad_nodes(p, 1.0, t.ref_nodes);
  }
  t.N_s = (int)t.ref_nodes.size() / D;
  const int N_s = t.N_s;
  // index of each node's coordinate in the 1D node list
  std::vector<int> ia(N_s), ib(N_s, 0);
  auto find1d = [&](double v) {
    for (int k = 0; k <= p; ++k)
      if (std::fabs(ln[k] - v) < 1e-12) return k;
    return -1;
  };
  for (int i = 0; i < N_s; ++i) {
    ia[i] = find1d(t.ref_nodes[i * D]);
    if (D == 2) ib[i] = find1d(t.ref_nodes[i * D + 1]);
  }

  // --- quadrature
  std::vector<double> gx, gw;
  gauss_legendre(p + 1, gx, gw);  // line rule of order 2p+1: p+1 points
  if (D == 1) {
    t.points = gx;
    t.weight = gw;
  } else if (p == 1) {
    // tabulated 4-point rule order (GaussQuadratureQuad.cc:25-35)
    const double a = 1.0 / std::sqrt(3.0);
    const double q[4][2] = {{a, a}, {-a, a}, {a, -a}, {-a, -a}};
    for (int g = 0; g < 4; ++g) { t.points.push_back(q[g][0]); t.points.push_back(q[g][1]); t.weight.push_back(1.0); }
  } else {
    int n = p + 2;
    std::vector<double> qx, qw;
    gauss_legendre(n, qx, qw);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        t.points.push_back(qx[i]);
        t.points.push_back(qx[j]);
        t.weight.push_back(qw[i] * qw[j]);
      }
  }
  t.N_G = (int)t.weight.size();
  const int N_G = t.N_G;
  if (D == 1) {
    t.M_s = 1; t.M_G = 1;
    t.pointsF = {0.0}; t.weightF = {1.0};
  } else {
    t.M_s = p + 1;
    t.pointsF = gx; t.weightF = gw;
    t.M_G = p + 1;
  }
  const int M_s = t.M_s, M_G = t.M_G;
  t.N_N = (D == 1) ? 2 : 4;
  const int N_N = t.N_N;
  const int nsides = (D == 1) ? 0 : 4;
  t.refArea = (D == 2) ? 4.0 : 2.0;

  // --- closures
  if (D == 1) {
    t.closures = {0, 1};
  } else {
    t.closures.assign(8 * M_s, 0);
    for (int j = 0; j < 4; ++j) {
      t.closures[j * M_s + 0] = j;
      t.closures[j * M_s + 1] = (j + 1) % 4;
      t.closures[(4 + j) * M_s + 0] = (j + 1) % 4;
      t.closures[(4 + j) * M_s + 1] = j;
      for (int i = 0; i < p - 1; ++i) {
        t.closures[j * M_s + 2 + i] = 4 + (p - 1) * j + i;
        t.closures[(4 + j) * M_s + 2 + i] = 4 + (p - 1) * (j + 1) - i - 1;
      }
    }
  }

  // --- basis tables (main.cc:372-441)
  auto phi_at = [&](int i, const double* x) {
    double v = lagrange1d(ln, ia[i], x[0]);
    if (D == 2) v *= lagrange1d(ln, ib[i], x[1]);
    return v;
  };
  auto dphi_at = [&](int i, const double* x, int a) {
    if (D == 1) return lagrange1d_deriv(ln, ia[i], x[0]);
    if (a == 0) return lagrange1d_deriv(ln, ia[i], x[0]) * lagrange1d(ln, ib[i], x[1]);
    return lagrange1d(ln, ia[i], x[0]) * lagrange1d_deriv(ln, ib[i], x[1]);
  };
  t.phi.assign((size_t)N_G * N_s, 0.0);
  t.phi_w.assign((size_t)N_G * N_s, 0.0);
  t.dphi.assign((size_t)N_G * D * N_s, 0.0);
  t.dphi_w.assign((size_t)N_G * D * N_s, 0.0);
  for (int g = 0; g < N_G; ++g)
    for (int i = 0; i < N_s; ++i) {
      double v = phi_at(i, &t.points[g * D]);
      t.phi[g + i * N_G] = v;
      t.phi_w[g + i * N_G] = v * t.weight[g];
      for (int a = 0; a < D; ++a) {
        double dv = dphi_at(i, &t.points[g * D], a);
        t.dphi[(g * D + a) + i * N_G * D] = dv;
        t.dphi_w[(g * D + a) + i * N_G * D] = dv * t.weight[g];
      }
    }
  t.psi.assign((size_t)M_G * M_s, 0.0);
  t.psi_w.assign((size_t)M_G * M_s, 0.0);
  std::vector<double> dpsi0(M_s, 0.0);  // face-basis derivative at the first face point
  if (D == 1) {
    t.psi[0] = 1.0;
    t.psi_w[0] = 1.0;
  } else {
    for (int g = 0; g < M_G; ++g)
      for (int j = 0; j < M_s; ++j) {
        double v = lagrange1d(ln, j, t.pointsF[g]);
        t.psi[g + j * M_G] = v;
        t.psi_w[g + j * M_G] = v * t.weightF[g];
      }
    for (int j = 0; j < M_s; ++j) dpsi0[j] = lagrange1d_deriv(ln, j, t.pointsF[0]);
  }

  // --- elements and physical nodes
  const int nx = s.nx, ny = t.ny;
  t.N_E = gm ? 0 : nx * ny;
  std::vector<const GmshMesh::Elem*> gvol;  // Gmsh: volume elements in file order
  if (gm) {
    const int vt = gmsh_element_type(D, p);
    for (const auto& el : gm->elems)
      if (el.type == vt) gvol.push_back(&el);
    if (gvol.empty()) return "no elements of the mesh's type";
    t.N_E = (int)gvol.size();
    t.nx = t.ny = 0;  // not a structured mesh
    for (int k = 0; k < 4; ++k) t.bc[k] = 0;
  }
  const int N_E = t.N_E;
  const double hx = (s.x1 - s.x0) / nx;
  const double hy = (D == 2) ? (s.y1 - s.y0) / ny : 0.0;
  const int W = p * nx + 1;  // lattice width (global node ids)
  std::vector<int64_t> gid((size_t)N_E * N_s);
  t.xyz.assign((size_t)N_E * N_s * D, 0.0);
  if (gm) {
    // node index = Gmsh node id - 1 (simpleMesh.cc:160-164, 188-190)
    for (int e = 0; e < N_E; ++e) {
      const GmshMesh::Elem& el = *gvol[e];
      if ((int)el.nodes.size() != N_s) return "element with the wrong number of nodes";
      for (int i = 0; i < N_s; ++i) {
        const int64_t id = el.nodes[i];
        auto it = gm->node_index.find(id);
        if (it == gm->node_index.end()) return "element refers to an undefined node";
        gid[(size_t)e * N_s + i] = id - 1;
        for (int a = 0; a < D; ++a) t.xyz[(e * N_s + i) * D + a] = gm->xyz[(size_t)it->second * 3 + a];
      }
    }
  }
  for (int e = 0; e < (gm ? 0 : N_E); ++e) {
    int ex = e % nx, ey = e / nx;
    for (int i = 0; i < N_s; ++i) {
      int la = (int)std::lround((t.ref_nodes[i * D] + 1.0) * p / 2.0);
      int I = p * ex + la;
      t.xyz[(e * N_s + i) * D + 0] = s.x0 + I * (hx / p);
      int64_t id = I;
      if (D == 2) {
        int lb = (int)std::lround((t.ref_nodes[i * D + 1] + 1.0) * p / 2.0);
        int Jn = p * ey + lb;
        t.xyz[(e * N_s + i) * D + 1] = s.y0 + Jn * (hy / p);
        id = (int64_t)Jn * W + I;
      }
      gid[(size_t)e * N_s + i] = id;
    }
  }

  // --- per-element Jacobians (dg_functions.cc:33-70): affine, evaluated at g=0
  t.J.assign(N_E, 0.0);
  t.invJac.assign((size_t)N_E * N_G * D * D, 0.0);
  for (int e = 0; e < N_E; ++e) {
    double Jac[2][2] = {{0, 0}, {0, 0}};  // Jac[alpha][a] = d x_alpha / d xi_a
    for (int al = 0; al < D; ++al)
      for (int a = 0; a < D; ++a) {
        double sum = 0.0;
        for (int i = 0; i < N_s; ++i)
          sum = sum + t.xyz[(e * N_s + i) * D + al] * t.dphi[(0 * D + a) + i * N_G * D];
        Jac[al][a] = sum;
      }
    double det = (D == 1) ? std::sqrt(Jac[0][0] * Jac[0][0]) : Jac[0][0] * Jac[1][1] - Jac[1][0] * Jac[0][1];
    t.J[e] = std::fabs(det);
    double inv[2][2];
    if (D == 1) {
      inv[0][0] = 1.0 / Jac[0][0];
    } else {
      inv[0][0] = 1.0 / det * Jac[1][1];
      inv[1][0] = -1.0 / det * Jac[1][0];
      inv[0][1] = -1.0 / det * Jac[0][1];
      inv[1][1] = 1.0 / det * Jac[0][0];
    }
    for (int g = 0; g < N_G; ++g)
      for (int al = 0; al < D; ++al)
        for (int a = 0; a < D; ++a) t.invJac[((size_t)(e * N_G + g) * D + al) * D + a] = inv[al][a];
  }

  if (gm && D == 2) {
    // The kernels keep one invJac per element, as the reference's affine
    // elements need (dg_functions.cc:33-70 takes det at the first point):
    // refuse elements whose corners are not a parallelogram.
    for (int e = 0; e < N_E; ++e) {
      const double* x = &t.xyz[(size_t)e * N_s * 2];
      const double sc = std::fabs(x[4] - x[0]) + std::fabs(x[5] - x[1]) + std::fabs(x[6] - x[2]) + std::fabs(x[7] - x[3]);
      for (int i = 0; i < N_s; ++i) {  // every node where the corners' affine map puts it
        const double xi = t.ref_nodes[2 * i], eta = t.ref_nodes[2 * i + 1];
        for (int a = 0; a < 2; ++a) {
          const double xa = x[a] + 0.5 * (xi + 1.0) * (x[2 + a] - x[a]) + 0.5 * (eta + 1.0) * (x[6 + a] - x[a]);
          if (std::fabs(xa - x[2 * i + a]) > 1e-10 * sc)
            return "element is not affine (a parallelogram with equispaced nodes); the kernels need one invJac per element";
        }
      }
    }
  }

  // --- inverse mass matrices (dg_functions.cc:98-174), shared across equal J
  t.Minv.assign((size_t)N_E * N_s * N_s, 0.0);
  {
    std::unordered_map<uint64_t, std::vector<double>> cache;
    for (int e = 0; e < N_E; ++e) {
      double jac = t.J[e];
      uint64_t key;
      std::memcpy(&key, &jac, 8);
      auto it = cache.find(key);
      if (it == cache.end()) {
        std::vector<double> M((size_t)N_s * N_s, 0.0), Mi;
        for (int k = 0; k < N_s; ++k)
          for (int i = 0; i < N_s; ++i) {
            double sum = 0.0;
            for (int g = 0; g < N_G; ++g) sum = sum + t.phi[g + i * N_G] * (t.phi_w[g + k * N_G] * jac);
            M[i + k * N_s] = sum;
          }
        if (!invert(N_s, M, Mi)) return "singular mass matrix";
        it = cache.emplace(key, Mi).first;
      }
      const std::vector<double>& Mi = it->second;
      for (int k = 0; k < N_s; ++k)
        for (int i = 0; i < N_s; ++i) t.Minv[((size_t)e * N_s + k) * N_s + i] = Mi[i + k * N_s];
    }
  }

  // --- interfaces (simpleMesh.cc:770-989)
  t.face_elem.clear(); t.face_closure.clear(); t.face_tag.clear();
  if (gm) {
    std::string err = gmsh_interfaces(*gm, D, p, N_E, N_s, M_s, nsides, gid, t);
    if (!err.empty()) return err;
  } else {
  struct Visit { std::array<int64_t, 4> key; int e, cl; };
  std::vector<Visit> visits;
  visits.reserve((size_t)N_E * (D == 1 ? 2 : 4));
  const int ncl = (D == 1) ? 2 : 4;
  for (int e = 0; e < N_E; ++e)
    for (int j = 0; j < ncl; ++j) {
      Visit v;
      v.key.fill(INT64_MAX);
      for (int k = 0; k < M_s; ++k) v.key[k] = gid[(size_t)e * N_s + t.closures[j * M_s + k]];
      std::sort(v.key.begin(), v.key.begin() + M_s);
      v.e = e;
      v.cl = j;
      visits.push_back(v);
    }
  std::stable_sort(visits.begin(), visits.end(), [](const Visit& a, const Visit& b) { return a.key < b.key; });
  // which domain side (0 L,1 R,2 B,3 T) a boundary closure sits on
  auto side_of = [&](int e, int cl) -> int {
    int ex = e % nx, ey = e / nx;
    if (D == 1) return (cl == 0 && ex == 0) ? 0 : ((cl == 1 && ex == nx - 1) ? 1 : -1);
    if (cl == 3 && ex == 0) return 0;
    if (cl == 1 && ex == nx - 1) return 1;
    if (cl == 0 && ey == 0) return 2;
    if (cl == 2 && ey == ny - 1) return 3;
    return -1;
  };
  // periodic partner element of a boundary element across its domain side
  auto partner = [&](int e, int side, int& pcl) -> int {
    int ex = e % nx, ey = e / nx;
    if (D == 1) { pcl = side == 0 ? 1 : 0; return side == 0 ? nx - 1 : 0; }
    switch (side) {
      case 0: pcl = 1; return ey * nx + (nx - 1);
      case 1: pcl = 3; return ey * nx + 0;
      case 2: pcl = 2; return (ny - 1) * nx + ex;
      default: pcl = 0; return 0 * nx + ex;
    }
  };
  for (size_t k = 0; k < visits.size();) {
    size_t k2 = k + 1;
    while (k2 < visits.size() && visits[k2].key == visits[k].key) ++k2;
    const Visit& a = visits[k];
    if (k2 - k == 2) {
      const Visit& b = visits[k + 1];
      t.face_elem.push_back(a.e); t.face_elem.push_back(b.e);
      t.face_closure.push_back(a.cl); t.face_closure.push_back(b.cl + nsides);
      t.face_tag.push_back(0);
    } else if (k2 - k == 1) {
      int side = side_of(a.e, a.cl);
      if (side < 0) return "internal error: unmatched interior face";
      int tag = s.bc[side];
      t.face_elem.push_back(a.e);
      t.face_closure.push_back(a.cl);
      if (tag == BC_PERIODIC) {
        int pcl = 0;
        int pe = partner(a.e, side, pcl);
        t.face_elem.push_back(pe);
        t.face_closure.push_back(pcl + nsides);
      } else {
        t.face_elem.push_back(a.e);
        t.face_closure.push_back(a.cl);
      }
      t.face_tag.push_back(tag);
    } else {
      return "internal error: face shared by more than two elements";
    }
    k = k2;
  }
  }
  t.M_T = (int)t.face_tag.size();
  const int M_T = t.M_T;
  // element ids for .pos files: the file's, or those write_gmsh gives (after
  // the boundary elements it writes first)
  t.elem_id.resize(N_E);
  if (gm) {
    for (int e = 0; e < N_E; ++e) t.elem_id[e] = (int)gvol[e]->id;
  } else {
    int nb = 0;
    for (int f = 0; f < M_T; ++f) nb += t.face_tag[f] != 0;
    for (int e = 0; e < N_E; ++e) t.elem_id[e] = nb + e + 1;
  }

  // --- normals (simpleMesh.cc:206-319): sum of physical gradients of the
  // closure nodes' own basis functions at their nodes, normalised.
  t.normals.assign((size_t)M_T * D, 0.0);
  for (int f = 0; f < M_T; ++f) {
    int e = t.face_elem[f * 2 + 0], cl = t.face_closure[f * 2 + 0];
    double n[2] = {0, 0};
    for (int k = 0; k < M_s; ++k) {
      int node = t.closures[cl * M_s + k];
      const double* xn = &t.ref_nodes[node * D];
      for (int al = 0; al < D; ++al)
        for (int a = 0; a < D; ++a)
          n[al] += t.invJac[((size_t)(e * N_G + 0) * D + a) * D + al] * dphi_at(node, xn, a);
    }
    double norm2 = 0.0;
    for (int al = 0; al < D; ++al) norm2 += n[al] * n[al];
    for (int al = 0; al < D; ++al) t.normals[f * D + al] = n[al] / std::sqrt(norm2);
  }

  // --- face Jacobians (dg_functions.cc:72-96)
  t.JF.assign((size_t)M_T * 2, 1.0);
  if (D == 2) {
    for (int f = 0; f < M_T; ++f)
      for (int d = 0; d < 2; ++d) {
        int e = t.face_elem[f * 2 + d], cl = t.face_closure[f * 2 + d];
        double det = 0.0;
        for (int al = 0; al < D; ++al) {
          double jf = 0.0;
          for (int j = 0; j < M_s; ++j) jf = jf + t.xyz[(e * N_s + t.closures[cl * M_s + j]) * D + al] * dpsi0[j];
          det += jf * jf;
        }
        t.JF[f * 2 + d] = std::sqrt(det);
      }
  }

  // --- map / invmap (dg_functions.cc:177-238)
  const int N_F = s.nfields;
  t.map.assign((size_t)M_s * M_T * N_F * 2, 0);
  t.invmap.assign((size_t)M_s * N_N * N_E * N_F * 2, 0);
  {
    std::vector<int> cnt((size_t)N_F * N_E, 0);
    for (int f = 0; f < M_T; ++f) {
      int e1 = t.face_elem[f * 2 + 0], e2 = t.face_elem[f * 2 + 1];
      const int* cl1 = &t.closures[t.face_closure[f * 2 + 0] * M_s];
      const int* cl2 = &t.closures[t.face_closure[f * 2 + 1] * M_s];
      for (int fc = 0; fc < N_F; ++fc)
        for (int j = 0; j < M_s; ++j) {
          t.map[((f * N_F + fc) * 2 + 0) * M_s + j] = (e1 * N_F + fc) * N_s + cl1[j];
          t.map[((f * N_F + fc) * 2 + 1) * M_s + j] = (e2 * N_F + fc) * N_s + cl2[j];
          int& c1 = cnt[e1 * N_F + fc];
          t.invmap[((size_t)(e1 * N_F + fc) * M_s * N_N + c1) * 2 + 0] = cl1[j];
          t.invmap[((size_t)(e1 * N_F + fc) * M_s * N_N + c1) * 2 + 1] = ((f * N_F + fc) * 2 + 0) * M_s + j;
          ++c1;
          if (t.face_tag[f] == 0) {
            int& c2 = cnt[e2 * N_F + fc];
            t.invmap[((size_t)(e2 * N_F + fc) * M_s * N_N + c2) * 2 + 0] = cl2[j];
            t.invmap[((size_t)(e2 * N_F + fc) * M_s * N_N + c2) * 2 + 1] = ((f * N_F + fc) * 2 + 1) * M_s + j;
            ++c2;
          }
        }
    }
  }

  // --- boundary map (simpleMesh.cc:462-532): reflective faces in face order
  t.boundaryMap.clear();
  for (int f = 0; f < M_T; ++f)
    if (t.face_tag[f] == BC_REFLECTIVE) t.boundaryMap.push_back(f);
  t.rflctiveIdx = (int)t.boundaryMap.size();
  t.M_B = (int)t.boundaryMap.size();

  // --- neighbours, Cartesian LRDU order (simpleMesh.cc:664-754)
  t.neighbors.assign((size_t)N_N * N_E, 0);
  const double eps = 1e-6;
  for (int f = 0; f < M_T; ++f) {
    int el1 = t.face_elem[f * 2 + 0], el2 = t.face_elem[f * 2 + 1];
    int tag = t.face_tag[f];
    if (el2 == el1 && (tag == 2 || tag == 3 || tag == 4)) el2 = -tag;
    double nxv = t.normals[f * D + 0], nyv = (D == 2) ? t.normals[f * D + 1] : 0.0;
    int s1, s2;
    if (std::fabs(nyv) < eps && nxv < 0) { s1 = 0; s2 = 1; }
    else if (std::fabs(nyv) < eps && nxv > 0) { s1 = 1; s2 = 0; }
    else if (std::fabs(nxv) < eps && nyv < 0) { s1 = 2; s2 = 3; }
    else if (std::fabs(nxv) < eps && nyv > 0) { s1 = 3; s2 = 2; }
    else return "neighbour creation: face normal is not axis aligned";
    t.neighbors[el1 * N_N + s1] = el2;
    if (el2 >= 0) t.neighbors[el2 * N_N + s2] = el1;
  }

  // --- minimum cell size (simpleMesh.cc:1140-1190)
  if (D == 1) {
    double dx = t.xyz[(0 * N_s + 1) * D] - t.xyz[(0 * N_s + 0) * D];
    for (int e = 0; e < N_E; ++e) {
      double d = t.xyz[(e * N_s + 1) * D] - t.xyz[(e * N_s + 0) * D];
      if (dx > d) dx = d;
    }
    t.Dx = dx;
  } else {
    double best = 0.0;
    bool first = true;
    for (int e = 0; e < N_E; ++e) {
      double cx = 0, cy = 0;
      for (int k = 0; k < 4; ++k) { cx += t.xyz[(e * N_s + k) * 2]; cy += t.xyz[(e * N_s + k) * 2 + 1]; }
      cx /= 4; cy /= 4;
      for (int k = 0; k < 4; ++k) {
        double x1 = t.xyz[(e * N_s + k) * 2], y1 = t.xyz[(e * N_s + k) * 2 + 1];
        double x2 = t.xyz[(e * N_s + (k + 1) % 4) * 2], y2 = t.xyz[(e * N_s + (k + 1) % 4) * 2 + 1];
        double d = std::fabs((x2 - x1) * (y1 - cy) - (x1 - cx) * (y2 - y1)) /
                   std::sqrt((x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1));
        if (first || best > d) { best = d; first = false; }
      }
    }
    t.Dx = 2 * best;
  }

  // --- limiter transforms (main.cc:557-583, 1348-1524; limiting.h:395-396)
  std::vector<double> T1 = lagrange_to_taylor(ln), T1inv = taylor_to_lagrange(ln);
  const int n1 = p + 1;
  if (D == 1) {
    t.Lag2Mono = T1;
    t.Mono2Lag = T1inv;
  } else {
    // Px: y-levels in order (min, max, interior ascending), nodes sorted by
    // x-level within each; Py likewise with the roles swapped.
    std::vector<int> px(N_s), py(N_s);
    int c = 0;
    for (int yl = 0; yl < n1; ++yl)
      for (int xl = 0; xl < n1; ++xl)
        for (int i = 0; i < N_s; ++i)
          if (ib[i] == yl && ia[i] == xl) px[c++] = i;
    c = 0;
    for (int xl = 0; xl < n1; ++xl)
      for (int yl = 0; yl < n1; ++yl)
        for (int i = 0; i < N_s; ++i)
          if (ia[i] == xl && ib[i] == yl) py[c++] = i;
    auto l2m = [&](const std::vector<int>& perm) {
      std::vector<double> L((size_t)N_s * N_s, 0.0);
      for (int sl = 0; sl < n1; ++sl)
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i) L[(sl * n1 + j) + perm[sl * n1 + i] * N_s] = T1[j + i * n1];
      return L;
    };
    auto m2l = [&](const std::vector<int>& perm) {
      std::vector<double> L((size_t)N_s * N_s, 0.0);
      for (int sl = 0; sl < n1; ++sl)
        for (int i = 0; i < n1; ++i)
          for (int j = 0; j < n1; ++j) L[perm[sl * n1 + i] + (sl * n1 + j) * N_s] = T1inv[i + j * n1];
      return L;
    };
    t.Lag2MonoX = l2m(px);
    std::vector<double> monoX2Lag_exact = m2l(px);
    std::vector<double> lag2MonoY_exact = l2m(py);
    t.MonoY2Lag = m2l(py);
    matmul(N_s, N_s, N_s, lag2MonoY_exact, monoX2Lag_exact, t.MonoX2MonoY);
    matmul(N_s, N_s, N_s, t.MonoY2Lag, t.MonoX2MonoY, t.MonoX2Lag);
    matmul(N_s, N_s, N_s, t.MonoX2MonoY, t.Lag2MonoX, t.Lag2MonoY);
  }
  return "";
}

std::string build_tables(const MeshSpec& s, Tables& t) { return build_tables_impl(s, nullptr, t); }

std::string build_tables_gmsh(const GmshMesh& g, int nfields, Tables& t) {
  MeshSpec s;
  if (!gmsh_dim_order(g, s.dim, s.order)) return "no line or quadrilateral elements of order 1..3";
  s.nfields = nfields;
  return build_tables_impl(s, &g, t);
}

}  // namespace dgx
