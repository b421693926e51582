// Structured-mesh and DG setup tables for the DG right-hand side.
//
// Builds everything the reference's driver computes once before time
// stepping (main.cc:312-441, 557-583, 712-740; dg_functions.cc:33-238;
// simpleMesh.cc:206-319, 462-532, 664-754, 770-989, 1102-1190), for
// Cartesian 1D line meshes and 2D quadrilateral meshes, in the reference's
// memory layouts so that the same arrays feed the reference-named launchers,
// the fused engine and the CPU oracle.
//
// Conventions (all kept from the reference):
//  * nodal Lagrange basis on Gmsh equispaced nodes.  Lines: xi = -1, 1, then
//    interior nodes ascending (polynomialBasis.cc:20-31).  Quads: 4 corners
//    counter-clockwise from (-1,-1), then edge nodes edge by edge, then the
//    interior nodes recursively (polynomialBasis.cc:713-751).
//  * face closures: closure j = {j, j+1, edge j nodes}, reversed closure
//    nsides+j = {j+1, j, edge j nodes reversed} (polynomialBasis.cc:911-926);
//    line closures {0},{1} (polynomialBasis.cc:928-934).
//  * element quadrature "order 2p+1": lines p+1 Gauss points; quads 4 points
//    at p=1 (tabulated order), n x n tensor points with x the slow index for
//    p>=2, n = p+2 (GaussQuadratureQuad.cc:113-149); faces use the line rule.
//  * interfaces ordered by their sorted node-id vectors (simpleMesh.cc:986-988);
//    the first element found owns side 0 and the outward normal of side 0.
//  * boundary physical tags: 1 periodic, 2 farfield (zero gradient),
//    3 reflective wall (simpleMesh.cc:855-868, 695-703).
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace dgx {

enum BoundaryTag { BC_PERIODIC = 1, BC_FARFIELD = 2, BC_REFLECTIVE = 3 };

struct MeshSpec {
  int dim = 1;          // 1 or 2
  int order = 2;        // polynomial order p (1..3)
  int nx = 1, ny = 1;   // elements per direction (ny ignored in 1D)
  double x0 = 0, x1 = 1, y0 = 0, y1 = 1;
  int bc[4] = {2, 2, 2, 2};  // 1D: left,right; 2D: left,right,bottom,top
  int nfields = 4;      // N_F (sizes the map/invmap tables)
};

struct Tables {
  // sizes (reference names)
  int D = 1, order = 1, N_s = 0, N_G = 0, M_s = 0, M_G = 0, N_N = 0;
  int N_E = 0, M_T = 0, M_B = 0, N_F = 0, N_s1D = 0, nx = 0, ny = 0;
  int rflctiveIdx = 0;  // number of reflective faces at the head of boundaryMap
  int bc[4] = {0, 0, 0, 0};  // boundary tags of the domain sides (MeshSpec::bc)
  double Dx = 0;        // minimum cell size used by the CFL step (simpleMesh.cc:1140-1190)
  double refArea = 0;

  // reference element
  std::vector<double> ref_nodes;   // [i*D+a] reference coordinates of the element nodes
  std::vector<double> points;      // [g*D+a] element quadrature points
  std::vector<double> weight;      // [g]
  std::vector<double> pointsF, weightF;  // face quadrature points/weights
  std::vector<int> closures;       // [c*M_s+j], c in [0, 2*nsides) (2D) or {0,1} (1D)

  // basis tables, column-major like fullMatrix::copyMatrixToPointer
  std::vector<double> phi, dphi, phi_w, dphi_w, psi, psi_w;

  // geometry
  std::vector<double> xyz;         // [(e*N_s+i)*D+a] physical node coordinates
  std::vector<double> J;           // [e]
  std::vector<double> invJac;      // [((e*N_G+g)*D+alpha)*D+a] = d xi_alpha / d x_a
  std::vector<double> JF;          // [t*2+d]
  std::vector<double> normals;     // [t*D+a]  outward normal of side 0
  std::vector<double> Minv;        // [(e*N_s+k)*N_s+i] = Minv_e(i,k)

  // connectivity
  std::vector<int> map;            // [((t*N_F+f)*2+d)*M_s+j] -> index into U
  std::vector<int> invmap;         // [((e*N_F+f)*M_s*N_N+k)*2+{0,1}]
  std::vector<int> boundaryMap;    // reflective faces first
  std::vector<int> neighbors;      // [e*N_N+{L,R,D,U}], negative = -physical tag
  std::vector<int> face_elem;      // [t*2+d] element on each side
  std::vector<int> face_closure;   // [t*2+d] closure id used by each side
  std::vector<int> face_tag;       // [t] physical tag (0 interior)
  std::vector<int> elem_id;        // [e] Gmsh element id (.pos files): the file's, or the one write_gmsh gives

  // limiter transforms (column-major N_s x N_s)
  std::vector<double> Lag2Mono, Mono2Lag;                       // 1D
  std::vector<double> Lag2MonoX, MonoX2MonoY, MonoY2Lag;         // 2D Cartesian
  std::vector<double> Lag2MonoY, MonoX2Lag;                      // 2D, derived (limiting.h:395-396)
};

// Builds the tables; returns an empty string on success or an error message.
std::string build_tables(const MeshSpec& spec, Tables& out);

// A Gmsh 2.2 ASCII mesh as the reference reads it (simpleMesh::load,
// simpleMesh.cc:141-204): nodes (x, y, z by file position, id -> position)
// and elements (Gmsh type, physical tag, node ids), single partition.
struct GmshMesh {
  struct Elem {
    long id = 0;
    int type = 0, tag = 0;
    std::vector<long> nodes;
  };
  std::vector<double> xyz;
  std::unordered_map<long, long> node_index;
  std::vector<Elem> elems;
};
std::string read_gmsh(const std::string& path, GmshMesh& out);
// Dimension and order from the element types present (quadrilaterals first).
bool gmsh_dim_order(const GmshMesh& g, int& dim, int& order);
int gmsh_element_type(int dim, int order);
// Tables of a Gmsh mesh: elements in file order, node index = id - 1,
// boundary physical tags 1 periodic, 2 farfield, 3 wall, 4 (as 2).
std::string build_tables_gmsh(const GmshMesh& g, int nfields, Tables& out);
// Writes a mesh built by build_tables as Gmsh 2.2 ASCII (boundary faces as
// tagged boundary elements); reading it back gives the same tables.
std::string write_gmsh(const Tables& t, const std::string& path);
std::string gmsh_interfaces(const GmshMesh& g, int D, int p, int N_E, int N_s, int M_s, int nsides,
                            const std::vector<int64_t>& gid, Tables& t);

// 1D Lagrange value/derivative helpers on arbitrary nodes (exposed for tests).
double lagrange1d(const std::vector<double>& nodes, int i, double x);
double lagrange1d_deriv(const std::vector<double>& nodes, int i, double x);

// Gauss-Legendre nodes (ascending) and weights by Newton iteration on P_n.
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w);

}  // namespace dgx
