// Host-side finite-element setup for the H1 Laplace hot path: meshes,
// topology, H1 dof tables, free dofs, element classes, reference-gradient
// tensors, CSR pattern and element-by-element assembly.
//
// This is the caller side of the device operators: it produces exactly the
// arrays the reference hands to its sparse matrix (comp/bilinearform.cpp:552-825,
// 1563-2215) and to its fused matrix-free operator (comp/matrixfree.cpp:1090-1464).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "intrule.hpp"

namespace ngsamd {

struct Mesh {
  int dim = 2;
  int nv = 0;
  std::vector<double> coords;   // nv * dim
  std::vector<int> els;         // ne * (dim+1), local vertex order as generated
  std::vector<int> bnd;         // nbnd * dim, boundary facets
  std::vector<int> bnd_index;   // 1-based boundary region per facet
  std::vector<std::string> bnd_names;

  // topology (BuildTopology)
  std::vector<int> edges;       // nedges * 2, ascending vertex pairs, sorted lexicographically
  std::vector<int> faces;       // nfaces * 3, ascending vertex triples, sorted (3D only)
  std::vector<int> el_edges;    // ne * (3 or 6), global edge of each local edge
  std::vector<int> el_faces;    // ne * 4 (3D)
  std::vector<int> bnd_edges;   // nbnd * (1 or 3)
  std::vector<int> bnd_faces;   // nbnd (3D)

  int NVE() const { return dim + 1; }
  int NE() const { return (int)(els.size() / (dim + 1)); }
  int NBnd() const { return (int)(bnd.size() / dim); }
  int NEdges() const { return (int)(edges.size() / 2); }
  int NFaces() const { return (int)(faces.size() / 3); }
  int NElEdges() const { return dim == 2 ? 3 : 6; }

  void BuildTopology();
};

// python/meshes.py:166-195 (quads=False, flip_triangles=False): (nx+1)(ny+1)
// points i*(nx+1)+j at (j/nx, i/ny); per square two triangles
// [b, b+1, b+nx+1] and [b+1, b+nx+2, b+nx+1]; boundary segments bottom(1),
// right(2), top(3), left(4).
Mesh MakeStructured2DMesh(int nx, int ny);

// python/meshes.py:351-410 (hexes=False): points i(ny+1)(nz+1)+j(nz+1)+k at
// (i/nx, j/ny, k/nz); per cube the six Kuhn tetrahedra of the reference's
// template list; boundary facets are the element faces owned by one element.
Mesh MakeStructured3DMesh(int nx, int ny, int nz);

// Netgen .vol text format (points, surfaceelements, volumeelements,
// edgesegmentsgi2).  2D meshes take their triangles from "surfaceelements"
// and boundary from "edgesegmentsgi2"; 3D meshes take tetrahedra from
// "volumeelements" and boundary triangles from "surfaceelements".
Mesh ReadVolMesh(const std::string& text);

struct H1Space {
  const Mesh* mesh = nullptr;
  int order = 1;
  int ndof = 0;
  int nd_el = 0;                      // local dofs per element
  std::vector<int> first_edge_dof;    // nedges+1
  std::vector<int> first_face_dof;    // nfaces+1 (3D)
  std::vector<int> first_el_dof;      // ne+1
  std::vector<int> el_dofs;           // ne * nd_el, reference GetDofNrs order
  std::vector<uint8_t> freedofs;      // ndof
  std::vector<int> classnr;           // ne

  // comp/h1hofespace.cpp:823-950 (UpdateDofTables) and :1389-1413 (GetDofNrs).
  // dirichlet_mask: bit r-1 set marks boundary region r as Dirichlet
  // (comp/fespace.cpp:765-810).
  H1Space(const Mesh& m, int order, uint64_t dirichlet_mask);
};

// Element colouring of FESpace::FinalizeUpdate (comp/fespace.cpp:518-760)
// for an H1 space (the colouring of its P1 space: elements of one colour
// share no vertex, hence no dof): rounds of 32 colours, each uncoloured
// element takes the lowest colour of the round free at all its vertices
// (32-bit masks), elements in order.  Returns the colour of every element.
std::vector<int> ElementColoring(const Mesh& m);

// Per-class reference data of the fused operator (comp/matrixfree.cpp:1159-1211).
struct ElementClass {
  int classnr = 0;
  std::vector<int> elements;   // global element numbers in this class, ascending
  std::vector<double> bmat;    // [Q][nd][dim]  d phi_i / d xi_d at ip q
  std::vector<double> shape;   // [Q][nd]       phi_i at ip q
};

struct ElementGeometry {
  std::vector<double> detj;    // ne, |det F|
  std::vector<double> ginv;    // ne * dim*dim, F^{-1} F^{-T} (row-major)
};

ElementGeometry ComputeGeometry(const Mesh& mesh);

std::vector<ElementClass> BuildClasses(const H1Space& fes, const IntegrationRule& ir);

struct CSRPattern {
  int64_t n = 0;
  std::vector<int64_t> rowptr;
  std::vector<int> cols;
};

// comp/bilinearform.cpp:552-825 + linalg/sparsematrix.cpp:450-975: row i holds
// the sorted union of the dof sets of all elements containing dof i
// (Dirichlet rows included, full non-symmetric storage).
CSRPattern BuildPattern(const H1Space& fes);

// Assemble A = sum_e (kappa_e K_e + rho_e M_e) into the pattern, where K_e is
// the grad.grad element matrix and M_e the mass matrix; either part can be
// switched off, and a null coefficient array means 1.  Row-parallel: each row
// sums its contributions in ascending element order, so the result does not
// depend on the thread count.
// fem/symbolicintegrator.cpp:1549-1860, linalg/sparsematrix_impl.hpp:150-189.
void AssembleCSR(const H1Space& fes, const CSRPattern& pat, const IntegrationRule& ir,
                 bool stiffness, const double* kappa, bool mass, const double* rho,
                 double* values);

// Row-sum / diagonal helpers.
void ExtractDiagonal(const CSRPattern& pat, const double* values, double* diag);

}  // namespace ngsamd
