#include "fe.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>

#include "h1shapes.hpp"

namespace ngsamd {

// ------------------------------------------------------------------ meshes

Mesh MakeStructured2DMesh(int nx, int ny) {
  if (nx < 1 || ny < 1) throw std::runtime_error("MakeStructured2DMesh: nx, ny must be >= 1");
  Mesh m;
  m.dim = 2;
  m.nv = (nx + 1) * (ny + 1);
  m.coords.resize(2 * (size_t)m.nv);
  for (int i = 0; i <= ny; i++)
    for (int j = 0; j <= nx; j++) {
      size_t v = (size_t)i * (nx + 1) + j;
      m.coords[2 * v] = double(j) / nx;
      m.coords[2 * v + 1] = double(i) / ny;
    }
  m.els.reserve((size_t)6 * nx * ny);
  for (int i = 0; i < ny; i++)
    for (int j = 0; j < nx; j++) {
      int b = i * (nx + 1) + j;
      int t1[3] = {b, b + 1, b + nx + 1};
      int t2[3] = {b + 1, b + nx + 2, b + nx + 1};
      m.els.insert(m.els.end(), t1, t1 + 3);
      m.els.insert(m.els.end(), t2, t2 + 3);
    }
  m.bnd_names = {"bottom", "right", "top", "left"};
  for (int i = 0; i < nx; i++) { m.bnd.push_back(i); m.bnd.push_back(i + 1); m.bnd_index.push_back(1); }
  for (int i = 0; i < ny; i++) {
    m.bnd.push_back(i * (nx + 1) + nx); m.bnd.push_back((i + 1) * (nx + 1) + nx); m.bnd_index.push_back(2);
  }
  for (int i = 0; i < nx; i++) {
    m.bnd.push_back(ny * (nx + 1) + i + 1); m.bnd.push_back(ny * (nx + 1) + i); m.bnd_index.push_back(3);
  }
  for (int i = 0; i < ny; i++) {
    m.bnd.push_back((i + 1) * (nx + 1)); m.bnd.push_back(i * (nx + 1)); m.bnd_index.push_back(4);
  }
  m.BuildTopology();
  return m;
}

Mesh MakeStructured3DMesh(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::runtime_error("MakeStructured3DMesh: sizes must be >= 1");
  Mesh m;
  m.dim = 3;
  m.nv = (nx + 1) * (ny + 1) * (nz + 1);
  m.coords.resize(3 * (size_t)m.nv);
  for (int i = 0; i <= nx; i++)
    for (int j = 0; j <= ny; j++)
      for (int k = 0; k <= nz; k++) {
        size_t v = ((size_t)i * (ny + 1) + j) * (nz + 1) + k;
        m.coords[3 * v] = double(i) / nx;
        m.coords[3 * v + 1] = double(j) / ny;
        m.coords[3 * v + 2] = double(k) / nz;
      }
  static const int kuhn[6][4] = {{0, 4, 5, 6}, {0, 6, 7, 4}, {0, 3, 7, 6},
                                 {0, 1, 6, 5}, {0, 1, 2, 6}, {0, 3, 6, 2}};
  m.els.reserve((size_t)24 * nx * ny * nz);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++) {
        int base = i * (ny + 1) * (nz + 1) + j * (nz + 1) + k;
        int up = base + (ny + 1) * (nz + 1);
        int p[8] = {base, base + 1, base + (nz + 1) + 1, base + (nz + 1),
                    up, up + 1, up + (nz + 1) + 1, up + (nz + 1)};
        for (auto& q : kuhn)
          for (int l = 0; l < 4; l++) m.els.push_back(p[q[l]]);
      }
  m.bnd_names = {"default"};
  m.BuildTopology();
  return m;
}

static std::string NextToken(std::istringstream& is) {
  std::string t;
  is >> t;
  return t;
}

Mesh ReadVolMesh(const std::string& text) {
  Mesh m;
  std::istringstream is(text);
  std::string line, tok;
  int dim = 3;
  std::vector<std::vector<int>> surf, vol, segs;
  std::vector<int> surf_bc, seg_bc;
  std::vector<double> pts;
  while (is >> tok) {
    if (tok == "dimension") {
      is >> dim;
    } else if (tok == "surfaceelements" || tok == "surfaceelementsgi" || tok == "surfaceelementsuv") {
      int n; is >> n;
      for (int e = 0; e < n; e++) {
        int surfnr, bcnr, domin, domout, np;
        is >> surfnr >> bcnr >> domin >> domout >> np;
        std::vector<int> v(np);
        for (int l = 0; l < np; l++) is >> v[l];
        std::getline(is, line);  // skip geometry info if present
        surf.push_back(v);
        surf_bc.push_back(bcnr);
      }
    } else if (tok == "volumeelements") {
      int n; is >> n;
      for (int e = 0; e < n; e++) {
        int matnr, np;
        is >> matnr >> np;
        std::vector<int> v(np);
        for (int l = 0; l < np; l++) is >> v[l];
        vol.push_back(v);
      }
    } else if (tok == "edgesegmentsgi2" || tok == "edgesegmentsgi" || tok == "edgesegments") {
      int n; is >> n;
      std::getline(is, line);
      for (int e = 0; e < n; e++) {
        std::getline(is, line);
        std::istringstream ls(line);
        int surfid, zero, p1, p2;
        ls >> surfid >> zero >> p1 >> p2;
        segs.push_back({p1, p2});
        seg_bc.push_back(surfid);
      }
    } else if (tok == "points") {
      int n; is >> n;
      pts.resize(3 * (size_t)n);
      for (int i = 0; i < 3 * n; i++) is >> pts[i];
    } else if (tok == "endmesh") {
      break;
    }
  }
  m.dim = dim;
  m.nv = (int)(pts.size() / 3);
  m.coords.resize((size_t)dim * m.nv);
  for (int v = 0; v < m.nv; v++)
    for (int d = 0; d < dim; d++) m.coords[(size_t)dim * v + d] = pts[3 * (size_t)v + d];
  if (dim == 2) {
    for (auto& t : surf) {
      if (t.size() != 3) throw std::runtime_error("ReadVolMesh: only triangles supported in 2D");
      for (int l = 0; l < 3; l++) m.els.push_back(t[l] - 1);
    }
    for (size_t s = 0; s < segs.size(); s++) {
      m.bnd.push_back(segs[s][0] - 1);
      m.bnd.push_back(segs[s][1] - 1);
      m.bnd_index.push_back(seg_bc[s]);
    }
  } else {
    for (auto& t : vol) {
      if (t.size() != 4) throw std::runtime_error("ReadVolMesh: only tetrahedra supported in 3D");
      for (int l = 0; l < 4; l++) m.els.push_back(t[l] - 1);
    }
    for (size_t s = 0; s < surf.size(); s++) {
      if (surf[s].size() != 3) throw std::runtime_error("ReadVolMesh: only triangular boundary supported");
      for (int l = 0; l < 3; l++) m.bnd.push_back(surf[s][l] - 1);
      m.bnd_index.push_back(surf_bc[s]);
    }
  }
  m.BuildTopology();
  return m;
}

// ---------------------------------------------------------------- topology
// Global edge/face numbers: ascending lexicographic order of the sorted
// vertex tuples.  (The reference takes them from netgen's mesh topology,
// which is not available here; every quantity on the hot path that does
// not depend on the global numbering is independent of this choice.)

void Mesh::BuildTopology() {
  const int ne = NE();
  const int nve = NVE();
  const int nle = NElEdges();
  const uint64_t NV = (uint64_t)nv;
  auto ekey = [NV](int a, int b) -> uint64_t {
    if (a > b) std::swap(a, b);
    return (uint64_t)a * NV + (uint64_t)b;
  };
  const int(*ledges)[2] = (dim == 2) ? kTrigEdges : kTetEdges;
  std::vector<uint64_t> keys;
  keys.reserve((size_t)ne * nle);
  for (int e = 0; e < ne; e++)
    for (int l = 0; l < nle; l++)
      keys.push_back(ekey(els[(size_t)e * nve + ledges[l][0]], els[(size_t)e * nve + ledges[l][1]]));
  std::vector<uint64_t> ukeys = keys;
  std::sort(ukeys.begin(), ukeys.end());
  ukeys.erase(std::unique(ukeys.begin(), ukeys.end()), ukeys.end());
  edges.resize(2 * ukeys.size());
  for (size_t i = 0; i < ukeys.size(); i++) {
    edges[2 * i] = (int)(ukeys[i] / NV);
    edges[2 * i + 1] = (int)(ukeys[i] % NV);
  }
  auto find_edge = [&](uint64_t k) -> int {
    auto it = std::lower_bound(ukeys.begin(), ukeys.end(), k);
    if (it == ukeys.end() || *it != k) throw std::runtime_error("BuildTopology: boundary edge not in mesh");
    return (int)(it - ukeys.begin());
  };
  el_edges.resize(keys.size());
#pragma omp parallel for
  for (long long i = 0; i < (long long)keys.size(); i++) el_edges[i] = find_edge(keys[i]);

  if (dim == 3) {
    if (nv > 2000000) throw std::runtime_error("BuildTopology: too many vertices for face keys");
    auto fkey = [NV](int a, int b, int c) -> uint64_t {
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
      return ((uint64_t)a * NV + (uint64_t)b) * NV + (uint64_t)c;
    };
    std::vector<uint64_t> fk;
    fk.reserve((size_t)ne * 4);
    for (int e = 0; e < ne; e++)
      for (int l = 0; l < 4; l++) {
        const int* v = &els[(size_t)e * 4];
        fk.push_back(fkey(v[kTetFaces[l][0]], v[kTetFaces[l][1]], v[kTetFaces[l][2]]));
      }
    std::vector<uint64_t> sk = fk;
    std::sort(sk.begin(), sk.end());
    std::vector<uint64_t> uk;
    std::vector<int> cnt;
    for (size_t i = 0; i < sk.size(); i++) {
      if (uk.empty() || uk.back() != sk[i]) { uk.push_back(sk[i]); cnt.push_back(1); }
      else cnt.back()++;
    }
    faces.resize(3 * uk.size());
    for (size_t i = 0; i < uk.size(); i++) {
      uint64_t k = uk[i];
      faces[3 * i + 2] = (int)(k % NV); k /= NV;
      faces[3 * i + 1] = (int)(k % NV);
      faces[3 * i] = (int)(k / NV);
    }
    auto find_face = [&](uint64_t k) -> int {
      auto it = std::lower_bound(uk.begin(), uk.end(), k);
      if (it == uk.end() || *it != k) throw std::runtime_error("BuildTopology: boundary face not in mesh");
      return (int)(it - uk.begin());
    };
    el_faces.resize(fk.size());
#pragma omp parallel for
    for (long long i = 0; i < (long long)fk.size(); i++) el_faces[i] = find_face(fk[i]);
    if (bnd.empty()) {
      // boundary = faces owned by exactly one element, in face-number order
      for (size_t i = 0; i < uk.size(); i++)
        if (cnt[i] == 1) {
          bnd.insert(bnd.end(), &faces[3 * i], &faces[3 * i] + 3);
          bnd_index.push_back(1);
        }
    }
    int nb = NBnd();
    bnd_faces.resize(nb);
    bnd_edges.resize((size_t)nb * 3);
    for (int b = 0; b < nb; b++) {
      const int* v = &bnd[(size_t)b * 3];
      bnd_faces[b] = find_face(fkey(v[0], v[1], v[2]));
      bnd_edges[3 * b] = find_edge(ekey(v[0], v[1]));
      bnd_edges[3 * b + 1] = find_edge(ekey(v[1], v[2]));
      bnd_edges[3 * b + 2] = find_edge(ekey(v[2], v[0]));
    }
  } else {
    int nb = NBnd();
    bnd_edges.resize(nb);
    for (int b = 0; b < nb; b++) bnd_edges[b] = find_edge(ekey(bnd[2 * b], bnd[2 * b + 1]));
  }
}

// ------------------------------------------------------------------ space

H1Space::H1Space(const Mesh& m, int p, uint64_t dirichlet_mask) : mesh(&m), order(p) {
  if (p < 1 || p > kMaxOrder) throw std::runtime_error("H1Space: order must be in 1..10");
  const int dim = m.dim, ne = m.NE(), ned = m.NEdges(), nfa = (dim == 3) ? m.NFaces() : 0;
  int per_edge = p - 1;
  int per_face = (dim == 3 && p > 2) ? (p - 1) * (p - 2) / 2 : 0;
  int per_cell = (dim == 2) ? (p > 2 ? (p - 1) * (p - 2) / 2 : 0) : (p > 3 ? (p - 1) * (p - 2) * (p - 3) / 6 : 0);
  int64_t hndof = m.nv;
  first_edge_dof.resize(ned + 1);
  for (int e = 0; e <= ned; e++) first_edge_dof[e] = (int)(hndof + (int64_t)e * per_edge);
  hndof += (int64_t)ned * per_edge;
  first_face_dof.resize(nfa + 1);
  for (int f = 0; f <= nfa; f++) first_face_dof[f] = (int)(hndof + (int64_t)f * per_face);
  hndof += (int64_t)nfa * per_face;
  first_el_dof.resize(ne + 1);
  for (int e = 0; e <= ne; e++) first_el_dof[e] = (int)(hndof + (int64_t)e * per_cell);
  hndof += (int64_t)ne * per_cell;
  if (hndof >= (int64_t)1 << 31) throw std::runtime_error("H1Space: too many dofs for int32");
  ndof = (int)hndof;
  nd_el = (dim == 2) ? TrigNDof(p) : TetNDof(p);

  el_dofs.resize((size_t)ne * nd_el);
  const int nve = dim + 1, nle = m.NElEdges();
#pragma omp parallel for
  for (int e = 0; e < ne; e++) {
    int* d = &el_dofs[(size_t)e * nd_el];
    int ii = 0;
    for (int l = 0; l < nve; l++) d[ii++] = m.els[(size_t)e * nve + l];
    for (int l = 0; l < nle; l++) {
      int g = m.el_edges[(size_t)e * nle + l];
      for (int k = 0; k < per_edge; k++) d[ii++] = first_edge_dof[g] + k;
    }
    if (dim == 3)
      for (int l = 0; l < 4; l++) {
        int g = m.el_faces[(size_t)e * 4 + l];
        for (int k = 0; k < per_face; k++) d[ii++] = first_face_dof[g] + k;
      }
    for (int k = 0; k < per_cell; k++) d[ii++] = first_el_dof[e] + k;
  }

  freedofs.assign(ndof, 1);
  const int nb = m.NBnd();
  for (int b = 0; b < nb; b++) {
    int r = m.bnd_index[b];
    if (r < 1 || r > 64 || !((dirichlet_mask >> (r - 1)) & 1)) continue;
    for (int l = 0; l < dim; l++) freedofs[m.bnd[(size_t)b * dim + l]] = 0;
    int nbe = (dim == 2) ? 1 : 3;
    for (int l = 0; l < nbe; l++) {
      int g = m.bnd_edges[(size_t)b * nbe + l];
      for (int k = 0; k < per_edge; k++) freedofs[first_edge_dof[g] + k] = 0;
    }
    if (dim == 3) {
      int g = m.bnd_faces[b];
      for (int k = 0; k < per_face; k++) freedofs[first_face_dof[g] + k] = 0;
    }
  }

  classnr.resize(ne);
  for (int e = 0; e < ne; e++) {
    const int* v = &m.els[(size_t)e * nve];
    classnr[e] = (dim == 2) ? TrigClassNr(v) : TetClassNr(v);
  }
}

// --------------------------------------------------------------- geometry
// comp/meshaccess.cpp:541-562: TET  F = [v0-v3 | v1-v3 | v2-v3],
//                              TRIG F = [v0-v2 | v1-v2].
ElementGeometry ComputeGeometry(const Mesh& m) {
  ElementGeometry g;
  const int dim = m.dim, ne = m.NE(), nve = dim + 1;
  g.detj.resize(ne);
  g.ginv.resize((size_t)ne * dim * dim);
#pragma omp parallel for
  for (int e = 0; e < ne; e++) {
    const int* v = &m.els[(size_t)e * nve];
    const double* p0 = &m.coords[(size_t)dim * v[dim]];
    double F[3][3] = {{0}};
    for (int j = 0; j < dim; j++) {
      const double* pj = &m.coords[(size_t)dim * v[j]];
      for (int k = 0; k < dim; k++) F[k][j] = pj[k] - p0[k];
    }
    double Fi[3][3] = {{0}};
    double det;
    if (dim == 2) {
      det = F[0][0] * F[1][1] - F[0][1] * F[1][0];
      Fi[0][0] = F[1][1] / det; Fi[0][1] = -F[0][1] / det;
      Fi[1][0] = -F[1][0] / det; Fi[1][1] = F[0][0] / det;
    } else {
      double c00 = F[1][1] * F[2][2] - F[1][2] * F[2][1];
      double c01 = F[1][2] * F[2][0] - F[1][0] * F[2][2];
      double c02 = F[1][0] * F[2][1] - F[1][1] * F[2][0];
      det = F[0][0] * c00 + F[0][1] * c01 + F[0][2] * c02;
      Fi[0][0] = c00 / det;
      Fi[0][1] = (F[0][2] * F[2][1] - F[0][1] * F[2][2]) / det;
      Fi[0][2] = (F[0][1] * F[1][2] - F[0][2] * F[1][1]) / det;
      Fi[1][0] = c01 / det;
      Fi[1][1] = (F[0][0] * F[2][2] - F[0][2] * F[2][0]) / det;
      Fi[1][2] = (F[0][2] * F[1][0] - F[0][0] * F[1][2]) / det;
      Fi[2][0] = c02 / det;
      Fi[2][1] = (F[0][1] * F[2][0] - F[0][0] * F[2][1]) / det;
      Fi[2][2] = (F[0][0] * F[1][1] - F[0][1] * F[1][0]) / det;
    }
    g.detj[e] = std::fabs(det);
    double* G = &g.ginv[(size_t)e * dim * dim];
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++) {
        double s = 0;
        for (int k = 0; k < dim; k++) s += Fi[a][k] * Fi[b][k];
        G[a * dim + b] = s;
      }
  }
  return g;
}

// ---------------------------------------------------------------- classes

std::vector<ElementClass> BuildClasses(const H1Space& fes, const IntegrationRule& ir) {
  const Mesh& m = *fes.mesh;
  const int dim = m.dim, ne = m.NE(), nd = fes.nd_el, nq = ir.Size(), p = fes.order;
  std::map<int, std::vector<int>> groups;
  for (int e = 0; e < ne; e++) groups[fes.classnr[e]].push_back(e);
  std::vector<ElementClass> out;
  for (auto& kv : groups) {
    ElementClass c;
    c.classnr = kv.first;
    c.elements = kv.second;
    const int* vn = &m.els[(size_t)c.elements[0] * (dim + 1)];
    c.bmat.assign((size_t)nq * nd * dim, 0.0);
    c.shape.assign((size_t)nq * nd, 0.0);
    for (int q = 0; q < nq; q++) {
      if (dim == 2) {
        std::vector<AD<2>> s(nd);
        TrigShapes(p, vn, AD<2>::var(ir.pts[2 * q], 0), AD<2>::var(ir.pts[2 * q + 1], 1), s.data());
        for (int i = 0; i < nd; i++) {
          c.shape[(size_t)q * nd + i] = s[i].v;
          for (int d = 0; d < 2; d++) c.bmat[((size_t)q * nd + i) * 2 + d] = s[i].d[d];
        }
      } else {
        std::vector<AD<3>> s(nd);
        TetShapes(p, vn, AD<3>::var(ir.pts[3 * q], 0), AD<3>::var(ir.pts[3 * q + 1], 1),
                  AD<3>::var(ir.pts[3 * q + 2], 2), s.data());
        for (int i = 0; i < nd; i++) {
          c.shape[(size_t)q * nd + i] = s[i].v;
          for (int d = 0; d < 3; d++) c.bmat[((size_t)q * nd + i) * 3 + d] = s[i].d[d];
        }
      }
    }
    out.push_back(std::move(c));
  }
  return out;
}

// ------------------------------------------------------------ CSR pattern

struct DofElementTable {
  std::vector<int64_t> start;  // ndof+1
  std::vector<int> el;         // element
  std::vector<int> loc;        // local index of the dof in that element
};

static DofElementTable BuildDofElementTable(const H1Space& fes) {
  const int ne = fes.mesh->NE(), nd = fes.nd_el, n = fes.ndof;
  DofElementTable t;
  t.start.assign(n + 1, 0);
  for (size_t i = 0; i < fes.el_dofs.size(); i++) t.start[fes.el_dofs[i] + 1]++;
  for (int i = 0; i < n; i++) t.start[i + 1] += t.start[i];
  t.el.resize(fes.el_dofs.size());
  t.loc.resize(fes.el_dofs.size());
  std::vector<int64_t> pos(t.start.begin(), t.start.end() - 1);
  for (int e = 0; e < ne; e++)
    for (int l = 0; l < nd; l++) {
      int d = fes.el_dofs[(size_t)e * nd + l];
      int64_t k = pos[d]++;
      t.el[k] = e;
      t.loc[k] = l;
    }
  return t;
}

CSRPattern BuildPattern(const H1Space& fes) {
  const int n = fes.ndof, nd = fes.nd_el;
  DofElementTable t = BuildDofElementTable(fes);
  CSRPattern pat;
  pat.n = n;
  pat.rowptr.assign(n + 1, 0);
  // pass 1: row lengths
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int r = 0; r < n; r++) {
      buf.clear();
      for (int64_t k = t.start[r]; k < t.start[r + 1]; k++) {
        const int* d = &fes.el_dofs[(size_t)t.el[k] * nd];
        buf.insert(buf.end(), d, d + nd);
      }
      std::sort(buf.begin(), buf.end());
      pat.rowptr[r + 1] = std::unique(buf.begin(), buf.end()) - buf.begin();
    }
  }
  for (int r = 0; r < n; r++) pat.rowptr[r + 1] += pat.rowptr[r];
  pat.cols.resize(pat.rowptr[n]);
  // pass 2: fill
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int r = 0; r < n; r++) {
      buf.clear();
      for (int64_t k = t.start[r]; k < t.start[r + 1]; k++) {
        const int* d = &fes.el_dofs[(size_t)t.el[k] * nd];
        buf.insert(buf.end(), d, d + nd);
      }
      std::sort(buf.begin(), buf.end());
      auto end = std::unique(buf.begin(), buf.end());
      std::copy(buf.begin(), end, pat.cols.begin() + pat.rowptr[r]);
    }
  }
  return pat;
}

// ---------------------------------------------------------------- assembly

void AssembleCSR(const H1Space& fes, const CSRPattern& pat, const IntegrationRule& ir,
                 bool stiffness, const double* kappa, bool mass, const double* rho,
                 double* values) {
  const Mesh& m = *fes.mesh;
  const int dim = m.dim, nd = fes.nd_el, nq = ir.Size(), n = fes.ndof;
  ElementGeometry geo = ComputeGeometry(m);
  std::vector<ElementClass> classes = BuildClasses(fes, ir);
  // Per class: Kref[a][b][i][j] = sum_q w_q dphi_i/dxi_a dphi_j/dxi_b,
  //            Mref[i][j]       = sum_q w_q phi_i phi_j
  std::map<int, int> class_index;
  std::vector<std::vector<double>> kref(classes.size()), mref(classes.size());
  for (size_t c = 0; c < classes.size(); c++) {
    class_index[classes[c].classnr] = (int)c;
    auto& K = kref[c];
    auto& M = mref[c];
    K.assign((size_t)dim * dim * nd * nd, 0.0);
    M.assign((size_t)nd * nd, 0.0);
    const auto& B = classes[c].bmat;
    const auto& S = classes[c].shape;
    for (int q = 0; q < nq; q++) {
      double w = ir.wts[q];
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++)
          for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++)
              K[(((size_t)a * dim + b) * nd + i) * nd + j] +=
                  w * B[((size_t)q * nd + i) * dim + a] * B[((size_t)q * nd + j) * dim + b];
      for (int i = 0; i < nd; i++)
        for (int j = 0; j < nd; j++) M[(size_t)i * nd + j] += w * S[(size_t)q * nd + i] * S[(size_t)q * nd + j];
    }
  }
  std::vector<int> elclass(m.NE());
  for (int e = 0; e < m.NE(); e++) elclass[e] = class_index[fes.classnr[e]];

  DofElementTable t = BuildDofElementTable(fes);
#pragma omp parallel
  {
    std::vector<double> krow(nd);
#pragma omp for schedule(dynamic, 1024)
    for (int r = 0; r < n; r++) {
      const int* rc = &pat.cols[pat.rowptr[r]];
      const int rlen = (int)(pat.rowptr[r + 1] - pat.rowptr[r]);
      double* rv = &values[pat.rowptr[r]];
      for (int j = 0; j < rlen; j++) rv[j] = 0.0;
      for (int64_t k = t.start[r]; k < t.start[r + 1]; k++) {
        int e = t.el[k], i = t.loc[k], c = elclass[e];
        for (int j = 0; j < nd; j++) krow[j] = 0.0;
        if (stiffness) {
          double s = geo.detj[e] * (kappa ? kappa[e] : 1.0);
          const double* G = &geo.ginv[(size_t)e * dim * dim];
          for (int ab = 0; ab < dim * dim; ab++) {
            double gs = s * G[ab];
            const double* K = &kref[c][((size_t)ab * nd + i) * nd];
            for (int j = 0; j < nd; j++) krow[j] += gs * K[j];
          }
        }
        if (mass) {
          double s = geo.detj[e] * (rho ? rho[e] : 1.0);
          const double* M = &mref[c][(size_t)i * nd];
          for (int j = 0; j < nd; j++) krow[j] += s * M[j];
        }
        const int* d = &fes.el_dofs[(size_t)e * nd];
        for (int j = 0; j < nd; j++) {
          const int* pos = std::lower_bound(rc, rc + rlen, d[j]);
          rv[pos - rc] += krow[j];
        }
      }
    }
  }
}

void ExtractDiagonal(const CSRPattern& pat, const double* values, double* diag) {
#pragma omp parallel for
  for (int64_t r = 0; r < pat.n; r++) {
    const int* rc = &pat.cols[pat.rowptr[r]];
    const int rlen = (int)(pat.rowptr[r + 1] - pat.rowptr[r]);
    const int* pos = std::lower_bound(rc, rc + rlen, (int)r);
    diag[r] = (pos != rc + rlen && *pos == r) ? values[pat.rowptr[r] + (pos - rc)] : 0.0;
  }
}

std::vector<int> ElementColoring(const Mesh& m) {
  const int ne = m.NE(), nve = m.NVE();
  std::vector<int> col(ne, -1);
  std::vector<uint32_t> mask(m.nv);
  int base = 0, found = 0;
  while (found < ne) {
    std::fill(mask.begin(), mask.end(), 0u);
    for (int e = 0; e < ne; e++) {
      if (col[e] >= 0) continue;
      const int* v = &m.els[(size_t)e * nve];
      uint32_t check = 0;
      for (int k = 0; k < nve; k++) check |= mask[v[k]];
      if (check == 0xFFFFFFFFu) continue;
      int c = 0;
      while (check & (1u << c)) c++;
      col[e] = base + c;
      found++;
      for (int k = 0; k < nve; k++) mask[v[k]] |= 1u << c;
    }
    base += 32;
  }
  return col;
}

}  // namespace ngsamd
