// Geometric multigrid for the P1 (vertex) part of an H1 problem on a
// structured mesh: V-cycles over the mesh hierarchy n, n/2, n/4, ...
//
// Reference: MultigridPreconditioner::Mult / MGM (multigrid/mgpre.cpp:205-318):
// pre-smoothing, residual, restriction, coarse-grid correction on the next
// coarser level, prolongation, post-smoothing, and an exact solve on the
// coarsest level; prolongation LinearProlongation (multigrid/prolongation.cpp:
// 77-144): a vertex created by refinement takes the mean of its two parent
// vertices, restriction is the transpose.
//
// Built here for the structured meshes of MakeStructured2DMesh /
// MakeStructured3DMesh, whose vertex grids nest: the mesh of n cells per
// direction is the regular refinement of the mesh of n/2 (every fine vertex
// with an odd grid coordinate is the midpoint of a coarse edge: in 3D the
// Kuhn cubes' edges run along every 0/1 direction; in 2D the squares' diagonal
// runs from (x+1, y) to (x, y+1)).  The reference assembles the form on every
// level of its mesh hierarchy; here the coarse matrices are the Galerkin
// products P^T A P (for a coefficient that is constant on coarse elements the
// same matrices; for a per-element coefficient of the fine mesh the P1 form
// with the coefficient averaged over each coarse element).  Smoothing is
// damped Jacobi (the device analogue of the reference's Gauss-Seidel
// smoother: every row at once), nu sweeps before and after the coarse
// correction, so the V-cycle is a symmetric preconditioner; the coarsest
// level (at most kMgDenseMax free dofs) is solved with its dense inverse.
//
// Layout: per level the free vertices in grid order, the level matrix, P (fine
// free x coarse free) and R = P^T as plain CSR; one thread per row (P1 rows
// hold at most 27 entries).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "csr.hpp"

namespace ngsamd {

constexpr int kMgBlock = 256;
constexpr int kMgDenseMax = 1024;  // the coarsest level is inverted densely up to this size

struct HostCsr {
  int64_t n = 0, m = 0;  // rows, columns
  std::vector<int64_t> rp{0};
  std::vector<int> c;
  std::vector<double> v;
};

// C = A B (sparse accumulator; columns sorted in each row)
static HostCsr SpGemm(const HostCsr& A, const HostCsr& B) {
  HostCsr C;
  C.n = A.n;
  C.m = B.m;
  C.rp.assign(A.n + 1, 0);
  std::vector<double> acc(B.m, 0.0);
  std::vector<int> mark(B.m, -1), cols;
  for (int64_t i = 0; i < A.n; i++) {
    cols.clear();
    for (int64_t ka = A.rp[i]; ka < A.rp[i + 1]; ka++) {
      const int j = A.c[ka];
      const double a = A.v[ka];
      for (int64_t kb = B.rp[j]; kb < B.rp[j + 1]; kb++) {
        const int q = B.c[kb];
        if (mark[q] != (int)i) {
          mark[q] = (int)i;
          acc[q] = 0.0;
          cols.push_back(q);
        }
        acc[q] += a * B.v[kb];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int q : cols) {
      C.c.push_back(q);
      C.v.push_back(acc[q]);
    }
    C.rp[i + 1] = (int64_t)C.c.size();
  }
  return C;
}

static HostCsr Transpose(const HostCsr& A) {
  HostCsr T;
  T.n = A.m;
  T.m = A.n;
  T.rp.assign(A.m + 1, 0);
  for (int q : A.c) T.rp[q + 1]++;
  for (int64_t j = 0; j < A.m; j++) T.rp[j + 1] += T.rp[j];
  T.c.resize(A.c.size());
  T.v.resize(A.v.size());
  std::vector<int64_t> pos(T.rp.begin(), T.rp.end() - 1);
  for (int64_t i = 0; i < A.n; i++)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; k++) {
      const int64_t d = pos[A.c[k]]++;
      T.c[d] = (int)i;
      T.v[d] = A.v[k];
    }
  return T;
}

// vertex grid of one level: dims[0..dim) cells per direction
struct MgGrid {
  int dim;
  int d[3];
  int64_t NV() const { return dim == 2 ? int64_t(d[0] + 1) * (d[1] + 1) : int64_t(d[0] + 1) * (d[1] + 1) * (d[2] + 1); }
  // 3D: vertex (i, j, k) of MakeStructured3DMesh(nx, ny, nz) = (i (ny+1) + j) (nz+1) + k
  // 2D: vertex (row i, column j) of MakeStructured2DMesh(nx, ny) = i (nx+1) + j
  int64_t Num(int a, int b, int c) const {
    if (dim == 2) return int64_t(a) * (d[0] + 1) + b;
    return (int64_t(a) * (d[1] + 1) + b) * (d[2] + 1) + c;
  }
  // grid extent of coordinate slot s (2D: slot 0 = row (ny), slot 1 = column (nx))
  int Ext(int s) const { return dim == 2 ? (s == 0 ? d[1] : d[0]) : d[s]; }
};

// Parents of fine vertex (a, b, c) in the coarse grid (coarse coordinates) and
// their weights; returns the count (1 or 2).
static int Parents(int dim, int a, int b, int c, int pa[2][3], double w[2]) {
  const int A = a >> 1, B = b >> 1, C = c >> 1;
  const int ea = a & 1, eb = b & 1, ec = dim == 3 ? (c & 1) : 0;
  if (!ea && !eb && !ec) {
    pa[0][0] = A; pa[0][1] = B; pa[0][2] = C;
    w[0] = 1.0;
    return 1;
  }
  w[0] = w[1] = 0.5;
  if (dim == 3 || !(ea && eb)) {
    // the coarse edge from (A, B, C) along the 0/1 direction (ea, eb, ec)
    pa[0][0] = A; pa[0][1] = B; pa[0][2] = C;
    pa[1][0] = A + ea; pa[1][1] = B + eb; pa[1][2] = C + ec;
  } else {
    // 2D, both odd: the square's diagonal from (row A, column B+1) to (row A+1, column B)
    pa[0][0] = A; pa[0][1] = B + 1; pa[0][2] = 0;
    pa[1][0] = A + 1; pa[1][1] = B; pa[1][2] = 0;
  }
  return 2;
}

struct MgLevelDev {
  int64_t n = 0;
  DevBuf<int> arp, ac, prp, pc, rrp, rc;
  DevBuf<double> av, invd, pv, rv;
  DevBuf<double> r, xa, xb;
};

struct MultiGridOp : ngsamd_op_s {
  int nlevels = 0, nu = 2;
  double omega = 2.0 / 3.0;
  std::vector<int64_t> level_n;
  std::vector<MgLevelDev> lev;
  DevBuf<int> fdof;         // finest level: free vertex -> vertex (dof) number
  DevBuf<double> coarse_inv;  // dense inverse of the coarsest level matrix (row-major)
  MultiGridOp(int dim, const int* dims, int64_t n, const int64_t* rp, const int* cols, const double* vals,
              const uint8_t* freedofs, int nu, double omega, int max_levels);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y (+)= s M x on the first height dofs; every kernel returns at once when *stop != 0
  void Apply(double s, const double* x, double* y, bool accumulate, const int* stop, hipStream_t st);
  void VCycle(int l, double* yf, double s, const int* stop, hipStream_t st, double* x_out_slot[]);
};

// ------------------------------------------------------------------ kernels

// xout = (first ? 0 : xin) + omega D^-1 (r - (first ? 0 : A xin))
__global__ __launch_bounds__(kMgBlock) void MgJacobiKernel(const int* __restrict__ rp, const int* __restrict__ c,
                                                           const double* __restrict__ v,
                                                           const double* __restrict__ invd, double omega,
                                                           const double* __restrict__ r,
                                                           const double* __restrict__ xin, double* __restrict__ xout,
                                                           int n, int first, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int i = blockIdx.x * kMgBlock + threadIdx.x;
  if (i >= n) return;
  if (first) {
    xout[i] = omega * invd[i] * r[i];
    return;
  }
  double ax = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) ax += v[k] * xin[c[k]];
  xout[i] = xin[i] + omega * invd[i] * (r[i] - ax);
}

// Pre-smoothing sweeps 1 and 2 from x = 0 in one pass (two == 0: sweep 1
// only): x1 = omega D^-1 r, xout = x1 + omega D^-1 (r - A x1), x1 of the
// neighbours formed on the fly (the same products as two separate sweeps)
__global__ __launch_bounds__(kMgBlock) void MgPreKernel(const int* __restrict__ rp, const int* __restrict__ c,
                                                        const double* __restrict__ v,
                                                        const double* __restrict__ invd, double omega,
                                                        const double* __restrict__ r, double* __restrict__ xout,
                                                        int n, int two, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int i = blockIdx.x * kMgBlock + threadIdx.x;
  if (i >= n) return;
  const double x1 = omega * invd[i] * r[i];
  if (!two) {
    xout[i] = x1;
    return;
  }
  double ax = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) {
    const int j = c[k];
    ax += v[k] * (omega * invd[j] * r[j]);
  }
  xout[i] = x1 + omega * invd[i] * (r[i] - ax);
}

// t = r - A x
__global__ __launch_bounds__(kMgBlock) void MgResidualKernel(const int* __restrict__ rp, const int* __restrict__ c,
                                                             const double* __restrict__ v,
                                                             const double* __restrict__ r,
                                                             const double* __restrict__ x, double* __restrict__ t,
                                                             int n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int i = blockIdx.x * kMgBlock + threadIdx.x;
  if (i >= n) return;
  double ax = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) ax += v[k] * x[c[k]];
  t[i] = r[i] - ax;
}

// y = (acc ? y : 0) + M x   (restriction and prolongation)
__global__ __launch_bounds__(kMgBlock) void MgSpmvKernel(const int* __restrict__ rp, const int* __restrict__ c,
                                                         const double* __restrict__ v, const double* __restrict__ x,
                                                         double* __restrict__ y, int n, int acc,
                                                         const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int i = blockIdx.x * kMgBlock + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[c[k]];
  y[i] = acc ? y[i] + s : s;
}

// The last post-smoothing sweep of the finest level added to the caller's
// vector: y[fdof[i]] += s (x + omega D^-1 (r - A x))_i
__global__ __launch_bounds__(kMgBlock) void MgSweepScatterKernel(const int* __restrict__ rp,
                                                                 const int* __restrict__ c,
                                                                 const double* __restrict__ v,
                                                                 const double* __restrict__ invd, double omega,
                                                                 const double* __restrict__ r,
                                                                 const double* __restrict__ x,
                                                                 const int* __restrict__ fdof,
                                                                 double* __restrict__ y, double s, int n,
                                                                 const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int i = blockIdx.x * kMgBlock + threadIdx.x;
  if (i >= n) return;
  double ax = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) ax += v[k] * x[c[k]];
  y[fdof[i]] += s * (x[i] + omega * invd[i] * (r[i] - ax));
}

// y = inv x, inv dense n x n row-major; one wavefront per row
__global__ __launch_bounds__(kMgBlock) void MgDenseKernel(const double* __restrict__ inv,
                                                          const double* __restrict__ x, double* __restrict__ y, int n,
                                                          const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int row = blockIdx.x * (kMgBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) s += inv[(int64_t)row * n + j] * x[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[row] = s;
}

// r[k] = x[fdof[k]]
__global__ void MgGatherKernel(const int* __restrict__ fdof, const double* __restrict__ x, double* __restrict__ r,
                               int n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) r[k] = x[fdof[k]];
}

// y[fdof[k]] += s z[k]
__global__ void MgScatterKernel(const int* __restrict__ fdof, const double* __restrict__ z, double* __restrict__ y,
                                int nf, double s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) y[fdof[k]] += s * z[k];
}

// y = 0
__global__ void MgZeroKernel(double* __restrict__ y, int n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = 0.0;
}

static int MgBlocks(int64_t n) { return (int)std::max<int64_t>(1, (n + kMgBlock - 1) / kMgBlock); }

// ------------------------------------------------------------------ set-up

// The level hierarchy on the host: free-vertex numbering of the finest level,
// level matrices (finest first) and prolongations P_l (level l free x level
// l+1 free).
struct MgHierarchy {
  std::vector<int> fdof;
  std::vector<HostCsr> A, P;
};

static MgHierarchy BuildMgHierarchy(int dim, const int* dims, int64_t n, const int64_t* rp, const int* cols,
                                    const double* vals, const uint8_t* freedofs, int max_levels) {
  NGSAMD_REQUIRE(dim == 2 || dim == 3, "ngsamd_multigrid_create: dim must be 2 or 3");
  MgGrid g{dim, {dims[0], dims[1], dim == 3 ? dims[2] : 0}};
  for (int s = 0; s < dim; s++) NGSAMD_REQUIRE(g.d[s] >= 1, "ngsamd_multigrid_create: grid sizes must be >= 1");
  NGSAMD_REQUIRE(n == g.NV(), "ngsamd_multigrid_create: matrix size is not the vertex count of the grid");
  NGSAMD_REQUIRE(max_levels >= 1 && max_levels <= 32, "ngsamd_multigrid_create: max_levels must be 1..32");
  NGSAMD_REQUIRE(rp[0] == 0, "ngsamd_multigrid_create: row pointer must start at 0");
  for (int64_t i = 0; i < n; i++) NGSAMD_REQUIRE(rp[i + 1] >= rp[i], "ngsamd_multigrid_create: row pointer not monotone");
  for (int64_t k = 0; k < rp[n]; k++)
    NGSAMD_REQUIRE(cols[k] >= 0 && cols[k] < n, "ngsamd_multigrid_create: column number out of range");
  MgHierarchy H;
  // finest level: the free block
  std::vector<uint8_t> free(n, 1);
  if (freedofs) for (int64_t i = 0; i < n; i++) free[i] = freedofs[i] ? 1 : 0;
  std::vector<int> idx(n, -1), fd;
  for (int64_t i = 0; i < n; i++)
    if (free[i]) {
      idx[i] = (int)fd.size();
      fd.push_back((int)i);
    }
  HostCsr A;
  A.n = A.m = (int64_t)fd.size();
  A.rp.assign(A.n + 1, 0);
  for (int64_t r = 0; r < A.n; r++) {
    const int64_t i = fd[r];
    std::vector<std::pair<int, double>> row;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
      if (idx[cols[k]] >= 0) row.push_back({idx[cols[k]], vals[k]});
    std::sort(row.begin(), row.end());
    for (auto& e : row) {
      if (!A.c.empty() && A.rp[r] < (int64_t)A.c.size() && A.c.back() == e.first) {
        A.v.back() += e.second;  // duplicate column entries are summed
      } else {
        A.c.push_back(e.first);
        A.v.push_back(e.second);
      }
    }
    A.rp[r + 1] = (int64_t)A.c.size();
  }

  // hierarchy: coarsen while every grid size is even, the level has more than
  // kMgDenseMax free dofs and max_levels allows
  std::vector<HostCsr> As{A}, Ps;
  std::vector<std::vector<uint8_t>> frees{free};
  std::vector<MgGrid> grids{g};
  while ((int)As.size() < max_levels && As.back().n > kMgDenseMax) {
    const MgGrid& gf = grids.back();
    bool even = true;
    for (int s = 0; s < dim; s++) even = even && gf.d[s] % 2 == 0;
    if (!even) break;
    MgGrid gc{dim, {gf.d[0] / 2, gf.d[1] / 2, gf.d[2] / 2}};
    const std::vector<uint8_t>& ff = frees.back();
    // coarse vertex free iff the fine vertex at the same place is free
    std::vector<uint8_t> fc(gc.NV(), 0);
    std::vector<int> cidx(gc.NV(), -1);
    int nc = 0;
    const int e0 = gc.Ext(0), e1 = gc.Ext(1), e2 = dim == 3 ? gc.Ext(2) : 0;
    for (int a = 0; a <= e0; a++)
      for (int b = 0; b <= e1; b++)
        for (int c = 0; c <= e2; c++) {
          const int64_t vc = gc.Num(a, b, c);
          fc[vc] = ff[gf.Num(2 * a, 2 * b, 2 * c)];
        }
    for (int64_t vc = 0; vc < gc.NV(); vc++)
      if (fc[vc]) cidx[vc] = nc++;
    if (nc == 0) break;
    // prolongation: rows = fine free vertices in order, columns = coarse free
    std::vector<int> fidx(gf.NV(), -1);
    {
      int k = 0;
      for (int64_t vf = 0; vf < gf.NV(); vf++)
        if (ff[vf]) fidx[vf] = k++;
    }
    HostCsr P;
    P.n = As.back().n;
    P.m = nc;
    P.rp.assign(P.n + 1, 0);
    const int f0 = gf.Ext(0), f1 = gf.Ext(1), f2 = dim == 3 ? gf.Ext(2) : 0;
    for (int a = 0; a <= f0; a++)
      for (int b = 0; b <= f1; b++)
        for (int c = 0; c <= f2; c++) {
          const int64_t vf = gf.Num(a, b, c);
          if (!ff[vf]) continue;
          int pa[2][3];
          double w[2];
          const int np = Parents(dim, a, b, c, pa, w);
          std::vector<std::pair<int, double>> row;
          for (int q = 0; q < np; q++) {
            const int ci = cidx[gc.Num(pa[q][0], pa[q][1], pa[q][2])];
            if (ci >= 0) row.push_back({ci, w[q]});
          }
          std::sort(row.begin(), row.end());
          const int r = fidx[vf];
          for (auto& e : row) {
            P.c.push_back(e.first);
            P.v.push_back(e.second);
          }
          P.rp[r + 1] = (int64_t)row.size();
        }
    for (int64_t r = 0; r < P.n; r++) P.rp[r + 1] += P.rp[r];
    // fine free vertices are visited in vertex order = row order, so the
    // column and value arrays are already in row order
    HostCsr Ac = SpGemm(Transpose(P), SpGemm(As.back(), P));
    Ps.push_back(std::move(P));
    As.push_back(std::move(Ac));
    frees.push_back(std::move(fc));
    grids.push_back(gc);
  }
  NGSAMD_REQUIRE(As.back().n <= kMgDenseMax,
                 "ngsamd_multigrid_create: the coarsest level has more than 1024 free dofs (grid sizes not "
                 "divisible by enough powers of 2, or max_levels too small)");
  H.fdof = std::move(fd);
  H.A = std::move(As);
  H.P = std::move(Ps);
  return H;
}

MultiGridOp::MultiGridOp(int dim, const int* dims, int64_t n, const int64_t* rp, const int* cols, const double* vals,
                         const uint8_t* freedofs, int nu_, double omega_, int max_levels)
    : ngsamd_op_s(OpKind::MultiGrid), nu(nu_), omega(omega_) {
  NGSAMD_REQUIRE(nu >= 1 && nu <= 16, "ngsamd_multigrid_create: smoothing steps must be 1..16");
  NGSAMD_REQUIRE(omega > 0.0 && omega < 2.0, "ngsamd_multigrid_create: omega must lie in (0, 2)");
  MgHierarchy H = BuildMgHierarchy(dim, dims, n, rp, cols, vals, freedofs, max_levels);
  height = width = n;
  const std::vector<HostCsr>& As = H.A;
  const std::vector<HostCsr>& Ps = H.P;
  nlevels = (int)As.size();
  fdof.alloc(std::max<size_t>(H.fdof.size(), 1));
  fdof.upload(H.fdof.data(), H.fdof.size());

  // device copies
  lev.resize(nlevels);
  for (int l = 0; l < nlevels; l++) {
    const HostCsr& M = As[l];
    MgLevelDev& L = lev[l];
    L.n = M.n;
    level_n.push_back(M.n);
    std::vector<int> rp32(M.rp.begin(), M.rp.end());
    std::vector<double> invd(M.n, 0.0);
    for (int64_t i = 0; i < M.n; i++)
      for (int64_t k = M.rp[i]; k < M.rp[i + 1]; k++)
        if (M.c[k] == (int)i) invd[i] = M.v[k] != 0.0 ? 1.0 / M.v[k] : 0.0;
    L.arp.alloc(rp32.size());
    L.arp.upload(rp32.data(), rp32.size());
    L.ac.alloc(std::max<size_t>(M.c.size(), 1));
    L.ac.upload(M.c.data(), M.c.size());
    L.av.alloc(std::max<size_t>(M.v.size(), 1));
    L.av.upload(M.v.data(), M.v.size());
    L.invd.alloc(std::max<int64_t>(M.n, 1));
    L.invd.upload(invd.data(), invd.size());
    const size_t nn = std::max<int64_t>(M.n, 1);
    L.r.alloc(nn);
    L.xa.alloc(nn);
    L.xb.alloc(nn);
    if (l + 1 < nlevels) {
      const HostCsr& P = Ps[l];
      const HostCsr R = Transpose(P);
      std::vector<int> prp(P.rp.begin(), P.rp.end()), rrp(R.rp.begin(), R.rp.end());
      L.prp.alloc(prp.size());
      L.prp.upload(prp.data(), prp.size());
      L.pc.alloc(std::max<size_t>(P.c.size(), 1));
      L.pc.upload(P.c.data(), P.c.size());
      L.pv.alloc(std::max<size_t>(P.v.size(), 1));
      L.pv.upload(P.v.data(), P.v.size());
      L.rrp.alloc(rrp.size());
      L.rrp.upload(rrp.data(), rrp.size());
      L.rc.alloc(std::max<size_t>(R.c.size(), 1));
      L.rc.upload(R.c.data(), R.c.size());
      L.rv.alloc(std::max<size_t>(R.v.size(), 1));
      L.rv.upload(R.v.data(), R.v.size());
    }
  }
  // dense inverse of the coarsest matrix (Cholesky, then L^-T L^-1)
  {
    const HostCsr& M = As.back();
    const int m = (int)M.n;
    std::vector<double> a((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++)
      for (int64_t k = M.rp[i]; k < M.rp[i + 1]; k++) a[(size_t)i * m + M.c[k]] = M.v[k];
    for (int j = 0; j < m; j++) {
      double d = a[(size_t)j * m + j];
      for (int k = 0; k < j; k++) d -= a[(size_t)j * m + k] * a[(size_t)j * m + k];
      NGSAMD_REQUIRE(d > 0.0, "ngsamd_multigrid_create: coarsest matrix is not positive definite");
      d = std::sqrt(d);
      a[(size_t)j * m + j] = d;
      for (int i = j + 1; i < m; i++) {
        double s = a[(size_t)i * m + j];
        for (int k = 0; k < j; k++) s -= a[(size_t)i * m + k] * a[(size_t)j * m + k];
        a[(size_t)i * m + j] = s / d;
      }
    }
    // inverse column by column: solve L L^T x = e_c
    std::vector<double> inv((size_t)m * m, 0.0), z(m);
    for (int col = 0; col < m; col++) {
      for (int i = 0; i < m; i++) {
        double s = i == col ? 1.0 : 0.0;
        for (int k = 0; k < i; k++) s -= a[(size_t)i * m + k] * z[k];
        z[i] = s / a[(size_t)i * m + i];
      }
      for (int i = m - 1; i >= 0; i--) {
        double s = z[i];
        for (int k = i + 1; k < m; k++) s -= a[(size_t)k * m + i] * z[k];
        z[i] = s / a[(size_t)i * m + i];
      }
      for (int i = 0; i < m; i++) inv[(size_t)i * m + col] = z[i];
    }
    coarse_inv.alloc(std::max<size_t>(inv.size(), 1));
    coarse_inv.upload(inv.data(), inv.size());
  }
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
}

// ------------------------------------------------------------------ apply

// V-cycle on level l for the right-hand side lev[l].r; levels l > 0 return
// the buffer that holds their result in x_out_slot[l], level 0 adds s times
// its result to yf[fdof].  Two passes are fused (the same arithmetic in the
// same order): pre-smoothing sweeps 1-2, and at level 0 the last sweep with
// the scatter.  (Fusing the residual into the restriction and the
// prolongation into the first post-smoothing sweep was measured slower: 240
// against 89 us per V-cycle at 32^3, the neighbour-of-neighbour gathers of one
// thread per row are serial.)
void MultiGridOp::VCycle(int l, double* yf, double s, const int* stop, hipStream_t st, double* x_out_slot[]) {
  MgLevelDev& L = lev[l];
  const int n = (int)L.n;
  if (l == nlevels - 1) {  // l > 0: a single level is handled by Apply
    MgDenseKernel<<<(n + kMgBlock / 64 - 1) / (kMgBlock / 64), kMgBlock, 0, st>>>(coarse_inv.p, L.r.p, L.xa.p, n,
                                                                                   stop);
    NGSAMD_CHECK_LAUNCH();
    x_out_slot[l] = L.xa.p;
    return;
  }
  double* x = L.xa.p;
  double* y = L.xb.p;
  const int two = nu >= 2 ? 1 : 0;
  MgPreKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.arp.p, L.ac.p, L.av.p, L.invd.p, omega, L.r.p, x, n, two, stop);
  NGSAMD_CHECK_LAUNCH();
  for (int k = two ? 2 : 1; k < nu; k++) {
    MgJacobiKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.arp.p, L.ac.p, L.av.p, L.invd.p, omega, L.r.p, x, y, n, 0,
                                                    stop);
    NGSAMD_CHECK_LAUNCH();
    std::swap(x, y);
  }
  // residual, restriction, coarse correction, prolongation (a level above the
  // coarsest always has a coarser one with free vertices)
  MgResidualKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.arp.p, L.ac.p, L.av.p, L.r.p, x, y, n, stop);
  NGSAMD_CHECK_LAUNCH();
  MgLevelDev& C = lev[l + 1];
  const int nc = (int)C.n;
  MgSpmvKernel<<<MgBlocks(nc), kMgBlock, 0, st>>>(L.rrp.p, L.rc.p, L.rv.p, y, C.r.p, nc, 0, stop);
  NGSAMD_CHECK_LAUNCH();
  VCycle(l + 1, nullptr, 0.0, stop, st, x_out_slot);
  MgSpmvKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.prp.p, L.pc.p, L.pv.p, x_out_slot[l + 1], x, n, 1, stop);
  NGSAMD_CHECK_LAUNCH();
  // post-smoothing; at level 0 the last sweep adds into the caller's vector
  for (int k = 0; k < nu; k++) {
    if (l == 0 && k == nu - 1) {
      MgSweepScatterKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.arp.p, L.ac.p, L.av.p, L.invd.p, omega, L.r.p, x,
                                                             fdof.p, yf, s, n, stop);
      NGSAMD_CHECK_LAUNCH();
      return;
    }
    MgJacobiKernel<<<MgBlocks(n), kMgBlock, 0, st>>>(L.arp.p, L.ac.p, L.av.p, L.invd.p, omega, L.r.p, x, y, n, 0,
                                                    stop);
    NGSAMD_CHECK_LAUNCH();
    std::swap(x, y);
  }
  x_out_slot[l] = x;
}

void MultiGridOp::Apply(double s, const double* x, double* y, bool accumulate, const int* stop, hipStream_t st) {
  if (height == 0) return;
  if (!accumulate) {
    MgZeroKernel<<<GridFor(height, 256, 4096), 256, 0, st>>>(y, (int)height, stop);
    NGSAMD_CHECK_LAUNCH();
  }
  const int nf = (int)lev[0].n;
  if (nf == 0) return;
  if (nlevels == 1) {
    // one level: the dense inverse of the free block
    MgGatherKernel<<<GridFor(nf, 256, 4096), 256, 0, st>>>(fdof.p, x, lev[0].r.p, nf, stop);
    NGSAMD_CHECK_LAUNCH();
    MgDenseKernel<<<(nf + kMgBlock / 64 - 1) / (kMgBlock / 64), kMgBlock, 0, st>>>(coarse_inv.p, lev[0].r.p,
                                                                                    lev[0].xa.p, nf, stop);
    NGSAMD_CHECK_LAUNCH();
    MgScatterKernel<<<GridFor(nf, 256, 4096), 256, 0, st>>>(fdof.p, lev[0].xa.p, y, nf, s, stop);
    NGSAMD_CHECK_LAUNCH();
    return;
  }
  MgGatherKernel<<<GridFor(nf, 256, 4096), 256, 0, st>>>(fdof.p, x, lev[0].r.p, nf, stop);
  NGSAMD_CHECK_LAUNCH();
  std::vector<double*> slot(nlevels, nullptr);
  VCycle(0, y, s, stop, st, slot.data());
}

void MultiGridOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  Apply(s, x, y, true, nullptr, st);
}
void MultiGridOp::Mult(const double* x, double* y, hipStream_t st) { Apply(1.0, x, y, false, nullptr, st); }

void MultiGridApply(ngsamd_op_t op, double s, const double* x, double* y, bool accumulate, const int* stop,
                    hipStream_t st) {
  static_cast<MultiGridOp*>(op)->Apply(s, x, y, accumulate, stop, st);
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_multigrid_create(ngsamd_op_t* op, int dim, const int* dims, int64_t n, const int64_t* rowptr,
                                       const int* cols, const double* vals, const uint8_t* freedofs,
                                       int smoothing_steps, double omega, int max_levels) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && dims && rowptr, "ngsamd_multigrid_create: null argument");
    NGSAMD_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "ngsamd_multigrid_create: size out of range");
    NGSAMD_REQUIRE(rowptr[n] == 0 || (cols && vals), "ngsamd_multigrid_create: null argument");
    *op = nullptr;
    *op = new MultiGridOp(dim, dims, n, rowptr, cols, vals, freedofs, smoothing_steps, omega, max_levels);
  });
}

extern "C" int ngsamd_multigrid_info(ngsamd_op_t op, int* h_nlevels, int64_t* h_level_dofs) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::MultiGrid && h_nlevels,
                   "ngsamd_multigrid_info: not a multigrid operator");
    const MultiGridOp* M = static_cast<const MultiGridOp*>(op);
    *h_nlevels = M->nlevels;
    if (h_level_dofs)
      for (int l = 0; l < M->nlevels; l++) h_level_dofs[l] = M->level_n[l];
  });
}
