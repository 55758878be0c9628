// Sparse Cholesky factorisation of the free block of a symmetric positive
// definite CSR matrix, with the triangular solves on the device, and the
// additive two-level preconditioner built on it.
//
// Reference: SparseCholesky (linalg/sparsecholesky.cpp; MultAdd = permute,
// solve with L, D and L^T, permute back) behind
// a.mat.Inverse(freedofs, inverse="sparsecholesky"), and its device solve
// DevSparseCholesky (ngscuda/dev_sparsecholesky.cpp:403-454), which walks a
// task graph of supernodal blocks with each warp spinning on the dependency
// counters of its task.  Here the factor is organised so that the device
// needs no inter-workgroup waiting at all:
//
// - ordering: nested dissection of the free graph (recursive bisection by
//   breadth-first level sets from a pseudo-peripheral vertex: the middle
//   level is the separator); every subtree of the dissection tree is a
//   contiguous range of the new numbering, leaves (<= kLeaf vertices) and
//   separators are the supernodes;
// - factorisation (host, FP64, multifrontal): the front of supernode s has
//   its own vertices I_s (m) and its boundary B_s (b: the vertices of
//   ancestor separators adjacent to the subtree of s); the dense partial
//   Cholesky of the front gives L11 (m x m), L21 (b x m) and the update
//   matrix passed to the parent; L11 is inverted explicitly;
// - solve L L^T z = x level by level: supernodes of one height (forward)
//   or one depth (backward) are independent, so each level is a few
//   kernels with one wavefront per row, and the stream order is the only
//   synchronisation:
//     forward, height h:   t_i = x_i - sum_c L21_c[row of i] . y_c  (the
//                          descendants c whose boundary holds i)
//                          y_s = L11_s^-1 t_s
//     backward, depth d:   r_s = y_s - L21_s^T x_(B_s)
//                          x_s = L11_s^-T r_s
//   Every sum runs in a fixed order: two solves give the same bits.
//
// Two-level preconditioner: y = S x + E C E^T x with S a Jacobi or
// block-Jacobi preconditioner and C an operator on the first C->height dofs
// (E their embedding): for a hierarchical H1 space those are the vertex
// dofs, i.e. the P1 space, and C the sparse Cholesky inverse of its block.
// It is the composite smoother + E @ a_lo.mat.Inverse(...) @ E.T of the
// reference's sum and product matrices (linalg/basematrix.hpp:413,592).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kScBlock = 256;  // 4 wavefronts, one row each
constexpr int kLeaf = 64;      // largest leaf of the dissection

// ---------------------------------------------------------------- kernels

__device__ __forceinline__ double ScWaveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// b[k] = x[fdof[k]] (the free dofs in dissection order)
__global__ void ScGatherKernel(const int* __restrict__ fdof, const double* __restrict__ x, double* __restrict__ b,
                               int nf, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nf; k += gridDim.x * blockDim.x) b[k] = x[fdof[k]];
}

// y[i] = (acc ? y[i] : 0) + s * z[pos[i]] (pos -1: not free, 0)
__global__ void ScScatterKernel(const int* __restrict__ pos, const double* __restrict__ z, double* __restrict__ y,
                                int n, double s, int acc, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int k = pos[i];
    const double v = k >= 0 ? s * z[k] : 0.0;
    y[i] = acc ? y[i] + v : v;
  }
}

// forward, level rows: t_i = x_i - sum over the row's segments (descendant
// c, row p of its L21): L21_c[p][:] . y[first_c ...]
__global__ __launch_bounds__(kScBlock) void ScFwdRowsKernel(
    const int* __restrict__ rows, int nrows, const int* __restrict__ seg_ptr, const int* __restrict__ seg_node,
    const int* __restrict__ seg_row, const ScNode* __restrict__ nodes, const double* __restrict__ lval,
    const double* __restrict__ xb, const double* __restrict__ y, double* __restrict__ t, const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * (kScBlock / 64) + (threadIdx.x >> 6); w < nrows; w += gridDim.x * (kScBlock / 64)) {
    const int i = rows[w];
    double acc = 0.0;
    for (int sg = seg_ptr[i]; sg < seg_ptr[i + 1]; sg++) {
      const ScNode c = nodes[seg_node[sg]];
      const double* lr = lval + c.l21 + (int64_t)seg_row[sg] * c.m;
      const double* yc = y + c.first;
      for (int k = lane; k < c.m; k += 64) acc += lr[k] * yc[k];
    }
    acc = ScWaveSum(acc);
    if (lane == 0) t[i] = xb[i] - acc;
  }
}

// forward, level rows: y_i = sum_{j <= i} L11_s^-1[i][j] t_j
__global__ __launch_bounds__(kScBlock) void ScFwdDiagKernel(const int* __restrict__ rows, int nrows,
                                                            const int* __restrict__ sn_of_row,
                                                            const ScNode* __restrict__ nodes,
                                                            const double* __restrict__ lval,
                                                            const double* __restrict__ t, double* __restrict__ y,
                                                            const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * (kScBlock / 64) + (threadIdx.x >> 6); w < nrows; w += gridDim.x * (kScBlock / 64)) {
    const int i = rows[w];
    const ScNode s = nodes[sn_of_row[i]];
    const int li = i - s.first;
    const double* ir = lval + s.inv + (int64_t)li * s.m;
    const double* ts = t + s.first;
    double acc = 0.0;
    for (int j = lane; j <= li; j += 64) acc += ir[j] * ts[j];
    acc = ScWaveSum(acc);
    if (lane == 0) y[i] = acc;
  }
}

// backward, level rows: r_j = y_j - sum_p L21_s[p][j] x[bnd_s[p]]
__global__ __launch_bounds__(kScBlock) void ScBwdRowsKernel(const int* __restrict__ rows, int nrows,
                                                            const int* __restrict__ sn_of_row,
                                                            const ScNode* __restrict__ nodes,
                                                            const double* __restrict__ lval,
                                                            const int* __restrict__ bnd, const double* __restrict__ y,
                                                            const double* __restrict__ x, double* __restrict__ r,
                                                            const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * (kScBlock / 64) + (threadIdx.x >> 6); w < nrows; w += gridDim.x * (kScBlock / 64)) {
    const int j = rows[w];
    const ScNode s = nodes[sn_of_row[j]];
    const double* lr = lval + s.l21t + (int64_t)(j - s.first) * s.b;
    const int* bs = bnd + s.bnd;
    double acc = 0.0;
    for (int p = lane; p < s.b; p += 64) acc += lr[p] * x[bs[p]];
    acc = ScWaveSum(acc);
    if (lane == 0) r[j] = y[j] - acc;
  }
}

// backward, level rows: x_j = sum_{i >= j} L11_s^-1[i][j] r_i (the
// transposed inverse stored row-major: row j holds column j)
__global__ __launch_bounds__(kScBlock) void ScBwdDiagKernel(const int* __restrict__ rows, int nrows,
                                                            const int* __restrict__ sn_of_row,
                                                            const ScNode* __restrict__ nodes,
                                                            const double* __restrict__ lval,
                                                            const int64_t* __restrict__ invt,
                                                            const double* __restrict__ r, double* __restrict__ x,
                                                            const int* __restrict__ stop) {
  if (stop && *stop) return;
  const int lane = threadIdx.x & 63;
  for (int w = blockIdx.x * (kScBlock / 64) + (threadIdx.x >> 6); w < nrows; w += gridDim.x * (kScBlock / 64)) {
    const int j = rows[w];
    const int sn = sn_of_row[j];
    const ScNode s = nodes[sn];
    const int lj = j - s.first;
    const double* cr = lval + invt[sn] + (int64_t)lj * s.m;
    const double* rs = r + s.first;
    double acc = 0.0;
    for (int i = lj + lane; i < s.m; i += 64) acc += cr[i] * rs[i];
    acc = ScWaveSum(acc);
    if (lane == 0) x[j] = acc;
  }
}

// ------------------------------------------------------------------- host

namespace {

struct Supernode {
  int first = 0, m = 0;
  int parent = -1;
  std::vector<int> children;
  std::vector<int> bnd;  // boundary rows (new numbering, ascending)
  int height = 0, depth = 0;
};

// Nested dissection of the graph (adj over local vertices 0..nv-1):
// fills order (position -> vertex) and the supernode tree (children before
// parents).
class Dissection {
 public:
  Dissection(const std::vector<int>& ap, const std::vector<int>& aj) : ap_(ap), aj_(aj) {
    const int nv = (int)ap.size() - 1;
    mark_.assign(nv, -1);
    level_.assign(nv, -1);
    part_.assign(nv, 0);
  }
  std::vector<int> order;
  std::vector<Supernode> nodes;

  // root supernode of the dissection of vertex set V (all with part_ == tag)
  int Dissect(std::vector<int> V, int depth_guard = 0) {
    if ((int)V.size() <= kLeaf || depth_guard > 60) return Leaf(V);
    const int tag = ++tag_;
    for (int v : V) part_[v] = tag;
    // pseudo-peripheral start: two sweeps from V[0]
    int start = V[0];
    for (int sweep = 0; sweep < 2; sweep++) {
      const std::vector<int> q = Bfs(start, tag);
      start = q.back();
    }
    const std::vector<int> q = Bfs(start, tag);
    if ((int)q.size() < (int)V.size()) {
      // disconnected: the reached component and the rest, no separator
      std::vector<int> A(q), B;
      for (int v : V)
        if (mark_[v] != tag2_) B.push_back(v);
      const int ia = Dissect(A, depth_guard + 1), ib = Dissect(B, depth_guard + 1);
      return Join({}, {ia, ib});
    }
    const int maxlev = level_[q.back()];
    std::vector<int> cnt(maxlev + 2, 0);
    for (int v : q) cnt[level_[v]]++;
    int acc = 0, mid = 0;
    for (int l = 0; l <= maxlev; l++) {
      if (2 * (acc + cnt[l]) >= (int)V.size()) {
        mid = l;
        break;
      }
      acc += cnt[l];
    }
    // the separator level strictly inside; a graph of diameter < 2 is one dense block
    if (maxlev < 2) {
      NGSAMD_REQUIRE(V.size() <= 16384, "ngsamd_sparsecholesky_create: dense block too large");
      return Leaf(V);
    }
    mid = std::min(std::max(mid, 1), maxlev - 1);
    std::vector<int> A, B, S;
    for (int v : q) (level_[v] < mid ? A : level_[v] > mid ? B : S).push_back(v);
    const int ia = Dissect(A, depth_guard + 1), ib = Dissect(B, depth_guard + 1);
    return Join(S, {ia, ib});
  }

 private:
  const std::vector<int>& ap_;
  const std::vector<int>& aj_;
  std::vector<int> mark_, level_, part_;
  int tag_ = 0, tag2_ = 0;

  std::vector<int> Bfs(int s, int tag) {
    const int t2 = ++tag2_;
    std::vector<int> q{s};
    mark_[s] = t2;
    level_[s] = 0;
    for (size_t h = 0; h < q.size(); h++) {
      const int v = q[h];
      for (int k = ap_[v]; k < ap_[v + 1]; k++) {
        const int w = aj_[k];
        if (part_[w] == tag && mark_[w] != t2) {
          mark_[w] = t2;
          level_[w] = level_[v] + 1;
          q.push_back(w);
        }
      }
    }
    return q;
  }
  int Leaf(const std::vector<int>& V) { return Join(V, {}); }
  int Join(const std::vector<int>& S, std::vector<int> ch) {
    Supernode sn;
    sn.first = (int)order.size();
    sn.m = (int)S.size();
    order.insert(order.end(), S.begin(), S.end());
    sn.children = std::move(ch);
    const int id = (int)nodes.size();
    for (int c : sn.children) nodes[c].parent = id;
    nodes.push_back(std::move(sn));
    return id;
  }
};

}  // namespace

ScFactor FactorSparseCholesky(int64_t n, const int64_t* rp, const int* cols, const double* vals,
                               const uint8_t* freedofs) {
  NGSAMD_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "ngsamd_sparsecholesky_create: bad size");
  NGSAMD_REQUIRE(rp[0] == 0, "ngsamd_sparsecholesky_create: row pointer must start at 0");
  ScFactor R;
  // free dofs and the free graph (symmetrised pattern, no diagonal)
  std::vector<int> loc(n, -1), glob;
  for (int64_t i = 0; i < n; i++)
    if (!freedofs || freedofs[i]) {
      loc[i] = (int)glob.size();
      glob.push_back((int)i);
    }
  const int nf = (int)glob.size();
  R.nf = nf;
  std::vector<int> ap(nf + 1, 0), aj;
  {
    std::vector<std::vector<int>> adj(nf);
    for (int a = 0; a < nf; a++) {
      const int i = glob[a];
      for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
        NGSAMD_REQUIRE(cols[k] >= 0 && cols[k] < n, "ngsamd_sparsecholesky_create: column out of range");
        const int bcol = loc[cols[k]];
        if (bcol < 0 || bcol == a) continue;
        adj[a].push_back(bcol);
        adj[bcol].push_back(a);
      }
    }
    for (int a = 0; a < nf; a++) {
      std::sort(adj[a].begin(), adj[a].end());
      adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
      ap[a + 1] = ap[a] + (int)adj[a].size();
    }
    aj.reserve(ap[nf]);
    for (int a = 0; a < nf; a++) aj.insert(aj.end(), adj[a].begin(), adj[a].end());
  }
  Dissection nd(ap, aj);
  if (nf) {
    std::vector<int> all(nf);
    std::iota(all.begin(), all.end(), 0);
    nd.Dissect(all);
  }
  std::vector<Supernode>& sn = nd.nodes;
  const int ns = (int)sn.size();
  R.nsuper = ns;
  std::vector<int> newidx(nf);
  for (int k = 0; k < nf; k++) newidx[nd.order[k]] = k;
  std::vector<int> sn_row(std::max(nf, 1), 0);
  for (int s = 0; s < ns; s++)
    for (int k = 0; k < sn[s].m; k++) sn_row[sn[s].first + k] = s;

  // symbolic: boundary of every supernode (children first)
  for (int s = 0; s < ns; s++) {
    const int last = sn[s].first + sn[s].m - 1;
    std::vector<int> b;
    for (int k = 0; k < sn[s].m; k++) {
      const int v = nd.order[sn[s].first + k];
      for (int e = ap[v]; e < ap[v + 1]; e++)
        if (newidx[aj[e]] > last) b.push_back(newidx[aj[e]]);
    }
    for (int c : sn[s].children)
      for (int r : sn[c].bnd)
        if (r > last) b.push_back(r);
    std::sort(b.begin(), b.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    sn[s].bnd = std::move(b);
  }
  // heights (forward levels) and depths (backward levels)
  for (int s = 0; s < ns; s++)
    for (int c : sn[s].children) sn[s].height = std::max(sn[s].height, sn[c].height + 1);
  for (int s = ns - 1; s >= 0; s--)
    if (sn[s].parent >= 0) sn[s].depth = sn[sn[s].parent].depth + 1;

  // numeric: multifrontal partial Cholesky of each front
  std::vector<ScNode>& tab = R.tab;
  tab.assign(std::max(ns, 1), ScNode{});
  std::vector<double>& lval = R.lval;  // L11^-1, L21, L21^T of every supernode
  std::vector<int64_t>& invt = R.invt;
  invt.assign(std::max(ns, 1), 0);
  std::vector<int>& bnd_all = R.bnd;
  std::vector<std::vector<double>> update(ns);  // pending update matrices (b x b, lower, row-major full)
  std::vector<int> fpos(std::max(nf, 1), -1);   // row -> position in the current front
  // matrix entry (a, c) of the free block in local numbering
  auto entries = [&](int a, auto&& f) {
    const int i = glob[a];
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) {
      const int c = loc[cols[k]];
      if (c >= 0) f(c, vals[k]);
    }
  };
  for (int s = 0; s < ns; s++) {
    const int m = sn[s].m, b = (int)sn[s].bnd.size(), F = m + b;
    for (int k = 0; k < m; k++) fpos[sn[s].first + k] = k;
    for (int p = 0; p < b; p++) fpos[sn[s].bnd[p]] = m + p;
    std::vector<double> front((size_t)F * F, 0.0);
    // A: rows of I_s, columns in I_s or B_s (the lower triangle of the front)
    for (int k = 0; k < m; k++) {
      const int row = sn[s].first + k;
      entries(nd.order[row], [&](int c, double v) {
        const int nc = newidx[c];
        if (nc < row) return;
        const int pc = fpos[nc];
        NGSAMD_REQUIRE(pc >= 0, "ngsamd_sparsecholesky_create: internal error (pattern outside the front)");
        front[(size_t)pc * F + k] += v;  // row pc >= k: lower triangle
      });
    }
    for (int c : sn[s].children) {
      const std::vector<int>& cb = sn[c].bnd;
      const int bc = (int)cb.size();
      const std::vector<double>& U = update[c];
      for (int p = 0; p < bc; p++) {
        const int rp_ = fpos[cb[p]];
        for (int q = 0; q <= p; q++) {
          const int cq = fpos[cb[q]];
          // cb ascending and fronts ordered: rp_ >= cq
          front[(size_t)rp_ * F + cq] += U[(size_t)p * bc + q];
        }
      }
      std::vector<double>().swap(update[c]);
    }
    // partial Cholesky of the first m columns (right-looking, lower triangle)
    std::vector<double> colk(F);
    for (int k = 0; k < m; k++) {
      double d = front[(size_t)k * F + k];
      NGSAMD_REQUIRE(d > 0.0, "ngsamd_sparsecholesky_create: matrix not positive definite");
      d = std::sqrt(d);
      front[(size_t)k * F + k] = d;
      for (int r = k + 1; r < F; r++) {
        front[(size_t)r * F + k] /= d;
        colk[r] = front[(size_t)r * F + k];
      }
      for (int r = k + 1; r < F; r++) {
        const double f = colk[r];
        if (f == 0.0) continue;
        double* fr = &front[(size_t)r * F];
        for (int c2 = k + 1; c2 <= r; c2++) fr[c2] -= f * colk[c2];
      }
    }
    // update matrix for the parent
    if (b) {
      std::vector<double>& U = update[s];
      U.assign((size_t)b * b, 0.0);
      for (int p = 0; p < b; p++)
        for (int q = 0; q <= p; q++) U[(size_t)p * b + q] = front[(size_t)(m + p) * F + m + q];
    }
    // L11^-1 (lower triangular, by columns of the identity)
    std::vector<double> inv((size_t)m * m, 0.0);
    for (int j = 0; j < m; j++) {
      inv[(size_t)j * m + j] = 1.0 / front[(size_t)j * F + j];
      for (int i = j + 1; i < m; i++) {
        double acc = 0.0;
        for (int k = j; k < i; k++) acc += front[(size_t)i * F + k] * inv[(size_t)k * m + j];
        inv[(size_t)i * m + j] = -acc / front[(size_t)i * F + i];
      }
    }
    ScNode& t = tab[s];
    t.first = sn[s].first;
    t.m = m;
    t.b = b;
    t.pad = 0;
    t.inv = (int64_t)lval.size();
    lval.insert(lval.end(), inv.begin(), inv.end());
    invt[s] = (int64_t)lval.size();
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) lval.push_back(inv[(size_t)i * m + j]);
    t.l21 = (int64_t)lval.size();
    for (int p = 0; p < b; p++)
      for (int k = 0; k < m; k++) lval.push_back(front[(size_t)(m + p) * F + k]);
    t.l21t = (int64_t)lval.size();
    for (int k = 0; k < m; k++)
      for (int p = 0; p < b; p++) lval.push_back(front[(size_t)(m + p) * F + k]);
    t.bnd = (int64_t)bnd_all.size();
    bnd_all.insert(bnd_all.end(), sn[s].bnd.begin(), sn[s].bnd.end());
    for (int k = 0; k < m; k++) fpos[sn[s].first + k] = -1;
    for (int p = 0; p < b; p++) fpos[sn[s].bnd[p]] = -1;
  }
  R.nnz_factor = 0;
  for (int s = 0; s < ns; s++) R.nnz_factor += (int64_t)sn[s].m * (sn[s].m + 1) / 2 + (int64_t)sn[s].m * tab[s].b;

  // forward segments: row i <- (descendant c, row p of L21_c), c ascending
  std::vector<int>& sp = R.seg_ptr;
  sp.assign(nf + 1, 0);
  for (int s = 0; s < ns; s++)
    for (int r : sn[s].bnd) sp[r + 1]++;
  for (int i = 0; i < nf; i++) sp[i + 1] += sp[i];
  std::vector<int>& sgn = R.seg_node;
  std::vector<int>& sgr = R.seg_row;
  sgn.assign(std::max(sp[nf], 1), 0);
  sgr.assign(std::max(sp[nf], 1), 0);
  {
    std::vector<int> fill(sp.begin(), sp.end() - 1);
    for (int s = 0; s < ns; s++)
      for (int p = 0; p < (int)sn[s].bnd.size(); p++) {
        const int r = sn[s].bnd[p];
        sgn[fill[r]] = s;
        sgr[fill[r]] = p;
        fill[r]++;
      }
  }
  // level row lists
  int maxh = 0, maxd = 0;
  for (int s = 0; s < ns; s++) maxh = std::max(maxh, sn[s].height), maxd = std::max(maxd, sn[s].depth);
  const int nlevels = ns ? maxh + 1 : 0;
  R.nlevels = nlevels;
  std::vector<int>& rows_f = R.rows_f;
  std::vector<int>& rows_b = R.rows_b;
  std::vector<int>& fwd_level = R.fwd_level;
  std::vector<int>& bwd_level = R.bwd_level;
  fwd_level.assign(nlevels + 1, 0);
  bwd_level.assign(ns ? maxd + 2 : 1, 0);
  for (int h = 0; h < nlevels; h++) {
    for (int s = 0; s < ns; s++)
      if (sn[s].height == h)
        for (int k = 0; k < sn[s].m; k++) rows_f.push_back(sn[s].first + k);
    fwd_level[h + 1] = (int)rows_f.size();
  }
  for (int d = 0; ns && d <= maxd; d++) {
    for (int s = 0; s < ns; s++)
      if (sn[s].depth == d)
        for (int k = 0; k < sn[s].m; k++) rows_b.push_back(sn[s].first + k);
    bwd_level[d + 1] = (int)rows_b.size();
  }
  R.fdof.assign(std::max(nf, 1), 0);
  R.pos.assign(std::max<int64_t>(n, 1), -1);
  for (int k = 0; k < nf; k++) {
    R.fdof[k] = glob[nd.order[k]];
    R.pos[R.fdof[k]] = k;
  }
  R.sn_of_row = std::move(sn_row);
  return R;
}

SparseCholeskyOp::SparseCholeskyOp(int64_t n, const int64_t* rp, const int* cols, const double* vals,
                                   const uint8_t* freedofs)
    : ngsamd_op_s(OpKind::SparseCholesky) {
  height = width = n;
  ScFactor R = FactorSparseCholesky(n, rp, cols, vals, freedofs);
  nf = R.nf;
  nsuper = R.nsuper;
  nlevels = R.nlevels;
  nnz_factor = R.nnz_factor;
  fwd_level = R.fwd_level;
  bwd_level = R.bwd_level;
  auto up = [&](auto& buf, const auto& v) {
    buf.alloc(std::max<size_t>(v.size(), 1));
    if (!v.empty()) buf.upload(v.data(), v.size());
  };
  up(d_nodes, R.tab);
  up(d_lval, R.lval);
  up(d_invt, R.invt);
  up(d_bnd, R.bnd);
  up(d_seg_ptr, R.seg_ptr);
  up(d_seg_node, R.seg_node);
  up(d_seg_row, R.seg_row);
  up(d_rows_f, R.rows_f);
  up(d_rows_b, R.rows_b);
  up(d_sn_of_row, R.sn_of_row);
  up(d_fdof, R.fdof);
  up(d_pos, R.pos);
  d_b.alloc(std::max(nf, 1));
  d_t.alloc(std::max(nf, 1));
  d_y.alloc(std::max(nf, 1));
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
}

static int ScGrid(int rows) { return GridFor(rows, kScBlock / 64, 8192); }

void SparseCholeskyOp::Apply(double s, const double* x, double* y, bool accumulate, const int* stop,
                             hipStream_t st) {
  if (height == 0) return;
  if (nf) {
    ScGatherKernel<<<GridFor(nf, 256, 4096), 256, 0, st>>>(d_fdof.p, x, d_b.p, nf, stop);
    NGSAMD_CHECK_LAUNCH();
  }
  // forward: L y = b, leaves first
  for (int h = 0; h < nlevels; h++) {
    const int r0 = fwd_level[h], nr = fwd_level[h + 1] - r0;
    if (nr == 0) continue;
    ScFwdRowsKernel<<<ScGrid(nr), kScBlock, 0, st>>>(d_rows_f.p + r0, nr, d_seg_ptr.p, d_seg_node.p, d_seg_row.p,
                                                     d_nodes.p, d_lval.p, d_b.p, d_y.p, d_t.p, stop);
    NGSAMD_CHECK_LAUNCH();
    ScFwdDiagKernel<<<ScGrid(nr), kScBlock, 0, st>>>(d_rows_f.p + r0, nr, d_sn_of_row.p, d_nodes.p, d_lval.p, d_t.p,
                                                     d_y.p, stop);
    NGSAMD_CHECK_LAUNCH();
  }
  // backward: L^T z = y, root first; z into d_b (the forward's input is spent)
  for (size_t d = 0; d + 1 < bwd_level.size(); d++) {
    const int r0 = bwd_level[d], nr = bwd_level[d + 1] - r0;
    if (nr == 0) continue;
    ScBwdRowsKernel<<<ScGrid(nr), kScBlock, 0, st>>>(d_rows_b.p + r0, nr, d_sn_of_row.p, d_nodes.p, d_lval.p,
                                                     d_bnd.p, d_y.p, d_b.p, d_t.p, stop);
    NGSAMD_CHECK_LAUNCH();
    ScBwdDiagKernel<<<ScGrid(nr), kScBlock, 0, st>>>(d_rows_b.p + r0, nr, d_sn_of_row.p, d_nodes.p, d_lval.p,
                                                     d_invt.p, d_t.p, d_b.p, stop);
    NGSAMD_CHECK_LAUNCH();
  }
  ScScatterKernel<<<GridFor(height, 256, 4096), 256, 0, st>>>(d_pos.p, d_b.p, y, (int)height, s,
                                                              accumulate ? 1 : 0, stop);
  NGSAMD_CHECK_LAUNCH();
}

void SparseCholeskyOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  Apply(s, x, y, true, nullptr, st);
}
void SparseCholeskyOp::Mult(const double* x, double* y, hipStream_t st) { Apply(1.0, x, y, false, nullptr, st); }

// ------------------------------------------------------- two-level operator

__global__ void ScJacobiStopKernel(const double* __restrict__ inv, const double* __restrict__ x,
                                   double* __restrict__ y, int64_t n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = inv[i] * x[i];
}

TwoLevelOp::TwoLevelOp(ngsamd_op_t sm, ngsamd_op_t co) : ngsamd_op_s(OpKind::TwoLevel), smoother(sm), coarse(co) {
  NGSAMD_REQUIRE(sm && co, "ngsamd_twolevel_create: null operator");
  NGSAMD_REQUIRE(sm->kind == OpKind::Jacobi || sm->kind == OpKind::BlockJacobi,
                 "ngsamd_twolevel_create: the smoother must be a Jacobi or block-Jacobi operator");
  NGSAMD_REQUIRE(co->kind == OpKind::SparseCholesky || co->kind == OpKind::MultiGrid,
                 "ngsamd_twolevel_create: the coarse operator must be a sparse Cholesky inverse or a multigrid operator");
  NGSAMD_REQUIRE(sm->height == sm->width && co->height == co->width && co->height <= sm->height,
                 "ngsamd_twolevel_create: the coarse space must be the first dofs of the smoother's");
  height = width = sm->height;
}

void TwoLevelOp::MultStop(const double* x, double* y, const int* stop, hipStream_t st) {
  if (height == 0) return;
  if (smoother->kind == OpKind::Jacobi) {
    ScJacobiStopKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(static_cast<JacobiOp*>(smoother)->invdiag.p, x, y,
                                                                     height, stop);
    NGSAMD_CHECK_LAUNCH();
  } else {
    static_cast<BlockJacobiOp*>(smoother)->MultStop(x, y, stop, st);
  }
  CoarseApply(1.0, x, y, stop, st);
}

void TwoLevelOp::CoarseApply(double s, const double* x, double* y, const int* stop, hipStream_t st) {
  if (coarse->kind == OpKind::MultiGrid) MultiGridApply(coarse, s, x, y, true, stop, st);
  else static_cast<SparseCholeskyOp*>(coarse)->Apply(s, x, y, true, stop, st);
}

void TwoLevelOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  smoother->MultAdd(s, x, y, st);
  CoarseApply(s, x, y, nullptr, st);
}

void TwoLevelOp::Mult(const double* x, double* y, hipStream_t st) {
  smoother->Mult(x, y, st);
  CoarseApply(1.0, x, y, nullptr, st);
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_sparsecholesky_create(ngsamd_op_t* op, int64_t n, const int64_t* rowptr, const int* cols,
                                            const double* vals, const uint8_t* freedofs) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && rowptr, "ngsamd_sparsecholesky_create: null argument");
    NGSAMD_REQUIRE(n >= 0, "ngsamd_sparsecholesky_create: negative size");
    NGSAMD_REQUIRE(rowptr[n] == 0 || (cols && vals), "ngsamd_sparsecholesky_create: null argument");
    *op = nullptr;
    *op = new SparseCholeskyOp(n, rowptr, cols, vals, freedofs);
  });
}

extern "C" int ngsamd_sparsecholesky_info(ngsamd_op_t op, int64_t* h_info) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::SparseCholesky && h_info,
                   "ngsamd_sparsecholesky_info: not a sparse Cholesky operator");
    const SparseCholeskyOp* A = static_cast<const SparseCholeskyOp*>(op);
    h_info[0] = A->nf;
    h_info[1] = A->nsuper;
    h_info[2] = A->nlevels;
    h_info[3] = A->nnz_factor;
  });
}

extern "C" int ngsamd_twolevel_create(ngsamd_op_t* op, ngsamd_op_t smoother, ngsamd_op_t coarse) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op, "ngsamd_twolevel_create: null argument");
    *op = nullptr;
    *op = new TwoLevelOp(smoother, coarse);
  });
}
