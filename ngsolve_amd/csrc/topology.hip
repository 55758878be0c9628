// Mesh topology and H1 dof tables built on the device.
//
// The reference numbers edges and faces through netgen's mesh topology and
// lays out the H1 dofs in H1HighOrderFESpace::Update
// (comp/h1hofespace.cpp:823-950): the vertex dofs, then p-1 dofs per edge,
// (p-1)(p-2)/2 per face and the cell dofs, each in global edge / face /
// element order; an element's dof table is its vertices, its edges' dofs,
// its faces' dofs and its own dofs in local order (GetDofNrs, :1389-1413);
// a dof on a boundary facet of a Dirichlet region is not free
// (FESpace::Update, comp/fespace.cpp:765-810).  The edge and face numbering
// rule here is the host setup's: ascending lexicographic order of the sorted
// vertex tuples.  On the device that order needs no global comparison sort:
// the element edges (faces) are bucketed by their smallest vertex with a
// counting sort, each bucket is sorted and made unique by one thread, and
//   number of the edge (a < b) = #distinct edges whose first vertex is < a
//                                + rank of b among the distinct partners of a,
// which is the position of (a, b) in the sorted list of all edges.  Bucket
// fills use atomic counters, so entries land in a bucket in any order; the
// per-bucket sort makes every result independent of that order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "common.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kTopoBlock = 256;

// local edge / face vertices (fem element conventions, as the host tables)
__constant__ int cTrigEdges[3][2] = {{2, 0}, {1, 2}, {0, 1}};
__constant__ int cTetEdges[6][2] = {{3, 0}, {3, 1}, {3, 2}, {0, 1}, {0, 2}, {1, 2}};
__constant__ int cTetFaces[4][3] = {{3, 1, 2}, {3, 2, 0}, {3, 0, 1}, {0, 2, 1}};

// Key of an edge or face in the bucket of its smallest vertex a: the other
// vertices in ascending order, packed so that integer order is
// lexicographic order.
struct Tuple {
  int a;
  uint64_t key;
};

__device__ __forceinline__ Tuple EdgeTuple(int u, int v) {
  return u < v ? Tuple{u, (uint64_t)(uint32_t)v} : Tuple{v, (uint64_t)(uint32_t)u};
}

__device__ __forceinline__ Tuple FaceTuple(int u, int v, int w) {
  int t;
  if (u > v) { t = u; u = v; v = t; }
  if (v > w) { t = v; v = w; w = t; }
  if (u > v) { t = u; u = v; v = t; }
  return Tuple{u, ((uint64_t)(uint32_t)v << 32) | (uint32_t)w};
}

// Tuple of entry i = (element, local edge or face) of the element table.
template <int DIM, bool FACE>
__device__ __forceinline__ Tuple ElementTuple(const int* __restrict__ els, int64_t i) {
  constexpr int nve = DIM + 1;
  constexpr int nloc = FACE ? 4 : (DIM == 2 ? 3 : 6);
  const int64_t e = i / nloc;
  const int l = (int)(i - e * nloc);
  const int* v = els + e * nve;
  if (FACE) return FaceTuple(v[cTetFaces[l][0]], v[cTetFaces[l][1]], v[cTetFaces[l][2]]);
  if (DIM == 2) return EdgeTuple(v[cTrigEdges[l][0]], v[cTrigEdges[l][1]]);
  return EdgeTuple(v[cTetEdges[l][0]], v[cTetEdges[l][1]]);
}

template <int DIM, bool FACE>
__global__ __launch_bounds__(kTopoBlock) void BucketCountKernel(const int* __restrict__ els, int64_t n,
                                                                int* __restrict__ cnt) {
  for (int64_t i = blockIdx.x * (int64_t)kTopoBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTopoBlock)
    atomicAdd(cnt + ElementTuple<DIM, FACE>(els, i).a, 1);
}

template <int DIM, bool FACE>
__global__ __launch_bounds__(kTopoBlock) void BucketFillKernel(const int* __restrict__ els, int64_t n,
                                                               int* __restrict__ fill, uint64_t* __restrict__ keys) {
  for (int64_t i = blockIdx.x * (int64_t)kTopoBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTopoBlock) {
    const Tuple t = ElementTuple<DIM, FACE>(els, i);
    keys[atomicAdd(fill + t.a, 1)] = t.key;
  }
}

// Ascending heap sort of k[0..m): O(m log m) for the rare large bucket (a
// vertex shared by very many elements).
__device__ void HeapSort(uint64_t* k, int m) {
  auto sift = [k](int i, int n) {
    const uint64_t x = k[i];
    for (int c = 2 * i + 1; c < n; c = 2 * i + 1) {
      if (c + 1 < n && k[c + 1] > k[c]) c++;
      if (k[c] <= x) break;
      k[i] = k[c];
      i = c;
    }
    k[i] = x;
  };
  for (int i = m / 2 - 1; i >= 0; i--) sift(i, m);
  for (int n = m - 1; n > 0; n--) {
    const uint64_t t = k[0];
    k[0] = k[n];
    k[n] = t;
    sift(0, n);
  }
}

// One thread per vertex: sort its bucket (insertion sort for the usual few
// dozen edges or faces around a vertex, heap sort above 64), keep the
// distinct keys at the front with their multiplicities, count them.
__global__ __launch_bounds__(kTopoBlock) void BucketUniqueKernel(const int* __restrict__ off, int nv,
                                                                 uint64_t* __restrict__ keys,
                                                                 int* __restrict__ mult, int* __restrict__ ucnt) {
  for (int a = blockIdx.x * kTopoBlock + threadIdx.x; a < nv; a += gridDim.x * kTopoBlock) {
    uint64_t* k = keys + off[a];
    const int m = off[a + 1] - off[a];
    if (m > 64) {
      HeapSort(k, m);
    } else {
      for (int i = 1; i < m; i++) {
        const uint64_t x = k[i];
        int j = i - 1;
        while (j >= 0 && k[j] > x) {
          k[j + 1] = k[j];
          j--;
        }
        k[j + 1] = x;
      }
    }
    int* mu = mult + off[a];
    int u = 0;
    for (int i = 0; i < m; i++) {
      if (u == 0 || k[i] != k[u - 1]) {
        k[u] = k[i];
        mu[u] = 1;
        u++;
      } else {
        mu[u - 1]++;
      }
    }
    ucnt[a] = u;
  }
}

// out[0..n] = exclusive prefix sums of in[0..n-1] (out[n] = total), one
// workgroup: each thread sums a contiguous chunk, thread 0 scans the chunk
// sums, each thread writes its chunk.
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void ExclusiveScanKernel(const int* __restrict__ in, int* __restrict__ out,
                                                                  int n) {
  __shared__ int part[kScanBlock];
  const int t = threadIdx.x;
  const int64_t chunk = ((int64_t)n + kScanBlock - 1) / kScanBlock;
  const int64_t lo = std::min<int64_t>(t * chunk, n), hi = std::min<int64_t>(lo + chunk, n);
  int s = 0;
  for (int64_t i = lo; i < hi; i++) s += in[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int k = 0; k < kScanBlock; k++) {
      const int v = part[k];
      part[k] = run;
      run += v;
    }
    out[n] = run;
  }
  __syncthreads();
  int run = part[t];
  for (int64_t i = lo; i < hi; i++) {
    const int v = in[i];
    out[i] = run;
    run += v;
  }
}

// edges[2 j] = (a, b), faces[3 j] = (a, b, c) for the distinct tuples of
// bucket a, numbered from first[a].
template <bool FACE>
__global__ __launch_bounds__(kTopoBlock) void WriteTuplesKernel(const int* __restrict__ off,
                                                                const int* __restrict__ first, int nv,
                                                                const uint64_t* __restrict__ keys,
                                                                int* __restrict__ verts) {
  for (int a = blockIdx.x * kTopoBlock + threadIdx.x; a < nv; a += gridDim.x * kTopoBlock) {
    const uint64_t* k = keys + off[a];
    for (int j = 0; j < first[a + 1] - first[a]; j++) {
      const int g = first[a] + j;
      if (FACE) {
        verts[3 * (int64_t)g] = a;
        verts[3 * (int64_t)g + 1] = (int)(k[j] >> 32);
        verts[3 * (int64_t)g + 2] = (int)(uint32_t)k[j];
      } else {
        verts[2 * (int64_t)g] = a;
        verts[2 * (int64_t)g + 1] = (int)k[j];
      }
    }
  }
}

// Global number of tuple t (binary search among the distinct keys of its
// bucket), -1 if absent.
__device__ __forceinline__ int FindTuple(const Tuple& t, const int* __restrict__ off, const int* __restrict__ first,
                                         const uint64_t* __restrict__ keys) {
  const uint64_t* k = keys + off[t.a];
  int lo = 0, hi = first[t.a + 1] - first[t.a];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (k[mid] < t.key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < first[t.a + 1] - first[t.a] && k[lo] == t.key) ? first[t.a] + lo : -1;
}

template <int DIM, bool FACE>
__global__ __launch_bounds__(kTopoBlock) void ElementNumbersKernel(const int* __restrict__ els, int64_t n,
                                                                   const int* __restrict__ off,
                                                                   const int* __restrict__ first,
                                                                   const uint64_t* __restrict__ keys,
                                                                   int* __restrict__ out) {
  for (int64_t i = blockIdx.x * (int64_t)kTopoBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTopoBlock)
    out[i] = FindTuple(ElementTuple<DIM, FACE>(els, i), off, first, keys);
}

// Faces held by exactly one element, per bucket: count, then write them in
// face-number order as boundary facets (sorted vertex triple, region 1).
__global__ __launch_bounds__(kTopoBlock) void OpenFacesCountKernel(const int* __restrict__ off,
                                                                   const int* __restrict__ first, int nv,
                                                                   const int* __restrict__ mult,
                                                                   int* __restrict__ cnt) {
  for (int a = blockIdx.x * kTopoBlock + threadIdx.x; a < nv; a += gridDim.x * kTopoBlock) {
    int c = 0;
    for (int j = 0; j < first[a + 1] - first[a]; j++) c += mult[off[a] + j] == 1;
    cnt[a] = c;
  }
}

__global__ __launch_bounds__(kTopoBlock) void OpenFacesWriteKernel(const int* __restrict__ off,
                                                                   const int* __restrict__ first, int nv,
                                                                   const int* __restrict__ mult,
                                                                   const uint64_t* __restrict__ keys,
                                                                   const int* __restrict__ boff,
                                                                   int* __restrict__ bnd, int* __restrict__ bnd_index) {
  for (int a = blockIdx.x * kTopoBlock + threadIdx.x; a < nv; a += gridDim.x * kTopoBlock) {
    int b = boff[a];
    for (int j = 0; j < first[a + 1] - first[a]; j++) {
      if (mult[off[a] + j] != 1) continue;
      const uint64_t k = keys[off[a] + j];
      bnd[3 * (int64_t)b] = a;
      bnd[3 * (int64_t)b + 1] = (int)(k >> 32);
      bnd[3 * (int64_t)b + 2] = (int)(uint32_t)k;
      bnd_index[b] = 1;
      b++;
    }
  }
}

// Edges (and in 3D the face) of each boundary facet; *missing = 1 if one is
// not an edge (face) of the mesh.
template <int DIM>
__global__ __launch_bounds__(kTopoBlock) void BoundaryNumbersKernel(
    const int* __restrict__ bnd, int nb, const int* __restrict__ eoff, const int* __restrict__ efirst,
    const uint64_t* __restrict__ ekeys, const int* __restrict__ foff, const int* __restrict__ ffirst,
    const uint64_t* __restrict__ fkeys, int* __restrict__ bnd_edges, int* __restrict__ bnd_faces,
    int* __restrict__ missing) {
  for (int b = blockIdx.x * kTopoBlock + threadIdx.x; b < nb; b += gridDim.x * kTopoBlock) {
    bool bad = false;
    if (DIM == 2) {
      const int g = FindTuple(EdgeTuple(bnd[2 * b], bnd[2 * b + 1]), eoff, efirst, ekeys);
      bnd_edges[b] = g;
      bad = g < 0;
    } else {
      const int* v = bnd + 3 * (int64_t)b;
      const int f = FindTuple(FaceTuple(v[0], v[1], v[2]), foff, ffirst, fkeys);
      bnd_faces[b] = f;
      bad = f < 0;
      for (int l = 0; l < 3; l++) {
        const int g = FindTuple(EdgeTuple(v[l], v[(l + 1) % 3]), eoff, efirst, ekeys);
        bnd_edges[3 * (int64_t)b + l] = g;
        bad = bad || g < 0;
      }
    }
    if (bad) *missing = 1;
  }
}

// Class of an element's vertex order: the comparisons of a fixed sorting
// network, one bit per swap (the host setup's TrigClassNr / TetClassNr).
template <int DIM>
__device__ __forceinline__ int ElementClass(const int* v) {
  int s[DIM + 1], c = 0, t;
#pragma unroll
  for (int i = 0; i <= DIM; i++) s[i] = v[i];
  if (DIM == 2) {
    if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; c += 1; }
    if (s[1] > s[2]) { t = s[1]; s[1] = s[2]; s[2] = t; c += 2; }
    if (s[0] > s[1]) { c += 2; }
  } else {
    if (s[0] > s[1]) { t = s[0]; s[0] = s[1]; s[1] = t; c += 1; }
    if (s[2] > s[3]) { t = s[2]; s[2] = s[3]; s[3] = t; c += 2; }
    if (s[0] > s[2]) { t = s[0]; s[0] = s[2]; s[2] = t; c += 4; }
    if (s[1] > s[3]) { t = s[1]; s[1] = s[3]; s[3] = t; c += 8; }
    if (s[1] > s[2]) { c += 16; }
  }
  return c;
}

struct DofLayout {
  int nd, per_edge, per_face, per_cell;
  int64_t first_edge, first_face, first_cell;
};

template <int DIM>
__global__ __launch_bounds__(kTopoBlock) void ElementDofsKernel(const int* __restrict__ els,
                                                                const int* __restrict__ el_edges,
                                                                const int* __restrict__ el_faces, int ne,
                                                                DofLayout L, int* __restrict__ el_dofs,
                                                                int* __restrict__ classnr) {
  constexpr int nve = DIM + 1, nle = DIM == 2 ? 3 : 6;
  for (int e = blockIdx.x * kTopoBlock + threadIdx.x; e < ne; e += gridDim.x * kTopoBlock) {
    int* d = el_dofs + (int64_t)e * L.nd;
    const int* v = els + (int64_t)e * nve;
    int ii = 0;
    for (int l = 0; l < nve; l++) d[ii++] = v[l];
    for (int l = 0; l < nle; l++) {
      const int64_t g0 = L.first_edge + (int64_t)el_edges[(int64_t)e * nle + l] * L.per_edge;
      for (int k = 0; k < L.per_edge; k++) d[ii++] = (int)(g0 + k);
    }
    if (DIM == 3)
      for (int l = 0; l < 4; l++) {
        const int64_t g0 = L.first_face + (int64_t)el_faces[(int64_t)e * 4 + l] * L.per_face;
        for (int k = 0; k < L.per_face; k++) d[ii++] = (int)(g0 + k);
      }
    const int64_t c0 = L.first_cell + (int64_t)e * L.per_cell;
    for (int k = 0; k < L.per_cell; k++) d[ii++] = (int)(c0 + k);
    classnr[e] = ElementClass<DIM>(v);
  }
}

__global__ __launch_bounds__(kTopoBlock) void FillBytesKernel(unsigned char* __restrict__ p, int64_t n,
                                                              unsigned char v) {
  for (int64_t i = blockIdx.x * (int64_t)kTopoBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kTopoBlock)
    p[i] = v;
}

// Dofs of the boundary facets of Dirichlet regions are not free.  Facets
// sharing a dof store the same 0, so the order of the stores does not matter.
template <int DIM>
__global__ __launch_bounds__(kTopoBlock) void DirichletKernel(const int* __restrict__ bnd,
                                                              const int* __restrict__ bnd_index,
                                                              const int* __restrict__ bnd_edges,
                                                              const int* __restrict__ bnd_faces, int nb,
                                                              uint64_t mask, DofLayout L,
                                                              unsigned char* __restrict__ free) {
  constexpr int nbe = DIM == 2 ? 1 : 3;
  for (int b = blockIdx.x * kTopoBlock + threadIdx.x; b < nb; b += gridDim.x * kTopoBlock) {
    const int r = bnd_index[b];
    if (r < 1 || r > 64 || !((mask >> (r - 1)) & 1)) continue;
    for (int l = 0; l < DIM; l++) free[bnd[(int64_t)b * DIM + l]] = 0;
    for (int l = 0; l < nbe; l++) {
      const int64_t g0 = L.first_edge + (int64_t)bnd_edges[(int64_t)b * nbe + l] * L.per_edge;
      for (int k = 0; k < L.per_edge; k++) free[g0 + k] = 0;
    }
    if (DIM == 3) {
      const int64_t g0 = L.first_face + (int64_t)bnd_faces[b] * L.per_face;
      for (int k = 0; k < L.per_face; k++) free[g0 + k] = 0;
    }
  }
}

// One numbering (edges or faces): buckets by smallest vertex, distinct keys
// with multiplicities, global numbers first[a] + rank.
struct TupleTable {
  DevBuf<int> off, first, mult, cnt;
  DevBuf<uint64_t> keys;
  int count = 0;
};

template <int DIM, bool FACE>
static void BuildTable(TupleTable& T, const int* d_els, int64_t nent, int nv, DevBuf<int>& verts,
                       DevBuf<int>& el_numbers, hipStream_t st) {
  T.cnt.alloc(nv + 1);
  T.off.alloc(nv + 1);
  T.first.alloc(nv + 1);
  T.keys.alloc(std::max<int64_t>(nent, 1));
  T.mult.alloc(std::max<int64_t>(nent, 1));
  NGSAMD_HIP(hipMemsetAsync(T.cnt.p, 0, sizeof(int) * (nv + 1), st));
  const int grid = GridFor(nent, kTopoBlock, 16384);
  const int vgrid = GridFor(nv, kTopoBlock, 16384);
  BucketCountKernel<DIM, FACE><<<grid, kTopoBlock, 0, st>>>(d_els, nent, T.cnt.p);
  NGSAMD_CHECK_LAUNCH();
  ExclusiveScanKernel<<<1, kScanBlock, 0, st>>>(T.cnt.p, T.off.p, nv);
  NGSAMD_CHECK_LAUNCH();
  // the fill counters start at the bucket offsets
  NGSAMD_HIP(hipMemcpyAsync(T.cnt.p, T.off.p, sizeof(int) * nv, hipMemcpyDeviceToDevice, st));
  BucketFillKernel<DIM, FACE><<<grid, kTopoBlock, 0, st>>>(d_els, nent, T.cnt.p, T.keys.p);
  NGSAMD_CHECK_LAUNCH();
  BucketUniqueKernel<<<vgrid, kTopoBlock, 0, st>>>(T.off.p, nv, T.keys.p, T.mult.p, T.cnt.p);
  NGSAMD_CHECK_LAUNCH();
  ExclusiveScanKernel<<<1, kScanBlock, 0, st>>>(T.cnt.p, T.first.p, nv);
  NGSAMD_CHECK_LAUNCH();
  NGSAMD_HIP(hipMemcpyAsync(&T.count, T.first.p + nv, sizeof(int), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  verts.alloc(std::max(T.count, 1) * (FACE ? 3 : 2));
  WriteTuplesKernel<FACE><<<vgrid, kTopoBlock, 0, st>>>(T.off.p, T.first.p, nv, T.keys.p, verts.p);
  NGSAMD_CHECK_LAUNCH();
  el_numbers.alloc(std::max<int64_t>(nent, 1));
  ElementNumbersKernel<DIM, FACE><<<grid, kTopoBlock, 0, st>>>(d_els, nent, T.off.p, T.first.p, T.keys.p,
                                                                el_numbers.p);
  NGSAMD_CHECK_LAUNCH();
}

}  // namespace ngsamd

using namespace ngsamd;

// The handle: every table stays on the device until downloaded.
struct ngsamd_h1_s {
  int dim = 0, nv = 0, ne = 0, nb = 0, nedges = 0, nfaces = 0, ndof = 0;
  DofLayout L{};
  DevBuf<int> els, bnd, bnd_index, edges, faces, el_edges, el_faces, bnd_edges, bnd_faces, el_dofs, classnr;
  DevBuf<unsigned char> freedofs;
};

template <int DIM>
static void BuildH1(ngsamd_h1_s& H, const int* h_els, const int* h_bnd, const int* h_bnd_index, int order,
                    uint64_t mask) {
  hipStream_t st = Stream();
  constexpr int nve = DIM + 1, nle = DIM == 2 ? 3 : 6;
  H.els.alloc(std::max<int64_t>((int64_t)H.ne * nve, 1));
  H.els.upload(h_els, (size_t)H.ne * nve);
  TupleTable E, F;
  BuildTable<DIM, false>(E, H.els.p, (int64_t)H.ne * nle, H.nv, H.edges, H.el_edges, st);
  H.nedges = E.count;
  if (DIM == 3) {
    BuildTable<3, true>(F, H.els.p, (int64_t)H.ne * 4, H.nv, H.faces, H.el_faces, st);
    H.nfaces = F.count;
  }
  if (DIM == 3 && H.nb == 0) {
    // boundary = the faces of exactly one element, in face-number order
    DevBuf<int> cnt(H.nv + 1), boff(H.nv + 1);
    const int vgrid = GridFor(H.nv, kTopoBlock, 16384);
    OpenFacesCountKernel<<<vgrid, kTopoBlock, 0, st>>>(F.off.p, F.first.p, H.nv, F.mult.p, cnt.p);
    NGSAMD_CHECK_LAUNCH();
    ExclusiveScanKernel<<<1, kScanBlock, 0, st>>>(cnt.p, boff.p, H.nv);
    NGSAMD_CHECK_LAUNCH();
    NGSAMD_HIP(hipMemcpyAsync(&H.nb, boff.p + H.nv, sizeof(int), hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipStreamSynchronize(st));
    H.bnd.alloc(std::max(H.nb, 1) * 3);
    H.bnd_index.alloc(std::max(H.nb, 1));
    OpenFacesWriteKernel<<<vgrid, kTopoBlock, 0, st>>>(F.off.p, F.first.p, H.nv, F.mult.p, F.keys.p, boff.p,
                                                       H.bnd.p, H.bnd_index.p);
    NGSAMD_CHECK_LAUNCH();
  } else {
    H.bnd.alloc(std::max<int64_t>((int64_t)H.nb * DIM, 1));
    H.bnd.upload(h_bnd, (size_t)H.nb * DIM);
    H.bnd_index.alloc(std::max(H.nb, 1));
    H.bnd_index.upload(h_bnd_index, H.nb);
  }
  const int nbe = DIM == 2 ? 1 : 3;
  H.bnd_edges.alloc(std::max<int64_t>((int64_t)H.nb * nbe, 1));
  H.bnd_faces.alloc(std::max(DIM == 3 ? H.nb : 0, 1));
  DevBuf<int> missing(1);
  NGSAMD_HIP(hipMemsetAsync(missing.p, 0, sizeof(int), st));
  if (H.nb) {
    BoundaryNumbersKernel<DIM><<<GridFor(H.nb, kTopoBlock, 16384), kTopoBlock, 0, st>>>(
        H.bnd.p, H.nb, E.off.p, E.first.p, E.keys.p, DIM == 3 ? F.off.p : nullptr,
        DIM == 3 ? F.first.p : nullptr, DIM == 3 ? F.keys.p : nullptr, H.bnd_edges.p, H.bnd_faces.p, missing.p);
    NGSAMD_CHECK_LAUNCH();
  }
  // dof layout (comp/h1hofespace.cpp:823-950)
  const int p = order;
  DofLayout& L = H.L;
  L.per_edge = p - 1;
  L.per_face = (DIM == 3 && p > 2) ? (p - 1) * (p - 2) / 2 : 0;
  L.per_cell = DIM == 2 ? (p > 2 ? (p - 1) * (p - 2) / 2 : 0) : (p > 3 ? (p - 1) * (p - 2) * (p - 3) / 6 : 0);
  L.nd = DIM == 2 ? (p + 1) * (p + 2) / 2 : (p + 1) * (p + 2) * (p + 3) / 6;
  L.first_edge = H.nv;
  L.first_face = L.first_edge + (int64_t)H.nedges * L.per_edge;
  L.first_cell = L.first_face + (int64_t)H.nfaces * L.per_face;
  const int64_t ndof = L.first_cell + (int64_t)H.ne * L.per_cell;
  NGSAMD_REQUIRE(ndof < (int64_t(1) << 31), "ngsamd_h1_create: too many dofs for int32");
  NGSAMD_REQUIRE((int64_t)H.ne * L.nd < (int64_t(1) << 40), "ngsamd_h1_create: element dof table too large");
  H.ndof = (int)ndof;
  H.el_dofs.alloc(std::max<int64_t>((int64_t)H.ne * L.nd, 1));
  H.classnr.alloc(std::max(H.ne, 1));
  if (H.ne) {
    ElementDofsKernel<DIM><<<GridFor(H.ne, kTopoBlock, 16384), kTopoBlock, 0, st>>>(
        H.els.p, H.el_edges.p, DIM == 3 ? H.el_faces.p : nullptr, H.ne, L, H.el_dofs.p, H.classnr.p);
    NGSAMD_CHECK_LAUNCH();
  }
  H.freedofs.alloc(std::max(H.ndof, 1));
  FillBytesKernel<<<GridFor(H.ndof, kTopoBlock, 16384), kTopoBlock, 0, st>>>(H.freedofs.p, H.ndof, 1);
  NGSAMD_CHECK_LAUNCH();
  int h_missing = 0;
  NGSAMD_HIP(hipMemcpyAsync(&h_missing, missing.p, sizeof(int), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  NGSAMD_REQUIRE(!h_missing, "ngsamd_h1_create: a boundary facet is not a facet of the mesh");
  if (H.nb) {
    DirichletKernel<DIM><<<GridFor(H.nb, kTopoBlock, 16384), kTopoBlock, 0, st>>>(
        H.bnd.p, H.bnd_index.p, H.bnd_edges.p, H.bnd_faces.p, H.nb, mask, L, H.freedofs.p);
    NGSAMD_CHECK_LAUNCH();
  }
  NGSAMD_HIP(hipStreamSynchronize(st));
}

extern "C" int ngsamd_h1_create(ngsamd_h1_t* h, int dim, int64_t nv, int64_t ne, const int* h_els, int64_t nbnd,
                                const int* h_bnd, const int* h_bnd_index, int order, uint64_t dirichlet_mask) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(h, "ngsamd_h1_create: null argument");
    *h = nullptr;
    NGSAMD_REQUIRE(dim == 2 || dim == 3, "ngsamd_h1_create: dim must be 2 or 3");
    NGSAMD_REQUIRE(order >= 1 && order <= 10, "ngsamd_h1_create: order must be in 1..10");
    NGSAMD_REQUIRE(nv >= 0 && ne >= 0 && nbnd >= 0, "ngsamd_h1_create: negative size");
    NGSAMD_REQUIRE(ne * 6 < (int64_t(1) << 31) && nv < (int64_t(1) << 31) && nbnd < (int64_t(1) << 29),
                   "ngsamd_h1_create: mesh too large for int32 tables");
    NGSAMD_REQUIRE(ne == 0 || h_els, "ngsamd_h1_create: null argument");
    NGSAMD_REQUIRE(nbnd == 0 || (h_bnd && h_bnd_index), "ngsamd_h1_create: null argument");
    // vertex numbers index the buckets: checked before anything is uploaded
    for (int64_t i = 0; i < ne * (dim + 1); i++)
      NGSAMD_REQUIRE(h_els[i] >= 0 && h_els[i] < nv, "ngsamd_h1_create: element vertex out of range");
    for (int64_t i = 0; i < nbnd * dim; i++)
      NGSAMD_REQUIRE(h_bnd[i] >= 0 && h_bnd[i] < nv, "ngsamd_h1_create: boundary vertex out of range");
    auto H = std::make_unique<ngsamd_h1_s>();
    H->dim = dim;
    H->nv = (int)nv;
    H->ne = (int)ne;
    H->nb = (int)nbnd;
    if (dim == 2) BuildH1<2>(*H, h_els, h_bnd, h_bnd_index, order, dirichlet_mask);
    else BuildH1<3>(*H, h_els, h_bnd, h_bnd_index, order, dirichlet_mask);
    *h = H.release();
  });
}

extern "C" int ngsamd_h1_sizes(ngsamd_h1_t h, int64_t* h_sizes) {
  return Guard([&] {
    NGSAMD_REQUIRE(h && h_sizes, "ngsamd_h1_sizes: null argument");
    const int64_t s[6] = {h->nedges, h->nfaces, h->nb, h->ndof, h->L.nd, h->ne};
    std::memcpy(h_sizes, s, sizeof(s));
  });
}

template <class T>
static void Download(T* h, const DevBuf<T>& d, int64_t count) {
  if (h && count > 0) NGSAMD_HIP(hipMemcpyAsync(h, d.p, sizeof(T) * count, hipMemcpyDeviceToHost, Stream()));
}

extern "C" int ngsamd_h1_topology(ngsamd_h1_t h, int* h_edges, int* h_faces, int* h_el_edges, int* h_el_faces,
                                  int* h_bnd, int* h_bnd_index, int* h_bnd_edges, int* h_bnd_faces) {
  return Guard([&] {
    NGSAMD_REQUIRE(h, "ngsamd_h1_topology: null handle");
    const int d = h->dim, nle = d == 2 ? 3 : 6;
    Download(h_edges, h->edges, 2 * (int64_t)h->nedges);
    Download(h_el_edges, h->el_edges, (int64_t)h->ne * nle);
    if (d == 3) {
      Download(h_faces, h->faces, 3 * (int64_t)h->nfaces);
      Download(h_el_faces, h->el_faces, 4 * (int64_t)h->ne);
      Download(h_bnd_faces, h->bnd_faces, (int64_t)h->nb);
    }
    Download(h_bnd, h->bnd, (int64_t)h->nb * d);
    Download(h_bnd_index, h->bnd_index, (int64_t)h->nb);
    Download(h_bnd_edges, h->bnd_edges, (int64_t)h->nb * (d == 2 ? 1 : 3));
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  });
}

extern "C" int ngsamd_h1_dofs(ngsamd_h1_t h, int* h_el_dofs, unsigned char* h_freedofs, int* h_classnr,
                              const int** d_el_dofs) {
  return Guard([&] {
    NGSAMD_REQUIRE(h, "ngsamd_h1_dofs: null handle");
    Download(h_el_dofs, h->el_dofs, (int64_t)h->ne * h->L.nd);
    Download(h_freedofs, h->freedofs, (int64_t)h->ndof);
    Download(h_classnr, h->classnr, (int64_t)h->ne);
    if (d_el_dofs) *d_el_dofs = h->el_dofs.p;
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  });
}

extern "C" int ngsamd_h1_destroy(ngsamd_h1_t h) {
  return Guard([&] { delete h; });
}
