// Fused element-batched matrix-free H1 Laplace operator on gfx950.
//
// Reference: MatrixFreeBTDTB::MultAdd (comp/matrixfree.cpp:681-1087) and its
// CUDA counterpart DevMatrixFreeBTDTB (ngscuda/cuda_matrix_free_btdtb.cpp:16-427).
// For every element e of a vertex-permutation class (shared reference
// gradients B, [q][i][d] = d phi_i / d xi_d at quadrature point q):
//     u_q = B_q^T x_e            (dim values per point)
//     v_q = w_q kappa_e G_e u_q  (G_e = |det F| F^-1 F^-T, symmetric)
//     y_e = sum_q B_q v_q
// followed by the scatter y[dofs_e] += y_e.
//
// MFMA path (FP64 v_mfma_f64_16x16x4f64, FP32 v_mfma_f32_16x16x4f32): one
// wavefront owns a tile of 16 elements.  Per block of 16 quadrature points it
// forms U^T (16 q x 16 el, one tile per direction d) = B^T X with the B block
// as the A operand (read from LDS) and the gathered element vectors as the B
// operand (held in registers for the whole tile), applies the 2x2/3x3 metric
// to the accumulators in registers, and feeds them straight back as the B
// operand of the second product Y^T (dof x el) += B V^T.  The output layout of
// a 16x16 MFMA (lane l holds 4 rows of column l%16: rows 4*(l/16)+r for FP32,
// (l/16)+4r for FP64, see AccRow) is used as the K index of the second
// product with the matching permutation of the A operand, so U/V never leave
// registers.  B of one class (nq padded to 16,
// nd padded to 4, one pad word per point against LDS bank conflicts) is loaded
// into LDS once per workgroup; every wavefront then walks element tiles.
//
// Scatter: GATHER (default) writes each element's result row to a
// contribution buffer; a second kernel sums, per dof, the contributions in a
// fixed order (deterministic, no atomics).  ATOMIC adds the element results
// with hardware floating-point atomics, as the reference kernel does.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "half.hpp"
#include "matfree.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

typedef double double4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int kMfWaves = 8;
constexpr int kMfBlock = 64 * kMfWaves;

__device__ __forceinline__ double4v Mfma(double a, double b, double4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4v Mfma(float a, float b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <class T> struct Acc4;
template <> struct Acc4<double> { typedef double4v type; };
template <> struct Acc4<float> { typedef float4v type; };

// Row of the 16x16 result held in accumulator element r of a lane in lane
// group grp = lane/16 (the column is lane%16).  Measured on gfx950:
// v_mfma_f32_16x16x4f32 holds rows 4*grp+r, v_mfma_f64_16x16x4f64 holds
// rows grp+4*r.
template <class T>
__device__ __forceinline__ constexpr int AccRow(int grp, int r) {
  return sizeof(T) == 8 ? grp + 4 * r : 4 * grp + r;
}

template <class V>
__device__ __forceinline__ void AtomicAddV(V* p, V v) {
  atomicAdd(p, v);
}

// Metric G (symmetric, packed) times vector: 2D (xx,yy,xy), 3D (xx,yy,zz,xy,xz,yz).
template <int DIM, class T>
__device__ __forceinline__ void ApplyMetric(const T* g, const T* u, T* v) {
  if constexpr (DIM == 2) {
    v[0] = g[0] * u[0] + g[2] * u[1];
    v[1] = g[2] * u[0] + g[1] * u[1];
  } else {
    v[0] = g[0] * u[0] + g[3] * u[1] + g[4] * u[2];
    v[1] = g[3] * u[0] + g[1] * u[1] + g[5] * u[2];
    v[2] = g[4] * u[0] + g[5] * u[1] + g[2] * u[2];
  }
}

// Work schedule shared by the MFMA kernels: the element tiles (16 elements)
// of every class are numbered class after class in rounds of kMfWaves tiles,
// one per wavefront, and a round never straddles two classes.  Workgroup w
// (after the XCD-contiguous remap) walks the rounds
// [w * nrounds / grid, (w + 1) * nrounds / grid) and reloads the class's B
// into LDS only when the class changes, so one launch covers all classes.
struct MfRounds {
  const int* round0;     // first round of each class, nclass + 1 entries
  const int64_t* range;  // element range [begin, end) of each class
  int nrounds;
  int nclass;
  const int* tbase = nullptr;  // first tile of each class (tile-major tables), nclass + 1
};

// classes whose schedule the reference-matrix kernels keep in LDS
constexpr int kMaxLdsClasses = 64;

// Copy n 16-byte words of a class's matrices from global memory into LDS.
// Each thread issues 4 loads before its first LDS write (from clamped
// indices, so no load sits behind a branch): a plain load-then-store loop
// waits for every load in turn.
template <int BLOCK = kMfBlock>
__device__ __forceinline__ void CopyToLds(void* dst, const void* src, int n) {
  constexpr int U = 4;
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (int base = 0; base < n; base += BLOCK * U) {
    uint4 v[U];
#pragma unroll
    for (int j = 0; j < U; j++) v[j] = s[min(base + j * BLOCK + (int)threadIdx.x, n - 1)];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int i = base + j * BLOCK + (int)threadIdx.x;
      if (i < n) d[i] = v[j];
    }
  }
}

// Class of round r, scanning on from the class of the previous round.
__device__ __forceinline__ int RoundClass(const MfRounds& R, int r, int from) {
  int c = from < 0 ? 0 : from;
  while (R.round0[c + 1] <= r) c++;
  return c;
}

// MODE 0: contributions (row-major [nel][nd], table row numbering)
// MODE 1: atomic add s*y_e into y
// MODE 2 (factored kernels): y += s*y_e with plain loads and stores, for a
//        launch over one colour (no two of its elements share a dof)
template <class T, int DIM, int NDP, int MODE, class V>
__global__ __launch_bounds__(kMfBlock) void MfMfmaKernel(
    const T* __restrict__ bmat, int64_t bstride, MfRounds R, const double* __restrict__ weights, int nqb,
    const int* __restrict__ dofs, const double* __restrict__ metric, const double* __restrict__ kappa,
    const int* __restrict__ elnr, int nd, const V* __restrict__ x, V* __restrict__ out, double s,
    const int* __restrict__ stop) {
  if (stop && *stop) return;
  typedef typename Acc4<T>::type acc_t;
  constexpr int QS = DIM * NDP + 1;  // LDS words per quadrature point
  constexpr int KS = NDP / 4;        // k-steps of the first product
  constexpr int MT = (NDP + 15) / 16;
  constexpr int NC = DIM * (DIM + 1) / 2;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lb = reinterpret_cast<T*>(lds_raw);
  const int nlds = nqb * 16 * QS;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    const int c = RoundClass(R, rd, cls);
    if (c != cls) {
      if (cls >= 0) __syncthreads();
      // nlds * sizeof(T) and the class offset are multiples of 64 bytes
      CopyToLds(lb, bmat + (int64_t)c * bstride, (int)(nlds * sizeof(T) / 16));
      __syncthreads();
      cls = c;
    }
    const int64_t end = R.range[2 * c + 1];
    const int64_t first = R.range[2 * c] + ((int64_t)(rd - R.round0[c]) * kMfWaves + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    // Gathers are unconditional, from indices clamped into range (element
    // `first`, dof nd-1), and masked afterwards: conditional loads would each
    // cost a round trip of their own.
    const int64_t ev = valid ? e : first;
    // element vector, K layout of the first product: lane holds dof 4s+grp
    int dv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) dv[st] = dofs[ev * nd + min(4 * st + grp, nd - 1)];
    const int en = elnr[ev];
    double mv[NC];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) mv[cc] = metric[ev * NC + cc];
    V xv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xv[st] = x[dv[st]];
    const double kap = kappa[en];
    T xk[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xk[st] = (valid && 4 * st + grp < nd) ? (T)xv[st] : (T)0;
    T g[NC];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) g[cc] = valid ? (T)(kap * mv[cc]) : (T)0;
    acc_t yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = acc_t{0, 0, 0, 0};

    for (int t = 0; t < nqb; t++) {
      acc_t u[DIM];
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        u[d] = acc_t{0, 0, 0, 0};
        const T* arow = lb + (16 * t + col) * QS + d * NDP + grp;
#pragma unroll
        for (int st = 0; st < KS; st++) u[d] = Mfma(arow[4 * st], xk[st], u[d]);
      }
      // metric at the 4 points this lane holds: q = 16t + AccRow(grp, r)
      T v[DIM][4];
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const T w = (T)weights[16 * t + AccRow<T>(grp, r)];
        T uu[DIM], vv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) uu[d] = u[d][r];
        ApplyMetric<DIM, T>(g, uu, vv);
#pragma unroll
        for (int d = 0; d < DIM; d++) v[d][r] = w * vv[d];
      }
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int dof = 16 * m + col;
        const bool inrange = dof < NDP;
#pragma unroll
        for (int d = 0; d < DIM; d++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const T a = inrange ? lb[(16 * t + AccRow<T>(grp, r)) * QS + d * NDP + dof] : (T)0;
            yacc[m] = Mfma(a, v[d][r], yacc[m]);
          }
      }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + AccRow<T>(grp, r);
          if (dof < nd) {
            const V val = (V)yacc[m][r];
            if (MODE == 0) out[e * nd + dof] = val;
            else AtomicAddV(&out[dofs[e * nd + dof]], (V)(s * (double)val));
          }
        }
    }
  }
}

// 16-bit paths, BF16 (v_mfma_f32_16x16x32_bf16) and FP16
// (v_mfma_f32_16x16x32_f16), FP32 accumulation.  Operand layout measured on
// gfx950 for the BF16 instruction, and the FP16 sibling is checked against it
// by the precision tests: lane l supplies A[l%16][8*(l/16)+j] and
// B[8*(l/16)+j][l%16], j = 0..7, and holds result rows 4*(l/16)+r.
// Quadrature points are taken in chunks of 32 so that the K dimension of the
// second product (one direction x 32 points) is one MFMA.  B is held in LDS
// twice, each copy laid out so that a lane's 8 A-operand values are one
// contiguous 16-byte read:
//   B1[q][d][K1P]            first product, dofs contiguous (K1P = nd -> 8k)
//   B2[chunk][d][MT*16][32]  second product, the 32 points of a chunk in the
//                            order the first product's accumulators hold them
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
template <class H> struct Vec8;
template <> struct Vec8<__bf16> { typedef bf16x8 type; };
template <> struct Vec8<_Float16> { typedef half8 type; };
__device__ __forceinline__ float4v Mfma16(bf16x8 a, bf16x8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float4v Mfma16(half8 a, half8 b, float4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

template <class H, int DIM, int KS1, int MT, int MODE, class V>
__global__ __launch_bounds__(kMfBlock) void MfBf16Kernel(
    const uint16_t* __restrict__ bmat, int64_t bstride, MfRounds R, int k1p, int nqc,
    const double* __restrict__ weights, const int* __restrict__ dofs, const double* __restrict__ metric,
    const double* __restrict__ kappa, const int* __restrict__ elnr, int nd, const V* __restrict__ x,
    V* __restrict__ out, double s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  constexpr int NC = DIM * (DIM + 1) / 2;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int size1 = nqc * 32 * DIM * k1p;
  typedef typename Vec8<H>::type h8;
  const H* B1 = reinterpret_cast<const H*>(lds_raw);
  const H* B2 = B1 + size1;

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    const int cl = RoundClass(R, rd, cls);
    if (cl != cls) {
      if (cls >= 0) __syncthreads();
      // B1 and B2 of the class: bstride 16-bit words, a multiple of 8
      CopyToLds(lds_raw, bmat + (int64_t)cl * bstride, (int)(bstride / 8));
      __syncthreads();
      cls = cl;
    }
    const int64_t end = R.range[2 * cl + 1];
    const int64_t first = R.range[2 * cl] + ((int64_t)(rd - R.round0[cl]) * kMfWaves + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    // unconditional gathers from clamped indices, masked afterwards (see
    // MfMfmaKernel)
    const int64_t ev = valid ? e : first;
    int dv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) dv[ks][j] = dofs[ev * nd + min(32 * ks + 8 * grp + j, nd - 1)];
    const int en = elnr[ev];
    double mv[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) mv[c] = metric[ev * NC + c];
    V xv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xv[ks][j] = x[dv[ks][j]];
    const double kap = kappa[en];
    float xf[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int dof = 32 * ks + 8 * grp + j;
        xf[ks][j] = (valid && dof < nd) ? (float)xv[ks][j] : 0.0f;
      }
    float g[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) g[c] = valid ? (float)(kap * mv[c]) : 0.0f;
    // FP16 has a 5-bit exponent: the element vector and the metric are
    // scaled by powers of two (exact) so that their largest entries lie in
    // [0.5, 1) before rounding, and the result is scaled back in FP32.
    // Without this the small vectors of a converging inner solve fall into
    // the FP16 subnormals and lose their significant bits.  BF16 keeps the
    // FP32 exponent range and needs no scaling.
    int escale = 0;
    if constexpr (sizeof(H) == 2 && !std::is_same<H, __bf16>::value) {
      float xm = 0.0f, gm = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xm = fmaxf(xm, fabsf(xf[ks][j]));
      // the 4 lanes holding one element's dofs: col, col+16, col+32, col+48
      xm = fmaxf(xm, __shfl_xor(xm, 16, 64));
      xm = fmaxf(xm, __shfl_xor(xm, 32, 64));
#pragma unroll
      for (int c = 0; c < NC; c++) gm = fmaxf(gm, fabsf(g[c]));
      int ex = 0, eg = 0;
      (void)frexpf(xm, &ex);
      (void)frexpf(gm, &eg);
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xf[ks][j] = ldexpf(xf[ks][j], -ex);
#pragma unroll
      for (int c = 0; c < NC; c++) g[c] = ldexpf(g[c], -eg);
      escale = ex + eg;
    }
    h8 xk[KS1];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xk[ks][j] = (H)xf[ks][j];
    float4v yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = float4v{0, 0, 0, 0};
    const h8 zero8 = {};

    for (int c = 0; c < nqc; c++) {
      float4v u[2][DIM];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int d = 0; d < DIM; d++) {
          u[h][d] = float4v{0, 0, 0, 0};
#pragma unroll
          for (int ks = 0; ks < KS1; ks++) {
            const int k0 = 32 * ks + 8 * grp;
            const h8 a = k0 < k1p ? *reinterpret_cast<const h8*>(B1 + ((32 * c + 16 * h + col) * DIM + d) * k1p + k0)
                                  : zero8;
            u[h][d] = Mfma16(a, xk[ks], u[h][d]);
          }
        }
      // metric at the 8 points this lane holds: q = 32c + 16h + 4grp + r
      float v[DIM][8];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const float w = (float)weights[32 * c + 16 * h + 4 * grp + r];
          float uu[DIM], vv[DIM];
#pragma unroll
          for (int d = 0; d < DIM; d++) uu[d] = u[h][d][r];
          ApplyMetric<DIM, float>(g, uu, vv);
#pragma unroll
          for (int d = 0; d < DIM; d++) v[d][4 * h + r] = w * vv[d];
        }
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        h8 vb;
#pragma unroll
        for (int j = 0; j < 8; j++) vb[j] = (H)v[d][j];
#pragma unroll
        for (int m = 0; m < MT; m++) {
          const h8 a = *reinterpret_cast<const h8*>(B2 + (((c * DIM + d) * MT * 16 + 16 * m + col) * 32 + 8 * grp));
          yacc[m] = Mfma16(a, vb, yacc[m]);
        }
      }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + 4 * grp + r;
          if (dof < nd) {
            const V val = (V)ldexpf(yacc[m][r], escale);
            if (MODE == 0) out[e * nd + dof] = val;
            else AtomicAddV(&out[dofs[e * nd + dof]], (V)(s * (double)val));
          }
        }
    }
  }
}

// ------------------------------------------------ reference-matrix kernels
// For an affine element with one coefficient value the quadrature sum
// collapses exactly:
//   sum_q w_q B_q^T (kappa G_e) B_q = sum_c g_{e,c} S_c,
//   S_(dd) = sum_q w_q b_d b_d^T,  S_(dd') = sum_q w_q (b_d b_d'^T + b_d' b_d^T),
// with g_e = kappa_e (G_xx, G_yy[, G_zz], G_xy[, G_xz, G_yz]) (the packed
// metric) and the symmetric S_c (3 in 2D, 6 in 3D) precomputed per class.  The
// kernel forms y_e = sum_c S_c (g_{e,c} x_e): the metric scales the B
// operand (one column per element) and every c accumulates into the same
// MFMA accumulators, so per 16-element tile the work is NC x MT x KS MFMAs
// (FP64 tet p4: 162 instead of 504 for the quadrature form) and no per-point
// arithmetic.  LDS holds S[c][i][j] for the rows i < nd (rows beyond nd read
// as 0), row stride NDP + 1 words (FP64/FP32) or K1P + 8 halves (16-bit).

template <class T, int DIM, int NDP, int MODE, class V, bool LSCHED>
__global__ __launch_bounds__(kMfBlock) void MfRefKernel(
    const T* __restrict__ smat, int64_t sstride, MfRounds R, const int* __restrict__ dofs,
    const double* __restrict__ metric, const double* __restrict__ kappa, const int* __restrict__ elnr, int nd,
    const V* __restrict__ x, V* __restrict__ out, double s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  typedef typename Acc4<T>::type acc_t;
  constexpr int NC = DIM * (DIM + 1) / 2;
  constexpr int KS = NDP / 4;
  constexpr int MT = (NDP + 15) / 16;
  constexpr int RS = NDP + 1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* ls = reinterpret_cast<T*>(lds_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  // LSCHED: the class schedule is copied to LDS once (looked up in global
  // memory, each round's lookup waits for every memory operation in flight,
  // the previous round's stores included)
  __shared__ int s_round0[LSCHED ? kMaxLdsClasses + 1 : 1];
  __shared__ int64_t s_range[LSCHED ? 2 * kMaxLdsClasses : 1];
  if constexpr (LSCHED) {
    for (int i = threadIdx.x; i <= R.nclass; i += kMfBlock) s_round0[i] = R.round0[i];
    for (int i = threadIdx.x; i < 2 * R.nclass; i += kMfBlock) s_range[i] = R.range[i];
    __syncthreads();
  }
  const int* sched_r0 = LSCHED ? s_round0 : R.round0;
  const int64_t* sched_rng = LSCHED ? s_range : R.range;
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    int c = cls < 0 ? 0 : cls;
    while (sched_r0[c + 1] <= rd) c++;
    if (c != cls) {
      if (cls >= 0) __syncthreads();
      // sstride * sizeof(T) is a multiple of 16 bytes
      CopyToLds(ls, smat + (int64_t)c * sstride, (int)(sstride * sizeof(T) / 16));
      __syncthreads();
      cls = c;
    }
    const int64_t end = sched_rng[2 * c + 1];
    const int64_t first = sched_rng[2 * c] + ((int64_t)(rd - sched_r0[c]) * kMfWaves + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    // unconditional gathers from clamped indices, masked afterwards (see
    // MfMfmaKernel)
    const int64_t ev = valid ? e : first;
    int dv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) dv[st] = dofs[ev * nd + min(4 * st + grp, nd - 1)];
    const int en = elnr[ev];
    double mv[NC];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) mv[cc] = metric[ev * NC + cc];
    V xv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xv[st] = x[dv[st]];
    const double kap = kappa[en];
    T xk[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xk[st] = (valid && 4 * st + grp < nd) ? (T)xv[st] : (T)0;
    acc_t yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
      const T gc = valid ? (T)(kap * mv[cc]) : (T)0;
      T xg[KS];
#pragma unroll
      for (int st = 0; st < KS; st++) xg[st] = gc * xk[st];
      if constexpr (std::is_same<T, float>::value) {
        // FP32: S in lane order [c][m][grp][col][KSP] (rows zero-padded to
        // the tile), so a lane's k-steps are contiguous: 16-byte LDS reads,
        // conflict-free, no row masking
        constexpr int KSP = (KS + 3) / 4 * 4;
        const T* lp = ls + ((cc * MT * 4 + grp) * 16 + col) * KSP;
#pragma unroll
        for (int q = 0; q < KSP / 4; q++) {
          float4v a4[MT];
#pragma unroll
          for (int m = 0; m < MT; m++) a4[m] = *reinterpret_cast<const float4v*>(lp + m * 64 * KSP + 4 * q);
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (4 * q + j < KS)
#pragma unroll
              for (int m = 0; m < MT; m++) yacc[m] = Mfma(a4[m][j], xg[4 * q + j], yacc[m]);
        }
      } else {
        const T* srow = ls + cc * nd * RS;
        // k-steps outside, dof tiles inside: consecutive MFMAs go to
        // independent accumulators.  Rows below NDP - 3 <= nd always exist.
#pragma unroll
        for (int st = 0; st < KS; st++)
#pragma unroll
          for (int m = 0; m < MT; m++) {
            const int row = 16 * m + col;
            const bool inrange = (16 * m + 16 <= NDP - 3) || row < nd;
            const T a = inrange ? srow[(inrange ? row : 0) * RS + grp + 4 * st] : (T)0;
            yacc[m] = Mfma(a, xg[st], yacc[m]);
          }
      }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + AccRow<T>(grp, r);
          if (dof < nd) {
            const V val = (V)yacc[m][r];
            if (MODE == 0) out[e * nd + dof] = val;
            else AtomicAddV(&out[dofs[e * nd + dof]], (V)(s * (double)val));
          }
        }
    }
  }
}

// 16-bit operands (BF16 / FP16), FP32 accumulation; operand layout of
// v_mfma_f32_16x16x32_{bf16,f16} as in MfBf16Kernel.  FP16 uses the same
// per-element power-of-two scaling of x and g as MfBf16Kernel.
template <class H, int DIM, int KS1, int MT, int MODE, class V, bool LSCHED>
__global__ __launch_bounds__(kMfBlock) void MfRef16Kernel(
    const uint16_t* __restrict__ smat, int64_t sstride, MfRounds R, const int* __restrict__ dofs,
    const double* __restrict__ metric, const double* __restrict__ kappa, const int* __restrict__ elnr, int nd,
    const V* __restrict__ x, V* __restrict__ out, double s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  typedef typename Vec8<H>::type h8;
  constexpr int NC = DIM * (DIM + 1) / 2;
  constexpr int RS = KS1 * 32 + 8;  // halves per row: 16-byte aligned rows
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const H* ls = reinterpret_cast<const H*>(lds_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  // LSCHED: the class schedule in LDS, as in MfRefKernel
  __shared__ int s_round0[LSCHED ? kMaxLdsClasses + 1 : 1];
  __shared__ int64_t s_range[LSCHED ? 2 * kMaxLdsClasses : 1];
  if constexpr (LSCHED) {
    for (int i = threadIdx.x; i <= R.nclass; i += kMfBlock) s_round0[i] = R.round0[i];
    for (int i = threadIdx.x; i < 2 * R.nclass; i += kMfBlock) s_range[i] = R.range[i];
    __syncthreads();
  }
  const int* sched_r0 = LSCHED ? s_round0 : R.round0;
  const int64_t* sched_rng = LSCHED ? s_range : R.range;
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    int cl = cls < 0 ? 0 : cls;
    while (sched_r0[cl + 1] <= rd) cl++;
    if (cl != cls) {
      if (cls >= 0) __syncthreads();
      // sstride 16-bit words, a multiple of 8
      CopyToLds(lds_raw, smat + (int64_t)cl * sstride, (int)(sstride / 8));
      __syncthreads();
      cls = cl;
    }
    const int64_t end = sched_rng[2 * cl + 1];
    const int64_t first = sched_rng[2 * cl] + ((int64_t)(rd - sched_r0[cl]) * kMfWaves + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    const int64_t ev = valid ? e : first;
    int dv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) dv[ks][j] = dofs[ev * nd + min(32 * ks + 8 * grp + j, nd - 1)];
    const int en = elnr[ev];
    double mv[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) mv[c] = metric[ev * NC + c];
    V xv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xv[ks][j] = x[dv[ks][j]];
    const double kap = kappa[en];
    float xf[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xf[ks][j] = (valid && 32 * ks + 8 * grp + j < nd) ? (float)xv[ks][j] : 0.0f;
    float g[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) g[c] = valid ? (float)(kap * mv[c]) : 0.0f;
    int escale = 0;
    if constexpr (!std::is_same<H, __bf16>::value) {
      float xm = 0.0f, gm = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xm = fmaxf(xm, fabsf(xf[ks][j]));
      xm = fmaxf(xm, __shfl_xor(xm, 16, 64));
      xm = fmaxf(xm, __shfl_xor(xm, 32, 64));
#pragma unroll
      for (int c = 0; c < NC; c++) gm = fmaxf(gm, fabsf(g[c]));
      int ex = 0, eg = 0;
      (void)frexpf(xm, &ex);
      (void)frexpf(gm, &eg);
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xf[ks][j] = ldexpf(xf[ks][j], -ex);
#pragma unroll
      for (int c = 0; c < NC; c++) g[c] = ldexpf(g[c], -eg);
      escale = ex + eg;
    }
    float4v yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = float4v{0, 0, 0, 0};
    const h8 zero8 = {};
#pragma unroll
    for (int c = 0; c < NC; c++) {
      h8 xg[KS1];
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xg[ks][j] = (H)(g[c] * xf[ks][j]);
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int m = 0; m < MT; m++) {
          const int row = 16 * m + col;
          const bool inrange = row < nd;
          const h8 a = inrange ? *reinterpret_cast<const h8*>(ls + (c * nd + (inrange ? row : 0)) * RS + 8 * grp + 32 * ks)
                               : zero8;
          yacc[m] = Mfma16(a, xg[ks], yacc[m]);
        }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + 4 * grp + r;
          if (dof < nd) {
            const V val = (V)ldexpf(yacc[m][r], escale);
            if (MODE == 0) out[e * nd + dof] = val;
            else AtomicAddV(&out[dofs[e * nd + dof]], (V)(s * (double)val));
          }
        }
    }
  }
}

// ------------------------------------------- factored reference matrices
// The weighted gradient components of a class's shape functions at the
// points, rows Bw[(d,i)][q] = b_{q,i,d} sqrt(w_q), span a space of dimension
// r (for H1 of degree p the gradients of degree-p polynomials, r = dim
// P_{p-1}: tet p4 r = 20 against 3 x 35 rows).  With an orthonormal basis Phi
// of that space Bw = D Phi (D = Bw Phi^T), hence
//   S_(dd') = sum_q w_q b_d b_d'^T = D_d D_d'^T,
//   y_e = sum_c g_{e,c} S_c x_e = sum_d D_d sum_d' G_{e,dd'} (D_d'^T x_e),
// exactly, for any rule.  Per 16-element tile: U = D^T X (DIM x r rows,
// MFMA), V = G U per element (VALU, lane-local), Y = D V (MFMA).  Tet p4,
// FP64: 9 x 4 + 15 x 3 = 81 MFMAs against 162 for sum_c S_c (g_c X).
//
// Layout of U: the accumulator slots of lane group grp, sigma = 4 t + r
// (tile t, accumulator element r; the row AccRow(grp, r) of tile t), hold
// component d = sigma % DIM of basis function k = grp * KG + sigma / DIM, so
// the DIM components of one k sit in one lane and the metric needs no lane
// exchange.  The backward product uses slot sigma of lane group grp as its K
// index 4 sigma + grp (FP64/FP32 16x16x4: lane l supplies K row l / 16 of
// each step), so V goes from the accumulators straight into the B operand.
// LDS, lane order so that each lane's A operands are 16-byte reads:
//   A1[t][grp][col][KSP]  row col of tile t (slot of AccRow^-1(col)), K = dof
//   A2[m][grp][col][JP]   row = dof 16 m + col, K = slot of lane group grp
// With partials (MODE 0, CG's <x, A x>): U . V summed over a lane's slots is
// its share of x_e^T K_e x_e (sum_d (D_d^T x_e) . V_d); each workgroup
// stores the sum over its tiles to partials[blockIdx.x].
// Register budget: unbounded (tet p4: FP64 124 VGPRs, FP32 93, 16-bit
// 120-128: two 8-wave workgroups per CU).  Holding the FP32 and 16-bit tiles to 80 VGPRs
// (three workgroups per CU, no spills) measured slower at 32^3 p=4: FP32
// 38.6 -> 42.9 us, 16-bit 39.9 -> 44 us, 16-bit with FP64 vectors 52 -> 73
// us -- more tiles in flight per CU only add pressure on L1 and L2; one
// workgroup per CU is slower too (FP32 vectors 64.6 against 62.6 us per
// apply).
template <class T, int DIM, int NDP, int KG, int MODE, class V>
__global__ __launch_bounds__(kMfBlock) void MfFacKernel(
    const T* __restrict__ fmat, int64_t fstride, MfRounds R, const int* __restrict__ dofs,
    const double* __restrict__ metric, const double* __restrict__ kappa, const int* __restrict__ elnr, int nd,
    const V* __restrict__ x, V* __restrict__ out, double s, const int* __restrict__ stop,
    double* __restrict__ partials) {
  if (stop && *stop) return;
  typedef typename Acc4<T>::type acc_t;
  constexpr int NC = DIM * (DIM + 1) / 2;
  constexpr int KS = NDP / 4;
  constexpr int MT = (NDP + 15) / 16;
  constexpr int NSL = DIM * KG;
  constexpr int TT = (NSL + 3) / 4;
  constexpr int VW = 16 / (int)sizeof(T);
  constexpr int KSP = (KS + VW - 1) / VW * VW;
  constexpr int JP = (NSL + VW - 1) / VW * VW;
  typedef T vec_t __attribute__((ext_vector_type(VW)));
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* ls = reinterpret_cast<T*>(lds_raw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  // MODE 0 / 1 read the tile-major tables, MODE 2 (a colour's) element-major ones
  constexpr bool TILED = MODE != 2;
  __shared__ int s_round0[kMaxLdsClasses + 1];
  __shared__ int s_tbase[kMaxLdsClasses + 1];
  __shared__ int64_t s_range[2 * kMaxLdsClasses];
  for (int i = threadIdx.x; i <= R.nclass; i += kMfBlock) {
    s_round0[i] = R.round0[i];
    if (TILED) s_tbase[i] = R.tbase[i];
  }
  for (int i = threadIdx.x; i < 2 * R.nclass; i += kMfBlock) s_range[i] = R.range[i];
  __syncthreads();
  const T* a1 = ls + lane * KSP;
  const T* a2 = ls + TT * 64 * KSP + lane * JP;
  double dacc = 0.0;  // with partials: sum of U . V = x_e^T K_e x_e over this lane's tiles
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    int c = cls < 0 ? 0 : cls;
    while (s_round0[c + 1] <= rd) c++;
    if (c != cls) {
      if (cls >= 0) __syncthreads();
      CopyToLds(ls, fmat + (int64_t)c * fstride, (int)(fstride * sizeof(T) / 16));
      __syncthreads();
      cls = c;
    }
    const int64_t end = s_range[2 * c + 1];
    const int64_t first = s_range[2 * c] + ((int64_t)(rd - s_round0[c]) * kMfWaves + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    // unconditional gathers from clamped indices, masked afterwards
    const int64_t ev = valid ? e : first;
    const int64_t tile = TILED ? s_tbase[c] + (first - s_range[2 * c]) / 16 : 0;
    int dv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) {
      const int k = min(4 * st + grp, nd - 1);
      dv[st] = TILED ? dofs[(tile * nd + k) * 16 + col] : dofs[ev * nd + k];
    }
    const int en = TILED ? elnr[tile * 16 + col] : elnr[ev];
    double mv[NC];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) mv[cc] = TILED ? metric[(tile * NC + cc) * 16 + col] : metric[ev * NC + cc];
    V xv[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xv[st] = x[dv[st]];
    const double kap = kappa[en];
    T xk[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) xk[st] = (valid && 4 * st + grp < nd) ? (T)xv[st] : (T)0;
    // U = D^T X
    acc_t u[TT];
#pragma unroll
    for (int t = 0; t < TT; t++) u[t] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < KSP / VW; q++) {
      vec_t a[TT];
#pragma unroll
      for (int t = 0; t < TT; t++) a[t] = *reinterpret_cast<const vec_t*>(a1 + t * 64 * KSP + VW * q);
#pragma unroll
      for (int j = 0; j < VW; j++)
        if (VW * q + j < KS)
#pragma unroll
          for (int t = 0; t < TT; t++) u[t] = Mfma(a[t][j], xk[VW * q + j], u[t]);
    }
    // V = (kappa G) U, per basis function k in this lane
    T g[NC];
#pragma unroll
    for (int cc = 0; cc < NC; cc++) g[cc] = valid ? (T)(kap * mv[cc]) : (T)0;
    T v[JP];
#pragma unroll
    for (int j = 0; j < JP; j++) v[j] = (T)0;
#pragma unroll
    for (int kl = 0; kl < KG; kl++) {
      T uu[DIM], vv[DIM];
#pragma unroll
      for (int d = 0; d < DIM; d++) uu[d] = u[(kl * DIM + d) / 4][(kl * DIM + d) % 4];
      ApplyMetric<DIM, T>(g, uu, vv);
#pragma unroll
      for (int d = 0; d < DIM; d++) v[kl * DIM + d] = vv[d];
    }
    if (MODE == 0 && partials) {
      T ld = (T)0;
#pragma unroll
      for (int j = 0; j < NSL; j++) ld += u[j / 4][j % 4] * v[j];
      dacc += (double)ld;
    }
    // Y = D V
    acc_t yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < JP / VW; q++) {
      vec_t a[MT];
#pragma unroll
      for (int m = 0; m < MT; m++) a[m] = *reinterpret_cast<const vec_t*>(a2 + m * 64 * JP + VW * q);
#pragma unroll
      for (int j = 0; j < VW; j++)
        if (VW * q + j < NSL)
#pragma unroll
          for (int m = 0; m < MT; m++) yacc[m] = Mfma(a[m][j], v[VW * q + j], yacc[m]);
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + AccRow<T>(grp, r);
          if (dof < nd) {
            const V val = (V)yacc[m][r];
            if (MODE == 0) out[e * nd + dof] = val;
            else if (MODE == 1) AtomicAddV(&out[dofs[(tile * nd + dof) * 16 + col]], (V)(s * (double)val));
            else {  // colour-ordered: no other element of the launch touches this dof
              V* yp = &out[dofs[e * nd + dof]];
              *yp = (V)((double)*yp + s * (double)val);
            }
          }
        }
    }
  }
  if (MODE == 0 && partials) {
    __shared__ double s_dot[kMfWaves];
    StorePartial<kMfBlock>(dacc, partials, s_dot);
  }
}

// 16-bit operands (BF16 / FP16), FP32 accumulation and FP32 metric, operand
// layout of v_mfma_f32_16x16x32_{bf16,f16} as in MfBf16Kernel (result rows
// 4 (l/16) + r).  Slots as in MfFacKernel with an even number TT of U tiles:
// the backward K step ks2 takes lane group grp's slots sigma = 8 ks2 + j
// (tiles 2 ks2 and 2 ks2 + 1) as its K rows 32 ks2 + 8 grp + j.  LDS:
//   A1[t][ks][grp][col][8]   row col of tile t (slot 4 t + col % 4 of lane
//                            group col / 4), K = dofs 32 ks + 8 grp + j
//   A2[m][ks2][grp][col][8]  row = dof 16 m + col, K = slots 8 ks2 + j
// FP16 scales each element's x and metric by powers of two as MfRef16Kernel.
// ND (dofs per element) is a template parameter: K steps and lane groups
// beyond it issue no loads (3D p=4: 10 of 40 load instructions per tile).
// SPLIT (FP16 only): three-term split products (Ootomo and Yokota 2022):
// D = D_hi + D_lo (both FP16, the A blocks of D_lo after those of D_hi in
// LDS, fstride counts both), x and V split the same way in registers, and
// each product is D_hi X_hi + D_hi X_lo + D_lo X_hi in the FP32
// accumulators: about 22 significant bits, the FP32 kernel's accuracy, at
// three times the 16-bit MFMA work (still a few percent of the MFMA pipe).
template <class H, int DIM, int KS1, int MT, int KG, int ND, int MODE, class V, bool SPLIT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void MfFac16Kernel(
    const uint16_t* __restrict__ fmat, int64_t fstride, MfRounds R, const int* __restrict__ dofs,
    const double* __restrict__ metric, const double* __restrict__ kappa, const int* __restrict__ elnr, int nd,
    const V* __restrict__ x, V* __restrict__ out, double s, const int* __restrict__ stop,
    double* __restrict__ partials) {
  if (stop && *stop) return;
  typedef typename Vec8<H>::type h8;
  constexpr int NC = DIM * (DIM + 1) / 2;
  constexpr int NSL = DIM * KG;
  constexpr int TT = ((NSL + 3) / 4 + 1) / 2 * 2;
  constexpr int KS2 = TT / 2;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, grp = lane >> 4;
  const int wg = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int r0 = (int)((int64_t)wg * R.nrounds / gridDim.x);
  const int r1 = (int)((int64_t)(wg + 1) * R.nrounds / gridDim.x);
  constexpr bool TILED = MODE != 2;  // as in MfFacKernel
  __shared__ int s_round0[kMaxLdsClasses + 1];
  __shared__ int s_tbase[kMaxLdsClasses + 1];
  __shared__ int64_t s_range[2 * kMaxLdsClasses];
  for (int i = threadIdx.x; i <= R.nclass; i += (64 * WAVES)) {
    s_round0[i] = R.round0[i];
    if (TILED) s_tbase[i] = R.tbase[i];
  }
  for (int i = threadIdx.x; i < 2 * R.nclass; i += (64 * WAVES)) s_range[i] = R.range[i];
  __syncthreads();
  const h8* a1 = reinterpret_cast<const h8*>(lds_raw) + lane;
  const h8* a2 = a1 + TT * KS1 * 64;
  // D_lo: the second half of the class's block
  const h8* a1l = a1 + (TT * KS1 + MT * KS2) * 64;
  const h8* a2l = a1l + TT * KS1 * 64;
  // with partials: sum of U . V (V as rounded for the backward product, the
  // FP16 scaling undone) = x_e^T K_e x_e over this lane's tiles
  double dacc = 0.0;
  int cls = -1;
  for (int rd = r0; rd < r1; rd++) {
    int cl = cls < 0 ? 0 : cls;
    while (s_round0[cl + 1] <= rd) cl++;
    if (cl != cls) {
      if (cls >= 0) __syncthreads();
      // fstride 16-bit words, a multiple of 8
      CopyToLds<64 * WAVES>(lds_raw, fmat + (int64_t)cl * fstride, (int)(fstride / 8));
      __syncthreads();
      cls = cl;
    }
    const int64_t end = s_range[2 * cl + 1];
    const int64_t first = s_range[2 * cl] + ((int64_t)(rd - s_round0[cl]) * WAVES + wave) * 16;
    if (first >= end) continue;
    const int64_t e = first + col;
    const bool valid = e < end;
    const int64_t ev = valid ? e : first;
    const int64_t tile = TILED ? s_tbase[cl] + (first - s_range[2 * cl]) / 16 : 0;
    // K row 32 ks + 8 grp + j; (ks, j) with 32 ks + j >= ND is padding in
    // every lane group and loads nothing
    int dv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        if (32 * ks + j >= ND) continue;
        const int k = min(32 * ks + 8 * grp + j, ND - 1);
        dv[ks][j] = TILED ? dofs[(tile * ND + k) * 16 + col] : dofs[ev * ND + k];
      }
    const int en = TILED ? elnr[tile * 16 + col] : elnr[ev];
    double mv[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) mv[c] = TILED ? metric[(tile * NC + c) * 16 + col] : metric[ev * NC + c];
    V xv[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xv[ks][j] = 32 * ks + j < ND ? x[dv[ks][j]] : (V)0;
    const double kap = kappa[en];
    float xf[KS1][8];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) xf[ks][j] = (valid && 32 * ks + 8 * grp + j < ND) ? (float)xv[ks][j] : 0.0f;
    float g[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) g[c] = valid ? (float)(kap * mv[c]) : 0.0f;
    int escale = 0, edot = 0;
    if constexpr (!std::is_same<H, __bf16>::value) {
      float xm = 0.0f, gm = 0.0f;
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xm = fmaxf(xm, fabsf(xf[ks][j]));
      xm = fmaxf(xm, __shfl_xor(xm, 16, 64));
      xm = fmaxf(xm, __shfl_xor(xm, 32, 64));
#pragma unroll
      for (int c = 0; c < NC; c++) gm = fmaxf(gm, fabsf(g[c]));
      int ex = 0, eg = 0;
      (void)frexpf(xm, &ex);
      (void)frexpf(gm, &eg);
#pragma unroll
      for (int ks = 0; ks < KS1; ks++)
#pragma unroll
        for (int j = 0; j < 8; j++) xf[ks][j] = ldexpf(xf[ks][j], -ex);
#pragma unroll
      for (int c = 0; c < NC; c++) g[c] = ldexpf(g[c], -eg);
      escale = ex + eg;
      edot = 2 * ex + eg;
    }
    h8 xg[KS1], xl[KS1];
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        xg[ks][j] = (H)xf[ks][j];
        if (SPLIT) xl[ks][j] = (H)(xf[ks][j] - (float)xg[ks][j]);
      }
    // U = D^T X
    float4v u[TT];
#pragma unroll
    for (int t = 0; t < TT; t++) u[t] = float4v{0, 0, 0, 0};
#pragma unroll
    for (int ks = 0; ks < KS1; ks++)
#pragma unroll
      for (int t = 0; t < TT; t++) {
        u[t] = Mfma16(a1[(t * KS1 + ks) * 64], xg[ks], u[t]);
        if (SPLIT) {
          u[t] = Mfma16(a1[(t * KS1 + ks) * 64], xl[ks], u[t]);
          u[t] = Mfma16(a1l[(t * KS1 + ks) * 64], xg[ks], u[t]);
        }
      }
    // V = (kappa G) U in FP32, rounded once into the B operand
    float v[4 * TT];
#pragma unroll
    for (int j = 0; j < 4 * TT; j++) v[j] = 0.0f;
#pragma unroll
    for (int kl = 0; kl < KG; kl++) {
      float uu[DIM], vv[DIM];
#pragma unroll
      for (int d = 0; d < DIM; d++) uu[d] = u[(kl * DIM + d) / 4][(kl * DIM + d) % 4];
      ApplyMetric<DIM, float>(g, uu, vv);
#pragma unroll
      for (int d = 0; d < DIM; d++) v[kl * DIM + d] = vv[d];
    }
    h8 vb[KS2], vl[KS2];
#pragma unroll
    for (int k2 = 0; k2 < KS2; k2++)
#pragma unroll
      for (int j = 0; j < 8; j++) {
        vb[k2][j] = (H)v[8 * k2 + j];
        if (SPLIT) vl[k2][j] = (H)(v[8 * k2 + j] - (float)vb[k2][j]);
      }
    if (MODE == 0 && partials) {
      float ld = 0.0f;
#pragma unroll
      for (int j = 0; j < NSL; j++)
        ld += u[j / 4][j % 4] * (SPLIT ? (float)vb[j / 8][j % 8] + (float)vl[j / 8][j % 8] : (float)vb[j / 8][j % 8]);
      dacc += ldexp((double)ld, edot);
    }
    // Y = D V
    float4v yacc[MT];
#pragma unroll
    for (int m = 0; m < MT; m++) yacc[m] = float4v{0, 0, 0, 0};
#pragma unroll
    for (int k2 = 0; k2 < KS2; k2++)
#pragma unroll
      for (int m = 0; m < MT; m++) {
        yacc[m] = Mfma16(a2[(m * KS2 + k2) * 64], vb[k2], yacc[m]);
        if (SPLIT) {
          yacc[m] = Mfma16(a2[(m * KS2 + k2) * 64], vl[k2], yacc[m]);
          yacc[m] = Mfma16(a2l[(m * KS2 + k2) * 64], vb[k2], yacc[m]);
        }
      }
    if (valid) {
#pragma unroll
      for (int m = 0; m < MT; m++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int dof = 16 * m + 4 * grp + r;
          if (dof < ND) {
            const V val = (V)ldexpf(yacc[m][r], escale);
            if (MODE == 0) out[(tile * ND + dof) * 16 + col] = val;
            else if (MODE == 1) AtomicAddV(&out[dofs[(tile * ND + dof) * 16 + col]], (V)(s * (double)val));
            else {  // colour-ordered: no other element of the launch touches this dof
              V* yp = &out[dofs[e * ND + dof]];
              *yp = (V)((double)*yp + s * (double)val);
            }
          }
        }
    }
  }
  if (MODE == 0 && partials) {
    __shared__ double s_dot[WAVES];
    StorePartial<(64 * WAVES)>(dacc, partials, s_dot);
  }
}

// Scalar fallback for shapes the MFMA path does not cover: one thread per
// element, FP64, element vectors staged in LDS columns.
constexpr int kScalarBlock = 64;
constexpr int kScalarMaxNd = 64;

template <int DIM, int MODE, class V>
__global__ __launch_bounds__(kScalarBlock) void MfScalarKernel(
    const double* __restrict__ bref /*[nq][nd][dim]*/, const double* __restrict__ weights, int nq,
    const int* __restrict__ dofs, const double* __restrict__ metric, const double* __restrict__ kappa,
    const int* __restrict__ elnr, int64_t begin, int64_t end, int nd, const V* __restrict__ x,
    V* __restrict__ out, double s, const int* __restrict__ stop) {
  if (stop && *stop) return;
  constexpr int NC = DIM * (DIM + 1) / 2;
  __shared__ double xe[kScalarMaxNd][kScalarBlock];
  __shared__ double ye[kScalarMaxNd][kScalarBlock];
  const int tid = threadIdx.x;
  for (int64_t base = begin + (int64_t)blockIdx.x * kScalarBlock; base < end; base += (int64_t)gridDim.x * kScalarBlock) {
    const int64_t e = base + tid;
    if (e < end) {
      for (int i = 0; i < nd; i++) { xe[i][tid] = (double)x[dofs[e * nd + i]]; ye[i][tid] = 0.0; }
      double g[NC];
      const double kap = kappa[elnr[e]];
      for (int c = 0; c < NC; c++) g[c] = kap * metric[e * NC + c];
      for (int q = 0; q < nq; q++) {
        double u[DIM] = {}, v[DIM];
        const double* Bq = bref + (size_t)q * nd * DIM;
        for (int i = 0; i < nd; i++)
          for (int d = 0; d < DIM; d++) u[d] += Bq[i * DIM + d] * xe[i][tid];
        ApplyMetric<DIM, double>(g, u, v);
        for (int d = 0; d < DIM; d++) v[d] *= weights[q];
        for (int i = 0; i < nd; i++) {
          double acc = 0.0;
          for (int d = 0; d < DIM; d++) acc += Bq[i * DIM + d] * v[d];
          ye[i][tid] += acc;
        }
      }
      for (int i = 0; i < nd; i++) {
        if (MODE == 0) out[e * nd + i] = (V)ye[i][tid];
        else AtomicAddV(&out[dofs[e * nd + i]], (V)(s * ye[i][tid]));
      }
    }
  }
}

// y[r] = (ACC ? y[r] : 0) + s * sum of the contributions of dof r, in slot
// order; with DOT also the per-workgroup partial sums of x[r] * y[r] (the
// <s, A s> of CG) into partials[].  With fin: the first min(64, grid)
// workgroups also fold the nfin partials fin[] (the contraction's) into
// fout[] (consecutive groups, summed in index order by one thread).
constexpr int kGatherBlock = 256;
constexpr int kGatherWidth = 4;  // 8 measured slower (FP32 gather 38.5 vs 34.7 us)

template <class C, class V, bool DOT>
__global__ __launch_bounds__(kGatherBlock) void MfGatherKernel(
    const int* __restrict__ rowptr, const int* __restrict__ slots, const int* __restrict__ order,
    const C* __restrict__ contrib, V* __restrict__ y, int64_t n, double s, int accumulate,
    const int* __restrict__ stop, const V* __restrict__ x, double* __restrict__ partials,
    const double* __restrict__ fin = nullptr, int nfin = 0, double* __restrict__ fout = nullptr) {
  if (stop && *stop) return;
  if (fin && threadIdx.x == 0) {
    const int nout = min(kFoldedPartials, (int)gridDim.x);
    if ((int)blockIdx.x < nout) {
      const int per = (nfin + nout - 1) / nout;
      const int b0 = (int)blockIdx.x * per, b1 = min(b0 + per, nfin);
      double v = 0.0;
      for (int i = b0; i < b1; i++) v += fin[i];
      fout[blockIdx.x] = v;
    }
  }
  double dotacc = 0.0;
  // contiguous ranges of the gather order per XCD
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = (int64_t)XcdContiguousBlock(blockIdx.x, gridDim.x) * per;
  const int64_t r1 = min(r0 + per, n);
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kGatherBlock) {
    double acc = 0.0;
    const int k0 = rowptr[r], k1 = rowptr[r + 1];
    // kGatherWidth slots at a time, loads unconditional (clamped to the last
    // slot) so they are in flight together; summed in slot order
    for (int k = k0; k < k1; k += kGatherWidth) {
      int sl[kGatherWidth];
      C cv[kGatherWidth];
#pragma unroll
      for (int j = 0; j < kGatherWidth; j++) sl[j] = slots[min(k + j, k1 - 1)];
#pragma unroll
      for (int j = 0; j < kGatherWidth; j++) cv[j] = contrib[sl[j]];
#pragma unroll
      for (int j = 0; j < kGatherWidth; j++)
        if (k + j < k1) acc += (double)cv[j];
    }
    const int d = order[r];
    const double v = s * acc;
    const V yv = accumulate ? (V)((double)y[d] + v) : (V)v;
    y[d] = yv;
    if (DOT) dotacc += (double)x[d] * (double)yv;
  }
  if (DOT) {
    __shared__ double lds[kGatherBlock / 64];
    StorePartial<kGatherBlock>(dotacc, partials, lds);
  }
}

template <class V>
__global__ void ScaleFillKernel(V* __restrict__ y, int64_t n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    y[r] = (V)0;
}

// Diagonal: per element, diag_i = kappa sum_q w_q B_qi^T G B_qi, into contributions.
template <int DIM>
__global__ void MfDiagKernel(const double* __restrict__ bref, const double* __restrict__ weights, int nq,
                             const double* __restrict__ metric, const double* __restrict__ kappa,
                             const int* __restrict__ elnr, int64_t begin, int64_t end, int nd,
                             double* __restrict__ contrib) {
  constexpr int NC = DIM * (DIM + 1) / 2;
  for (int64_t e = begin + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < end;
       e += (int64_t)gridDim.x * blockDim.x) {
    double g[NC];
    const double kap = kappa[elnr[e]];
    for (int c = 0; c < NC; c++) g[c] = kap * metric[e * NC + c];
    for (int i = 0; i < nd; i++) {
      double acc = 0.0;
      for (int q = 0; q < nq; q++) {
        const double* b = bref + ((size_t)q * nd + i) * DIM;
        double v[DIM];
        ApplyMetric<DIM, double>(g, b, v);
        double dd = 0.0;
        for (int d = 0; d < DIM; d++) dd += b[d] * v[d];
        acc += weights[q] * dd;
      }
      contrib[e * nd + i] = acc;
    }
  }
}

// g_e = -int_e grad(lam).grad(u) = -sum_q w_q (B_q^T lam_e)^T G_e (B_q^T u_e)
template <int DIM>
__global__ void MfElementGradKernel(const double* __restrict__ bref, const double* __restrict__ weights, int nq,
                                    const int* __restrict__ dofs, const double* __restrict__ metric,
                                    const int* __restrict__ elnr, int64_t begin, int64_t end, int nd,
                                    const double* __restrict__ lam, const double* __restrict__ u,
                                    double* __restrict__ grad) {
  constexpr int NC = DIM * (DIM + 1) / 2;
  for (int64_t e = begin + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < end;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double* g = metric + e * NC;
    double acc = 0.0;
    for (int q = 0; q < nq; q++) {
      double a[DIM] = {}, b[DIM] = {}, gb[DIM];
      for (int i = 0; i < nd; i++) {
        const int dof = dofs[e * nd + i];
        const double li = lam[dof], ui = u[dof];
        const double* B = bref + ((size_t)q * nd + i) * DIM;
        for (int d = 0; d < DIM; d++) { a[d] += B[d] * li; b[d] += B[d] * ui; }
      }
      ApplyMetric<DIM, double>(g, b, gb);
      double dd = 0.0;
      for (int d = 0; d < DIM; d++) dd += a[d] * gb[d];
      acc += weights[q] * dd;
    }
    grad[elnr[e]] = -acc;
  }
}

// ----------------------------------------------------------------- host side

static bool SupportedNdp(int ndp) {
  return ndp == 4 || ndp == 8 || ndp == 12 || ndp == 16 || ndp == 20 || ndp == 24 || ndp == 36;
}

static bool SupportedBf16(int ks1, int mt) {
  return (ks1 == 1 && (mt == 1 || mt == 2)) || (ks1 == 2 && mt == 3);
}

// Shapes with an instance of the factored kernels (H1 p = 1..5 in 2D, 1..4
// in 3D): padded dofs NDP (16-bit: K steps KS1, dof tiles MT and the dofs per
// element ND) and basis functions per lane group KG.
static bool SupportedFac(int dim, int ndp, int kg) {
  if (dim == 2)
    return (ndp == 4 && kg == 1) || (ndp == 8 && kg == 1) || (ndp == 12 && kg == 2) || (ndp == 16 && kg == 3) ||
           (ndp == 24 && kg == 4);
  return (ndp == 4 && kg == 1) || (ndp == 12 && kg == 1) || (ndp == 20 && kg == 3) || (ndp == 36 && kg == 5);
}
static bool SupportedFac16(int dim, int ks1, int mt, int kg, int nd) {
  auto is = [&](int k, int m, int g, int n) { return ks1 == k && mt == m && kg == g && nd == n; };
  if (dim == 2)
    return is(1, 1, 1, 3) || is(1, 1, 1, 6) || is(1, 1, 2, 10) || is(1, 1, 3, 15) || is(1, 2, 4, 21);
  return is(1, 1, 1, 4) || is(1, 1, 1, 10) || is(1, 2, 3, 20) || is(2, 3, 5, 35);
}

// Factor of each class's reference matrices (see MfFacKernel): the rows
// Bw[(d,i)][q] = b_{q,i,d} sqrt(w_q) are orthonormalised by Gram-Schmidt with
// pivoting on the largest remaining row (each new vector orthogonalised twice
// against the basis), stopping when every remaining row is below 1e-10 of
// the largest row; D[(d,i)][k] = Bw[(d,i)] . phi_k.  Returns the largest rank
// over the classes (D has that many columns, zero beyond a class's own rank),
// or 0 if a weight is not positive or D Phi misses Bw by more than 1e-12
// relative.  D is [class][d][i][rank].
static int FactorGradients(int dim, int nd, int nq, int nclass, const double* bmat, const double* w,
                           std::vector<double>& D) {
  for (int q = 0; q < nq; q++)
    if (!(w[q] > 0.0)) return 0;
  const int nr = dim * nd;
  const size_t bsz = (size_t)nq * nd * dim;
  std::vector<std::vector<double>> phis(nclass);
  std::vector<double> Bw((size_t)nr * nq);
  int rank = 0;
  auto dot = [&](const double* a, const double* b) {
    double t = 0.0;
    for (int q = 0; q < nq; q++) t += a[q] * b[q];
    return t;
  };
  for (int c = 0; c < nclass; c++) {
    for (int d = 0; d < dim; d++)
      for (int i = 0; i < nd; i++)
        for (int q = 0; q < nq; q++)
          Bw[((size_t)d * nd + i) * nq + q] = bmat[c * bsz + ((size_t)q * nd + i) * dim + d] * std::sqrt(w[q]);
    std::vector<double> res(Bw);
    std::vector<double>& phi = phis[c];
    double norm0 = 0.0;
    for (int r = 0; r < nr; r++) norm0 = std::max(norm0, std::sqrt(dot(&Bw[(size_t)r * nq], &Bw[(size_t)r * nq])));
    if (norm0 == 0.0) continue;
    for (int k = 0; k < std::min(nr, nq); k++) {
      int best = -1;
      double bn = 0.0;
      for (int r = 0; r < nr; r++) {
        const double t = std::sqrt(dot(&res[(size_t)r * nq], &res[(size_t)r * nq]));
        if (t > bn) bn = t, best = r;
      }
      if (best < 0 || bn <= 1e-10 * norm0) break;
      std::vector<double> v(res.begin() + (size_t)best * nq, res.begin() + (size_t)(best + 1) * nq);
      for (int pass = 0; pass < 2; pass++) {
        for (size_t j = 0; j < phi.size() / nq; j++) {
          const double t = dot(v.data(), &phi[j * nq]);
          for (int q = 0; q < nq; q++) v[q] -= t * phi[j * nq + q];
        }
        const double vn = std::sqrt(dot(v.data(), v.data()));
        for (int q = 0; q < nq; q++) v[q] /= vn;
      }
      phi.insert(phi.end(), v.begin(), v.end());
      for (int r = 0; r < nr; r++) {
        double* rr = &res[(size_t)r * nq];
        const double t = dot(rr, v.data());
        for (int q = 0; q < nq; q++) rr[q] -= t * v[q];
      }
    }
    rank = std::max(rank, (int)(phi.size() / nq));
  }
  if (rank == 0) return 0;
  D.assign((size_t)nclass * nr * rank, 0.0);
  for (int c = 0; c < nclass; c++) {
    const std::vector<double>& phi = phis[c];
    const int rc = (int)(phi.size() / nq);
    double norm0 = 0.0, err = 0.0;
    for (int d = 0; d < dim; d++)
      for (int i = 0; i < nd; i++) {
        std::vector<double> row(nq), rec(nq, 0.0);
        for (int q = 0; q < nq; q++) row[q] = bmat[c * bsz + ((size_t)q * nd + i) * dim + d] * std::sqrt(w[q]);
        double* Dr = &D[((size_t)c * nr + (size_t)d * nd + i) * rank];
        for (int k = 0; k < rc; k++) {
          Dr[k] = dot(row.data(), &phi[(size_t)k * nq]);
          for (int q = 0; q < nq; q++) rec[q] += Dr[k] * phi[(size_t)k * nq + q];
        }
        for (int q = 0; q < nq; q++) {
          norm0 = std::max(norm0, std::fabs(row[q]));
          err = std::max(err, std::fabs(row[q] - rec[q]));
        }
      }
    if (err > 1e-12 * norm0) return 0;
  }
  return rank;
}


template <int MODE, class V>
static void LaunchElements(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                           bool prepare);

MatFreeOp::MatFreeOp(int dim_, int nd_, int nq_, int64_t ndof, int nclass, const int64_t* off, const double* bmat,
                     const double* w, const int* hdofs, const double* hmetric, const double* hkappa,
                     const int* helnr, int prec, int scat)
    : ngsamd_op_s(OpKind::MatrixFree) {
  NGSAMD_REQUIRE(dim_ == 2 || dim_ == 3, "ngsamd_mf_create: dim must be 2 or 3");
  NGSAMD_REQUIRE(nd_ >= 1 && nd_ <= kScalarMaxNd, "ngsamd_mf_create: nd out of range (1..64)");
  NGSAMD_REQUIRE(nq_ >= 1 && nclass >= 1, "ngsamd_mf_create: empty rule or no classes");
  NGSAMD_REQUIRE(prec == NGSAMD_PREC_F64 || prec == NGSAMD_PREC_F32 || prec == NGSAMD_PREC_BF16 ||
                     prec == NGSAMD_PREC_F16 || prec == NGSAMD_PREC_F16X3,
                 "ngsamd_mf_create: unknown precision");
  NGSAMD_REQUIRE(scat == NGSAMD_SCATTER_GATHER || scat == NGSAMD_SCATTER_ATOMIC, "ngsamd_mf_create: unknown scatter");
  NGSAMD_REQUIRE(ndof >= 0 && ndof < (int64_t(1) << 31), "ngsamd_mf_create: ndof out of range");
  dim = dim_;
  nd = nd_;
  nq = nq_;
  ncomp = dim * (dim + 1) / 2;
  precision = prec;
  scatter = scat;
  height = width = ndof;
  nel = off[nclass];
  NGSAMD_REQUIRE(off[0] == 0 && nel >= 0 && nel * nd < (int64_t(1) << 31), "ngsamd_mf_create: bad class offsets");
  for (int c = 0; c < nclass; c++) {
    NGSAMD_REQUIRE(off[c + 1] >= off[c], "ngsamd_mf_create: class offsets not monotone");
    classes.push_back({off[c], off[c + 1]});
  }
  for (int64_t i = 0; i < nel * nd; i++)
    NGSAMD_REQUIRE(hdofs[i] >= 0 && hdofs[i] < ndof, "ngsamd_mf_create: dof number out of range");
  std::vector<int> el(nel);
  ne_global = 0;
  for (int64_t e = 0; e < nel; e++) {
    el[e] = helnr ? helnr[e] : (int)e;
    NGSAMD_REQUIRE(el[e] >= 0, "ngsamd_mf_create: negative element number");
    ne_global = std::max<int64_t>(ne_global, el[e] + 1);
  }

  ndp = (nd + 3) / 4 * 4;
  nqp = (nq + 15) / 16 * 16;
  const int qs = dim * ndp + 1;
  bstride = (int64_t)nqp * qs;
  // MFMA path if the padded shape is instantiated and B fits in LDS
  // the 16-bit operand formats (split FP16 rounds its FP16 part like FP16)
  const bool half = prec == NGSAMD_PREC_BF16 || prec == NGSAMD_PREC_F16 || prec == NGSAMD_PREC_F16X3;
  if (half) {
    nqc = (nq + 31) / 32;
    k1p = (nd + 7) / 8 * 8;
    ks1 = (nd + 31) / 32;
    mt16 = (nd + 15) / 16;
    bstride = (int64_t)nqc * 32 * dim * k1p + (int64_t)nqc * dim * mt16 * 16 * 32;
    use_mfma = SupportedBf16(ks1, mt16) && (size_t)bstride * 2 <= (size_t)DeviceMaxLds();
  } else {
    const size_t esize = prec == NGSAMD_PREC_F64 ? 8 : 4;
    use_mfma = SupportedNdp(ndp) && (size_t)bstride * esize <= (size_t)DeviceMaxLds();
  }

  // Reference-matrix kernels by default (affine elements, one coefficient
  // value per element: exact), in their factored form where an instance
  // exists; NGSAMD_MF_KERNEL=ref keeps the sum over the S_c, =quad selects the
  // per-quadrature-point kernels.
  {
    const char* kenv = getenv("NGSAMD_MF_KERNEL");
    NGSAMD_REQUIRE(!kenv || std::string(kenv) == "quad" || std::string(kenv) == "ref" || std::string(kenv) == "fac",
                   "NGSAMD_MF_KERNEL must be fac, ref or quad");
    ref_kernel = use_mfma && !(kenv && std::string(kenv) == "quad");
  }
  // S[c][k][i][j] in FP64, k over the packed metric components: the operand
  // of the reference-matrix kernels (in their layouts below) and of the CSR
  // assembly (smat_asm, as is)
  const int nc = ncomp;
  std::vector<double> S((size_t)nclass * nc * nd * nd, 0.0);
  {
    const int pd[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    const int pd2[3][2] = {{0, 0}, {1, 1}, {0, 1}};
    const size_t bsz0 = (size_t)nq * nd * dim;
    for (int c = 0; c < nclass; c++)
      for (int k = 0; k < nc; k++) {
        const int d0 = dim == 3 ? pd[k][0] : pd2[k][0], d1 = dim == 3 ? pd[k][1] : pd2[k][1];
        double* Sk = S.data() + ((size_t)c * nc + k) * nd * nd;
        for (int q = 0; q < nq; q++) {
          const double* bq = bmat + c * bsz0 + (size_t)q * nd * dim;
          for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++) {
              const double v = d0 == d1 ? bq[i * dim + d0] * bq[j * dim + d0]
                                        : bq[i * dim + d0] * bq[j * dim + d1] + bq[i * dim + d1] * bq[j * dim + d0];
              Sk[(size_t)i * nd + j] += w[q] * v;
            }
        }
      }
  }
  smat_asm.alloc(std::max<size_t>(S.size(), 1));
  if (!S.empty()) smat_asm.upload(S.data(), S.size());
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  if (ref_kernel) {
    if (half) {
      const int rs = ks1 * 32 + 8;
      sstride = (int64_t)nc * nd * rs;
      std::vector<uint16_t> lay((size_t)sstride * nclass, 0);
      for (int c = 0; c < nclass; c++)
        for (int k = 0; k < nc; k++)
          for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++) {
              const double v = S[(((size_t)c * nc + k) * nd + i) * nd + j];
              lay[(size_t)c * sstride + ((size_t)k * nd + i) * rs + j] =
                  prec != NGSAMD_PREC_BF16 ? ToF16(v) : ToBf16(v);
            }
      smat16.alloc(lay.size());
      smat16.upload(lay.data(), lay.size());
    } else {
      if (prec == NGSAMD_PREC_F64) {
        const int rs = ndp + 1;
        sstride = ((int64_t)nc * nd * rs + 1) / 2 * 2;  // whole 16-byte words
        std::vector<double> lay((size_t)sstride * nclass, 0.0);
        for (int c = 0; c < nclass; c++)
          for (int k = 0; k < nc; k++)
            for (int i = 0; i < nd; i++)
              for (int j = 0; j < nd; j++)
                lay[(size_t)c * sstride + ((size_t)k * nd + i) * rs + j] = S[(((size_t)c * nc + k) * nd + i) * nd + j];
        smat64.alloc(lay.size());
        smat64.upload(lay.data(), lay.size());
      } else {
        // lane order [c][m][grp][col][KSP]: lane (col, grp) of dof tile m
        // reads S[16m + col][4 st + grp] for st < KS contiguously
        const int ks = ndp / 4, mt = (ndp + 15) / 16, ksp = (ks + 3) / 4 * 4;
        sstride = (int64_t)nc * mt * 64 * ksp;
        std::vector<float> lay((size_t)sstride * nclass, 0.0f);
        for (int c = 0; c < nclass; c++)
          for (int k = 0; k < nc; k++)
            for (int m = 0; m < mt; m++)
              for (int grp = 0; grp < 4; grp++)
                for (int col = 0; col < 16; col++)
                  for (int st = 0; st < ks; st++) {
                    const int i = 16 * m + col, j = 4 * st + grp;
                    if (i < nd && j < nd)
                      lay[(size_t)c * sstride + ((((size_t)k * mt + m) * 4 + grp) * 16 + col) * ksp + st] =
                          (float)S[(((size_t)c * nc + k) * nd + i) * nd + j];
                  }
        smat32.alloc(lay.size());
        smat32.upload(lay.data(), lay.size());
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
      }
    }
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  }
  // factored kernels (default where an instance exists; NGSAMD_MF_KERNEL=ref
  // keeps the sum over S_c)
  if (ref_kernel && nclass <= kMaxLdsClasses && !(getenv("NGSAMD_MF_KERNEL") &&
                                                  std::string(getenv("NGSAMD_MF_KERNEL")) == "ref")) {
    std::vector<double> D;
    const int rank = FactorGradients(dim, nd, nq, nclass, bmat, w, D);
    const int kg = (rank + 3) / 4;
    const int nsl = dim * kg;
    auto dval = [&](int c, int g, int sigma, int i) -> double {
      if (sigma >= nsl || i >= nd) return 0.0;
      const int k = g * kg + sigma / dim, d = sigma % dim;
      if (k >= rank) return 0.0;
      return D[(((size_t)c * dim + d) * nd + i) * rank + k];
    };
    if (rank > 0 && half && SupportedFac16(dim, ks1, mt16, kg, nd)) {
      const int tt = ((nsl + 3) / 4 + 1) / 2 * 2, ks2 = tt / 2;
      // split FP16: the block of D_hi, then the block of D_lo
      const bool split = prec == NGSAMD_PREC_F16X3;
      const int64_t blk = (int64_t)(tt * ks1 + mt16 * ks2) * 64 * 8;
      fstride = split ? 2 * blk : blk;
      if ((size_t)fstride * 2 <= (size_t)DeviceMaxLds()) {
        std::vector<uint16_t> lay((size_t)fstride * nclass, 0);
        // part 0: the 16-bit value (D_hi for split FP16), part 1: D_lo
        auto h = [&](double v, int part) -> uint16_t {
          if (!split) return prec == NGSAMD_PREC_F16 ? ToF16(v) : ToBf16(v);
          uint16_t hi, lo;
          SplitF16(v, &hi, &lo);
          return part ? lo : hi;
        };
        for (int c = 0; c < nclass; c++)
          for (int part = 0; part < (split ? 2 : 1); part++) {
            uint16_t* a1 = lay.data() + (size_t)c * fstride + (size_t)part * blk;
            uint16_t* a2 = a1 + (size_t)tt * ks1 * 64 * 8;
            for (int t = 0; t < tt; t++)
              for (int k = 0; k < ks1; k++)
                for (int l = 0; l < 64; l++)
                  for (int j = 0; j < 8; j++) {
                    const int gl = l / 16, cl = l % 16;
                    a1[(((size_t)t * ks1 + k) * 64 + l) * 8 + j] =
                        h(dval(c, cl / 4, 4 * t + cl % 4, 32 * k + 8 * gl + j), part);
                  }
            for (int m = 0; m < mt16; m++)
              for (int k = 0; k < ks2; k++)
                for (int l = 0; l < 64; l++)
                  for (int j = 0; j < 8; j++) {
                    const int gl = l / 16, cl = l % 16;
                    a2[(((size_t)m * ks2 + k) * 64 + l) * 8 + j] = h(dval(c, gl, 8 * k + j, 16 * m + cl), part);
                  }
          }
        fmat16.alloc(lay.size());
        fmat16.upload(lay.data(), lay.size());
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
        fac_kernel = true;
      }
    } else if (rank > 0 && !half && SupportedFac(dim, ndp, kg)) {
      const bool f64 = prec == NGSAMD_PREC_F64;
      const int vw = f64 ? 2 : 4;
      const int ks = ndp / 4, mt = (ndp + 15) / 16, tt = (nsl + 3) / 4;
      const int ksp = (ks + vw - 1) / vw * vw, jp = (nsl + vw - 1) / vw * vw;
      fstride = (int64_t)tt * 64 * ksp + (int64_t)mt * 64 * jp;
      if ((size_t)fstride * (f64 ? 8 : 4) <= (size_t)DeviceMaxLds()) {
        std::vector<double> lay((size_t)fstride * nclass, 0.0);
        for (int c = 0; c < nclass; c++) {
          double* a1 = lay.data() + (size_t)c * fstride;
          double* a2 = a1 + (size_t)tt * 64 * ksp;
          for (int t = 0; t < tt; t++)
            for (int l = 0; l < 64; l++)
              for (int st = 0; st < ks; st++) {
                const int gl = l / 16, cl = l % 16;
                // the slot of accumulator row cl (AccRow: FP64 grp + 4 r, FP32 4 grp + r)
                const int g = f64 ? cl % 4 : cl / 4, r = f64 ? cl / 4 : cl % 4;
                a1[((size_t)t * 64 + l) * ksp + st] = dval(c, g, 4 * t + r, 4 * st + gl);
              }
          for (int m = 0; m < mt; m++)
            for (int l = 0; l < 64; l++)
              for (int j = 0; j < nsl; j++) {
                const int gl = l / 16, cl = l % 16;
                a2[((size_t)m * 64 + l) * jp + j] = dval(c, gl, j, 16 * m + cl);
              }
        }
        if (f64) {
          fmat64.alloc(lay.size());
          fmat64.upload(lay.data(), lay.size());
          NGSAMD_HIP(hipStreamSynchronize(Stream()));
        } else {
          std::vector<float> l32(lay.begin(), lay.end());
          fmat32.alloc(l32.size());
          fmat32.upload(l32.data(), l32.size());
          NGSAMD_HIP(hipStreamSynchronize(Stream()));
        }
        fac_kernel = true;
      }
    }
    fac_rank = fac_kernel ? rank : 0;
    fac_kg = fac_kernel ? kg : 0;
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  }
  NGSAMD_REQUIRE(prec != NGSAMD_PREC_F16X3 || (use_mfma && fac_kernel),
                 "ngsamd_mf_create: NGSAMD_PREC_F16X3 needs the factored kernels (H1 triangles p <= 5, "
                 "tetrahedra p <= 4)");

  // FP64 reference gradients [class][nq][nd][dim] for the scalar kernels
  const size_t bsz = (size_t)nq * nd * dim;
  bref.alloc(bsz * nclass);
  bref.upload(bmat, bsz * nclass);
  const int nw = (nq + 31) / 32 * 32;  // covers the FP64/FP32 (16) and BF16 (32) point padding
  std::vector<double> wp(nw, 0.0);
  for (int q = 0; q < nq; q++) wp[q] = w[q];
  weights.alloc(nw);
  weights.upload(wp.data(), nw);
  if (use_mfma && half && !ref_kernel) {
    std::vector<uint16_t> lay((size_t)bstride * nclass, 0);
    const int64_t size1 = (int64_t)nqc * 32 * dim * k1p;
    for (int c = 0; c < nclass; c++) {
      uint16_t* b1 = lay.data() + (size_t)c * bstride;
      uint16_t* b2 = b1 + size1;
      for (int q = 0; q < nq; q++)
        for (int i = 0; i < nd; i++)
          for (int d = 0; d < dim; d++) {
            const double bv = bmat[c * bsz + ((size_t)q * nd + i) * dim + d];
            const uint16_t v = prec != NGSAMD_PREC_BF16 ? ToF16(bv) : ToBf16(bv);
            b1[((size_t)q * dim + d) * k1p + i] = v;
            // position of point q inside its chunk in the accumulator order
            const int ch = q / 32, ql = q % 32, h = ql / 16, grp = (ql % 16) / 4, r = ql % 4;
            const int kk = 8 * grp + 4 * h + r;
            b2[(((size_t)ch * dim + d) * mt16 * 16 + i) * 32 + kk] = v;
          }
    }
    bmat16.alloc(lay.size());
    bmat16.upload(lay.data(), lay.size());
  } else if (use_mfma && !ref_kernel) {
    std::vector<double> lay((size_t)bstride * nclass, 0.0);
    for (int c = 0; c < nclass; c++)
      for (int q = 0; q < nq; q++)
        for (int i = 0; i < nd; i++)
          for (int d = 0; d < dim; d++)
            lay[(size_t)c * bstride + (size_t)q * qs + d * ndp + i] = bmat[c * bsz + ((size_t)q * nd + i) * dim + d];
    if (prec == NGSAMD_PREC_F64) {
      bmat64.alloc(lay.size());
      bmat64.upload(lay.data(), lay.size());
    } else {
      std::vector<float> l32(lay.begin(), lay.end());
      bmat32.alloc(l32.size());
      bmat32.upload(l32.data(), l32.size());
    }
  }
  dofs.alloc((size_t)nel * nd);
  dofs.upload(hdofs, (size_t)nel * nd);
  metric.alloc((size_t)nel * ncomp);
  metric.upload(hmetric, (size_t)nel * ncomp);
  elnr.alloc(nel);
  elnr.upload(el.data(), nel);
  std::vector<double> kap(ne_global, 1.0);
  if (hkappa)
    for (int64_t e = 0; e < nel; e++) kap[el[e]] = hkappa[e];
  kappa.alloc(ne_global);
  kappa.upload(kap.data(), ne_global);

  std::vector<int> gather_slots;  // element-major slots in gather order (kept for the tile-major copy)
  // dof -> contribution slots, ascending slot order (the order of each sum)
  std::vector<int> cnt(ndof + 1, 0);
  for (int64_t i = 0; i < nel * nd; i++) cnt[hdofs[i] + 1]++;
  for (int64_t r = 0; r < ndof; r++) cnt[r + 1] += cnt[r];
  std::vector<int> sl(nel * nd), pos(cnt.begin(), cnt.end() - 1);
  for (int64_t i = 0; i < nel * nd; i++) sl[pos[hdofs[i]]++] = (int)i;
  // Gather order: dofs by the lowest global element number among their
  // elements (counting sort, stable).  The dofs of one element (vertex, edge,
  // face and cell dofs lie far apart in the numbering) are then gathered close
  // together in time and, with the XCD-contiguous grid, on one XCD, so each
  // element row of the contribution buffer is fetched into L2 about once
  // instead of once per node that reads it.  (By the lowest row of the
  // class-sorted tables instead: FP64 apply 81 -> 93 us.)
  {
    // the 16-bit factored kernels store their rows tile-major: there the key
    // is the lowest tile-major position ((tile * nd + local dof) * 16 + lane)
    // among a dof's contributions, so the dofs reading one 16-element line
    // are gathered together (BF16 apply with FP64 vectors 75 -> 71 us, split
    // FP16 with FP32 vectors 56 -> 54 us at 32^3 p=4)
    const bool tiled16 = use_mfma && fac_kernel && half;
    int64_t kmax = ne_global;
    std::vector<int> rowkey;
    if (tiled16) {
      rowkey.resize((size_t)nel * nd);
      int64_t tb = 0;
      for (int c = 0; c < nclass; c++) {
        for (int64_t pe = off[c]; pe < off[c + 1]; pe++) {
          const int64_t T = tb + (pe - off[c]) / 16, col = (pe - off[c]) % 16;
          for (int k = 0; k < nd; k++) rowkey[(size_t)pe * nd + k] = (int)((T * nd + k) * 16 + col);
        }
        tb += (off[c + 1] - off[c] + 15) / 16;
      }
      kmax = tb * nd * 16;
    }
    std::vector<int> key(ndof, (int)kmax);
    for (int64_t i = 0; i < nel * nd; i++)
      key[hdofs[i]] = std::min(key[hdofs[i]], tiled16 ? rowkey[i] : el[i / nd]);
    std::vector<int> kc(kmax + 2, 0);
    for (int64_t d = 0; d < ndof; d++) kc[key[d] + 1]++;
    for (int64_t k = 0; k <= kmax; k++) kc[k + 1] += kc[k];
    std::vector<int> order(ndof);
    for (int64_t d = 0; d < ndof; d++) order[kc[key[d]]++] = (int)d;
    // Inside each workgroup's range of the gather kernel (the launch grid
    // depends on ndof only), dofs by decreasing contribution count (stable):
    // the wavefronts of a range then hold dofs of like count, so a vertex dof
    // (up to 24 contributions on tets) no longer holds up a wavefront of
    // edge, face and cell dofs (2 to 6).  Which dofs a workgroup gathers and
    // each dof's summation order are unchanged.
    if (ndof > 0) {
      const int64_t grid = GridFor(ndof, kGatherBlock, kMaxWideBlocks);
      const int64_t per = (ndof + grid - 1) / grid;
      for (int64_t a = 0; a < ndof; a += per) {
        const int64_t b = std::min<int64_t>(ndof, a + per);
        std::stable_sort(order.begin() + a, order.begin() + b,
                         [&](int u, int v) { return cnt[u + 1] - cnt[u] > cnt[v + 1] - cnt[v]; });
      }
    }
    std::vector<int> prp(ndof + 1, 0), psl(sl.size());
    for (int64_t r = 0; r < ndof; r++) {
      const int d = order[r];
      std::copy(sl.begin() + cnt[d], sl.begin() + cnt[d + 1], psl.begin() + prp[r]);
      prp[r + 1] = prp[r] + (cnt[d + 1] - cnt[d]);
    }
    g_rowptr.alloc(ndof + 1);
    g_rowptr.upload(prp.data(), ndof + 1);
    for (int64_t r = 0; r < ndof; r++) max_row_slots = std::max(max_row_slots, prp[r + 1] - prp[r]);
    g_slots.alloc(std::max<size_t>(psl.size(), 1));
    if (!psl.empty()) g_slots.upload(psl.data(), psl.size());
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
    g_order.alloc(std::max<int64_t>(ndof, 1));
    if (ndof) g_order.upload(order.data(), ndof);
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
    gather_slots = std::move(psl);
  }
  // at least one entry (operators without elements)
  contrib64.alloc(std::max<size_t>((size_t)nel * nd, 1));
  contrib32.alloc(std::max<size_t>((size_t)nel * nd, 1));
  // round schedule of the MFMA kernels: 16-element tiles, mf_waves tiles per
  // round (one per wavefront of a workgroup), rounds never straddle a class.
  // Split FP16 runs 4-wave workgroups: at 132 VGPRs an 8-wave one fits once
  // per CU, 4-wave ones three times (55.5 against 57.8 us per apply at
  // 32^3 p=4, FP32 vectors); the other kernels measured slower that way.
  mf_waves = prec == NGSAMD_PREC_F16X3 ? 4 : kMfWaves;
  {
    std::vector<int> r0(nclass + 1, 0);
    std::vector<int64_t> rng(2 * (size_t)nclass);
    for (int c = 0; c < nclass; c++) {
      const int64_t tiles = (off[c + 1] - off[c] + 15) / 16;
      r0[c + 1] = r0[c] + (int)((tiles + mf_waves - 1) / mf_waves);
      rng[2 * c] = off[c];
      rng[2 * c + 1] = off[c + 1];
    }
    nrounds = r0[nclass];
    round0.alloc(nclass + 1);
    round0.upload(r0.data(), nclass + 1);
    cls_range.alloc(2 * (size_t)nclass);
    cls_range.upload(rng.data(), 2 * (size_t)nclass);
  }
  // tile-major tables of the factored kernels: a wavefront's load of one dof
  // row (or metric component) of its 16 elements covers one or two lines
  // instead of 16
  if (fac_kernel) {
    std::vector<int> tb(nclass + 1, 0);
    for (int c = 0; c < nclass; c++) tb[c + 1] = tb[c] + (int)((off[c + 1] - off[c] + 15) / 16);
    const int64_t nt = tb[nclass];
    std::vector<int> td((size_t)nt * nd * 16), te((size_t)nt * 16);
    std::vector<double> tm((size_t)nt * ncomp * 16);
    for (int c = 0; c < nclass; c++)
      for (int64_t t = 0; t < tb[c + 1] - tb[c]; t++) {
        const int64_t T = tb[c] + t, first = off[c] + 16 * t;
        for (int col = 0; col < 16; col++) {
          const int64_t pe = first + col < off[c + 1] ? first + col : first;
          for (int k = 0; k < nd; k++) td[((size_t)T * nd + k) * 16 + col] = hdofs[pe * nd + k];
          for (int q = 0; q < ncomp; q++) tm[((size_t)T * ncomp + q) * 16 + col] = hmetric[pe * ncomp + q];
          te[(size_t)T * 16 + col] = el[pe];
        }
      }
    t_dofs.alloc(std::max<size_t>(td.size(), 1));
    if (!td.empty()) t_dofs.upload(td.data(), td.size());
    t_metric.alloc(std::max<size_t>(tm.size(), 1));
    if (!tm.empty()) t_metric.upload(tm.data(), tm.size());
    t_elnr.alloc(std::max<size_t>(te.size(), 1));
    if (!te.empty()) t_elnr.upload(te.data(), te.size());
    tile_base.alloc(tb.size());
    tile_base.upload(tb.data(), tb.size());
    // the factored 16-bit kernels store their element rows tile-major too
    // ([tile][dof][16]: one store instruction covers whole lines); the gather
    // reads them through the same slots in the same order, renumbered.  At
    // 32^3 p=4 that took the 16-bit contraction from 35 to 27 us and the
    // FP32 gather from 27 to 30 us (a net 4 us per apply with FP32 vectors);
    // for the FP32 / FP64 kernels the gather lost what the contraction
    // gained (FP32 33.7 + 30.1 against 37.3 + 27.2 us, FP64 gather 36 -> 44
    // us), so they keep element-major rows.
    std::vector<int> tslot((size_t)nel * nd);
    for (int c = 0; c < nclass; c++)
      for (int64_t pe = off[c]; pe < off[c + 1]; pe++) {
        const int64_t T = tb[c] + (pe - off[c]) / 16, col = (pe - off[c]) % 16;
        for (int k = 0; k < nd; k++) tslot[(size_t)pe * nd + k] = (int)((T * nd + k) * 16 + col);
      }
    std::vector<int> gts(gather_slots.size());
    for (size_t i = 0; i < gts.size(); i++) gts[i] = tslot[gather_slots[i]];
    g_tslots.alloc(std::max<size_t>(gts.size(), 1));
    if (!gts.empty()) g_tslots.upload(gts.data(), gts.size());
    const size_t need = std::max<size_t>((size_t)nt * nd * 16, (size_t)nel * nd);
    if (need > contrib64.n) {
      contrib64.alloc(need);
      contrib32.alloc(need);
    }
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  }
  // enable the dynamic LDS size of every kernel variant this operator uses
  // and record its occupancy
  LaunchElements<0, double>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true);
  LaunchElements<1, double>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true);
  LaunchElements<0, float>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true);
  LaunchElements<1, float>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true);
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
}

// Workgroups of one launch: every round's worth if it fits, else as many as
// are resident at once (occupancy queried when the operator is built).
static int MfGrid(const MatFreeOp& A, int mode, bool f64_vectors, int nrounds = -1) {
  const int per_cu = std::max(1, A.occ[mode][f64_vectors ? 1 : 0]);
  if (nrounds < 0) nrounds = A.nrounds;
  return (int)std::max<int64_t>(1, std::min<int64_t>(nrounds, (int64_t)DeviceCUs() * per_cu));
}

// The element tables and round schedule one launch of the factored kernels
// walks: the operator's own (class-sorted, all elements) or one colour's.
// The operator's own schedule comes with the tile-major tables (MODE 0 and
// 1 of the factored kernels), a colour's with element-major ones (MODE 2).
struct MfSched {
  const int* round0;
  const int64_t* range;
  int nrounds;
  const int* dofs;
  const double* metric;
  const int* elnr;
  const int* tbase = nullptr;
  // MODE 0 only: one partial of <x, A x> per workgroup (the factored kernels)
  double* partials = nullptr;
};
static MfSched OwnSched(const MatFreeOp& A) {
  return {A.round0.p, A.cls_range.p, A.nrounds, A.t_dofs.p, A.t_metric.p, A.t_elnr.p, A.tile_base.p};
}

// prepare: enable the kernel's dynamic LDS size and record its occupancy
// (once, outside any stream capture)
template <class K>
static void PrepareKernel(const MatFreeOp& A, K kern, size_t lds, int mode, bool f64_vectors,
                          int block = kMfBlock) {
  NGSAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int nb = 0;
  NGSAMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, block, lds));
  A.occ[mode][f64_vectors ? 1 : 0] = nb;
}

template <class T, int DIM, int NDP, int MODE, class V>
static void LaunchMfma(const MatFreeOp& A, const T* b, const V* x, V* out, double s, const int* stop, hipStream_t st,
                       bool prepare) {
  const int nqb = A.nqp / 16;
  const size_t lds = (size_t)A.bstride * sizeof(T);
  auto kern = MfMfmaKernel<T, DIM, NDP, MODE, V>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8);
  const MfRounds R{A.round0.p, A.cls_range.p, A.nrounds, (int)A.classes.size()};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8), kMfBlock, lds, st>>>(b, A.bstride, R, A.weights.p, nqb, A.dofs.p,
                                                               A.metric.p, A.Kappa(), A.elnr.p, A.nd, x, out, s, stop);
  NGSAMD_CHECK_LAUNCH();
}

template <class T, int DIM, int MODE, class V>
static void DispatchNdp(const MatFreeOp& A, const T* b, const V* x, V* out, double s, const int* stop,
                        hipStream_t st, bool prepare) {
  switch (A.ndp) {
    case 4: LaunchMfma<T, DIM, 4, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 8: LaunchMfma<T, DIM, 8, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 12: LaunchMfma<T, DIM, 12, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 16: LaunchMfma<T, DIM, 16, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 20: LaunchMfma<T, DIM, 20, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 24: LaunchMfma<T, DIM, 24, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    case 36: LaunchMfma<T, DIM, 36, MODE, V>(A, b, x, out, s, stop, st, prepare); break;
    default: throw Error(NGSAMD_E_STATE, "matrix-free: unsupported padded dof count");
  }
}

template <class H, int DIM, int KS1, int MT, int MODE, class V>
static void LaunchBf16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                       bool prepare) {
  const size_t lds = (size_t)A.bstride * 2;
  auto kern = MfBf16Kernel<H, DIM, KS1, MT, MODE, V>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8);
  const MfRounds R{A.round0.p, A.cls_range.p, A.nrounds, (int)A.classes.size()};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8), kMfBlock, lds, st>>>(A.bmat16.p, A.bstride, R, A.k1p, A.nqc, A.weights.p,
                                                               A.dofs.p, A.metric.p, A.Kappa(), A.elnr.p, A.nd, x,
                                                               out, s, stop);
  NGSAMD_CHECK_LAUNCH();
}

template <class H, int DIM, int MODE, class V>
static void DispatchBf16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                         bool prepare) {
  if (A.ks1 == 1 && A.mt16 == 1) LaunchBf16<H, DIM, 1, 1, MODE, V>(A, x, out, s, stop, st, prepare);
  else if (A.ks1 == 1 && A.mt16 == 2) LaunchBf16<H, DIM, 1, 2, MODE, V>(A, x, out, s, stop, st, prepare);
  else if (A.ks1 == 2 && A.mt16 == 3) LaunchBf16<H, DIM, 2, 3, MODE, V>(A, x, out, s, stop, st, prepare);
  else throw Error(NGSAMD_E_STATE, "matrix-free: unsupported 16-bit shape");
}

template <class T, int DIM, int NDP, int MODE, class V>
static void LaunchRef(const MatFreeOp& A, const T* sm, const V* x, V* out, double s, const int* stop,
                      hipStream_t st, bool prepare) {
  const size_t lds = (size_t)A.sstride * sizeof(T);
  auto kern = A.classes.size() <= (size_t)kMaxLdsClasses ? MfRefKernel<T, DIM, NDP, MODE, V, true>
                                                         : MfRefKernel<T, DIM, NDP, MODE, V, false>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8);
  const MfRounds R{A.round0.p, A.cls_range.p, A.nrounds, (int)A.classes.size()};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8), kMfBlock, lds, st>>>(sm, A.sstride, R, A.dofs.p, A.metric.p, A.Kappa(),
                                                               A.elnr.p, A.nd, x, out, s, stop);
  NGSAMD_CHECK_LAUNCH();
}

template <class T, int DIM, int MODE, class V>
static void DispatchRefNdp(const MatFreeOp& A, const T* sm, const V* x, V* out, double s, const int* stop,
                           hipStream_t st, bool prepare) {
  switch (A.ndp) {
    case 4: LaunchRef<T, DIM, 4, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 8: LaunchRef<T, DIM, 8, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 12: LaunchRef<T, DIM, 12, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 16: LaunchRef<T, DIM, 16, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 20: LaunchRef<T, DIM, 20, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 24: LaunchRef<T, DIM, 24, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    case 36: LaunchRef<T, DIM, 36, MODE, V>(A, sm, x, out, s, stop, st, prepare); break;
    default: throw Error(NGSAMD_E_STATE, "matrix-free: unsupported padded dof count");
  }
}

template <class H, int DIM, int KS1, int MT, int MODE, class V>
static void LaunchRef16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                        bool prepare) {
  const size_t lds = (size_t)A.sstride * 2;
  auto kern = A.classes.size() <= (size_t)kMaxLdsClasses ? MfRef16Kernel<H, DIM, KS1, MT, MODE, V, true>
                                                         : MfRef16Kernel<H, DIM, KS1, MT, MODE, V, false>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8);
  const MfRounds R{A.round0.p, A.cls_range.p, A.nrounds, (int)A.classes.size()};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8), kMfBlock, lds, st>>>(A.smat16.p, A.sstride, R, A.dofs.p, A.metric.p,
                                                               A.Kappa(), A.elnr.p, A.nd, x, out, s, stop);
  NGSAMD_CHECK_LAUNCH();
}

template <class H, int DIM, int MODE, class V>
static void DispatchRef16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                          bool prepare) {
  if (A.ks1 == 1 && A.mt16 == 1) LaunchRef16<H, DIM, 1, 1, MODE, V>(A, x, out, s, stop, st, prepare);
  else if (A.ks1 == 1 && A.mt16 == 2) LaunchRef16<H, DIM, 1, 2, MODE, V>(A, x, out, s, stop, st, prepare);
  else if (A.ks1 == 2 && A.mt16 == 3) LaunchRef16<H, DIM, 2, 3, MODE, V>(A, x, out, s, stop, st, prepare);
  else throw Error(NGSAMD_E_STATE, "matrix-free: unsupported 16-bit shape");
}

template <class T, int DIM, int NDP, int KG, int MODE, class V>
static void LaunchFac(const MatFreeOp& A, const T* fm, const V* x, V* out, double s, const int* stop,
                      hipStream_t st, bool prepare, const MfSched& S) {
  const size_t lds = (size_t)A.fstride * sizeof(T);
  auto kern = MfFacKernel<T, DIM, NDP, KG, MODE, V>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8);
  const MfRounds R{S.round0, S.range, S.nrounds, (int)A.classes.size(), S.tbase};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8, S.nrounds), kMfBlock, lds, st>>>(
      fm, A.fstride, R, S.dofs, S.metric, A.Kappa(), S.elnr, A.nd, x, out, s, stop, S.partials);
  NGSAMD_CHECK_LAUNCH();
}

template <class T, int DIM, int MODE, class V>
static void DispatchFac(const MatFreeOp& A, const T* fm, const V* x, V* out, double s, const int* stop,
                        hipStream_t st, bool prepare, const MfSched& S) {
#define NGSAMD_FAC(NDP, KG)                                                                   \
  if (A.ndp == NDP && A.fac_kg == KG) return LaunchFac<T, DIM, NDP, KG, MODE, V>(A, fm, x, out, s, stop, st, prepare, S);
  if constexpr (DIM == 2) {
    NGSAMD_FAC(4, 1) NGSAMD_FAC(8, 1) NGSAMD_FAC(12, 2) NGSAMD_FAC(16, 3) NGSAMD_FAC(24, 4)
  } else {
    NGSAMD_FAC(4, 1) NGSAMD_FAC(12, 1) NGSAMD_FAC(20, 3) NGSAMD_FAC(36, 5)
  }
#undef NGSAMD_FAC
  throw Error(NGSAMD_E_STATE, "matrix-free: unsupported factored shape");
}

template <class H, int DIM, int KS1, int MT, int KG, int ND, int MODE, class V, bool SPLIT>
static void LaunchFac16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                        bool prepare, const MfSched& S) {
  const size_t lds = (size_t)A.fstride * 2;
  // split FP16: 4-wave workgroups (the operator's rounds hold 4 tiles)
  constexpr int W = SPLIT ? 4 : kMfWaves;
  auto kern = MfFac16Kernel<H, DIM, KS1, MT, KG, ND, MODE, V, SPLIT, W>;
  if (prepare) return PrepareKernel(A, kern, lds, MODE, sizeof(V) == 8, 64 * W);
  NGSAMD_REQUIRE(A.mf_waves == W, "matrix-free: round schedule does not match the kernel's workgroup");
  const MfRounds R{S.round0, S.range, S.nrounds, (int)A.classes.size(), S.tbase};
  kern<<<MfGrid(A, MODE, sizeof(V) == 8, S.nrounds), 64 * W, lds, st>>>(
      A.fmat16.p, A.fstride, R, S.dofs, S.metric, A.Kappa(), S.elnr, A.nd, x, out, s, stop, S.partials);
  NGSAMD_CHECK_LAUNCH();
}

template <class H, int DIM, int MODE, class V, bool SPLIT = false>
static void DispatchFac16(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                          bool prepare, const MfSched& S) {
#define NGSAMD_FAC16(KS1, MT, KG, ND)                                                  \
  if (A.ks1 == KS1 && A.mt16 == MT && A.fac_kg == KG && A.nd == ND)                      \
    return LaunchFac16<H, DIM, KS1, MT, KG, ND, MODE, V, SPLIT>(A, x, out, s, stop, st, prepare, S);
  // H1 triangles p = 1..5, tetrahedra p = 1..4
  if constexpr (DIM == 2) {
    NGSAMD_FAC16(1, 1, 1, 3) NGSAMD_FAC16(1, 1, 1, 6) NGSAMD_FAC16(1, 1, 2, 10) NGSAMD_FAC16(1, 1, 3, 15)
    NGSAMD_FAC16(1, 2, 4, 21)
  } else {
    NGSAMD_FAC16(1, 1, 1, 4) NGSAMD_FAC16(1, 1, 1, 10) NGSAMD_FAC16(1, 2, 3, 20) NGSAMD_FAC16(2, 3, 5, 35)
  }
#undef NGSAMD_FAC16
  throw Error(NGSAMD_E_STATE, "matrix-free: unsupported factored 16-bit shape");
}

// All elements: one launch of the MFMA kernel (every class), or the scalar
// kernel class by class.
// The factored kernels over schedule S (MODE 2 only here: the colour
// launches).
template <int MODE, class V>
static void LaunchFacElements(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                              bool prepare, const MfSched& S) {
  if (S.nrounds == 0 && !prepare) return;
  if (A.precision == NGSAMD_PREC_BF16) {
    if (A.dim == 2) DispatchFac16<__bf16, 2, MODE, V>(A, x, out, s, stop, st, prepare, S);
    else DispatchFac16<__bf16, 3, MODE, V>(A, x, out, s, stop, st, prepare, S);
  } else if (A.precision == NGSAMD_PREC_F16) {
    if (A.dim == 2) DispatchFac16<_Float16, 2, MODE, V>(A, x, out, s, stop, st, prepare, S);
    else DispatchFac16<_Float16, 3, MODE, V>(A, x, out, s, stop, st, prepare, S);
  } else if (A.precision == NGSAMD_PREC_F16X3) {
    if (A.dim == 2) DispatchFac16<_Float16, 2, MODE, V, true>(A, x, out, s, stop, st, prepare, S);
    else DispatchFac16<_Float16, 3, MODE, V, true>(A, x, out, s, stop, st, prepare, S);
  } else if (A.precision == NGSAMD_PREC_F64) {
    if (A.dim == 2) DispatchFac<double, 2, MODE, V>(A, A.fmat64.p, x, out, s, stop, st, prepare, S);
    else DispatchFac<double, 3, MODE, V>(A, A.fmat64.p, x, out, s, stop, st, prepare, S);
  } else {
    if (A.dim == 2) DispatchFac<float, 2, MODE, V>(A, A.fmat32.p, x, out, s, stop, st, prepare, S);
    else DispatchFac<float, 3, MODE, V>(A, A.fmat32.p, x, out, s, stop, st, prepare, S);
  }
}

template <int MODE, class V>
static void LaunchElements(const MatFreeOp& A, const V* x, V* out, double s, const int* stop, hipStream_t st,
                           bool prepare) {
  if (A.use_mfma && A.fac_kernel) return LaunchFacElements<MODE, V>(A, x, out, s, stop, st, prepare, OwnSched(A));
  if (A.use_mfma && A.ref_kernel) {
    if (A.nrounds == 0 && !prepare) return;
    if (A.precision == NGSAMD_PREC_BF16) {
      if (A.dim == 2) DispatchRef16<__bf16, 2, MODE, V>(A, x, out, s, stop, st, prepare);
      else DispatchRef16<__bf16, 3, MODE, V>(A, x, out, s, stop, st, prepare);
    } else if (A.precision == NGSAMD_PREC_F16) {
      if (A.dim == 2) DispatchRef16<_Float16, 2, MODE, V>(A, x, out, s, stop, st, prepare);
      else DispatchRef16<_Float16, 3, MODE, V>(A, x, out, s, stop, st, prepare);
    } else if (A.precision == NGSAMD_PREC_F64) {
      if (A.dim == 2) DispatchRefNdp<double, 2, MODE, V>(A, A.smat64.p, x, out, s, stop, st, prepare);
      else DispatchRefNdp<double, 3, MODE, V>(A, A.smat64.p, x, out, s, stop, st, prepare);
    } else {
      if (A.dim == 2) DispatchRefNdp<float, 2, MODE, V>(A, A.smat32.p, x, out, s, stop, st, prepare);
      else DispatchRefNdp<float, 3, MODE, V>(A, A.smat32.p, x, out, s, stop, st, prepare);
    }
    return;
  }
  if (A.use_mfma) {
    if (A.nrounds == 0 && !prepare) return;
    if (A.precision == NGSAMD_PREC_BF16) {
      if (A.dim == 2) DispatchBf16<__bf16, 2, MODE, V>(A, x, out, s, stop, st, prepare);
      else DispatchBf16<__bf16, 3, MODE, V>(A, x, out, s, stop, st, prepare);
    } else if (A.precision == NGSAMD_PREC_F16) {
      if (A.dim == 2) DispatchBf16<_Float16, 2, MODE, V>(A, x, out, s, stop, st, prepare);
      else DispatchBf16<_Float16, 3, MODE, V>(A, x, out, s, stop, st, prepare);
    } else if (A.precision == NGSAMD_PREC_F64) {
      if (A.dim == 2) DispatchNdp<double, 2, MODE, V>(A, A.bmat64.p, x, out, s, stop, st, prepare);
      else DispatchNdp<double, 3, MODE, V>(A, A.bmat64.p, x, out, s, stop, st, prepare);
    } else {
      if (A.dim == 2) DispatchNdp<float, 2, MODE, V>(A, A.bmat32.p, x, out, s, stop, st, prepare);
      else DispatchNdp<float, 3, MODE, V>(A, A.bmat32.p, x, out, s, stop, st, prepare);
    }
    return;
  }
  if (prepare) return;
  for (size_t ci = 0; ci < A.classes.size(); ci++) {
    const MfClass& c = A.classes[ci];
    if (c.end <= c.begin) continue;
    const double* b = A.bref.p + ci * A.nq * A.nd * A.dim;
    const int grid = GridFor(c.end - c.begin, kScalarBlock, 4096);
    if (A.dim == 2)
      MfScalarKernel<2, MODE, V><<<grid, kScalarBlock, 0, st>>>(b, A.weights.p, A.nq, A.dofs.p, A.metric.p, A.Kappa(),
                                                               A.elnr.p, c.begin, c.end, A.nd, x, out, s, stop);
    else
      MfScalarKernel<3, MODE, V><<<grid, kScalarBlock, 0, st>>>(b, A.weights.p, A.nq, A.dofs.p, A.metric.p, A.Kappa(),
                                                               A.elnr.p, c.begin, c.end, A.nd, x, out, s, stop);
    NGSAMD_CHECK_LAUNCH();
  }
}

void MatFreeOp::SetColoring(const int* colors) {
  NGSAMD_REQUIRE(use_mfma && fac_kernel, "ngsamd_mf_set_coloring: the colour scatter needs the factored kernels");
  const int nclass = (int)classes.size();
  int nc = 0;
  for (int64_t e = 0; e < nel; e++) {
    NGSAMD_REQUIRE(colors[e] >= 0 && colors[e] < (1 << 20), "ngsamd_mf_set_coloring: colour out of range");
    nc = std::max(nc, colors[e] + 1);
  }
  std::vector<int> hd((size_t)nel * nd), he(nel);
  std::vector<double> hm((size_t)nel * ncomp);
  NGSAMD_HIP(hipMemcpy(hd.data(), dofs.p, sizeof(int) * hd.size(), hipMemcpyDeviceToHost));
  NGSAMD_HIP(hipMemcpy(he.data(), elnr.p, sizeof(int) * he.size(), hipMemcpyDeviceToHost));
  NGSAMD_HIP(hipMemcpy(hm.data(), metric.p, sizeof(double) * hm.size(), hipMemcpyDeviceToHost));
  // rows by colour, row order kept (so each colour's rows stay class-sorted)
  std::vector<int64_t> cptr(nc + 1, 0);
  for (int64_t e = 0; e < nel; e++) cptr[colors[e] + 1]++;
  for (int c = 0; c < nc; c++) cptr[c + 1] += cptr[c];
  std::vector<int64_t> order(nel);
  {
    std::vector<int64_t> fill(cptr.begin(), cptr.end() - 1);
    for (int64_t e = 0; e < nel; e++) order[fill[colors[e]]++] = e;
  }
  // no two elements of one colour share a dof
  std::vector<int> stamp(height, -1);
  for (int c = 0; c < nc; c++)
    for (int64_t k = cptr[c]; k < cptr[c + 1]; k++)
      for (int i = 0; i < nd; i++) {
        const int d = hd[(size_t)order[k] * nd + i];
        NGSAMD_REQUIRE(stamp[d] != c, "ngsamd_mf_set_coloring: two elements of one colour share a dof");
        stamp[d] = c;
      }
  std::vector<int> cd((size_t)nel * nd), ce(nel);
  std::vector<double> cm((size_t)nel * ncomp);
  for (int64_t k = 0; k < nel; k++) {
    const int64_t e = order[k];
    std::copy(&hd[(size_t)e * nd], &hd[(size_t)e * nd] + nd, &cd[(size_t)k * nd]);
    std::copy(&hm[(size_t)e * ncomp], &hm[(size_t)e * ncomp] + ncomp, &cm[(size_t)k * ncomp]);
    ce[k] = he[e];
  }
  // per colour: class segments and rounds (a round never straddles a class)
  std::vector<int> r0((size_t)nc * (nclass + 1), 0);
  std::vector<int64_t> rng((size_t)nc * 2 * nclass, 0);
  c_nrounds.assign(nc, 0);
  for (int c = 0; c < nc; c++) {
    int64_t k = cptr[c];
    int* rc = &r0[(size_t)c * (nclass + 1)];
    for (int cl = 0; cl < nclass; cl++) {
      const int64_t b = k;
      while (k < cptr[c + 1] && order[k] < classes[cl].end) k++;
      rng[(size_t)c * 2 * nclass + 2 * cl] = b;
      rng[(size_t)c * 2 * nclass + 2 * cl + 1] = k;
      const int64_t tiles = (k - b + 15) / 16;
      rc[cl + 1] = rc[cl] + (int)((tiles + mf_waves - 1) / mf_waves);
    }
    c_nrounds[c] = rc[nclass];
  }
  c_round0.alloc(r0.size());
  c_round0.upload(r0.data(), r0.size());
  c_range.alloc(std::max<size_t>(rng.size(), 1));
  if (!rng.empty()) c_range.upload(rng.data(), rng.size());
  c_dofs.alloc(std::max<size_t>(cd.size(), 1));
  if (!cd.empty()) c_dofs.upload(cd.data(), cd.size());
  c_metric.alloc(std::max<size_t>(cm.size(), 1));
  if (!cm.empty()) c_metric.upload(cm.data(), cm.size());
  c_elnr.alloc(std::max<size_t>(ce.size(), 1));
  if (!ce.empty()) c_elnr.upload(ce.data(), ce.size());
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  ncolors = nc;
  scatter = NGSAMD_SCATTER_COLOR;
  const MfSched S0 = OwnSched(*this);
  LaunchFacElements<2, double>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true, S0);
  LaunchFacElements<2, float>(*this, nullptr, nullptr, 1.0, nullptr, nullptr, true, S0);
}

template <class V>
int MatFreeOp::Apply(double s, const V* x, V* y, bool accumulate, const int* stop, hipStream_t st,
                     double* partials, double* wide) {
  if (height == 0) return 0;
  if (scatter == NGSAMD_SCATTER_COLOR) {
    // colour after colour, each element's result added straight into y
    if (!accumulate) {
      ScaleFillKernel<V><<<GridFor(height, 256, 8192), 256, 0, st>>>(y, height, stop);
      NGSAMD_CHECK_LAUNCH();
    }
    const int nclass = (int)classes.size();
    for (int c = 0; c < ncolors; c++) {
      const MfSched S{c_round0.p + (size_t)c * (nclass + 1), c_range.p + (size_t)c * 2 * nclass, c_nrounds[c],
                      c_dofs.p, c_metric.p, c_elnr.p};
      LaunchFacElements<2, V>(*this, x, y, s, stop, st, false, S);
    }
    return 0;  // no gather kernel: the caller forms <x, A x> itself
  }
  if (scatter == NGSAMD_SCATTER_ATOMIC) {
    if (!accumulate) {
      ScaleFillKernel<V><<<GridFor(height, 256, 8192), 256, 0, st>>>(y, height, stop);
      NGSAMD_CHECK_LAUNCH();
    }
    LaunchElements<1, V>(*this, x, y, s, stop, st, false);
    return 0;
  }
  // contribution buffers are allocated by the constructor: nothing here
  // allocates, so the apply can be captured into a graph
  V* contrib;
  if constexpr (sizeof(V) == 8) contrib = reinterpret_cast<V*>(contrib64.p);
  else contrib = reinterpret_cast<V*>(contrib32.p);
  // with partials (CG): one partial of <x, y> per workgroup.  For y = A x
  // the factored kernels form it themselves as sum_e U_e . V_e =
  // sum_e x_e^T K_e x_e (no second read of x by dof), one per workgroup of
  // the contraction's resident grid, which the gather's first workgroups
  // fold into partials[] (at most kFoldedPartials); otherwise the gather
  // forms sum_r x[r] y[r] on its wide grid, folded to kFoldedPartials by a
  // kernel of its own.
  const bool fac_dot = partials && use_mfma && fac_kernel && !accumulate && s == 1.0 &&
                       MfGrid(*this, 0, sizeof(V) == 8, nrounds) <= kMaxWideBlocks;
  double* dotp = partials ? (wide ? wide : State().wide_partials) : nullptr;
  int ndot = 0;
  if (fac_dot) {
    MfSched S = OwnSched(*this);
    S.partials = dotp;
    LaunchFacElements<0, V>(*this, x, contrib, 1.0, stop, st, false, S);
    ndot = nrounds ? MfGrid(*this, 0, sizeof(V) == 8, nrounds) : 0;
    if (ndot == 0) {  // no elements: <x, A x> = 0
      LaunchFill(dotp, 1, 0.0, st);
      ndot = 1;
    }
  } else {
    LaunchElements<0, V>(*this, x, contrib, 1.0, stop, st, false);
  }
  // one thread per dof: a wide grid hides each dof's load chain
  const int np = GridFor(height, kGatherBlock, kMaxWideBlocks);
  // the factored 16-bit kernels wrote tile-major rows
  const bool half = precision == NGSAMD_PREC_BF16 || precision == NGSAMD_PREC_F16 || precision == NGSAMD_PREC_F16X3;
  const int* slots = use_mfma && fac_kernel && half ? g_tslots.p : g_slots.p;
  if (fac_dot) {
    MfGatherKernel<V, V, false><<<np, kGatherBlock, 0, st>>>(g_rowptr.p, slots, g_order.p, contrib, y, height, s,
                                                             0, stop, nullptr, nullptr, dotp, ndot, partials);
    NGSAMD_CHECK_LAUNCH();
    return std::min(kFoldedPartials, np);
  }
  if (partials)
    MfGatherKernel<V, V, true><<<np, kGatherBlock, 0, st>>>(g_rowptr.p, slots, g_order.p, contrib, y, height, s,
                                                            accumulate ? 1 : 0, stop, x, dotp);
  else
    MfGatherKernel<V, V, false><<<np, kGatherBlock, 0, st>>>(g_rowptr.p, slots, g_order.p, contrib, y, height, s,
                                                             accumulate ? 1 : 0, stop, nullptr, nullptr);
  NGSAMD_CHECK_LAUNCH();
  if (!partials) return 0;
  LaunchFoldPartials(dotp, np, partials, stop, st);
  return kFoldedPartials;
}

void MatFreeOp::MultAdd(double s, const double* x, double* y, hipStream_t st) { Apply<double>(s, x, y, true, nullptr, st); }
void MatFreeOp::Mult(const double* x, double* y, hipStream_t st) { Apply<double>(1.0, x, y, false, nullptr, st); }
void MatFreeOp::MultStop(const double* x, double* y, const int* stop, hipStream_t st) {
  Apply<double>(1.0, x, y, false, stop, st);
}
void MatFreeOp::MultF32(const float* x, float* y, const int* stop, hipStream_t st) {
  Apply<float>(1.0, x, y, false, stop, st);
}
int MatFreeOp::MultDot(const double* x, double* y, double* partials, const int* stop, hipStream_t st,
                       double* wide) {
  return Apply<double>(1.0, x, y, false, stop, st, partials, wide);
}
int MatFreeOp::MultDotF32(const float* x, float* y, double* partials, const int* stop, hipStream_t st,
                          double* wide) {
  return Apply<float>(1.0, x, y, false, stop, st, partials, wide);
}

void MatFreeOp::Diagonal(double* d, hipStream_t st) {
  if (height == 0) return;
  for (size_t ci = 0; ci < classes.size(); ci++) {
    const MfClass& c = classes[ci];
    if (c.end <= c.begin) continue;
    const double* b = bref.p + ci * (size_t)nq * nd * dim;
    const int grid = GridFor(c.end - c.begin, 128, 8192);
    if (dim == 2)
      MfDiagKernel<2><<<grid, 128, 0, st>>>(b, weights.p, nq, metric.p, Kappa(), elnr.p, c.begin, c.end, nd, contrib64.p);
    else
      MfDiagKernel<3><<<grid, 128, 0, st>>>(b, weights.p, nq, metric.p, Kappa(), elnr.p, c.begin, c.end, nd, contrib64.p);
    NGSAMD_CHECK_LAUNCH();
  }
  MfGatherKernel<double, double, false><<<GridFor(height, kGatherBlock, 16384), kGatherBlock, 0, st>>>(
      g_rowptr.p, g_slots.p, g_order.p, contrib64.p, d, height, 1.0, 0, nullptr, nullptr, nullptr);
  NGSAMD_CHECK_LAUNCH();
}

// f_e[i] = |det F_e| sum_q w_q f(e, q) phi_i(x_q) for the elements of one
// class (table rows begin..end), into the contribution buffer; the gather
// then sums each dof's contributions in slot order.
__global__ void MfLinearFormKernel(const double* __restrict__ shape, const double* __restrict__ w, int nq,
                                   const double* __restrict__ detj, const double* __restrict__ fvals,
                                   const int* __restrict__ elnr, int64_t begin, int64_t end, int nd,
                                   double* __restrict__ contrib) {
  const int64_t total = (end - begin) * nd;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = begin + t / nd;
    const int i = (int)(t % nd);
    const int g = elnr[e];
    const double* fe = fvals + (int64_t)g * nq;
    double acc = 0.0;
    for (int q = 0; q < nq; q++) acc += w[q] * fe[q] * shape[(size_t)q * nd + i];
    contrib[e * nd + i] = fabs(detj[g]) * acc;
  }
}

void MatFreeOp::LinearForm(int nq_lf, const double* h_w, const double* h_shape, const double* h_detj,
                           const double* d_fvals, double* d_out, hipStream_t st) {
  if (height == 0) return;
  NGSAMD_REQUIRE(nq_lf >= 1, "ngsamd_mf_linear_form: empty rule");
  DevBuf<double> w(nq_lf), sh((size_t)classes.size() * nq_lf * nd), dj(std::max<int64_t>(ne_global, 1));
  w.upload(h_w, nq_lf);
  sh.upload(h_shape, (size_t)classes.size() * nq_lf * nd);
  if (ne_global) dj.upload(h_detj, ne_global);
  for (size_t ci = 0; ci < classes.size(); ci++) {
    const MfClass& c = classes[ci];
    if (c.end <= c.begin) continue;
    MfLinearFormKernel<<<GridFor((c.end - c.begin) * nd, 256, 8192), 256, 0, st>>>(
        sh.p + ci * (size_t)nq_lf * nd, w.p, nq_lf, dj.p, d_fvals, elnr.p, c.begin, c.end, nd, contrib64.p);
    NGSAMD_CHECK_LAUNCH();
  }
  MfGatherKernel<double, double, false><<<GridFor(height, kGatherBlock, 16384), kGatherBlock, 0, st>>>(
      g_rowptr.p, g_slots.p, g_order.p, contrib64.p, d_out, height, 1.0, 0, nullptr, nullptr, nullptr);
  NGSAMD_CHECK_LAUNCH();
  NGSAMD_HIP(hipStreamSynchronize(st));  // the uploaded tables are freed on return
}

// ------------------------------------------------------ CSR assembly on device
// One 64-lane workgroup per row (dofs in gather order, grid-stride): the row's
// column list into LDS; then, for each contribution (e, i) of the row in slot
// order, lane j adds K_e[i][j] = sum_k g_{e,k} S_k[i][j] (k over the packed
// metric components, g_e = kappa_e times the packed |det F| F^-1 F^-T) at the
// position of dofs[e][j] in the list (binary search); the row's values are
// written once at the end.  Deterministic: every value is summed in slot
// order.  RUNS: the target is the runs layout (the row's list is its run's
// shared list, values at val_ptr[u] + position * R + row in run).  A column
// missing from the pattern sets *bad.  NC = 1 with metric NULL is the mass
// form: K_e = kappa_e S (S the class's weighted shape products, kappa_e the
// element's coefficient times |det F|).
template <int NC, bool RUNS>
__global__ __launch_bounds__(64) void AssembleCsrKernel(
    const int* __restrict__ g_rowptr, const int* __restrict__ g_slots, const int* __restrict__ g_order,
    int64_t nrows, const int* __restrict__ dofs, const double* __restrict__ metric,
    const double* __restrict__ kappa, const int* __restrict__ elnr, int nd, const double* __restrict__ smat,
    const int64_t* __restrict__ cls_range, int nclass, const int* __restrict__ lptr,
    const int* __restrict__ lcols, const int* __restrict__ row_run, const int* __restrict__ run_row,
    const int* __restrict__ val_ptr, double* __restrict__ vals, int max_len, int* __restrict__ bad) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double* lval = reinterpret_cast<double*>(lds_raw);
  int* lcol = reinterpret_cast<int*>(lds_raw + sizeof(double) * (size_t)max_len);
  __shared__ double cg[64 * NC];  // per contribution of a chunk: kappa_e times the packed metric
  __shared__ int ce[64], ci[64];  // its element (table row) and local dof | class << 16
  const int lane = threadIdx.x;
  for (int64_t t = blockIdx.x; t < nrows; t += gridDim.x) {
    const int r = g_order[t];
    int base, L, R = 1;
    int64_t obase;
    if constexpr (RUNS) {
      const int u = row_run[r];
      base = lptr[u];
      L = lptr[u + 1] - base;
      R = run_row[u + 1] - run_row[u];
      obase = (int64_t)val_ptr[u] + (r - run_row[u]);
    } else {
      base = lptr[r];
      L = lptr[r + 1] - base;
      obase = base;
    }
    for (int k = lane; k < L; k += 64) {
      lcol[k] = lcols[base + k];
      lval[k] = 0.0;
    }
    __syncthreads();
    // contributions of the row in chunks of 64: lane q fetches contribution
    // q's element data (slot, class, kappa, metric) into LDS, all lanes at
    // once; then the chunk's contributions are added one after another
    const int s0 = g_rowptr[t], s1 = g_rowptr[t + 1];
    for (int c0 = s0; c0 < s1; c0 += 64) {
      const int cn = min(64, s1 - c0);
      if (lane < cn) {
        const int slot = g_slots[c0 + lane];
        const int e = slot / nd;
        int c = 0;
        while (c + 1 < nclass && cls_range[2 * c + 1] <= e) c++;
        const double kap = kappa[elnr[e]];
        ce[lane] = e;
        ci[lane] = (slot - e * nd) | (c << 16);
#pragma unroll
        for (int k = 0; k < NC; k++) cg[lane * NC + k] = metric ? kap * metric[(int64_t)e * NC + k] : kap;
      }
      __syncthreads();
      // lane j < nd (nd <= 64) adds entry j of each contribution's row; the
      // global loads of contribution q + 1 (S rows, dof number) are issued
      // before contribution q is added
      const int j = min(lane, nd - 1);
      double sv[NC];
      int col = 0;
      auto fetch = [&](int q) {
        const int e = ce[q], i = ci[q] & 0xFFFF, c = ci[q] >> 16;
        const double* sp = smat + ((int64_t)c * NC * nd + i) * nd + j;
#pragma unroll
        for (int k = 0; k < NC; k++) sv[k] = sp[(int64_t)k * nd * nd];
        col = dofs[(int64_t)e * nd + j];
      };
      fetch(0);
      for (int q = 0; q < cn; q++) {
        double cur[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) cur[k] = sv[k];
        const int ccol = col;
        if (q + 1 < cn) fetch(q + 1);
        if (lane < nd) {
          double v = 0.0;
#pragma unroll
          for (int k = 0; k < NC; k++) v += cg[q * NC + k] * cur[k];
          int lo = 0, hi = L;  // first position whose column is >= ccol
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (lcol[mid] < ccol) lo = mid + 1;
            else hi = mid;
          }
          if (lo < L && lcol[lo] == ccol) lval[lo] += v;
          else *bad = 1;
        }
        // the next contribution may add to the same positions
        __syncthreads();
      }
    }
    for (int k = lane; k < L; k += 64) vals[obase + (int64_t)k * R] = lval[k];
    __syncthreads();
  }
}

void MatFreeOp::AssembleCsr(CsrOp& A, hipStream_t st) {
  AssembleCsrWith(A, smat_asm.p, metric.p, Kappa(), dim * (dim + 1) / 2, st);
}

// Mass form rho_e |det F_e| sum_q w_q phi_i phi_j: the class matrices
// S_c[i][j] = sum_q w_q phi_c[q][i] phi_c[q][j] (host, FP64, q ascending)
// and the element factors go through the same assembly kernel with one
// component.
void MatFreeOp::AssembleMassCsr(CsrOp& A, int nq_m, const double* h_w, const double* h_shape,
                                const double* h_scale, hipStream_t st) {
  NGSAMD_REQUIRE(nq_m >= 1, "ngsamd_mf_assemble_mass_csr: empty rule");
  const size_t nc = classes.size();
  std::vector<double> sm(std::max<size_t>(nc * nd * nd, 1), 0.0);
  for (size_t c = 0; c < nc; c++) {
    const double* sh = h_shape + c * (size_t)nq_m * nd;
    double* S = sm.data() + c * (size_t)nd * nd;
    for (int i = 0; i < nd; i++)
      for (int j = 0; j < nd; j++) {
        double v = 0.0;
        for (int q = 0; q < nq_m; q++) v += h_w[q] * sh[(size_t)q * nd + i] * sh[(size_t)q * nd + j];
        S[(size_t)i * nd + j] = v;
      }
  }
  DevBuf<double> smass(sm.size()), scale(std::max<int64_t>(ne_global, 1));
  smass.upload(sm.data(), sm.size());
  if (ne_global) scale.upload(h_scale, ne_global);
  AssembleCsrWith(A, smass.p, nullptr, scale.p, 1, st);  // synchronises before the buffers are freed
}

void MatFreeOp::AssembleCsrWith(CsrOp& A, const double* smat, const double* met, const double* kap, int nc,
                                hipStream_t st) {
  NGSAMD_REQUIRE(A.height == height && A.width == width, "ngsamd_mf_assemble_csr: pattern size differs");
  if (height == 0 || nel == 0 || A.nnz == 0) return;
  const size_t lds = (size_t)A.max_row_len * (sizeof(double) + sizeof(int));
  // plus the kernel's static LDS (contribution data of one chunk, <= 4 KB)
  NGSAMD_REQUIRE(lds + 4096 <= (size_t)DeviceMaxLds(), "ngsamd_mf_assemble_csr: a row is too long for LDS");
  DevBuf<int> bad;
  bad.alloc(1);
  NGSAMD_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
  const bool runs = A.algo == 2;
  auto kern = nc == 1   ? (runs ? AssembleCsrKernel<1, true> : AssembleCsrKernel<1, false>)
              : nc == 3 ? (runs ? AssembleCsrKernel<3, true> : AssembleCsrKernel<3, false>)
                        : (runs ? AssembleCsrKernel<6, true> : AssembleCsrKernel<6, false>);
  NGSAMD_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = (int)std::min<int64_t>(height, 65536);
  kern<<<grid, 64, lds, st>>>(g_rowptr.p, g_slots.p, g_order.p, height, dofs.p, met, kap, elnr.p, nd,
                              smat, cls_range.p, (int)classes.size(), runs ? A.run_ptr.p : A.rowptr.p,
                              runs ? A.rcols.p : A.cols.p, A.row_run.p, A.run_row.p, A.val_ptr.p,
                              runs ? A.rvals.p : A.vals.p, (int)A.max_row_len, bad.p);
  NGSAMD_CHECK_LAUNCH();
  int hbad = 0;
  NGSAMD_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  NGSAMD_REQUIRE(!hbad, "ngsamd_mf_assemble_csr: the pattern lacks a pair of dofs of one element");
}

// ------------------------------------------------- sparsity pattern on device
// One workgroup per dof r (gather order, grid-stride): the dofs of every
// element holding r go to LDS (padded with INT_MAX to a power of two, at
// most cap), a bitonic sort orders them, and the distinct ones are row r of
// the pattern
// (MatrixGraph: the sorted union of the element dofs).  MODE 0 stores the
// row length in len[r]; MODE 1 writes the columns at cols[rowptr[r]...]
// (each thread counts its contiguous chunk, a serial scan of the 256 counts
// gives the offsets, so the order is the sorted order).
constexpr int kPatBlock = 256;

template <int MODE>
__global__ __launch_bounds__(kPatBlock) void PatternKernel(const int* __restrict__ g_rowptr,
                                                           const int* __restrict__ g_slots,
                                                           const int* __restrict__ g_order, int64_t nrows,
                                                           const int* __restrict__ dofs, int nd, int cap,
                                                           int* __restrict__ len,
                                                           const int64_t* __restrict__ rowptr,
                                                           int* __restrict__ cols) {
  extern __shared__ int buf[];
  __shared__ int cnt[kPatBlock + 1];
  const int tid = threadIdx.x;
  for (int64_t t = blockIdx.x; t < nrows; t += gridDim.x) {
    const int r = g_order[t];
    const int s0 = g_rowptr[t];
    const int m = (g_rowptr[t + 1] - s0) * nd;
    int rc = 2;  // this row's sort length: a power of two >= m, <= cap
    while (rc < m) rc <<= 1;
    rc = min(rc, cap);  // m <= cap holds: cap covers the most slots of any dof
    const int per = (rc + kPatBlock - 1) / kPatBlock;
    for (int k = tid; k < rc; k += kPatBlock) {
      int v = INT_MAX;
      if (k < m) {
        const int q = k / nd, j = k - q * nd;
        v = dofs[(int64_t)(g_slots[s0 + q] / nd) * nd + j];
      }
      buf[k] = v;
    }
    __syncthreads();
    for (int size = 2; size <= rc; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int k = tid; k < rc / 2; k += kPatBlock) {
          const int i = 2 * k - (k & (stride - 1));
          const int a = buf[i], b = buf[i + stride];
          if ((a > b) == ((i & size) == 0)) {
            buf[i] = b;
            buf[i + stride] = a;
          }
        }
        __syncthreads();
      }
    // distinct entries of this thread's chunk [tid * per, (tid + 1) * per)
    const int k0 = min(tid * per, rc), k1 = min(k0 + per, rc);
    int c = 0;
    for (int k = k0; k < k1; k++) c += (buf[k] != INT_MAX && (k == 0 || buf[k] != buf[k - 1])) ? 1 : 0;
    cnt[tid] = c;
    __syncthreads();
    if (tid == 0) {
      int acc = 0;
      for (int i = 0; i < kPatBlock; i++) {
        const int v = cnt[i];
        cnt[i] = acc;
        acc += v;
      }
      cnt[kPatBlock] = acc;
    }
    __syncthreads();
    if (MODE == 0) {
      if (tid == 0) len[r] = cnt[kPatBlock];
    } else {
      int* out = cols + rowptr[r] + cnt[tid];
      for (int k = k0; k < k1; k++)
        if (buf[k] != INT_MAX && (k == 0 || buf[k] != buf[k - 1])) *out++ = buf[k];
    }
    __syncthreads();
  }
}

void MatFreeOp::Pattern(int64_t* h_rowptr, int* h_cols, hipStream_t st) {
  const int64_t n = height;
  if (!h_cols) {
    std::vector<int64_t> rp(n + 1, 0);
    if (n > 0 && nel > 0) {
      int cap = 2;
      while (cap < max_row_slots * nd) cap <<= 1;
      NGSAMD_REQUIRE((size_t)cap * sizeof(int) + 2048 <= (size_t)DeviceMaxLds(),
                     "ngsamd_mf_pattern: too many element dofs around one dof for LDS");
      DevBuf<int> len(n);
      NGSAMD_HIP(hipFuncSetAttribute((const void*)PatternKernel<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     cap * (int)sizeof(int)));
      PatternKernel<0><<<(int)std::min<int64_t>(n, 8192), kPatBlock, cap * sizeof(int), st>>>(
          g_rowptr.p, g_slots.p, g_order.p, n, dofs.p, nd, cap, len.p, nullptr, nullptr);
      NGSAMD_CHECK_LAUNCH();
      std::vector<int> hl(n);
      NGSAMD_HIP(hipMemcpyAsync(hl.data(), len.p, sizeof(int) * n, hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      for (int64_t r = 0; r < n; r++) rp[r + 1] = rp[r] + hl[r];
    }
    pat_rowptr.alloc(n + 1);
    pat_rowptr.upload(rp.data(), n + 1);
    NGSAMD_HIP(hipStreamSynchronize(st));
    std::copy(rp.begin(), rp.end(), h_rowptr);
    return;
  }
  NGSAMD_REQUIRE(pat_rowptr.n == (size_t)(n + 1), "ngsamd_mf_pattern: call with h_cols NULL first");
  int64_t nnz = 0;
  NGSAMD_HIP(hipMemcpy(&nnz, pat_rowptr.p + n, sizeof(int64_t), hipMemcpyDeviceToHost));
  if (nnz == 0) return;
  int cap = 2;
  while (cap < max_row_slots * nd) cap <<= 1;
  DevBuf<int> dcols(nnz);
  NGSAMD_HIP(hipFuncSetAttribute((const void*)PatternKernel<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 cap * (int)sizeof(int)));
  PatternKernel<1><<<(int)std::min<int64_t>(n, 8192), kPatBlock, cap * sizeof(int), st>>>(
      g_rowptr.p, g_slots.p, g_order.p, n, dofs.p, nd, cap, nullptr, pat_rowptr.p, dcols.p);
  NGSAMD_CHECK_LAUNCH();
  NGSAMD_HIP(hipMemcpyAsync(h_cols, dcols.p, sizeof(int) * nnz, hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
}

void MatFreeOp::ElementGrad(const double* lam, const double* u, double* grad, hipStream_t st) {
  for (size_t ci = 0; ci < classes.size(); ci++) {
    const MfClass& c = classes[ci];
    if (c.end <= c.begin) continue;
    const double* b = bref.p + ci * (size_t)nq * nd * dim;
    const int grid = GridFor(c.end - c.begin, 128, 8192);
    if (dim == 2)
      MfElementGradKernel<2><<<grid, 128, 0, st>>>(b, weights.p, nq, dofs.p, metric.p, elnr.p, c.begin, c.end, nd,
                                                   lam, u, grad);
    else
      MfElementGradKernel<3><<<grid, 128, 0, st>>>(b, weights.p, nq, dofs.p, metric.p, elnr.p, c.begin, c.end, nd,
                                                   lam, u, grad);
    NGSAMD_CHECK_LAUNCH();
  }
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_mf_create(ngsamd_op_t* op, int dim, int nd, int nq, int64_t ndof, int nclass,
                                const int64_t* off, const double* bmat, const double* w, const int* dofs,
                                const double* metric, const double* kappa, const int* elnr, int precision,
                                int scatter) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(nclass >= 1, "ngsamd_mf_create: no classes");
    NGSAMD_REQUIRE(op && off && bmat && w && (off[nclass] == 0 || (dofs && metric)), "ngsamd_mf_create: null argument");
    *op = nullptr;
    *op = new MatFreeOp(dim, nd, nq, ndof, nclass, off, bmat, w, dofs, metric, kappa, elnr, precision, scatter);
  });
}

extern "C" int ngsamd_mf_num_elements(ngsamd_op_t op, int64_t* ne) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_num_elements: not a matrix-free operator");
    NGSAMD_REQUIRE(ne, "ngsamd_mf_num_elements: null argument");
    *ne = static_cast<MatFreeOp*>(op)->ne_global;
  });
}

extern "C" int ngsamd_mf_set_coloring(ngsamd_op_t op, const int* h_colors) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_set_coloring: not a matrix-free operator");
    MatFreeOp* M = static_cast<MatFreeOp*>(op);
    NGSAMD_REQUIRE(h_colors || M->nel == 0, "ngsamd_mf_set_coloring: null argument");
    M->SetColoring(h_colors);
  });
}

extern "C" int ngsamd_mf_set_kappa(ngsamd_op_t op, const double* d_kappa, int64_t n) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_set_kappa: not a matrix-free operator");
    MatFreeOp* M = static_cast<MatFreeOp*>(op);
    NGSAMD_REQUIRE(n == M->ne_global, "ngsamd_mf_set_kappa: coefficient count does not match the element count");
    NGSAMD_REQUIRE(d_kappa || n == 0, "ngsamd_mf_set_kappa: null argument");
    M->kappa_ext = d_kappa;
  });
}

extern "C" int ngsamd_mf_mult_f32(ngsamd_op_t op, const float* x, float* y) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_mult_f32: not a matrix-free operator");
    static_cast<MatFreeOp*>(op)->MultF32(x, y, nullptr, Stream());
  });
}

extern "C" int ngsamd_mf_diagonal(ngsamd_op_t op, double* d) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_diagonal: not a matrix-free operator");
    static_cast<MatFreeOp*>(op)->Diagonal(d, Stream());
  });
}

extern "C" int ngsamd_mf_assemble_csr(ngsamd_op_t mf, const int64_t* rowptr, const int* cols, ngsamd_op_t* csr) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(mf && mf->kind == OpKind::MatrixFree, "ngsamd_mf_assemble_csr: not a matrix-free operator");
    NGSAMD_REQUIRE(csr && rowptr, "ngsamd_mf_assemble_csr: null argument");
    *csr = nullptr;
    MatFreeOp* M = static_cast<MatFreeOp*>(mf);
    NGSAMD_REQUIRE(rowptr[M->height] == 0 || cols, "ngsamd_mf_assemble_csr: null argument");
    std::unique_ptr<CsrOp> A(new CsrOp(M->height, M->width, rowptr, cols, nullptr));
    M->AssembleCsr(*A, Stream());
    *csr = A.release();
  });
}

extern "C" int ngsamd_mf_assemble_mass_csr(ngsamd_op_t mf, int nclass, int nq, int nd, int64_t ne,
                                           const double* h_weights, const double* h_shape, const double* h_scale,
                                           const int64_t* rowptr, const int* cols, ngsamd_op_t* csr) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(mf && mf->kind == OpKind::MatrixFree,
                   "ngsamd_mf_assemble_mass_csr: not a matrix-free operator");
    NGSAMD_REQUIRE(csr && rowptr && h_weights && h_shape && (h_scale || ne == 0),
                   "ngsamd_mf_assemble_mass_csr: null argument");
    *csr = nullptr;
    MatFreeOp* M = static_cast<MatFreeOp*>(mf);
    NGSAMD_REQUIRE(nclass == (int)M->classes.size() && nd == M->nd && ne == M->ne_global,
                   "ngsamd_mf_assemble_mass_csr: shape or element count does not match the operator");
    NGSAMD_REQUIRE(rowptr[M->height] == 0 || cols, "ngsamd_mf_assemble_mass_csr: null argument");
    std::unique_ptr<CsrOp> A(new CsrOp(M->height, M->width, rowptr, cols, nullptr));
    M->AssembleMassCsr(*A, nq, h_weights, h_shape, h_scale, Stream());
    *csr = A.release();
  });
}

extern "C" int ngsamd_mf_pattern(ngsamd_op_t mf, int64_t* h_rowptr, int* h_cols) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(mf && mf->kind == OpKind::MatrixFree, "ngsamd_mf_pattern: not a matrix-free operator");
    NGSAMD_REQUIRE(h_rowptr, "ngsamd_mf_pattern: null argument");
    static_cast<MatFreeOp*>(mf)->Pattern(h_rowptr, h_cols, Stream());
  });
}

extern "C" int ngsamd_mf_linear_form(ngsamd_op_t mf, int nclass, int nq, int nd, int64_t ne, const double* h_weights,
                                     const double* h_shape, const double* h_detj, const double* d_fvals,
                                     double* d_out) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(mf && mf->kind == OpKind::MatrixFree, "ngsamd_mf_linear_form: not a matrix-free operator");
    NGSAMD_REQUIRE(h_weights && h_shape && h_detj && d_out && (d_fvals || mf->height == 0),
                   "ngsamd_mf_linear_form: null argument");
    MatFreeOp* M = static_cast<MatFreeOp*>(mf);
    NGSAMD_REQUIRE(nclass == (int)M->classes.size() && nd == M->nd && ne == M->ne_global,
                   "ngsamd_mf_linear_form: shape or element count does not match the operator");
    M->LinearForm(nq, h_weights, h_shape, h_detj, d_fvals, d_out, Stream());
  });
}

extern "C" int ngsamd_mf_element_grad(ngsamd_op_t op, const double* lam, const double* u, double* grad) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::MatrixFree, "ngsamd_mf_element_grad: not a matrix-free operator");
    static_cast<MatFreeOp*>(op)->ElementGrad(lam, u, grad, Stream());
  });
}
