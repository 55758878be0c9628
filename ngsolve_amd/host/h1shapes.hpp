// H1 high-order shape functions on the reference triangle and tetrahedron,
// with exact first derivatives by forward-mode automatic differentiation.
//
// Follows the reference's hierarchical basis:
//   fem/h1hofe_impl.hpp:154-185   (ET_TRIG T_CalcShape)
//   fem/h1hofe_impl.hpp:510-570   (ET_TET T_CalcShape)
//   fem/recursive_pol.hpp:1344-1373 (IntLegNoBubble, edge polynomials)
//   fem/recursive_pol.hpp:1641-1676 (JacobiPolynomialFix<1,1>)
//   fem/recursive_pol.hpp:1857-1940 (IntegratedJacobiPolynomialAlpha)
//   fem/recursive_pol.hpp:2820-2860 (DubinerBasisOrthoBub)
//   fem/recursive_pol.hpp:2974-3000 (DubinerBasis3DOrthoBub)
//   fem/elementtopology.hpp:960-1007, 1195-1252 (local edges/faces, sorting, classes)
//   fem/tscalarfe_impl.hpp:718-730 (gradients by AutoDiff)
#pragma once
#include <array>
#include <cstddef>
#include <utility>

namespace ngsamd {

// ---------------------------------------------------------------- AutoDiff
template <int D>
struct AD {
  double v;
  double d[D];
  AD() : v(0) { for (int i = 0; i < D; i++) d[i] = 0; }
  AD(double c) : v(c) { for (int i = 0; i < D; i++) d[i] = 0; }
  static AD var(double val, int dir) {
    AD r(val);
    r.d[dir] = 1.0;
    return r;
  }
};
template <int D> inline AD<D> operator+(const AD<D>& a, const AD<D>& b) {
  AD<D> r; r.v = a.v + b.v; for (int i = 0; i < D; i++) r.d[i] = a.d[i] + b.d[i]; return r;
}
template <int D> inline AD<D> operator-(const AD<D>& a, const AD<D>& b) {
  AD<D> r; r.v = a.v - b.v; for (int i = 0; i < D; i++) r.d[i] = a.d[i] - b.d[i]; return r;
}
template <int D> inline AD<D> operator-(const AD<D>& a) {
  AD<D> r; r.v = -a.v; for (int i = 0; i < D; i++) r.d[i] = -a.d[i]; return r;
}
template <int D> inline AD<D> operator*(const AD<D>& a, const AD<D>& b) {
  AD<D> r; r.v = a.v * b.v; for (int i = 0; i < D; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r;
}
template <int D> inline AD<D> operator*(double a, const AD<D>& b) {
  AD<D> r; r.v = a * b.v; for (int i = 0; i < D; i++) r.d[i] = a * b.d[i]; return r;
}
template <int D> inline AD<D> operator*(const AD<D>& b, double a) { return a * b; }
template <int D> inline AD<D> operator+(double a, const AD<D>& b) { AD<D> r = b; r.v += a; return r; }
template <int D> inline AD<D> operator+(const AD<D>& b, double a) { AD<D> r = b; r.v += a; return r; }
template <int D> inline AD<D> operator-(double a, const AD<D>& b) { AD<D> r = -b; r.v += a; return r; }
template <int D> inline AD<D> operator-(const AD<D>& b, double a) { AD<D> r = b; r.v -= a; return r; }

// ------------------------------------------------------- 1D recursions
// IntLegNoBubble, scaled form: P0 = -1/2, P1 = -x/2,
// P_i = (2i+1)/(i+2) x P_{i-1} - (i-1)/(i+2) y^2 P_{i-2}
// (recursive_pol.hpp:1344-1373: CalcA/CalcC with the index shifted by 2).
template <class T>
inline void IntLegNoBubbleScaled(int n, T x, T y, T c, T* out) {
  if (n < 0) return;
  T p0 = c * (-0.5);
  out[0] = p0;
  if (n == 0) return;
  T p1 = c * (-0.5 * x);
  out[1] = p1;
  T yy = y * y;
  for (int i = 2; i <= n; i++) {
    double a = (2.0 * i + 1) / (i + 2.0);
    double cc = -(i - 1.0) / (i + 2.0);
    T pn = (a * x) * p1 + cc * (yy * p0);
    out[i] = pn;
    p0 = p1;
    p1 = pn;
  }
}

// Jacobi polynomials P^(1,1), scaled form (recursive_pol.hpp:1641-1676 with al=be=1):
// P0 = 1, P1 = 2x, P_i = A_i x P_{i-1} + C_i y^2 P_{i-2} (B_i = 0 since al = be).
inline double Jac11A(int i) {
  double j = i - 1;
  return (2 * j + 2) * (2 * j + 3) * (2 * j + 4) / (2 * (j + 1) * (j + 3) * (2 * j + 2));
}
inline double Jac11C(int i) {
  double j = i - 1;
  return -2.0 * (j + 1) * (j + 1) * (2 * j + 4) / (2 * (j + 1) * (j + 3) * (2 * j + 2));
}
template <class T>
inline void Jacobi11Scaled(int n, T x, T y, T c, T* out) {
  if (n < 0) return;
  T p0 = c;
  out[0] = p0;
  if (n == 0) return;
  T p1 = c * (2.0 * x);
  out[1] = p1;
  T yy = y * y;
  for (int i = 2; i <= n; i++) {
    T pn = (Jac11A(i) * x) * p1 + Jac11C(i) * (yy * p0);
    out[i] = pn;
    p0 = p1;
    p1 = pn;
  }
}

// Integrated Jacobi polynomials with parameter alpha (recursive_pol.hpp:1857-1940,
// coefficients recursive_pol.cpp:128-157):
// P0 = 1, P1 = (alpha+2)/4 x + (alpha-2)/4 [y],
// P_i = (A_i x + B_i [y]) P_{i-1} + C_i [y^2] P_{i-2}.
struct IntJacCoef {
  double a, b, c;
};
inline IntJacCoef IntJacABC(int n, double al) {
  double den = (2.0 * n + 2) * (n + al) * (2.0 * n + al - 2);
  return {(2.0 * n + al - 1) * (2.0 * n + al - 2) * (2.0 * n + al) / den,
          (2.0 * n + al - 1) * al * (al - 2) / den,
          -2.0 * (n - 1) * (n + al - 2) * (2.0 * n + al) / den};
}
template <class T>
inline void IntJacobiAlpha(int n, int alpha, T x, T c, T* out) {
  if (n < 0) return;
  T p0 = c;
  out[0] = p0;
  if (n == 0) return;
  T p1 = c * (0.25 * (alpha + 2) * x + 0.25 * (alpha - 2));
  out[1] = p1;
  for (int i = 2; i <= n; i++) {
    IntJacCoef k = IntJacABC(i, alpha);
    T pn = (k.a * x + k.b) * p1 + k.c * p0;
    out[i] = pn;
    p0 = p1;
    p1 = pn;
  }
}
template <class T>
inline void IntJacobiAlphaScaled(int n, int alpha, T x, T y, T c, T* out) {
  if (n < 0) return;
  T p0 = c;
  out[0] = p0;
  if (n == 0) return;
  T p1 = c * (0.25 * (alpha + 2) * x + 0.25 * (alpha - 2) * y);
  out[1] = p1;
  T yy = y * y;
  for (int i = 2; i <= n; i++) {
    IntJacCoef k = IntJacABC(i, alpha);
    T pn = (k.a * x + k.b * y) * p1 + k.c * (yy * p0);
    out[i] = pn;
    p0 = p1;
    p1 = pn;
  }
}

// -------------------------------------------------- 2D/3D bubble bases
constexpr int kMaxOrder = 10;

// DubinerBasisOrthoBub::EvalMult (recursive_pol.hpp:2829-2841): trig interior.
template <class T>
inline void DubinerOrthoBub(int n, T x, T y, T c, T* out) {
  if (n < 0) return;
  T leg[kMaxOrder + 1];
  Jacobi11Scaled(n, y - (1.0 - x - y), 1.0 - x, c, leg);
  int ii = 0;
  for (int i = 0; i <= n; i++) {
    IntJacobiAlpha(n - i, 4 + 2 * i, 2.0 * x - 1.0, leg[i], out + ii);
    ii += n - i + 1;
  }
}

// DubinerBasisOrthoBub::EvalScaledMult (recursive_pol.hpp:2849-2861): tet faces.
template <class T>
inline void DubinerOrthoBubScaled(int n, T x, T y, T t, T c, T* out) {
  if (n < 0) return;
  T leg[kMaxOrder + 1];
  Jacobi11Scaled(n, y - (t - x - y), t - x, c, leg);
  int ii = 0;
  for (int i = 0; i <= n; i++) {
    IntJacobiAlphaScaled(n - i, 4 + 2 * i, 2.0 * x - t, t, leg[i], out + ii);
    ii += n - i + 1;
  }
}

// DubinerBasis3DOrthoBub::EvalMult (recursive_pol.hpp:2983-3005): tet interior.
template <class T>
inline void Dubiner3DOrthoBub(int n, T x, T y, T z, T c, T* out) {
  if (n < 0) return;
  T lam4 = 1.0 - x - y - z;
  T polz[kMaxOrder + 1], polsy[kMaxOrder + 1];
  Jacobi11Scaled(n, z - lam4, z + lam4, c, polz);
  int ii = 0;
  for (int k = 0; k <= n; k++) {
    IntJacobiAlphaScaled(n - k, 4 + 2 * k, y - z - lam4, 1.0 - x, polz[k], polsy);
    for (int j = 0; j <= n - k; j++) {
      IntJacobiAlpha(n - k - j, 6 + 2 * k + 2 * j, 2.0 * x - 1.0, polsy[j], out + ii);
      ii += n - k - j + 1;
    }
  }
}

// ------------------------------------------------------ element topology
// Local edges/faces (elementtopology.hpp:960-968, 1195-1225).
constexpr int kTrigEdges[3][2] = {{2, 0}, {1, 2}, {0, 1}};
constexpr int kTetEdges[6][2] = {{3, 0}, {3, 1}, {3, 2}, {0, 1}, {0, 2}, {1, 2}};
constexpr int kTetFaces[4][3] = {{3, 1, 2}, {3, 2, 0}, {3, 0, 1}, {0, 2, 1}};

// Vertex-permutation class of an element (elementtopology.hpp:999-1007, 1241-1252).
inline int TrigClassNr(const int* v) {
  int c = 0;
  int s[3] = {0, 1, 2};
  if (v[s[0]] > v[s[1]]) { std::swap(s[0], s[1]); c += 1; }
  if (v[s[1]] > v[s[2]]) { std::swap(s[1], s[2]); c += 2; }
  if (v[s[0]] > v[s[1]]) { std::swap(s[0], s[1]); c += 2; }
  return c;
}
inline int TetClassNr(const int* v) {
  int c = 0;
  int s[4] = {0, 1, 2, 3};
  if (v[s[0]] > v[s[1]]) { std::swap(s[0], s[1]); c += 1; }
  if (v[s[2]] > v[s[3]]) { std::swap(s[2], s[3]); c += 2; }
  if (v[s[0]] > v[s[2]]) { std::swap(s[0], s[2]); c += 4; }
  if (v[s[1]] > v[s[3]]) { std::swap(s[1], s[3]); c += 8; }
  if (v[s[1]] > v[s[2]]) { std::swap(s[1], s[2]); c += 16; }
  return c;
}

// Orient a local edge so that its first vertex has the smaller global number.
inline void SortEdge(const int* vnums, int& e0, int& e1) {
  if (vnums[e0] > vnums[e1]) std::swap(e0, e1);
}
// Orient a local face by ascending global vertex number (3-comparator network).
inline void SortFace(const int* vnums, int& f0, int& f1, int& f2) {
  if (vnums[f0] > vnums[f1]) std::swap(f0, f1);
  if (vnums[f1] > vnums[f2]) std::swap(f1, f2);
  if (vnums[f0] > vnums[f1]) std::swap(f0, f1);
}

inline int TrigNDof(int p) { return (p + 1) * (p + 2) / 2; }
inline int TetNDof(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// All H1 shape functions of a uniform-order-p triangle at a point.
// Local order: 3 vertices, 3 edges * (p-1), (p-1)(p-2)/2 interior.
template <class T>
inline void TrigShapes(int p, const int* vnums, T x, T y, T* shape) {
  T lam[3] = {x, y, 1.0 - x - y};
  for (int i = 0; i < 3; i++) shape[i] = lam[i];
  int ii = 3;
  if (p >= 2)
    for (int i = 0; i < 3; i++) {
      int e0 = kTrigEdges[i][0], e1 = kTrigEdges[i][1];
      SortEdge(vnums, e0, e1);
      IntLegNoBubbleScaled(p - 2, lam[e1] - lam[e0], lam[e0] + lam[e1], lam[e0] * lam[e1], shape + ii);
      ii += p - 1;
    }
  if (p >= 3) {
    int f0 = 0, f1 = 1, f2 = 2;
    SortFace(vnums, f0, f1, f2);
    DubinerOrthoBub(p - 3, lam[f0], lam[f1], lam[f0] * lam[f1] * lam[f2], shape + ii);
  }
}

// All H1 shape functions of a uniform-order-p tetrahedron at a point.
// Local order: 4 vertices, 6 edges * (p-1), 4 faces * (p-1)(p-2)/2, interior.
template <class T>
inline void TetShapes(int p, const int* vnums, T x, T y, T z, T* shape) {
  T lam[4] = {x, y, z, 1.0 - x - y - z};
  for (int i = 0; i < 4; i++) shape[i] = lam[i];
  int ii = 4;
  if (p >= 2)
    for (int i = 0; i < 6; i++) {
      int e0 = kTetEdges[i][0], e1 = kTetEdges[i][1];
      SortEdge(vnums, e0, e1);
      IntLegNoBubbleScaled(p - 2, lam[e1] - lam[e0], lam[e0] + lam[e1], lam[e0] * lam[e1], shape + ii);
      ii += p - 1;
    }
  if (p >= 3)
    for (int i = 0; i < 4; i++) {
      int f0 = kTetFaces[i][0], f1 = kTetFaces[i][1], f2 = kTetFaces[i][2];
      SortFace(vnums, f0, f1, f2);
      int vop = 6 - f0 - f1 - f2;
      DubinerOrthoBubScaled(p - 3, lam[f0], lam[f1], 1.0 - lam[vop], lam[f0] * lam[f1] * lam[f2], shape + ii);
      ii += (p - 2) * (p - 1) / 2;
    }
  if (p >= 4) Dubiner3DOrthoBub(p - 4, lam[0], lam[1], lam[2], lam[0] * lam[1] * lam[2] * lam[3], shape + ii);
}

}  // namespace ngsamd
