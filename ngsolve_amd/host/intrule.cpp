#include "intrule.hpp"

#include <cmath>
#include <stdexcept>

namespace ngsamd {

// Number of eigenvalues of the symmetric tridiagonal matrix (diag a, offdiag b)
// that are smaller than lam (Sturm sequence count).
static int SturmCount(const std::vector<double>& a, const std::vector<double>& b, double lam) {
  int n = (int)a.size(), cnt = 0;
  double q = 1.0;
  for (int i = 0; i < n; i++) {
    double off = (i == 0) ? 0.0 : b[i] * b[i] / q;
    q = (a[i] - lam) - off;
    if (q == 0.0) q = -1e-300;
    if (q < 0) cnt++;
  }
  return cnt;
}

void GaussJacobi01(int n, int alpha, std::vector<double>& x, std::vector<double>& w) {
  if (n < 1) throw std::runtime_error("GaussJacobi01: n must be >= 1");
  const double al = alpha, be = 0.0;
  // Jacobi matrix of the orthonormal Jacobi polynomials on [-1,1].
  std::vector<double> a(n), b(n, 0.0);
  for (int k = 0; k < n; k++) {
    double s = 2.0 * k + al + be;
    a[k] = (k == 0) ? (be - al) / (al + be + 2.0) : (be * be - al * al) / (s * (s + 2.0));
    if (k >= 1)
      b[k] = std::sqrt(4.0 * k * (k + al) * (k + be) * (k + al + be) / (s * s * (s + 1.0) * (s - 1.0)));
  }
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  double mu0 = std::pow(2.0, al + 1.0) / (al + 1.0);
  for (int k = 0; k < n; k++) {
    double lo = -1.0, hi = 1.0;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      if (SturmCount(a, b, mid) > k) hi = mid; else lo = mid;
    }
    double xi = 0.5 * (lo + hi);
    // orthonormal polynomial values -> Christoffel weight
    double pm1 = 0.0, p0 = 1.0, sum = 1.0;
    for (int j = 0; j + 1 < n; j++) {
      double pn = ((xi - a[j]) * p0 - b[j] * pm1) / b[j + 1];
      sum += pn * pn;
      pm1 = p0;
      p0 = pn;
    }
    x[k] = 0.5 * (xi + 1.0);
    w[k] = mu0 / sum * std::pow(0.5, al + 1.0);
  }
}

static void AddPoint(IntegrationRule& ir, double x, double y, double w) {
  ir.pts.push_back(x);
  ir.pts.push_back(y);
  ir.wts.push_back(w);
}

IntegrationRule TrigRule(int order) {
  IntegrationRule ir;
  ir.dim = 2;
  ir.order = order;
  if (order < 0) order = 0;
  if (order <= 1) {
    AddPoint(ir, 1.0 / 3.0, 1.0 / 3.0, 0.5);
  } else if (order == 2) {
    AddPoint(ir, 0, 0.5, 1.0 / 6.0);
    AddPoint(ir, 0.5, 0, 1.0 / 6.0);
    AddPoint(ir, 0.5, 0.5, 1.0 / 6.0);
  } else if (order <= 4) {
    // Strang-Fix / Dunavant degree-4 rule, 6 points (fem/intrule.cpp:2524-2540).
    const double a = 0.816847572980459, b = 0.091576213509771;
    const double c = 0.108103018168070, d = 0.445948490915965;
    const double wa = 0.054975871827661, wc = 0.111690794839005;
    AddPoint(ir, a, b, wa); AddPoint(ir, b, a, wa); AddPoint(ir, b, b, wa);
    AddPoint(ir, c, d, wc); AddPoint(ir, d, c, wc); AddPoint(ir, d, d, wc);
  } else if (order <= 6) {
    // Strang-Fix / Dunavant degree-6 rule, 12 points (fem/intrule.cpp:2545-2571).
    const double a = 0.873821971016996, b = 0.063089014491502;
    const double c = 0.501426509658179, d = 0.249286745170910;
    const double e = 0.636502499121399, f = 0.310352451033785, g = 0.053145049844816;
    const double wa = 0.025422453185103, wc = 0.058393137863189, we = 0.041425537809187;
    AddPoint(ir, a, b, wa); AddPoint(ir, b, a, wa); AddPoint(ir, b, b, wa);
    AddPoint(ir, c, d, wc); AddPoint(ir, d, c, wc); AddPoint(ir, d, d, wc);
    AddPoint(ir, e, f, we); AddPoint(ir, f, g, we); AddPoint(ir, g, e, we);
    AddPoint(ir, e, g, we); AddPoint(ir, f, e, we); AddPoint(ir, g, f, we);
  } else {
    int n = order / 2 + 1;
    std::vector<double> xx, wx, xy, wy;
    GaussJacobi01(n, 1, xx, wx);
    GaussJacobi01(n, 0, xy, wy);
    for (int j = 0; j < n; j++)
      for (int i = n - 1; i >= 0; i--)  // Gauss-Jacobi nodes descending
        AddPoint(ir, xx[i], xy[j] * (1 - xx[i]), wx[i] * wy[j]);
  }
  return ir;
}

IntegrationRule TetRule(int order) {
  IntegrationRule ir;
  ir.dim = 3;
  ir.order = order;
  if (order < 0) order = 0;
  int n = order / 2 + 1;
  std::vector<double> xx, wx, xy, wy, xz, wz;
  GaussJacobi01(n, 2, xx, wx);
  GaussJacobi01(n, 1, xy, wy);
  GaussJacobi01(n, 0, xz, wz);
  for (int k = 0; k < n; k++)
    for (int j = n - 1; j >= 0; j--)
      for (int i = n - 1; i >= 0; i--) {
        ir.pts.push_back(xx[i]);
        ir.pts.push_back(xy[j] * (1 - xx[i]));
        ir.pts.push_back(xz[k] * (1 - xx[i]) * (1 - xy[j]));
        ir.wts.push_back(wx[i] * wy[j] * wz[k]);
      }
  return ir;
}

}  // namespace ngsamd
