// Drives the C++ operator layer (include/ngsamd_la.hpp) through the C ABI on
// the device and checks it against host arithmetic in this program:
//   - DevSparseMatrix Mult / MultAdd on a 2D 5-point Laplacian with a
//     varying coefficient (1e-14 relative, inf-norm);
//   - the BaseMatrix defaults (linalg/basematrix.cpp:120-154): an operator
//     overriding only MultAdd gets Mult = SetZero + MultAdd, one overriding
//     only Mult gets MultAdd = temporary + Mult, one overriding neither
//     throws instead of recursing;
//   - SumMatrix and ScaleMatrix composites;
//   - CGSolver with DevJacobi and DevBlockJacobi against a host
//     Jacobi-PCG with the CGSolver recurrence (python/krylovspace.py:263-291):
//     steps within 1, solution within 1e-10;
//   - an invalid CSR (column out of range) raises ngsamd::Exception.
// Prints "basematrix adapter: OK" and returns 0 on success.
//
// Usage: basematrix_adapter [--no-device]   (--no-device: construct nothing
// on the GPU, only check that the layer links and reports a state error)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "ngsamd_la.hpp"

using namespace ngsamd;

static int g_fail = 0;
#define EXPECT(cond, ...)                              \
  do {                                                 \
    if (!(cond)) {                                     \
      std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      std::fprintf(stderr, __VA_ARGS__);               \
      std::fprintf(stderr, "\n");                      \
      g_fail++;                                        \
    }                                                  \
  } while (0)

struct HostCsr {
  int n = 0;
  std::vector<size_t> firsti;
  std::vector<int> colnr;
  std::vector<double> data;
  void Mult(const std::vector<double>& x, std::vector<double>& y) const {
    y.assign(n, 0.0);
    for (int r = 0; r < n; r++) {
      double s = 0;
      for (size_t k = firsti[r]; k < firsti[r + 1]; k++) s += data[k] * x[colnr[k]];
      y[r] = s;
    }
  }
  double Diag(int r) const {
    for (size_t k = firsti[r]; k < firsti[r + 1]; k++)
      if (colnr[k] == r) return data[k];
    return 0.0;
  }
};

// -div(c grad u) on an m x m grid of interior points, 5-point stencil,
// c(x, y) = 1 + 0.9 sin(3x) cos(2y) at the edge midpoints
static HostCsr Laplacian(int m) {
  HostCsr A;
  A.n = m * m;
  A.firsti.push_back(0);
  auto c = [&](double x, double y) { return 1.0 + 0.9 * std::sin(3 * x) * std::cos(2 * y); };
  const double h = 1.0 / (m + 1);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < m; j++) {
      const double x = (j + 1) * h, y = (i + 1) * h;
      const double cw = c(x - h / 2, y), ce = c(x + h / 2, y), cs = c(x, y - h / 2), cn = c(x, y + h / 2);
      auto add = [&](int col, double v) {
        A.colnr.push_back(col);
        A.data.push_back(v);
      };
      if (i > 0) add((i - 1) * m + j, -cs);
      if (j > 0) add(i * m + j - 1, -cw);
      add(i * m + j, cw + ce + cs + cn);
      if (j < m - 1) add(i * m + j + 1, -ce);
      if (i < m - 1) add((i + 1) * m + j, -cn);
      A.firsti.push_back(A.colnr.size());
    }
  return A;
}

static double RelInf(const double* a, const std::vector<double>& b) {
  double d = 0, m = 0;
  for (size_t i = 0; i < b.size(); i++) {
    d = std::max(d, std::fabs(a[i] - b[i]));
    m = std::max(m, std::fabs(b[i]));
  }
  return d / (m > 0 ? m : 1);
}

// host PCG, CGSolver recurrence; pre(d, w) computes w = C d
template <class Pre>
static int HostPcg(const HostCsr& A, Pre pre, const std::vector<double>& f, std::vector<double>& u, double tol,
                   int maxsteps) {
  const int n = A.n;
  u.assign(n, 0.0);
  std::vector<double> d = f, w(n), s(n), t(n);
  pre(d, w);
  s = w;
  auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
    double r = 0;
    for (int i = 0; i < n; i++) r += a[i] * b[i];
    return r;
  };
  double wdn = dot(w, d);
  int checks = 1;
  const double stop = std::sqrt(std::fabs(wdn)) * tol;
  if (std::sqrt(std::fabs(wdn)) <= stop || checks >= maxsteps) return checks;
  while (true) {
    A.Mult(s, t);
    const double wd = wdn, as = dot(s, t);
    if (as == 0 || wd == 0) break;
    const double al = wd / as;
    for (int i = 0; i < n; i++) {
      u[i] += al * s[i];
      d[i] -= al * t[i];
    }
    pre(d, w);
    wdn = dot(w, d);
    checks++;
    if (checks >= maxsteps || std::sqrt(std::fabs(wdn)) <= stop) break;
    const double be = wdn / wd;
    for (int i = 0; i < n; i++) s[i] = be * s[i] + w[i];
  }
  return checks;
}

// operators defined through the BaseMatrix defaults
struct OnlyMultAdd : BaseMatrix {  // 2 A, through MultAdd only
  std::shared_ptr<BaseMatrix> a;
  int VHeight() const override { return a->Height(); }
  int VWidth() const override { return a->Width(); }
  void MultAdd(double s, const BaseVector& x, BaseVector& y) const override { a->MultAdd(2 * s, x, y); }
};
struct OnlyMult : BaseMatrix {  // 3 A, through Mult only
  std::shared_ptr<BaseMatrix> a;
  int VHeight() const override { return a->Height(); }
  int VWidth() const override { return a->Width(); }
  void Mult(const BaseVector& x, BaseVector& y) const override {
    a->Mult(x, y);
    y.Scale(3.0);
  }
};
struct Neither : BaseMatrix {
  int VHeight() const override { return 4; }
  int VWidth() const override { return 4; }
};

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "--no-device") == 0) {
    // no ngsamd_init: any entry point reports a state error, as an exception
    bool threw = false;
    try {
      Check(ngsamd_synchronize());
    } catch (const Exception& e) {
      threw = std::strstr(e.what(), "ngsamd_init") != nullptr;
    }
    EXPECT(threw, "state error expected before ngsamd_init");
    std::printf(g_fail ? "basematrix adapter (no device): FAILED\n" : "basematrix adapter (no device): OK\n");
    return g_fail ? 1 : 0;
  }
  Check(ngsamd_init(0));
  const int m = 48;
  HostCsr H = Laplacian(m);
  const int n = H.n;
  auto A = std::make_shared<DevSparseMatrix>(n, n, H.firsti.data(), H.colnr.data(), H.data.data());
  EXPECT(A->Height() == n && A->Width() == n, "shape");

  std::mt19937_64 rng(5);
  std::uniform_real_distribution<double> U(-1, 1);
  AutoVector x = A->CreateRowVector(), y = A->CreateColVector();
  std::vector<double> hx(n), hy, hy0(n);
  for (int i = 0; i < n; i++) hx[i] = U(rng);
  std::memcpy(x->FVDouble(), hx.data(), n * sizeof(double));
  x->HostModified();
  H.Mult(hx, hy);

  A->Mult(*x, *y);
  EXPECT(RelInf(y->FVDouble(), hy) < 1e-14, "Mult: %g", RelInf(y->FVDouble(), hy));
  for (int i = 0; i < n; i++) hy0[i] = U(rng);
  std::memcpy(y->FVDouble(), hy0.data(), n * sizeof(double));
  y->HostModified();
  A->MultAdd(-0.5, *x, *y);
  std::vector<double> ref(n);
  for (int i = 0; i < n; i++) ref[i] = hy0[i] - 0.5 * hy[i];
  EXPECT(RelInf(y->FVDouble(), ref) < 1e-14, "MultAdd: %g", RelInf(y->FVDouble(), ref));

  // BaseMatrix defaults
  auto oma = std::make_shared<OnlyMultAdd>();
  oma->a = A;
  oma->Mult(*x, *y);  // SetZero + MultAdd
  for (int i = 0; i < n; i++) ref[i] = 2 * hy[i];
  EXPECT(RelInf(y->FVDouble(), ref) < 1e-14, "default Mult");
  auto om = std::make_shared<OnlyMult>();
  om->a = A;
  y->SetScalar(1.0);
  om->MultAdd(2.0, *x, *y);  // temporary + Mult
  for (int i = 0; i < n; i++) ref[i] = 1.0 + 6 * hy[i];
  EXPECT(RelInf(y->FVDouble(), ref) < 1e-14, "default MultAdd");
  Neither nei;
  DeviceVector x4(4), y4(4);
  bool threw = false;
  try {
    nei.Mult(x4, y4);
  } catch (const Exception& e) {
    threw = std::strstr(e.what(), "neither Mult nor MultAdd") != nullptr;
  }
  EXPECT(threw, "an operator overriding neither Mult nor MultAdd must throw");

  // composites: (A + 2A) - 0.5 A = 2.5 A
  auto sum = std::make_shared<SumMatrix>(A, oma);
  SumMatrix comb(sum, std::make_shared<ScaleMatrix>(-0.5, A));
  comb.Mult(*x, *y);
  for (int i = 0; i < n; i++) ref[i] = 2.5 * hy[i];
  EXPECT(RelInf(y->FVDouble(), ref) < 1e-13, "composites: %g", RelInf(y->FVDouble(), ref));
  EXPECT(comb.IsSymmetric() == false, "OnlyMultAdd is not declared symmetric");

  // CG with Jacobi and with block Jacobi (blocks of 4 consecutive dofs)
  std::vector<double> diag(n), f(n);
  for (int i = 0; i < n; i++) {
    diag[i] = H.Diag(i);
    f[i] = 1.0 + 0.5 * std::sin(0.01 * i);
  }
  AutoVector df = A->CreateColVector(), du = A->CreateRowVector();
  std::memcpy(df->FVDouble(), f.data(), n * sizeof(double));
  df->HostModified();
  auto J = std::make_shared<DevJacobi>(n, diag.data(), nullptr);
  CGSolver inv(A, J, 1e-10, 5000);
  inv.Mult(*df, *du);
  std::vector<double> u_ref;
  const int steps_ref = HostPcg(
      H, [&](const std::vector<double>& d, std::vector<double>& w) {
        for (int i = 0; i < n; i++) w[i] = d[i] / diag[i];
      },
      f, u_ref, 1e-10, 5000);
  EXPECT(std::abs(inv.GetSteps() - steps_ref) <= 1, "Jacobi CG steps %d vs host %d", inv.GetSteps(), steps_ref);
  {
    const double* u = du->FVDouble();
    double e = 0, nr = 0;
    for (int i = 0; i < n; i++) {
      e += (u[i] - u_ref[i]) * (u[i] - u_ref[i]);
      nr += u_ref[i] * u_ref[i];
    }
    EXPECT(std::sqrt(e / nr) <= 1e-10, "Jacobi CG solution %g", std::sqrt(e / nr));
  }
  std::vector<int64_t> bptr;
  std::vector<int> bdofs;
  for (int b = 0; b * 4 < n; b++) {
    bptr.push_back(bdofs.size());
    for (int k = b * 4; k < std::min(n, b * 4 + 4); k++) bdofs.push_back(k);
  }
  bptr.push_back(bdofs.size());
  auto BJ = std::make_shared<DevBlockJacobi>(n, H.firsti.data(), H.colnr.data(), H.data.data(),
                                             (int64_t)bptr.size() - 1, bptr.data(), bdofs.data());
  CGSolver invb(A, BJ, 1e-10, 5000);
  invb.Mult(*df, *du);
  // host block Jacobi: the 4x4 diagonal blocks solved by Gaussian elimination
  auto block_pre = [&](const std::vector<double>& d, std::vector<double>& w) {
    for (size_t b = 0; b + 1 < bptr.size(); b++) {
      const int b0 = (int)bptr[b], bs = (int)(bptr[b + 1] - bptr[b]);
      double M[4][5];
      for (int r = 0; r < bs; r++) {
        const int row = bdofs[b0 + r];
        for (int c = 0; c < bs; c++) {
          M[r][c] = 0.0;
          for (size_t k = H.firsti[row]; k < H.firsti[row + 1]; k++)
            if (H.colnr[k] == bdofs[b0 + c]) M[r][c] = H.data[k];
        }
        M[r][bs] = d[row];
      }
      for (int k = 0; k < bs; k++)
        for (int r = k + 1; r < bs; r++) {
          const double l = M[r][k] / M[k][k];
          for (int c = k; c <= bs; c++) M[r][c] -= l * M[k][c];
        }
      for (int r = bs - 1; r >= 0; r--) {
        double v = M[r][bs];
        for (int c = r + 1; c < bs; c++) v -= M[r][c] * w[bdofs[b0 + c]];
        w[bdofs[b0 + r]] = v / M[r][r];
      }
    }
  };
  const int steps_b = HostPcg(H, block_pre, f, u_ref, 1e-10, 5000);
  EXPECT(std::abs(invb.GetSteps() - steps_b) <= 1, "block CG steps %d vs host %d", invb.GetSteps(), steps_b);
  {
    const double* u = du->FVDouble();
    double e = 0, nr = 0;
    for (int i = 0; i < n; i++) {
      e += (u[i] - u_ref[i]) * (u[i] - u_ref[i]);
      nr += u_ref[i] * u_ref[i];
    }
    EXPECT(std::sqrt(e / nr) <= 1e-10, "block CG solution %g", std::sqrt(e / nr));
  }
  EXPECT((int)inv.Residuals().size() == inv.GetSteps(), "residual history length");

  // errors: a column number beyond the width
  threw = false;
  try {
    std::vector<size_t> fi = {0, 1};
    std::vector<int> cn = {7};
    std::vector<double> dv = {1.0};
    DevSparseMatrix bad(1, 3, fi.data(), cn.data(), dv.data());
  } catch (const Exception& e) {
    threw = true;
  }
  EXPECT(threw, "column out of range must throw");

  std::printf("steps: Jacobi %d (host %d), block Jacobi %d (host %d)\n", inv.GetSteps(), steps_ref, invb.GetSteps(),
              steps_b);
  std::printf(g_fail ? "basematrix adapter: FAILED (%d)\n" : "basematrix adapter: OK\n", g_fail);
  return g_fail ? 1 : 0;
}
