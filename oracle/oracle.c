/* CPU oracle kernels (test infrastructure only; used by tests/, the smoke
 * check and bench.py's cpu_baseline leg, never by the product path).
 *
 *   oracle_csr_spmv   SparseMatrix<double>::MultAdd, linalg/sparsematrix_impl.hpp:264-306:
 *                     row-parallel y += s * sum_k vals[k] x[cols[k]]
 *   oracle_mf_apply   MatrixFreeBTDTB::MultAdd, comp/matrixfree.cpp:681-1087: per element
 *                     and quadrature point px = B^T x_e (reference gradient),
 *                     hx = F^-T px, hy = kappa w_q |det F| hx, py = F^-1 hy,
 *                     y_e += B py, scattered with atomic adds (MFOpts atomic=true)
 *   oracle_pcg_csr    CGSolver._SolveImpl, python/krylovspace.py:263-291, with the
 *                     Jacobi preconditioner of linalg/jacobi.cpp:72-90
 *
 * Built with:  gcc -O2 -fopenmp -shared -fPIC oracle.c -o liboracle.so
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { omp_set_num_threads(n); }

void oracle_csr_spmv(int64_t n, const int64_t* rowptr, const int* cols, const double* vals, const double* x,
                     double* y, double s, int accumulate) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; r++) {
    double sum = 0.0;
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; k++) sum += vals[k] * x[cols[k]];
    y[r] = accumulate ? y[r] + s * sum : s * sum;
  }
}

/* bref: [nclass][nq][nd][dim]; el_class: class index per element;
 * finv: [nel][dim][dim] (F^{-1}); detf: [nel] (|det F|); kappa: [nel] or NULL */
void oracle_mf_apply(int dim, int nd, int nq, int64_t nel, const double* bref, const int* el_class, const double* w,
                     const int* dofs, const double* finv, const double* detf, const double* kappa, const double* x,
                     double* y, double s) {
#pragma omp parallel
  {
    double xe[128], ye[128];
#pragma omp for schedule(static)
    for (int64_t e = 0; e < nel; e++) {
      const double* B = bref + (size_t)el_class[e] * nq * nd * dim;
      const double* Fi = finv + (size_t)e * dim * dim;
      const double D = (kappa ? kappa[e] : 1.0) * detf[e];
      for (int i = 0; i < nd; i++) {
        xe[i] = x[dofs[e * nd + i]];
        ye[i] = 0.0;
      }
      for (int q = 0; q < nq; q++) {
        const double* Bq = B + (size_t)q * nd * dim;
        double px[3] = {0, 0, 0}, hx[3] = {0, 0, 0}, py[3] = {0, 0, 0};
        for (int i = 0; i < nd; i++)
          for (int a = 0; a < dim; a++) px[a] += Bq[i * dim + a] * xe[i];
        /* hx = F^{-T} px */
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++) hx[a] += Fi[b * dim + a] * px[b];
        const double c = D * w[q];
        /* py = F^{-1} (c hx) */
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++) py[a] += Fi[a * dim + b] * c * hx[b];
        for (int i = 0; i < nd; i++)
          for (int a = 0; a < dim; a++) ye[i] += Bq[i * dim + a] * py[a];
      }
      for (int i = 0; i < nd; i++) {
#pragma omp atomic
        y[dofs[e * nd + i]] += s * ye[i];
      }
    }
  }
}

static double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

typedef struct {
  int it, maxiter;
  double tol, atol, final_res;
  double* residuals;
} Check;

/* LinearSolver.CheckResidual, python/krylovspace.py:129-150 */
static int check_residual(Check* c, double res) {
  c->it++;
  if (c->it <= c->maxiter) c->residuals[c->it - 1] = res;
  if (c->it == 1) {
    if (c->tol < 0) c->final_res = c->atol;
    else {
      c->final_res = res * c->tol;
      if (c->atol >= 0 && c->atol > c->final_res) c->final_res = c->atol;
    }
  }
  return c->it >= c->maxiter || res <= c->final_res;
}

/* Returns the number of residual checks (the reference's `iterations`);
 * residuals[] receives up to maxiter entries. tol < 0: not given;
 * atol < 0: not given.  u is initialised to zero. */
int oracle_pcg_csr(int64_t n, const int64_t* rowptr, const int* cols, const double* vals, const double* invdiag,
                   const double* f, double* u, double tol, double atol, int maxiter, double* residuals) {
  double* d = (double*)malloc(n * sizeof(double));
  double* w = (double*)malloc(n * sizeof(double));
  double* s = (double*)malloc(n * sizeof(double));
  Check c = {0, maxiter, tol, atol, 0.0, residuals};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    u[i] = 0.0;
    d[i] = f[i];
    w[i] = invdiag[i] * d[i];
    s[i] = w[i];
  }
  double wdn = dot(n, w, d);
  if (!check_residual(&c, sqrt(fabs(wdn)))) {
    for (;;) {
      oracle_csr_spmv(n, rowptr, cols, vals, s, w, 1.0, 0);
      const double wd = wdn;
      const double as_s = dot(n, s, w);
      if (as_s == 0.0 || wd == 0.0) break;
      const double alpha = wd / as_s;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) {
        u[i] += alpha * s[i];
        d[i] -= alpha * w[i];
        w[i] = invdiag[i] * d[i];
      }
      wdn = dot(n, w, d);
      if (check_residual(&c, sqrt(fabs(wdn)))) break;
      const double beta = wdn / wd;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; i++) s[i] = beta * s[i] + w[i];
    }
  }
  free(d);
  free(w);
  free(s);
  return c.it;
}
