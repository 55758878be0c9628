// Integration rules on the reference triangle {x,y>=0, x+y<=1} and
// tetrahedron {x,y,z>=0, x+y+z<=1}.
//
// Triangle orders 0..6 use the fixed symmetric rules the reference tabulates
// (fem/intrule.cpp:2490-2571): 1 point (order 0-1), 3 edge midpoints (order 2),
// Strang-Fix/Dunavant 6 points (orders 3-4) and 12 points (orders 5-6).
// Higher triangle orders and all tetrahedron orders use the collapsed
// (Duffy) tensor rules of fem/intrule.cpp:3143-3166 and 3198-3232:
//   trig:  n = order/2+1, Gauss-Jacobi(alpha=1) in x, Gauss in y;
//          ip = (x_i, y_j (1-x_i)), w = W1_i * W0_j
//   tet:   Gauss-Jacobi(alpha=2) in x, Gauss-Jacobi(alpha=1) in y, Gauss in z,
//          x innermost; ip = (x, y(1-x), z(1-x)(1-y)), w = W2_i W1_j W0_k
// where W_alpha are the Gauss-Jacobi weights of int_0^1 (1-t)^alpha f(t) dt.
// Node ordering follows the reference: Gauss nodes ascending, Gauss-Jacobi
// nodes descending.  The nodes and weights are computed here with a
// Sturm-sequence bisection on the Jacobi matrix of the orthonormal Jacobi
// polynomials (Golub-Welsch recurrence coefficients), not with the
// reference's Newton iteration; both give the unique Gauss-Jacobi rule.
// Tetrahedron orders 0..6 in the reference use tabulated Keast-type rules;
// here the collapsed rule serves every tetrahedron order (exact to the
// same polynomial degree, more points).
#pragma once
#include <vector>

namespace ngsamd {

struct IntegrationRule {
  int dim = 0;
  int order = 0;
  std::vector<double> pts;  // npts * dim
  std::vector<double> wts;  // npts
  int Size() const { return (int)wts.size(); }
};

// Gauss-Jacobi rule for int_0^1 (1-t)^alpha f(t) dt with n points;
// nodes returned in ascending order.
void GaussJacobi01(int n, int alpha, std::vector<double>& x, std::vector<double>& w);

IntegrationRule TrigRule(int order);
IntegrationRule TetRule(int order);

}  // namespace ngsamd
