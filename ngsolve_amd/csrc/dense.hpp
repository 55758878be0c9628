// Host dense kernels for the preconditioner setup.  Plain C++: no device
// code, testable on the host.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace ngsamd {

// In-place inverse of the m x m row-major matrix a: Gauss-Jordan elimination
// with partial pivoting, the inverse built in the storage of a (each step
// turns pivot column k into column k of the inverse of the rows swapped so
// far; the row swaps become column swaps of the inverse, undone at the end
// in reverse order).  Returns false when a pivot is 0 (singular matrix).
inline bool InvertInPlace(double* a, int m) {
  std::vector<int> swap_with(m);
  for (int k = 0; k < m; k++) {
    int p = k;
    double best = std::fabs(a[(size_t)k * m + k]);
    for (int i = k + 1; i < m; i++)
      if (std::fabs(a[(size_t)i * m + k]) > best) {
        best = std::fabs(a[(size_t)i * m + k]);
        p = i;
      }
    if (best == 0.0) return false;
    swap_with[k] = p;
    if (p != k)
      for (int j = 0; j < m; j++) std::swap(a[(size_t)k * m + j], a[(size_t)p * m + j]);
    const double inv_p = 1.0 / a[(size_t)k * m + k];
    a[(size_t)k * m + k] = 1.0;
    for (int j = 0; j < m; j++) a[(size_t)k * m + j] *= inv_p;
    for (int i = 0; i < m; i++) {
      if (i == k) continue;
      const double f = a[(size_t)i * m + k];
      if (f == 0.0) continue;
      a[(size_t)i * m + k] = 0.0;
      for (int j = 0; j < m; j++) a[(size_t)i * m + j] -= f * a[(size_t)k * m + j];
    }
  }
  for (int k = m - 1; k >= 0; k--)
    if (swap_with[k] != k)
      for (int i = 0; i < m; i++) std::swap(a[(size_t)i * m + k], a[(size_t)i * m + swap_with[k]]);
  return true;
}

}  // namespace ngsamd
