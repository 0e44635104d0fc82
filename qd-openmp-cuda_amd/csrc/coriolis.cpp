// coriolis.cpp - body-fixed Coriolis matrices and their eigendecomposition.
//
// Reference: setup_coriolis_matrix (CoriolisMatrix.C:63-99) builds the
// tridiagonal matrix over omega = omega_min..min(J, l) and diagonalises it
// with LAPACK dstev; setup_coriolis_matrices_and_copy_to_device
// (cudaOpenMP2.cu:158-209) packs [eigenvalues | eigenvectors] per l.
// dstev is not available here; tridiag_eigen below is a symmetric
// tridiagonal QL iteration with implicit Wilkinson shifts, accumulating the
// plane rotations into the eigenvector matrix.
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <vector>

#include "host_internal.h"

namespace qd {

// sqrt(a(a+1) - omega(omega+1)) for a >= omega, else 0 (CoriolisMatrix.C:23-33)
static double lambda_up(int a, int omega) {
  return a >= omega ? std::sqrt(a * (a + 1.0) - omega * (omega + 1.0)) : 0.0;
}

void coriolis_tridiagonal(int J, int l, int omega_min, int n, double *diag, double *off) {
  for (int k = 0; k < n; k++) {
    const int om = k + omega_min;
    diag[k] = J * (J + 1) - 2 * om * om;                        // CoriolisMatrix.C:76-79
  }
  for (int k = 0; k + 1 < n; k++) {
    const int om = k + omega_min;
    off[k] = -lambda_up(J, om) * lambda_up(l, om) * std::sqrt(om == 0 ? 2.0 : 1.0);   // :81-86
  }
}

// Symmetric tridiagonal eigenproblem.  d[0..n-1] diagonal (overwritten by the
// eigenvalues), e[0..n-2] off-diagonal (destroyed).  z is n x n column-major,
// column alpha = eigenvector alpha, i.e. z[alpha * n + i].
void tridiag_eigen(int n, double *d, double *e_in, double *z) {
  std::vector<double> e(n, 0.0);
  for (int i = 0; i + 1 < n; i++) e[i] = e_in[i];
  for (int i = 0; i < n * n; i++) z[i] = 0.0;
  for (int i = 0; i < n; i++) z[i * n + i] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; l++) {
    for (int iter = 0;; iter++) {
      int m = l;
      for (; m < n - 1; m++) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (iter > 60) throw std::runtime_error("tridiagonal eigensolver did not converge");
      // Wilkinson shift from the leading 2x2 block at l
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool deflated = false;
      for (; i >= l; i--) {
        const double f = s * e[i], bb = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {        // underflow: split the problem
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * bb;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - bb;
        for (int k = 0; k < n; k++) {   // rotate eigenvector columns i, i+1
          const double zi = z[i * n + k], zi1 = z[(i + 1) * n + k];
          z[(i + 1) * n + k] = s * zi + c * zi1;
          z[i * n + k] = c * zi - s * zi1;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
}

}  // namespace qd
