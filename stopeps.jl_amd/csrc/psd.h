// Spectrum of a symmetric tridiagonal matrix with the projections of given
// vectors onto its eigenvectors, for the PSD-projected site metric of
// physical-kind Born sampling (`_positive_metric`, src/states/tensors.jl:65-94,
// used at src/sampling/born.jl:117-134).
//
// Implicit QL iteration with Wilkinson shifts (Bowdler, Martin, Reinsch and
// Wilkinson, Numer. Math. 11 (1968) 293-306): every plane rotation that acts on
// columns (i, i+1) of the accumulated eigenvector matrix Z acts instead on the
// entries (i, i+1) of y = Z^T w, so with y starting as w the result holds
// z_j^T w for every eigenvector z_j without forming Z (O(n^2) work in all).
#pragma once
#include <hip/hip_runtime.h>
#include <cfloat>
#include <cmath>

namespace spp {

// d[0..n-1]: diagonal, overwritten by the eigenvalues (unordered);
// e[0..n-1]: e[i] = T[i][i+1] for i < n-1, e[n-1] unused, overwritten;
// y: ny vectors of length n (vector v at y + v * ystride), overwritten by the
// projections. Returns 0, or 1 when an eigenvalue needed more than maxit
// iterations (the values are then incomplete).
__host__ __device__ inline int ql_project(int n, double* __restrict__ d, double* __restrict__ e,
                                          double* __restrict__ y, long long ystride, int ny, int maxit) {
  if (n <= 0) return 0;
  e[n - 1] = 0.0;
  // A coupling is negligible below eps times either its two diagonal entries or
  // the matrix's norm (a row-sum bound, invariant up to rounding under the
  // rotations): the second keeps rank-deficient matrices, whose null space ends
  // up as blocks of entries at the rounding level of the norm, from iterating on
  // couplings that no local test can call negligible (absolute accuracy
  // eps * |T|, that of a backward-stable eigensolver).
  double anorm = 0.0;
  for (int i = 0; i < n; ++i) anorm = fmax(anorm, fabs(d[i]) + fabs(e[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0));
  const double tol = DBL_EPSILON * anorm;
  double* __restrict__ y0 = y;
  double* __restrict__ y1 = y + ystride;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    for (;;) {
      // the active block starts at l and ends at the first negligible coupling
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= DBL_EPSILON * dd || fabs(e[m]) <= tol) break;
      }
      if (m == l) break;
      if (iter++ >= maxit) return 1;
      // shift: the eigenvalue of the leading 2x2 of the block closer to d[l]
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      // the rotated entry i of each vector is carried to the next step in a
      // register (it is that step's entry i + 1)
      double c0 = ny > 0 ? y0[m] : 0.0, c1 = ny > 1 ? y1[m] : 0.0;
      int i = m - 1;
      for (; i >= l; --i) {
        const double f = s * e[i];
        const double b = c * e[i];
        // sqrt(f^2 + g^2), by hypot only where the squares could over- or underflow
        const double big = fmax(fabs(f), fabs(g));
        r = (big > 1e-150 && big < 1e150) ? sqrt(f * f + g * g) : hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          // the coupling underflowed: the block splits at i + 1
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        const double rinv = 1.0 / r;
        s = f * rinv;
        c = g * rinv;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (ny > 0) {
          const double lo = y0[i];
          y0[i + 1] = s * lo + c * c0;
          c0 = c * lo - s * c0;
        }
        if (ny > 1) {
          const double lo = y1[i];
          y1[i + 1] = s * lo + c * c1;
          c1 = c * lo - s * c1;
        }
      }
      // the last rotated entry (the step's entry l, or i + 1 where it split)
      if (ny > 0) y0[i + 1] = c0;
      if (ny > 1) y1[i + 1] = c1;
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  return 0;
}

}  // namespace spp
