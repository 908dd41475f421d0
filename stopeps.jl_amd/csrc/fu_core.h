// Per-bond solve of the full update (src/evolution/full_update.jl:31-117,
// 342-482): environment gauge, gauged truncated SVD start, ALS with
// Jacobi-preconditioned CG, bond-centre balance. FP64, one sample per call;
// compiled for the device (one thread per sample, arrays interleaved by
// sample so that a wavefront's accesses coalesce) and for the host (selftest).
// All indices are parity-sorted (even entries first on every leg).
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define FU_HD __host__ __device__
#else
#define FU_HD
#endif

namespace spp {
namespace fu {

// Strided view: element i at p[i * s].
struct View {
  double* p;
  long long s;
  FU_HD double& operator[](long long i) const { return p[i * s]; }
};

struct Work {
  double* base;
  long long stride;
  long long top;
  long long cap;
  FU_HD View take(long long n) {
    View v{base + top * stride, stride};
    top += n;
    return v;
  }
};

struct Dims {
  int na, ea;   // A-side reduced core dimension and its even count
  int nb, eb;   // B-side
  int d, ed;    // physical
  int D, eD;    // bond (device capacity, parity-sorted)
};

struct Opts {
  int max_iterations;
  double error_tolerance;
  double diagonal_shift;
  int linear_max_iterations;
  double linear_tolerance;
};

enum Stop { kDisabled = 0, kIterationLimit = 1, kErrorTolerance = 2, kResidualFallback = 3 };
enum Status { kOk = 0, kNonfiniteMetric = 1, kZeroMetric = 2, kZeroTarget = 3, kRepairNeeded = 4,
              kNonfiniteCost = 5, kSingularGauge = 6, kWorkspace = 7 };

struct Result {
  double initial_cost, final_cost, residual, target_norm;
  int iterations, limit_reached, stop, status;
  int kept_even, kept_odd, floored;
};

constexpr double kEps = 2.220446049250313e-16;
constexpr double kTiny = 2.2250738585072014e-308;
constexpr double kRoundoff = 1e-8;  // RESIDUAL_ROUNDOFF_TOLERANCE, full_update.jl:18

FU_HD inline int par(int i, int e) { return i >= e ? 1 : 0; }
FU_HD inline bool finite(double x) { return fabs(x) <= 1.7976931348623157e308; }  // false for inf and NaN

// ---------------------------------------------------------------------------
// dense kernels on small row-major matrices
// ---------------------------------------------------------------------------

// Cyclic Jacobi eigensolver of the symmetric n x n matrix A (destroyed):
// eigenvalues in w, eigenvectors as columns of V (V[i*n+k] = component i of k).
FU_HD inline void jacobi_eigh(View A, int n, View w, View V) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[(long long)i * n + j] = i == j ? 1.0 : 0.0;
  double fro = 0.0;
  for (long long i = 0; i < (long long)n * n; ++i) fro += A[i] * A[i];
  for (int sweep = 0; sweep < 80; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[(long long)p * n + q] * A[(long long)p * n + q];
    if (off <= 1e-34 * fro || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[(long long)p * n + q];
        if (apq == 0.0) continue;
        const double app = A[(long long)p * n + p], aqq = A[(long long)q * n + q];
        if (fabs(apq) <= 1e-300) continue;
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = A[(long long)k * n + p], akq = A[(long long)k * n + q];
          A[(long long)k * n + p] = c * akp - s * akq;
          A[(long long)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[(long long)p * n + k], aqk = A[(long long)q * n + k];
          A[(long long)p * n + k] = c * apk - s * aqk;
          A[(long long)q * n + k] = s * apk + c * aqk;
        }
        A[(long long)p * n + q] = 0.0;
        A[(long long)q * n + p] = 0.0;
        for (int k = 0; k < n; ++k) {
          const double vkp = V[(long long)k * n + p], vkq = V[(long long)k * n + q];
          V[(long long)k * n + p] = c * vkp - s * vkq;
          V[(long long)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) w[i] = A[(long long)i * n + i];
}

// Householder QR of the m x n matrix M (row-major, destroyed): Q (m x k),
// R (k x n), k = min(m, n), R with a nonnegative diagonal. v: m scratch.
FU_HD inline void householder_qr(View M, int m, int n, View Q, View R, View vs, View tau) {
  const int k = m < n ? m : n;
  for (int j = 0; j < k; ++j) {
    double nrm = 0.0;
    for (int i = j; i < m; ++i) nrm += M[(long long)i * n + j] * M[(long long)i * n + j];
    nrm = sqrt(nrm);
    if (nrm == 0.0) {
      tau[j] = 0.0;
      for (int i = 0; i < m; ++i) vs[(long long)i * k + j] = 0.0;
      continue;
    }
    const double x0 = M[(long long)j * n + j];
    const double alpha = x0 >= 0 ? -nrm : nrm;
    // v = x - alpha e_j, normalised so that v_j = 1
    const double v0 = x0 - alpha;
    for (int i = 0; i < m; ++i) vs[(long long)i * k + j] = i < j ? 0.0 : (i == j ? 1.0 : M[(long long)i * n + j] / v0);
    double vv = 0.0;
    for (int i = j; i < m; ++i) vv += vs[(long long)i * k + j] * vs[(long long)i * k + j];
    tau[j] = 2.0 / vv;
    for (int c = j; c < n; ++c) {
      double dot = 0.0;
      for (int i = j; i < m; ++i) dot += vs[(long long)i * k + j] * M[(long long)i * n + c];
      dot *= tau[j];
      for (int i = j; i < m; ++i) M[(long long)i * n + c] -= dot * vs[(long long)i * k + j];
    }
  }
  for (int i = 0; i < k; ++i)
    for (int c = 0; c < n; ++c) R[(long long)i * n + c] = c >= i ? M[(long long)i * n + c] : 0.0;
  // Q = H_0 ... H_{k-1} I[:, :k]
  for (int i = 0; i < m; ++i)
    for (int c = 0; c < k; ++c) Q[(long long)i * k + c] = i == c ? 1.0 : 0.0;
  for (int j = k - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    for (int c = 0; c < k; ++c) {
      double dot = 0.0;
      for (int i = j; i < m; ++i) dot += vs[(long long)i * k + j] * Q[(long long)i * k + c];
      dot *= tau[j];
      for (int i = j; i < m; ++i) Q[(long long)i * k + c] -= dot * vs[(long long)i * k + j];
    }
  }
  for (int i = 0; i < k; ++i)
    if (R[(long long)i * n + i] < 0.0) {
      for (int c = 0; c < n; ++c) R[(long long)i * n + c] = -R[(long long)i * n + c];
      for (int r = 0; r < m; ++r) Q[(long long)r * k + i] = -Q[(long long)r * k + i];
    }
}

// One-sided Jacobi SVD of the m x n matrix M (row-major, destroyed) with
// m >= n: U (m x n), S (n, descending), Vt (n x n). Vw: n*n scratch.
FU_HD inline void jacobi_svd_tall(View M, int m, int n, View U, View S, View Vt, View Vw) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Vw[(long long)i * n + j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = 0; i < m; ++i) {
          const double x = M[(long long)i * n + p], y = M[(long long)i * n + q];
          a += x * x;
          b += y * y;
          g += x * y;
        }
        if (g == 0.0 || fabs(g) <= 1e-15 * sqrt(a * b)) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
        for (int i = 0; i < m; ++i) {
          const double x = M[(long long)i * n + p], y = M[(long long)i * n + q];
          M[(long long)i * n + p] = c * x - s * y;
          M[(long long)i * n + q] = s * x + c * y;
        }
        for (int i = 0; i < n; ++i) {
          const double x = Vw[(long long)i * n + p], y = Vw[(long long)i * n + q];
          Vw[(long long)i * n + p] = c * x - s * y;
          Vw[(long long)i * n + q] = s * x + c * y;
        }
      }
    if (!rotated) break;
  }
  // column norms, then a selection sort into descending order
  for (int j = 0; j < n; ++j) {
    double a = 0.0;
    for (int i = 0; i < m; ++i) a += M[(long long)i * n + j] * M[(long long)i * n + j];
    S[j] = sqrt(a);
  }
  for (int j = 0; j < n; ++j) {
    int best = j;
    for (int c = j + 1; c < n; ++c)
      if (S[c] > S[best]) best = c;
    if (best != j) {
      const double t = S[j];
      S[j] = S[best];
      S[best] = t;
      for (int i = 0; i < m; ++i) {
        const double x = M[(long long)i * n + j];
        M[(long long)i * n + j] = M[(long long)i * n + best];
        M[(long long)i * n + best] = x;
      }
      for (int i = 0; i < n; ++i) {
        const double x = Vw[(long long)i * n + j];
        Vw[(long long)i * n + j] = Vw[(long long)i * n + best];
        Vw[(long long)i * n + best] = x;
      }
    }
  }
  for (int j = 0; j < n; ++j) {
    const double inv = S[j] > 0.0 ? 1.0 / S[j] : 0.0;
    for (int i = 0; i < m; ++i) U[(long long)i * n + j] = M[(long long)i * n + j] * inv;
    for (int i = 0; i < n; ++i) Vt[(long long)j * n + i] = Vw[(long long)i * n + j];
  }
}

// Compact SVD of an m x n matrix (any shape): U (m x k), S (k), Vt (k x n),
// k = min(m, n). work: m*n + k*k + max(m,n)*k doubles.
FU_HD inline void svd_compact(View M, int m, int n, View U, View S, View Vt, Work& w) {
  const long long mark = w.top;
  if (m >= n) {
    View T = w.take((long long)m * n);
    for (long long i = 0; i < (long long)m * n; ++i) T[i] = M[i];
    View Vw = w.take((long long)n * n);
    jacobi_svd_tall(T, m, n, U, S, Vt, Vw);
  } else {
    // M^T = U' S V'^T  =>  M = V' S U'^T
    View T = w.take((long long)m * n);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j) T[(long long)j * m + i] = M[(long long)i * n + j];
    View Uu = w.take((long long)n * m);
    View Vv = w.take((long long)m * m);
    View Vw = w.take((long long)m * m);
    jacobi_svd_tall(T, n, m, Uu, S, Vv, Vw);
    for (int i = 0; i < m; ++i)
      for (int k = 0; k < m; ++k) U[(long long)i * m + k] = Vv[(long long)k * m + i];
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < n; ++j) Vt[(long long)k * n + j] = Uu[(long long)j * m + k];
  }
  w.top = mark;
}

// ---------------------------------------------------------------------------
// parity-block factorizations of a matrix whose rows/columns carry parities
// given by rpar(i) / cpar(j); the new bond (k) is parity-sorted with
// capacities kcap[0], kcap[1] (slots [0, kcap0) even, [kcap0, kcap0+kcap1) odd);
// a block producing fewer columns than its capacity leaves zeros.
// ---------------------------------------------------------------------------

template <typename RP, typename CP>
FU_HD inline void block_qr(View M, int m, int n, RP rpar, CP cpar, const int kcap[2], View Q, View R, Work& w) {
  const int K = kcap[0] + kcap[1];
  for (long long i = 0; i < (long long)m * K; ++i) Q[i] = 0.0;
  for (long long i = 0; i < (long long)K * n; ++i) R[i] = 0.0;
  for (int q = 0; q < 2; ++q) {
    const long long mark = w.top;
    int mq = 0, nq = 0;
    for (int i = 0; i < m; ++i) mq += rpar(i) == q;
    for (int j = 0; j < n; ++j) nq += cpar(j) == q;
    if (mq == 0 || nq == 0) continue;
    View B = w.take((long long)mq * nq);
    {
      int bi = 0;
      for (int i = 0; i < m; ++i) {
        if (rpar(i) != q) continue;
        int bj = 0;
        for (int j = 0; j < n; ++j) {
          if (cpar(j) != q) continue;
          B[(long long)bi * nq + bj] = M[(long long)i * n + j];
          ++bj;
        }
        ++bi;
      }
    }
    const int kq = mq < nq ? mq : nq;
    View Qb = w.take((long long)mq * kq), Rb = w.take((long long)kq * nq);
    View vs = w.take((long long)mq * kq), tau = w.take(kq);
    householder_qr(B, mq, nq, Qb, Rb, vs, tau);
    const int off = q == 0 ? 0 : kcap[0];
    const int kput = kq < kcap[q] ? kq : kcap[q];
    int bi = 0;
    for (int i = 0; i < m; ++i) {
      if (rpar(i) != q) continue;
      for (int c = 0; c < kput; ++c) Q[(long long)i * K + off + c] = Qb[(long long)bi * kq + c];
      ++bi;
    }
    for (int c = 0; c < kput; ++c) {
      int bj = 0;
      for (int j = 0; j < n; ++j) {
        if (cpar(j) != q) continue;
        R[(long long)(off + c) * n + j] = Rb[(long long)c * nq + bj];
        ++bj;
      }
    }
    w.top = mark;
  }
}

// Inverse of a block-diagonal (by parity) upper-triangular n x n matrix with
// parity-sorted indices (even count e). Returns false on a zero pivot.
FU_HD inline bool triangular_inverse(View R, int n, int e, View Ri) {
  for (long long i = 0; i < (long long)n * n; ++i) Ri[i] = 0.0;
  for (int q = 0; q < 2; ++q) {
    const int lo = q == 0 ? 0 : e, hi = q == 0 ? e : n;
    for (int c = hi - 1; c >= lo; --c) {
      // solve R x = e_c for the column c of Ri, rows lo..c
      for (int i = c; i >= lo; --i) {
        double s = i == c ? 1.0 : 0.0;
        for (int k = i + 1; k <= c; ++k) s -= R[(long long)i * n + k] * Ri[(long long)k * n + c];
        const double piv = R[(long long)i * n + i];
        if (piv == 0.0 || !finite(piv)) return false;
        Ri[(long long)i * n + c] = s / piv;
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// pair tensors and metrics. N4[rb][sb][rk][sk] (na, nb, na, nb);
// theta[r][p][q][s] (na, d, d, nb); cores A[r][p][k] (na, d, D), B[k][q][s].
// ---------------------------------------------------------------------------

// out[r,p,q,s] = sum_{t,u} N[r,s,t,u] theta[t,p,q,u]
FU_HD inline void apply_metric(View N, View th, View out, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d;
  for (int r = 0; r < na; ++r)
    for (int s = 0; s < nb; ++s)
      for (int p = 0; p < d; ++p)
        for (int q = 0; q < d; ++q) {
          double acc = 0.0;
          for (int t = 0; t < na; ++t)
            for (int u = 0; u < nb; ++u)
              acc += N[(((long long)r * nb + s) * na + t) * nb + u] * th[(((long long)t * d + p) * d + q) * nb + u];
          out[(((long long)r * d + p) * d + q) * nb + s] = acc;
        }
}

FU_HD inline void pair_tensor(View A, View B, View out, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  for (int r = 0; r < na; ++r)
    for (int p = 0; p < d; ++p)
      for (int q = 0; q < d; ++q)
        for (int s = 0; s < nb; ++s) {
          double acc = 0.0;
          for (int k = 0; k < D; ++k) acc += A[((long long)r * d + p) * D + k] * B[((long long)k * d + q) * nb + s];
          out[(((long long)r * d + p) * d + q) * nb + s] = acc;
        }
}

FU_HD inline double dot(View a, View b, long long n) {
  double s = 0.0;
  for (long long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// res . N res with res = pair(A, B) - target
FU_HD inline double residual_cost(View N, View A, View B, View target, const Dims& g, Work& w) {
  const long long mark = w.top;
  const long long n = (long long)g.na * g.d * g.d * g.nb;
  View res = w.take(n), nr = w.take(n);
  pair_tensor(A, B, res, g);
  for (long long i = 0; i < n; ++i) res[i] -= target[i];
  apply_metric(N, res, nr, g);
  const double v = dot(res, nr, n);
  w.top = mark;
  return v;
}

// A'[i,p,k] = sum_x Li[i][x] A[x,p,k];  B'[k,q,i] = sum_y Ri[i][y] B[k,q,y]
FU_HD inline void original_cores(View A, View B, View Li, View Ri, View oA, View oB, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  for (int i = 0; i < na; ++i)
    for (int p = 0; p < d; ++p)
      for (int k = 0; k < D; ++k) {
        double acc = 0.0;
        for (int x = 0; x < na; ++x) acc += Li[(long long)i * na + x] * A[((long long)x * d + p) * D + k];
        oA[((long long)i * d + p) * D + k] = acc;
      }
  for (int k = 0; k < D; ++k)
    for (int q = 0; q < d; ++q)
      for (int i = 0; i < nb; ++i) {
        double acc = 0.0;
        for (int y = 0; y < nb; ++y) acc += Ri[(long long)i * nb + y] * B[((long long)k * d + q) * nb + y];
        oB[((long long)k * d + q) * nb + i] = acc;
      }
}

// Jacobi-preconditioned CG with a diagonal shift (_solve_core, full_update.jl:371-400)
// on the n x n matrix M for nc right-hand-side columns (x, rhs: n x nc).
FU_HD inline double solve_core(View M, int n, int nc, View rhs, View x, const Opts& o, Work& w) {
  const long long mark = w.top;
  const long long nv = (long long)n * nc;
  View pre = w.take(n), r = w.take(nv), z = w.take(nv), p = w.take(nv), Ap = w.take(nv);
  double scale = 0.0;
  for (int i = 0; i < n; ++i) {
    const double dg = M[(long long)i * n + i] > 0.0 ? M[(long long)i * n + i] : 0.0;
    pre[i] = dg;
    if (dg > scale) scale = dg;
  }
  const double shift = o.diagonal_shift * (scale > kTiny ? scale : kTiny);
  double floor = shift;
  if (kEps * scale > floor) floor = kEps * scale;
  if (kTiny > floor) floor = kTiny;
  for (int i = 0; i < n; ++i) pre[i] = (pre[i] + shift) > floor ? (pre[i] + shift) : floor;
  auto action = [&](View v, View out) {
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < nc; ++c) {
        double acc = 0.0;
        for (int k = 0; k < n; ++k) acc += M[(long long)i * n + k] * v[(long long)k * nc + c];
        out[(long long)i * nc + c] = acc + shift * v[(long long)i * nc + c];
      }
  };
  action(x, Ap);
  double rhsn = 0.0;
  for (long long i = 0; i < nv; ++i) {
    r[i] = rhs[i] - Ap[i];
    rhsn += rhs[i] * rhs[i];
  }
  rhsn = sqrt(rhsn);
  if (rhsn < kTiny) rhsn = kTiny;
  double residual = sqrt(dot(r, r, nv)) / rhsn;
  if (residual <= o.linear_tolerance) {
    w.top = mark;
    return residual;
  }
  for (long long i = 0; i < nv; ++i) {
    z[i] = r[i] / pre[i / nc];
    p[i] = z[i];
  }
  double rz = dot(r, z, nv);
  for (int it = 0; it < o.linear_max_iterations; ++it) {
    action(p, Ap);
    const double curv = dot(p, Ap, nv);
    const double pn = sqrt(dot(p, p, nv)), an = sqrt(dot(Ap, Ap, nv));
    const double lim = pn * an > kTiny ? pn * an : kTiny;
    if (!(curv > kEps * lim)) break;
    const double alpha = rz / curv;
    for (long long i = 0; i < nv; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
    }
    residual = sqrt(dot(r, r, nv)) / rhsn;
    if (residual <= o.linear_tolerance) break;
    for (long long i = 0; i < nv; ++i) z[i] = r[i] / pre[i / nc];
    const double nxt = dot(r, z, nv);
    if (!(nxt >= 0)) break;
    const double beta = nxt / (rz > kTiny ? rz : kTiny);
    for (long long i = 0; i < nv; ++i) p[i] = z[i] + beta * p[i];
    rz = nxt;
  }
  w.top = mark;
  return residual;
}

// Doubles of workspace fu_bond needs for these dimensions.
FU_HD inline long long workspace_doubles(const Dims& g) {
  const long long n = (long long)g.na * g.nb;
  const long long t = (long long)g.na * g.d * g.d * g.nb;
  const long long ca = (long long)g.na * g.d * g.D, cb = (long long)g.D * g.d * g.nb;
  const long long nl = (long long)g.na * g.D, nr = (long long)g.D * g.nb;
  long long big = 9 * n * n + 4 * n + 4 * (long long)g.na * g.na + 4 * (long long)g.nb * g.nb;
  big += 8 * t + 16 * (ca > cb ? ca : cb) + 4 * (long long)g.D * g.D;
  big += 2 * (nl > nr ? nl : nr) * (nl > nr ? nl : nr) + 12 * (nl > nr ? nl : nr) * g.d;
  big += 4 * (n * (g.na > g.nb ? g.na : g.nb)) + 4 * t;
  return big + 1024;
}

// ---------------------------------------------------------------------------
// the bond solve
// ---------------------------------------------------------------------------

// N4: the bond metric N[rb,sb,rk,sk]; target[r,p,q,s]; outputs the new cores
// in original coordinates, bond zero-padded to D. Returns the record.
FU_HD inline Result fu_bond(const Dims& g, const Opts& o, View N4, View target, View Aout, View Bout, Work& w) {
  Result res{};
  res.stop = o.max_iterations == 0 ? kDisabled : kIterationLimit;
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  const long long n = (long long)na * nb;
  const long long tn = (long long)na * d * d * nb;
  const long long ca = (long long)na * d * D, cb = (long long)D * d * nb;
  if (workspace_doubles(g) > w.cap) {
    res.status = kWorkspace;
    return res;
  }
  auto pa = [&](int i) { return par(i, g.ea); };
  auto pb = [&](int i) { return par(i, g.eb); };
  auto pd = [&](int i) { return par(i, g.ed); };
  // ---- environment gauge (full_update.jl:64-104) ----
  double scale = 0.0;
  for (long long i = 0; i < n * n; ++i) {
    const double v = N4[i];
    if (!finite(v)) {
      res.status = kNonfiniteMetric;
      return res;
    }
    scale += v * v;
  }
  scale = sqrt(scale);
  if (!(scale > 0.0)) {
    res.status = kZeroMetric;
    return res;
  }
  View Z = w.take(n * n);       // rows: eigenvectors (k), columns (rk, sk)
  View zpar = w.take(n);        // parity of each row of Z
  double spectral = 0.0;
  {
    const long long mark = w.top;
    View wv = w.take(n);
    View blockpar = w.take(n);
    int cnt[2] = {0, 0};
    for (long long I = 0; I < n; ++I) {
      const int q = pa((int)(I / nb)) ^ pb((int)(I % nb));
      blockpar[I] = q;
      ++cnt[q];
    }
    // eigenvalues of both blocks first (for the spectral scale)
    long long row = 0;
    View Hq = w.take(n * n), Vq = w.take(n * n);
    View evals = w.take(n), evecs = w.take(n * n);  // block-major storage
    long long eoff = 0, voff = 0;
    for (int q = 0; q < 2; ++q) {
      const int m = cnt[q];
      if (m == 0) continue;
      int bi = 0;
      for (long long I = 0; I < n; ++I) {
        if ((int)blockpar[I] != q) continue;
        int bj = 0;
        for (long long K = 0; K < n; ++K) {
          if ((int)blockpar[K] != q) continue;
          Hq[(long long)bi * m + bj] = 0.5 * (N4[I * n + K] + N4[K * n + I]) / scale;
          ++bj;
        }
        ++bi;
      }
      View wq{&evals[eoff], evals.s};
      View vq{&evecs[voff], evecs.s};
      jacobi_eigh(Hq, m, wq, vq);
      for (int k = 0; k < m; ++k)
        if (fabs(wq[k]) > spectral) spectral = fabs(wq[k]);
      eoff += m;
      voff += (long long)m * m;
    }
    if (spectral == 0.0) spectral = 1.0;
    double rel = sqrt(kEps);
    if ((double)n * kEps > rel) rel = (double)n * kEps;
    const double floorv = rel * spectral;
    eoff = 0;
    voff = 0;
    for (int q = 0; q < 2; ++q) {
      const int m = cnt[q];
      if (m == 0) continue;
      for (int k = 0; k < m; ++k) {
        double lam = evals[eoff + k];
        if (lam < floorv) ++res.floored;
        const double root = sqrt(lam > floorv ? lam : floorv);
        for (long long K = 0; K < n; ++K) Z[row * n + K] = 0.0;
        int bi = 0;
        for (long long K = 0; K < n; ++K) {
          if ((int)blockpar[K] != q) continue;
          Z[row * n + K] = root * evecs[voff + (long long)bi * m + k];
          ++bi;
        }
        zpar[row] = q;
        ++row;
      }
      eoff += m;
      voff += (long long)m * m;
    }
    (void)wv;
    (void)Vq;
    w.top = mark;
  }
  // left = R of ((k, sk) | rk), right = R of ((k, rk) | sk)
  View left = w.take((long long)na * na), right = w.take((long long)nb * nb);
  View Li = w.take((long long)na * na), Ri = w.take((long long)nb * nb);
  {
    const long long mark = w.top;
    View Ml = w.take(n * nb * na);
    for (long long k = 0; k < n; ++k)
      for (int l = 0; l < nb; ++l)
        for (int a = 0; a < na; ++a) Ml[(k * nb + l) * na + a] = Z[k * n + (long long)a * nb + l];
    View Qd = w.take(n * nb * na);
    int kc[2] = {g.ea, na - g.ea};
    block_qr(Ml, (int)(n * nb), na, [&](int i) { return ((int)zpar[i / nb]) ^ pb(i % nb); }, pa, kc, Qd, left, w);
    w.top = mark;
    View Mr = w.take(n * na * nb);
    for (long long k = 0; k < n; ++k)
      for (int a = 0; a < na; ++a)
        for (int l = 0; l < nb; ++l) Mr[(k * na + a) * nb + l] = Z[k * n + (long long)a * nb + l];
    View Qe = w.take(n * na * nb);
    int kd[2] = {g.eb, nb - g.eb};
    block_qr(Mr, (int)(n * na), nb, [&](int i) { return ((int)zpar[i / na]) ^ pa(i % na); }, pb, kd, Qe, right, w);
    w.top = mark;
  }
  if (!triangular_inverse(left, na, g.ea, Li) || !triangular_inverse(right, nb, g.eb, Ri)) {
    res.status = kSingularGauge;
    return res;
  }
  // metric[x,y,z,w] = sum N[i,j,k,l] Li[i,x] Ri[j,y] Li[k,z] Ri[l,w]
  View metric = w.take(n * n);
  {
    const long long mark = w.top;
    View T1 = w.take(n * n);
    for (int x = 0; x < na; ++x)
      for (long long rest = 0; rest < (long long)nb * n; ++rest) {
        double acc = 0.0;
        for (int i = 0; i < na; ++i) acc += Li[(long long)i * na + x] * N4[(long long)i * nb * n + rest];
        T1[(long long)x * nb * n + rest] = acc;
      }
    for (int x = 0; x < na; ++x)
      for (int y = 0; y < nb; ++y)
        for (long long rest = 0; rest < n; ++rest) {
          double acc = 0.0;
          for (int j = 0; j < nb; ++j) acc += Ri[(long long)j * nb + y] * T1[((long long)x * nb + j) * n + rest];
          metric[((long long)x * nb + y) * n + rest] = acc;
        }
    for (long long xy = 0; xy < n; ++xy)
      for (int z = 0; z < na; ++z)
        for (int l = 0; l < nb; ++l) {
          double acc = 0.0;
          for (int k = 0; k < na; ++k) acc += Li[(long long)k * na + z] * metric[xy * n + (long long)k * nb + l];
          T1[xy * n + (long long)z * nb + l] = acc;
        }
    for (long long xyz = 0; xyz < n * na; ++xyz)
      for (int ww = 0; ww < nb; ++ww) {
        double acc = 0.0;
        for (int l = 0; l < nb; ++l) acc += Ri[(long long)l * nb + ww] * T1[xyz * nb + l];
        metric[xyz * nb + ww] = acc;
      }
    w.top = mark;
  }
  // working target wt[x,p,q,y] = sum target[r,p,q,s] left[x,r] right[y,s]
  View wt = w.take(tn);
  for (int x = 0; x < na; ++x)
    for (int p = 0; p < d; ++p)
      for (int q = 0; q < d; ++q)
        for (int y = 0; y < nb; ++y) {
          double acc = 0.0;
          for (int r = 0; r < na; ++r) {
            const double lx = left[(long long)x * na + r];
            if (lx == 0.0) continue;
            for (int s = 0; s < nb; ++s)
              acc += lx * right[(long long)y * nb + s] * target[(((long long)r * d + p) * d + q) * nb + s];
          }
          wt[(((long long)x * d + p) * d + q) * nb + y] = acc;
        }
  // ---- gauged truncated SVD start (full_update.jl:31-42) ----
  // The cores live in the kept bond (ke even + ko odd, ke + ko <= D) from here
  // on, as the reference's bond space does; the output is expanded to the
  // device's padded capacity (even slots [0, eD), odd slots [eD, D)) at the end.
  View A = w.take(ca), B = w.take(cb);
  int ke = 0, ko = 0;
  {
    const long long mark = w.top;
    for (long long i = 0; i < ca; ++i) A[i] = 0.0;
    for (long long i = 0; i < cb; ++i) B[i] = 0.0;
    const int mr = na * d, mc = d * nb;
    double wn = 0.0;
    for (long long i = 0; i < tn; ++i) wn += wt[i] * wt[i];
    wn = sqrt(wn);
    int mxd = na > nb ? na : nb;
    if (d > mxd) mxd = d;
    const double atol = kEps * mxd * wn;
    int mq[2] = {0, 0}, nq[2] = {0, 0};
    for (int i = 0; i < mr; ++i) ++mq[pa(i / d) ^ pd(i % d)];
    for (int j = 0; j < mc; ++j) ++nq[pd(j / nb) ^ pb(j % nb)];
    View Sv[2], Uv[2], Vv[2];
    int kq[2] = {0, 0};
    for (int q = 0; q < 2; ++q) {
      kq[q] = mq[q] < nq[q] ? mq[q] : nq[q];
      Sv[q] = w.take(kq[q] > 0 ? kq[q] : 1);
      Uv[q] = w.take((long long)mq[q] * (kq[q] > 0 ? kq[q] : 1));
      Vv[q] = w.take((long long)(kq[q] > 0 ? kq[q] : 1) * nq[q]);
      if (kq[q] == 0) continue;
      View Bm = w.take((long long)mq[q] * nq[q]);
      int bi = 0;
      for (int i = 0; i < mr; ++i) {
        if ((pa(i / d) ^ pd(i % d)) != q) continue;
        int bj = 0;
        for (int j = 0; j < mc; ++j) {
          if ((pd(j / nb) ^ pb(j % nb)) != q) continue;
          Bm[(long long)bi * nq[q] + bj] = wt[(long long)i * mc + j];
          ++bj;
        }
        ++bi;
      }
      svd_compact(Bm, mq[q], nq[q], Uv[q], Sv[q], Vv[q], w);
    }
    // truncrank(D) over both sectors (ties: sector 0 first, earlier first),
    // truncspace caps (the bond capacity), trunctol(atol)
    int take_q[2] = {0, 0};
    for (int t = 0; t < D; ++t) {
      int pick = -1;
      double bestv = -1.0;
      for (int q = 0; q < 2; ++q)
        if (take_q[q] < kq[q] && Sv[q][take_q[q]] > bestv) {
          bestv = Sv[q][take_q[q]];
          pick = q;
        }
      if (pick < 0) break;
      ++take_q[pick];
    }
    const int capq[2] = {g.eD, D - g.eD};
    for (int q = 0; q < 2; ++q) {
      if (take_q[q] > capq[q]) take_q[q] = capq[q];
      int above = 0;
      for (int k = 0; k < kq[q]; ++k) above += Sv[q][k] > atol;
      if (take_q[q] > above) take_q[q] = above;
    }
    res.kept_even = take_q[0];
    res.kept_odd = take_q[1];
    if (take_q[0] + take_q[1] == 0) {
      res.status = kZeroTarget;
      w.top = mark;
      return res;
    }
    ke = take_q[0];
    ko = take_q[1];
    const int Dk = ke + ko;
    for (int q = 0; q < 2; ++q) {
      const int off = q == 0 ? 0 : ke;
      int bi = 0;
      for (int i = 0; i < mr; ++i) {
        if ((pa(i / d) ^ pd(i % d)) != q) continue;
        for (int k = 0; k < take_q[q]; ++k)
          A[(long long)i * Dk + off + k] = Uv[q][(long long)bi * kq[q] + k] * sqrt(Sv[q][k]);
        ++bi;
      }
      for (int k = 0; k < take_q[q]; ++k) {
        int bj = 0;
        for (int j = 0; j < mc; ++j) {
          if ((pd(j / nb) ^ pb(j % nb)) != q) continue;
          B[(long long)(off + k) * mc + j] = sqrt(Sv[q][k]) * Vv[q][(long long)k * nq[q] + bj];
          ++bj;
        }
      }
    }
    w.top = mark;
  }
  // ALS dimensions: the kept bond
  Dims gk = g;
  gk.D = ke + ko;
  gk.eD = ke;
  // ---- ALS (full_update.jl:418-482) ----
  {
  const int D = gk.D;
  const long long ca = (long long)na * d * D, cb = (long long)D * d * nb;
  auto pD = [&](int i) { return par(i, gk.eD); };
  View Nt = w.take(tn);
  apply_metric(N4, target, Nt, g);
  const double target_norm = dot(target, Nt, tn);
  res.target_norm = target_norm;
  View oA = w.take(ca), oB = w.take(cb), bestA = w.take(ca), bestB = w.take(cb);
  original_cores(A, B, Li, Ri, oA, oB, gk);
  auto evaluate = [&](double raw, double& cost, bool& fallback) -> bool {
    const double sc = fabs(target_norm) > kTiny ? fabs(target_norm) : kTiny;
    if (!finite(raw) || !finite(target_norm)) {
      res.status = kNonfiniteCost;
      return false;
    }
    if (raw / sc < -kRoundoff || target_norm < 0.0) {
      res.status = kRepairNeeded;
      return false;
    }
    const double relv = raw / sc;
    cost = fabs(relv);
    fallback = relv < 0.0;
    return true;
  };
  double best = 0.0;
  bool fb = false;
  if (!evaluate(residual_cost(N4, oA, oB, target, gk, w), best, fb)) return res;
  res.initial_cost = best;
  for (long long i = 0; i < ca; ++i) bestA[i] = oA[i];
  for (long long i = 0; i < cb; ++i) bestB[i] = oB[i];
  View fQA = w.take(ca), fC = w.take((long long)D * D), fQB = w.take(cb);
  bool have = false;
  if (fb) {
    res.stop = kResidualFallback;
  } else if (o.max_iterations > 0 && best <= o.error_tolerance) {
    res.stop = kErrorTolerance;
  } else if (o.max_iterations > 0) {
    View wtw = w.take(tn);
    apply_metric(metric, wt, wtw, g);
    const int nl = na * D, nr = D * nb;
    View NL = w.take((long long)nl * nl), yl = w.take((long long)nl * d), xl = w.take((long long)nl * d);
    View NR = w.take((long long)nr * nr), yr = w.take((long long)nr * d), xr = w.take((long long)nr * d);
    View QA = w.take(ca), Rm = w.take((long long)D * D), rinit = w.take(cb), C = w.take((long long)D * D);
    View QB = w.take(cb), solm = w.take(ca), solt = w.take(cb), Qt = w.take(cb);
    const int kcD[2] = {gk.eD, D - gk.eD};
    for (int it = 1; it <= o.max_iterations; ++it) {
      // left normal equation (full_update.jl:402-408)
      for (int f = 0; f < na; ++f)
        for (int b = 0; b < D; ++b)
          for (int gg = 0; gg < na; ++gg)
            for (int oo = 0; oo < D; ++oo) {
              double acc = 0.0;
              for (int s = 0; s < nb; ++s)
                for (int nn = 0; nn < nb; ++nn) {
                  const double mv = metric[(((long long)f * nb + s) * na + gg) * nb + nn];
                  if (mv == 0.0) continue;
                  double bb = 0.0;
                  for (int p = 0; p < d; ++p)
                    bb += B[((long long)b * d + p) * nb + s] * B[((long long)oo * d + p) * nb + nn];
                  acc += mv * bb;
                }
              NL[((long long)f * D + b) * nl + (long long)gg * D + oo] = acc;
            }
      for (int f = 0; f < na; ++f)
        for (int b = 0; b < D; ++b)
          for (int p = 0; p < d; ++p) {
            double acc = 0.0;
            for (int q = 0; q < d; ++q)
              for (int s = 0; s < nb; ++s)
                acc += wtw[(((long long)f * d + p) * d + q) * nb + s] * B[((long long)b * d + q) * nb + s];
            yl[((long long)f * D + b) * d + p] = acc;
            xl[((long long)f * D + b) * d + p] = A[((long long)f * d + p) * D + b];
          }
      const double ra = solve_core(NL, nl, d, yl, xl, o, w);
      // QA, R = left_orth(sol as ((r, p) | k))
      for (int r = 0; r < na; ++r)
        for (int p = 0; p < d; ++p)
          for (int k = 0; k < D; ++k) solm[((long long)r * d + p) * D + k] = xl[((long long)r * D + k) * d + p];
      block_qr(solm, na * d, D, [&](int i) { return pa(i / d) ^ pd(i % d); }, pD, kcD, QA, Rm, w);
      // right_initial = R B
      for (int k = 0; k < D; ++k)
        for (long long j = 0; j < (long long)d * nb; ++j) {
          double acc = 0.0;
          for (int m = 0; m < D; ++m) acc += Rm[(long long)k * D + m] * B[(long long)m * d * nb + j];
          rinit[(long long)k * d * nb + j] = acc;
        }
      // right normal equation (full_update.jl:410-416) with A = QA
      for (int b = 0; b < D; ++b)
        for (int s = 0; s < nb; ++s)
          for (int oo = 0; oo < D; ++oo)
            for (int nn = 0; nn < nb; ++nn) {
              double acc = 0.0;
              for (int ww = 0; ww < na; ++ww)
                for (int e = 0; e < na; ++e) {
                  const double mv = metric[(((long long)ww * nb + s) * na + e) * nb + nn];
                  if (mv == 0.0) continue;
                  double aa = 0.0;
                  for (int p = 0; p < d; ++p)
                    aa += QA[((long long)ww * d + p) * D + b] * QA[((long long)e * d + p) * D + oo];
                  acc += mv * aa;
                }
              NR[((long long)b * nb + s) * nr + (long long)oo * nb + nn] = acc;
            }
      for (int b = 0; b < D; ++b)
        for (int s = 0; s < nb; ++s)
          for (int p = 0; p < d; ++p) {
            double acc = 0.0;
            for (int ww = 0; ww < na; ++ww)
              for (int q = 0; q < d; ++q)
                acc += wtw[(((long long)ww * d + q) * d + p) * nb + s] * QA[((long long)ww * d + q) * D + b];
            yr[((long long)b * nb + s) * d + p] = acc;
            xr[((long long)b * nb + s) * d + p] = rinit[((long long)b * d + p) * nb + s];
          }
      const double rb = solve_core(NR, nr, d, yr, xr, o, w);
      // C, QB = right_orth(sol as (k | (q, s))) via QR of its transpose
      for (int k = 0; k < D; ++k)
        for (int q = 0; q < d; ++q)
          for (int s = 0; s < nb; ++s) solt[((long long)q * nb + s) * D + k] = xr[((long long)k * nb + s) * d + q];
      block_qr(solt, d * nb, D, [&](int i) { return pd(i / nb) ^ pb(i % nb); }, pD, kcD, Qt, Rm, w);
      for (int k = 0; k < D; ++k)
        for (int m = 0; m < D; ++m) C[(long long)k * D + m] = Rm[(long long)m * D + k];
      for (int m = 0; m < D; ++m)
        for (long long j = 0; j < (long long)d * nb; ++j) QB[(long long)m * d * nb + j] = Qt[j * D + m];
      // A = QA C, B = QB
      for (long long rp = 0; rp < (long long)na * d; ++rp)
        for (int m = 0; m < D; ++m) {
          double acc = 0.0;
          for (int k = 0; k < D; ++k) acc += QA[rp * D + k] * C[(long long)k * D + m];
          A[rp * D + m] = acc;
        }
      for (long long i = 0; i < cb; ++i) B[i] = QB[i];
      original_cores(A, B, Li, Ri, oA, oB, gk);
      double cost = 0.0;
      bool fb2 = false;
      if (!evaluate(residual_cost(N4, oA, oB, target, gk, w), cost, fb2)) return res;
      res.residual = ra > rb ? ra : rb;
      res.iterations = it;
      if (fb2) {
        res.stop = kResidualFallback;
        break;
      }
      if (cost <= best) {
        best = cost;
        for (long long i = 0; i < ca; ++i) bestA[i] = oA[i];
        for (long long i = 0; i < cb; ++i) bestB[i] = oB[i];
        for (long long i = 0; i < ca; ++i) fQA[i] = QA[i];
        for (long long i = 0; i < (long long)D * D; ++i) fC[i] = C[i];
        for (long long i = 0; i < cb; ++i) fQB[i] = QB[i];
        have = true;
      }
      if (best <= o.error_tolerance) {
        res.stop = kErrorTolerance;
        break;
      }
      res.limit_reached = it == o.max_iterations ? 1 : 0;
    }
  }
  res.final_cost = best;
  if (have) {
    // _fu_balance_bond (full_update.jl:44-52): compact SVD of C per parity block
    const long long mark = w.top;
    View Us = w.take((long long)D * D), Vs = w.take((long long)D * D);
    for (long long i = 0; i < (long long)D * D; ++i) Us[i] = Vs[i] = 0.0;
    for (int q = 0; q < 2; ++q) {
      const int lo = q == 0 ? 0 : gk.eD, m = q == 0 ? gk.eD : D - gk.eD;
      if (m == 0) continue;
      View Cm = w.take((long long)m * m), U = w.take((long long)m * m), S = w.take(m), Vt = w.take((long long)m * m);
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) Cm[(long long)i * m + j] = fC[(long long)(lo + i) * D + lo + j];
      svd_compact(Cm, m, m, U, S, Vt, w);
      for (int i = 0; i < m; ++i)
        for (int k = 0; k < m; ++k) {
          Us[(long long)(lo + i) * D + lo + k] = U[(long long)i * m + k] * sqrt(S[k]);
          Vs[(long long)(lo + k) * D + lo + i] = sqrt(S[k]) * Vt[(long long)k * m + i];
        }
    }
    for (long long rp = 0; rp < (long long)na * d; ++rp)
      for (int m = 0; m < D; ++m) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc += fQA[rp * D + k] * Us[(long long)k * D + m];
        A[rp * D + m] = acc;
      }
    for (int m = 0; m < D; ++m)
      for (long long j = 0; j < (long long)d * nb; ++j) {
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc += Vs[(long long)m * D + k] * fQB[(long long)k * d * nb + j];
        B[(long long)m * d * nb + j] = acc;
      }
    original_cores(A, B, Li, Ri, bestA, bestB, gk);
    w.top = mark;
  }
  // expand the kept bond into the padded capacity
  for (long long i = 0; i < (long long)na * d * g.D; ++i) Aout[i] = 0.0;
  for (long long i = 0; i < (long long)g.D * d * nb; ++i) Bout[i] = 0.0;
  for (long long rp = 0; rp < (long long)na * d; ++rp)
    for (int k = 0; k < ke + ko; ++k) Aout[rp * g.D + (k < ke ? k : g.eD + k - ke)] = bestA[rp * (ke + ko) + k];
  for (int k = 0; k < ke + ko; ++k)
    for (long long j = 0; j < (long long)d * nb; ++j)
      Bout[(long long)(k < ke ? k : g.eD + k - ke) * d * nb + j] = bestB[(long long)k * d * nb + j];
  }
  return res;
}

}  // namespace fu
}  // namespace spp
