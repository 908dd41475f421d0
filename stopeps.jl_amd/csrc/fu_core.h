// Per-bond solve of the full update (src/evolution/full_update.jl:31-117,
// 342-482): environment gauge, gauged truncated SVD start, ALS with
// Jacobi-preconditioned CG, bond-centre balance. FP64, one sample per call,
// solved by a Team of lanes: on the device one workgroup per sample (loops
// split over the lanes, reductions through LDS), on the host a single lane
// (the selftest against the oracle runs the same source).
// All indices are parity-sorted (even entries first on every leg).
#pragma once
#include <cmath>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
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

// The lanes that solve one sample together: a workgroup on the device (lane =
// thread, size = a multiple of the 64-lane wavefront; red: LDS scratch of
// 3 * size / 64 doubles), one lane on the host. All lanes run the same control
// flow and take the same workspace; a matrix row i is "owned" by lane
// i % size wherever a routine keeps rows private to a lane between its
// reductions. Hot working sets (the gauge eigenproblem, the normal-matrix
// operands) are staged in the LDS pool when they fit, in the workspace
// otherwise. Reductions are xor butterflies inside a wavefront (every lane
// ends with the bitwise same value, since each step adds the same two
// operands on both partners), then a fixed-order sum over wavefronts.
struct Team {
  int lane, size;
  double* red;
  double* lds;        // the workgroup's LDS pool (device), null on the host
  long long lds_cap;  // its size in doubles (0: stage nothing)
  FU_HD void sync() const {
#ifdef __HIP_DEVICE_COMPILE__
    if (size > 1) __syncthreads();
#endif
  }
#ifdef __HIP_DEVICE_COMPILE__
  __device__ void waves(double* v, int nv, bool is_max) const {
    if (size == 1) return;
    for (int m = 32; m >= 1; m >>= 1)
      for (int k = 0; k < nv; ++k) {
        const double o = __shfl_xor(v[k], m, 64);
        v[k] = is_max ? (o > v[k] ? o : v[k]) : v[k] + o;
      }
    if (size > 64) {
      const int nw = size >> 6;
      __syncthreads();  // earlier readers of red are done
      if ((lane & 63) == 0)
        for (int k = 0; k < nv; ++k) red[k * nw + (lane >> 6)] = v[k];
      __syncthreads();
      for (int k = 0; k < nv; ++k) {
        double r = red[k * nw];
        for (int w = 1; w < nw; ++w) r = is_max ? (red[k * nw + w] > r ? red[k * nw + w] : r) : r + red[k * nw + w];
        v[k] = r;
      }
    }
  }
#endif
  // exclusive prefix sum of v over lanes (lane order); total: the full sum
  FU_HD int scan(int v, int& total) const {
#ifdef __HIP_DEVICE_COMPILE__
    if (size > 1) {
      const int wl = lane & 63;
      int x = v;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (wl >= o) x += y;
      }
      if (size > 64) {
        const int nw = size >> 6, me = lane >> 6;
        __syncthreads();  // earlier readers of red are done
        if (wl == 63) red[me] = (double)x;
        __syncthreads();
        int before = 0, tot = 0;
        for (int w = 0; w < nw; ++w) {
          const int c = (int)red[w];
          if (w < me) before += c;
          tot += c;
        }
        __syncthreads();
        total = tot;
        return before + x - v;
      }
      total = __shfl(x, 63, 64);
      return x - v;
    }
#endif
    total = v;
    return 0;
  }
  // all lanes receive the sum of v over lanes
  FU_HD double sum(double v) const {
#ifdef __HIP_DEVICE_COMPILE__
    waves(&v, 1, false);
#endif
    return v;
  }
  FU_HD double max(double v) const {
#ifdef __HIP_DEVICE_COMPILE__
    waves(&v, 1, true);
#endif
    return v;
  }
  // two or three sums in one pass
  FU_HD void sum2(double& a, double& b) const {
#ifdef __HIP_DEVICE_COMPILE__
    double v[2] = {a, b};
    waves(v, 2, false);
    a = v[0];
    b = v[1];
#endif
  }
  FU_HD void sum3(double& a, double& b, double& c) const {
#ifdef __HIP_DEVICE_COMPILE__
    double v[3] = {a, b, c};
    waves(v, 3, false);
    a = v[0];
    b = v[1];
    c = v[2];
#endif
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
  int rank = 0;  // truncrank(D) of the SVD start; 0: the bond capacity's dimension
};

enum Stop { kDisabled = 0, kIterationLimit = 1, kErrorTolerance = 2, kResidualFallback = 3 };
enum Status { kOk = 0, kNonfiniteMetric = 1, kZeroMetric = 2, kZeroTarget = 3, kRepairNeeded = 4,
              kNonfiniteCost = 5, kSingularGauge = 6, kWorkspace = 7, kNonconverged = 8 };

// Sweep cap of the Jacobi eigensolver and SVD. Both converge quadratically
// once rotations are small; a solve that reaches the cap is reported
// (Result::nonconverged, status kNonconverged), as the reference raises when
// its SVD fails to converge after the gesdd -> gesvd retry (src/states/svd.jl:1-11).
constexpr int kJacobiSweeps = 80;

struct Result {
  double initial_cost, final_cost, residual, target_norm;
  int iterations, limit_reached, stop, status;
  int kept_even, kept_odd, floored;
  int nonconverged;  // eigen/SVD solves of this bond that reached the sweep cap
};

constexpr double kEps = 2.220446049250313e-16;
constexpr double kTiny = 2.2250738585072014e-308;
constexpr double kRoundoff = 1e-8;  // RESIDUAL_ROUNDOFF_TOLERANCE, full_update.jl:18

FU_HD inline int par(int i, int e) { return i >= e ? 1 : 0; }
FU_HD inline bool finite(double x) { return fabs(x) <= 1.7976931348623157e308; }  // false for inf and NaN

// ---------------------------------------------------------------------------
// dense kernels on small row-major matrices
// ---------------------------------------------------------------------------

// Jacobi eigensolver of the symmetric n x n matrix A (destroyed) in
// round-robin (tournament) order: each round rotates n/2 disjoint index pairs
// at once, so the lanes share a round's rotations and every round costs three
// barriers. Eigenvalues in w, eigenvectors as columns of V (V[i*n+k] =
// component i of k), unordered. cs: 3 * (n / 2 + 1) scratch. The bond gauge
// uses only V max(w, floor) V^T, which does not depend on the order or on the
// basis chosen inside a degenerate eigenspace.
constexpr int kU = 4;  // tasks whose loads a lane issues together

FU_HD inline bool jacobi_eigh(const Team& t, View A, int n, View w, View V, View cs,
                               int max_sweeps = kJacobiSweeps) {
  for (int i = t.lane; i < (int)n * n; i += t.size) V[i] = (i / n) == (i % n) ? 1.0 : 0.0;
  double fl = 0.0;
  for (int i = t.lane; i < (int)n * n; i += t.size) fl += A[i] * A[i];
  const double fro = t.sum(fl);
  const int m = n + (n & 1);  // one idle slot for odd n
  const int np = m / 2;
  // pair i of round r: slots (i, m-1-i) of the sequence [0, r rotations of 1..m-1]
  auto slot = [&](int r, int j) { return j == 0 ? 0 : 1 + (j - 1 + r) % (m - 1); };
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double ol = 0.0;
    for (int i = t.lane; i < (int)n * n; i += t.size) {
      const int p = (int)(i / n), q = (int)(i % n);
      if (q > p) ol += A[i] * A[i];
    }
    const double off = t.sum(ol);
    if (off <= 1e-34 * fro || off == 0.0) {
      converged = true;
      break;
    }
    for (int r = 0; r < m - 1; ++r) {
      // rotation of each pair (c = 2 marks "no rotation")
      for (int i = t.lane; i < np; i += t.size) {
        int p = slot(r, i), q = slot(r, m - 1 - i);
        if (p > q) {
          const int x = p;
          p = q;
          q = x;
        }
        double c = 2.0, sn = 0.0;
        if (q < n) {
          const double apq = A[(int)p * n + q];
          if (!(apq == 0.0 || fabs(apq) <= 1e-300)) {
            const double app = A[(int)p * n + p], aqq = A[(int)q * n + q];
            const double theta = (aqq - app) / (2.0 * apq);
            const double tt = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(tt * tt + 1.0);
            sn = tt * c;
          }
        }
        cs[3 * i] = c;
        cs[3 * i + 1] = sn;
        cs[3 * i + 2] = p + q * 65536.0;
      }
      t.sync();
      // columns p, q of A and V: A <- A J, V <- V J. Tasks touch disjoint
      // entries; each lane loads kU tasks before it stores any, so that
      // their loads are in flight together.
      for (int t0 = t.lane; t0 < (int)np * n; t0 += (int)kU * t.size) {
        double c[kU], sn[kU], akp[kU], akq[kU], vkp[kU], vkq[kU];
        int ip[kU], iq[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int task = t0 + (int)u * t.size;
          c[u] = 2.0;
          if (task < (int)np * n) {
            const int i = (int)(task / n), k = (int)(task % n);
            c[u] = cs[3 * i];
            sn[u] = cs[3 * i + 1];
            const int pq = (int)cs[3 * i + 2];
            ip[u] = (int)k * n + (pq & 65535);
            iq[u] = (int)k * n + (pq >> 16);
            if (c[u] != 2.0) {
              akp[u] = A[ip[u]];
              akq[u] = A[iq[u]];
              vkp[u] = V[ip[u]];
              vkq[u] = V[iq[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (c[u] != 2.0) {
            A[ip[u]] = c[u] * akp[u] - sn[u] * akq[u];
            A[iq[u]] = sn[u] * akp[u] + c[u] * akq[u];
            V[ip[u]] = c[u] * vkp[u] - sn[u] * vkq[u];
            V[iq[u]] = sn[u] * vkp[u] + c[u] * vkq[u];
          }
      }
      t.sync();
      // rows p, q: A <- J^T A; the rotated pair's off-diagonal entries become 0
      for (int t0 = t.lane; t0 < (int)np * n; t0 += (int)kU * t.size) {
        double c[kU], sn[kU], apk[kU], aqk[kU];
        int ip[kU], iq[kU];
        bool zp[kU], zq[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int task = t0 + (int)u * t.size;
          c[u] = 2.0;
          if (task < (int)np * n) {
            const int i = (int)(task / n), k = (int)(task % n);
            c[u] = cs[3 * i];
            sn[u] = cs[3 * i + 1];
            const int pq = (int)cs[3 * i + 2], p = pq & 65535, q = pq >> 16;
            ip[u] = (int)p * n + k;
            iq[u] = (int)q * n + k;
            zp[u] = k == q;
            zq[u] = k == p;
            if (c[u] != 2.0) {
              apk[u] = A[ip[u]];
              aqk[u] = A[iq[u]];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (c[u] != 2.0) {
            A[ip[u]] = zp[u] ? 0.0 : c[u] * apk[u] - sn[u] * aqk[u];
            A[iq[u]] = zq[u] ? 0.0 : sn[u] * apk[u] + c[u] * aqk[u];
          }
      }
      t.sync();
    }
  }
  for (int i = t.lane; i < n; i += t.size) w[i] = A[(int)i * n + i];
  t.sync();
  return converged;
}

// Householder QR of the m x n matrix M (row-major, destroyed): Q (m x k; not
// formed when Q.p is null), R (k x n), k = min(m, n), R with a nonnegative
// diagonal. vs: m x k scratch, tau: k. Rows of M, vs and Q are lane-owned.
FU_HD inline void householder_qr(const Team& t, View M, int m, int n, View Q, View R, View vs, View tau) {
  const int k = m < n ? m : n;
  for (int j = 0; j < k; ++j) {
    double nl = 0.0;
    for (int i = t.lane; i < m; i += t.size)
      if (i >= j) nl += M[(int)i * n + j] * M[(int)i * n + j];
    const double nrm = sqrt(t.sum(nl));  // the sum's barrier also publishes M[j][j]
    if (nrm == 0.0) {
      if (t.lane == 0) tau[j] = 0.0;
      for (int i = t.lane; i < m; i += t.size) vs[(int)i * k + j] = 0.0;
      continue;
    }
    const double x0 = M[(int)j * n + j];
    const double alpha = x0 >= 0 ? -nrm : nrm;
    // v = x - alpha e_j, normalised so that v_j = 1
    const double v0 = x0 - alpha;
    double vl = 0.0;
    for (int i = t.lane; i < m; i += t.size) {
      const double v = i < j ? 0.0 : (i == j ? 1.0 : M[(int)i * n + j] / v0);
      vs[(int)i * k + j] = v;
      vl += v * v;
    }
    const double tj = 2.0 / t.sum(vl);
    if (t.lane == 0) tau[j] = tj;
    for (int c = j; c < n; ++c) {
      double dl = 0.0;
      for (int i = t.lane; i < m; i += t.size)
        if (i >= j) dl += vs[(int)i * k + j] * M[(int)i * n + c];
      const double dt = t.sum(dl) * tj;
      for (int i = t.lane; i < m; i += t.size)
        if (i >= j) M[(int)i * n + c] -= dt * vs[(int)i * k + j];
    }
  }
  t.sync();
  // the sign of each R diagonal entry, read from the final M, flips its row of
  // R and its column of Q
  for (int idx = t.lane; idx < (int)k * n; idx += t.size) {
    const int i = (int)(idx / n), c = (int)(idx % n);
    const double sg = M[(int)i * n + i] < 0.0 ? -1.0 : 1.0;
    R[idx] = c >= i ? sg * M[idx] : 0.0;
  }
  if (Q.p != nullptr) {
    // Q = H_0 ... H_{k-1} I[:, :k]
    for (int i = t.lane; i < m; i += t.size)
      for (int c = 0; c < k; ++c) Q[(int)i * k + c] = i == c ? 1.0 : 0.0;
    for (int j = k - 1; j >= 0; --j) {
      const double tj = tau[j];
      if (tj == 0.0) continue;
      for (int c = 0; c < k; ++c) {
        double dl = 0.0;
        for (int i = t.lane; i < m; i += t.size)
          if (i >= j) dl += vs[(int)i * k + j] * Q[(int)i * k + c];
        const double dt = t.sum(dl) * tj;
        for (int i = t.lane; i < m; i += t.size)
          if (i >= j) Q[(int)i * k + c] -= dt * vs[(int)i * k + j];
      }
    }
    for (int i = t.lane; i < m; i += t.size)
      for (int c = 0; c < k; ++c)
        if (M[(int)c * n + c] < 0.0) Q[(int)i * k + c] = -Q[(int)i * k + c];
  }
  t.sync();
}

// One-sided Jacobi SVD of the m x n matrix M (row-major, destroyed) with
// m >= n: U (m x n), S (n, descending), Vt (n x n). Vw: n*n scratch. Rows of M
// and Vw are lane-owned; the three column products of a pair are one reduction.
FU_HD inline bool jacobi_svd_tall(const Team& t, View M, int m, int n, View U, View S, View Vt, View Vw,
                                   int max_sweeps = kJacobiSweeps) {
  for (int i = t.lane; i < n; i += t.size)
    for (int j = 0; j < n; ++j) Vw[(int)i * n + j] = i == j ? 1.0 : 0.0;
  bool converged = false;
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double a = 0.0, b = 0.0, g = 0.0;
        for (int i = t.lane; i < m; i += t.size) {
          const double x = M[(int)i * n + p], y = M[(int)i * n + q];
          a += x * x;
          b += y * y;
          g += x * y;
        }
        t.sum3(a, b, g);
        if (g == 0.0 || fabs(g) <= 1e-15 * sqrt(a * b)) continue;
        rotated = true;
        const double zeta = (b - a) / (2.0 * g);
        const double tt = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + tt * tt), s = c * tt;
        for (int i = t.lane; i < m; i += t.size) {
          const double x = M[(int)i * n + p], y = M[(int)i * n + q];
          M[(int)i * n + p] = c * x - s * y;
          M[(int)i * n + q] = s * x + c * y;
        }
        for (int i = t.lane; i < n; i += t.size) {
          const double x = Vw[(int)i * n + p], y = Vw[(int)i * n + q];
          Vw[(int)i * n + p] = c * x - s * y;
          Vw[(int)i * n + q] = s * x + c * y;
        }
      }
    if (!rotated) {
      converged = true;
      break;
    }
  }
  // column norms, then a selection sort into descending order
  for (int j = 0; j < n; ++j) {
    double al = 0.0;
    for (int i = t.lane; i < m; i += t.size) al += M[(int)i * n + j] * M[(int)i * n + j];
    const double a = sqrt(t.sum(al));
    if (t.lane == 0) S[j] = a;
  }
  t.sync();
  for (int j = 0; j < n; ++j) {
    int best = j;
    for (int c = j + 1; c < n; ++c)
      if (S[c] > S[best]) best = c;
    t.sync();  // every lane has read S before lane 0 swaps it
    if (best != j) {
      if (t.lane == 0) {
        const double x = S[j];
        S[j] = S[best];
        S[best] = x;
      }
      for (int i = t.lane; i < m; i += t.size) {
        const double x = M[(int)i * n + j];
        M[(int)i * n + j] = M[(int)i * n + best];
        M[(int)i * n + best] = x;
      }
      for (int i = t.lane; i < n; i += t.size) {
        const double x = Vw[(int)i * n + j];
        Vw[(int)i * n + j] = Vw[(int)i * n + best];
        Vw[(int)i * n + best] = x;
      }
    }
    t.sync();
  }
  for (int j = 0; j < n; ++j) {
    const double inv = S[j] > 0.0 ? 1.0 / S[j] : 0.0;
    for (int i = t.lane; i < m; i += t.size) U[(int)i * n + j] = M[(int)i * n + j] * inv;
    for (int i = t.lane; i < n; i += t.size) Vt[(int)j * n + i] = Vw[(int)i * n + j];
  }
  t.sync();
  return converged;
}

// Compact SVD of an m x n matrix (any shape): U (m x k), S (k), Vt (k x n),
// k = min(m, n). work: m*n + k*k + max(m,n)*k doubles. Returns false when the
// Jacobi sweeps reached their cap.
FU_HD inline bool svd_compact(const Team& t, View M, int m, int n, View U, View S, View Vt, Work& w) {
  const long long mark = w.top;
  bool converged;
  if (m >= n) {
    View T = w.take((int)m * n);
    for (int i = t.lane; i < (int)m * n; i += t.size) T[i] = M[i];
    View Vw = w.take((int)n * n);
    t.sync();
    converged = jacobi_svd_tall(t, T, m, n, U, S, Vt, Vw);
  } else {
    // M^T = U' S V'^T  =>  M = V' S U'^T
    View T = w.take((int)m * n);
    for (int idx = t.lane; idx < (int)m * n; idx += t.size) {
      const int i = (int)(idx / n), j = (int)(idx % n);
      T[(int)j * m + i] = M[idx];
    }
    View Uu = w.take((int)n * m);
    View Vv = w.take((int)m * m);
    View Vw = w.take((int)m * m);
    t.sync();
    converged = jacobi_svd_tall(t, T, n, m, Uu, S, Vv, Vw);
    for (int idx = t.lane; idx < (int)m * m; idx += t.size) {
      const int i = (int)(idx / m), k = (int)(idx % m);
      U[idx] = Vv[(int)k * m + i];
    }
    for (int idx = t.lane; idx < (int)m * n; idx += t.size) {
      const int k = (int)(idx / n), j = (int)(idx % n);
      Vt[idx] = Uu[(int)j * m + k];
    }
    t.sync();
  }
  w.top = mark;
  return converged;
}

// ---------------------------------------------------------------------------
// parity-block factorizations of a matrix whose rows/columns carry parities
// given by rpar(i) / cpar(j); the new bond (k) is parity-sorted with
// capacities kcap[0], kcap[1] (slots [0, kcap0) even, [kcap0, kcap0+kcap1) odd);
// a block producing fewer columns than its capacity leaves zeros. Q is not
// formed when Q.p is null.
// ---------------------------------------------------------------------------

template <typename RP, typename CP>
FU_HD inline void block_qr(const Team& t, View M, int m, int n, RP rpar, CP cpar, const int kcap[2], View Q, View R,
                           Work& w) {
  const int K = kcap[0] + kcap[1];
  const bool want_q = Q.p != nullptr;
  if (want_q)
    for (int i = t.lane; i < m * K; i += t.size) Q[i] = 0.0;
  for (int i = t.lane; i < K * n; i += t.size) R[i] = 0.0;
  // each lane walks a contiguous range of rows; a prefix count over lanes
  // gives the block row index of its first row of parity q
  const int per = (m + t.size - 1) / t.size;
  const int r0 = t.lane * per < m ? t.lane * per : m;
  const int r1 = r0 + per < m ? r0 + per : m;
  for (int q = 0; q < 2; ++q) {
    const long long mark = w.top;
    int cnt = 0;
    for (int i = r0; i < r1; ++i) cnt += rpar(i) == q;
    int mq = 0;
    const int first = t.scan(cnt, mq);
    int nq = 0;
    for (int j = 0; j < n; ++j) nq += cpar(j) == q;
    if (mq == 0 || nq == 0) continue;
    View B = w.take(mq * nq);
    {
      int bi = first;
      for (int i = r0; i < r1; ++i) {
        if (rpar(i) != q) continue;
        int bj = 0;
        for (int j = 0; j < n; ++j) {
          if (cpar(j) != q) continue;
          B[bi * nq + bj] = M[i * n + j];
          ++bj;
        }
        ++bi;
      }
    }
    const int kq = mq < nq ? mq : nq;
    View Qb = want_q ? w.take(mq * kq) : View{nullptr, 0};
    View Rb = w.take(kq * nq);
    View vs = w.take(mq * kq), tau = w.take(kq);
    t.sync();
    householder_qr(t, B, mq, nq, Qb, Rb, vs, tau);
    const int off = q == 0 ? 0 : kcap[0];
    const int kput = kq < kcap[q] ? kq : kcap[q];
    if (want_q) {
      int bi = first;
      for (int i = r0; i < r1; ++i) {
        if (rpar(i) != q) continue;
        for (int c = 0; c < kput; ++c) Q[i * K + off + c] = Qb[bi * kq + c];
        ++bi;
      }
    }
    for (int c = t.lane; c < kput; c += t.size) {
      int bj = 0;
      for (int j = 0; j < n; ++j) {
        if (cpar(j) != q) continue;
        R[(off + c) * n + j] = Rb[c * nq + bj];
        ++bj;
      }
    }
    t.sync();
    w.top = mark;
  }
  t.sync();
}

// Inverse of a block-diagonal (by parity) upper-triangular n x n matrix with
// parity-sorted indices (even count e), one column per lane. Returns false
// (on every lane) on a zero or nonfinite pivot.
FU_HD inline bool triangular_inverse(const Team& t, View R, int n, int e, View Ri) {
  for (int i = t.lane; i < (int)n * n; i += t.size) Ri[i] = 0.0;
  t.sync();
  double bad = 0.0;
  for (int c = t.lane; c < n; c += t.size) {
    const int lo = c < e ? 0 : e;
    // solve R x = e_c for the column c of Ri, rows lo..c
    for (int i = c; i >= lo; --i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = i + 1; k <= c; ++k) s -= R[(int)i * n + k] * Ri[(int)k * n + c];
      const double piv = R[(int)i * n + i];
      if (piv == 0.0 || !finite(piv)) {
        bad = 1.0;
        break;
      }
      Ri[(int)i * n + c] = s / piv;
    }
  }
  return t.sum(bad) == 0.0;
}

// ---------------------------------------------------------------------------
// pair tensors and metrics. N4[rb][sb][rk][sk] (na, nb, na, nb);
// theta[r][p][q][s] (na, d, d, nb); cores A[r][p][k] (na, d, D), B[k][q][s].
// ---------------------------------------------------------------------------

// out[r,p,q,s] = sum_{t,u} N[r,s,t,u] theta[t,p,q,u] for a Z2-symmetric N
FU_HD inline void apply_metric(const Team& tm, View N, View th, View out, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d;
  const int tot = (int)na * nb * d * d;
  for (int idx = tm.lane; idx < tot; idx += tm.size) {
    int r_ = idx;
    const int q = (int)(r_ % d);
    r_ /= d;
    const int p = (int)(r_ % d);
    r_ /= d;
    const int s = (int)(r_ % nb);
    const int r = (int)(r_ / nb);
    // N[r,s,t,u] vanishes unless par(r) + par(s) = par(t) + par(u): u runs
    // over the one parity sector each t allows
    const int prs = par(r, g.ea) ^ par(s, g.eb);
    double acc = 0.0;
    for (int t = 0; t < na; ++t) {
      const int pu = prs ^ par(t, g.ea);
      const int u0 = pu == 0 ? 0 : g.eb, u1 = pu == 0 ? g.eb : nb;
      for (int u = u0; u < u1; ++u)
        acc += N[(((int)r * nb + s) * na + t) * nb + u] * th[(((int)t * d + p) * d + q) * nb + u];
    }
    out[(((int)r * d + p) * d + q) * nb + s] = acc;
  }
  tm.sync();
}

FU_HD inline void pair_tensor(const Team& tm, View A, View B, View out, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  const int tot = (int)na * d * d * nb;
  for (int idx = tm.lane; idx < tot; idx += tm.size) {
    const int s = (int)(idx % nb);
    const int rpq = idx / nb;
    const int q = (int)(rpq % d);
    const int rp = rpq / d;
    double acc = 0.0;
    for (int k = 0; k < D; ++k) acc += A[rp * D + k] * B[((int)k * d + q) * nb + s];
    out[idx] = acc;
  }
  tm.sync();
}

FU_HD inline double dot(const Team& tm, View a, View b, int n) {
  double s = 0.0;
  for (int i = tm.lane; i < n; i += tm.size) s += a[i] * b[i];
  return tm.sum(s);
}

// res . N res with res = pair(A, B) - target
FU_HD inline double residual_cost(const Team& tm, View N, View A, View B, View target, const Dims& g, Work& w) {
  const long long mark = w.top;
  const int n = (int)g.na * g.d * g.d * g.nb;
  View res = w.take(n), nr = w.take(n);
  pair_tensor(tm, A, B, res, g);
  for (int i = tm.lane; i < n; i += tm.size) res[i] -= target[i];
  tm.sync();
  apply_metric(tm, N, res, nr, g);
  const double v = dot(tm, res, nr, n);
  w.top = mark;
  return v;
}

// A'[i,p,k] = sum_x Li[i][x] A[x,p,k];  B'[k,q,i] = sum_y Ri[i][y] B[k,q,y]
FU_HD inline void original_cores(const Team& tm, View A, View B, View Li, View Ri, View oA, View oB, const Dims& g) {
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  for (int idx = tm.lane; idx < (int)na * d * D; idx += tm.size) {
    const int pk = idx % ((int)d * D);
    const int i = (int)(idx / ((int)d * D));
    double acc = 0.0;
    for (int x = 0; x < na; ++x) acc += Li[(int)i * na + x] * A[(int)x * d * D + pk];
    oA[idx] = acc;
  }
  for (int idx = tm.lane; idx < (int)D * d * nb; idx += tm.size) {
    const int i = (int)(idx % nb);
    const int kq = idx / nb;
    double acc = 0.0;
    for (int y = 0; y < nb; ++y) acc += Ri[(int)i * nb + y] * B[kq * nb + y];
    oB[idx] = acc;
  }
  tm.sync();
}

// Jacobi-preconditioned CG with a diagonal shift (_solve_core, full_update.jl:371-400)
// on the n x n matrix M for nc right-hand-side columns (x, rhs: n x nc).
FU_HD inline double solve_core(const Team& tm, View M, int n, int nc, View rhs, View x, const Opts& o, Work& w) {
  const long long mark = w.top;
  const int nv = (int)n * nc;
  View pre = w.take(n), r = w.take(nv), z = w.take(nv), p = w.take(nv), Ap = w.take(nv);
  double sl = 0.0;
  for (int i = tm.lane; i < n; i += tm.size) {
    const double dg = M[(int)i * n + i] > 0.0 ? M[(int)i * n + i] : 0.0;
    pre[i] = dg;
    if (dg > sl) sl = dg;
  }
  const double scale = tm.max(sl);
  const double shift = o.diagonal_shift * (scale > kTiny ? scale : kTiny);
  double floor = shift;
  if (kEps * scale > floor) floor = kEps * scale;
  if (kTiny > floor) floor = kTiny;
  for (int i = tm.lane; i < n; i += tm.size) pre[i] = (pre[i] + shift) > floor ? (pre[i] + shift) : floor;
  auto action = [&](View v, View out) {
    tm.sync();
    for (int idx = tm.lane; idx < nv; idx += tm.size) {
      const int i = (int)(idx / nc), c = (int)(idx % nc);
      double acc = 0.0;
      for (int k = 0; k < n; ++k) acc += M[(int)i * n + k] * v[(int)k * nc + c];
      out[idx] = acc + shift * v[idx];
    }
    tm.sync();
  };
  action(x, Ap);
  double rl = 0.0;
  for (int i = tm.lane; i < nv; i += tm.size) {
    r[i] = rhs[i] - Ap[i];
    rl += rhs[i] * rhs[i];
  }
  double rhsn = sqrt(tm.sum(rl));
  if (rhsn < kTiny) rhsn = kTiny;
  double residual = sqrt(dot(tm, r, r, nv)) / rhsn;
  if (residual <= o.linear_tolerance) {
    w.top = mark;
    return residual;
  }
  for (int i = tm.lane; i < nv; i += tm.size) {
    z[i] = r[i] / pre[i / nc];
    p[i] = z[i];
  }
  double rz = dot(tm, r, z, nv);
  for (int it = 0; it < o.linear_max_iterations; ++it) {
    action(p, Ap);
    double curv = 0.0, pp = 0.0, aa = 0.0;
    for (int i = tm.lane; i < nv; i += tm.size) {
      curv += p[i] * Ap[i];
      pp += p[i] * p[i];
      aa += Ap[i] * Ap[i];
    }
    tm.sum3(curv, pp, aa);
    const double lim = sqrt(pp) * sqrt(aa) > kTiny ? sqrt(pp) * sqrt(aa) : kTiny;
    if (!(curv > kEps * lim)) break;
    const double alpha = rz / curv;
    double rr = 0.0, nxt = 0.0;
    for (int i = tm.lane; i < nv; i += tm.size) {
      x[i] += alpha * p[i];
      r[i] -= alpha * Ap[i];
      z[i] = r[i] / pre[i / nc];
      rr += r[i] * r[i];
      nxt += r[i] * z[i];
    }
    tm.sum2(rr, nxt);
    residual = sqrt(rr) / rhsn;
    if (residual <= o.linear_tolerance) break;
    if (!(nxt >= 0)) break;
    const double beta = nxt / (rz > kTiny ? rz : kTiny);
    for (int i = tm.lane; i < nv; i += tm.size) p[i] = z[i] + beta * p[i];
    rz = nxt;
  }
  tm.sync();
  w.top = mark;
  return residual;
}

// One ALS normal matrix from a Z2-symmetric metric (full_update.jl:402-416),
// written for both sides:
//   out[(x,b),(y,o)] = sum_{xi,yi} M[x,xi,y,yi] P[(b,xi),(o,yi)]
// with (x, y) the outer pair (no entries, even count eo), (xi, yi) the summed
// pair (ni, ei), b, o the kept bond (D, eD). The metric vanishes unless
// par(x) + par(xi) = par(y) + par(yi), so yi runs over one parity sector, and
// the result vanishes unless par(x) + par(b) = par(y) + par(o). One lane per
// (x, y, b) keeps its row of D entries in registers; P is read by all lanes
// (staged in LDS by the caller when it fits). The terms of every entry are
// added in the order (xi outer, yi inner).
struct NormalShape {
  int no, eo, ni, ei;
  int mx, my, mxi, myi;  // metric strides
  int pxi, pyi, pb, po;  // P strides
  int ox, oy, ob, oo;    // output strides
};

constexpr int kTileD = 8;  // register rows cover kept bonds up to this size

FU_HD inline void normal_matrix(const Team& t, View M, View P, View out, const NormalShape& h, int D, int eD) {
  for (int task = t.lane; task < (int)h.no * h.no * D; task += t.size) {
    const int b = (int)(task % D);
    const int xy = task / D;
    const int x = (int)(xy / h.no), y = (int)(xy % h.no);
    const int px = par(x, h.eo), py = par(y, h.eo);
    const int obase = x * h.ox + y * h.oy + b * h.ob;
    if (D <= kTileD) {
      double acc[kTileD];
#pragma unroll
      for (int o = 0; o < kTileD; ++o) acc[o] = 0.0;
      for (int xi = 0; xi < h.ni; ++xi) {
        const int pyi = px ^ py ^ par(xi, h.ei);
        const int y0 = pyi == 0 ? 0 : h.ei, y1 = pyi == 0 ? h.ei : h.ni;
        for (int yi = y0; yi < y1; ++yi) {
          const double m = M[x * h.mx + y * h.my + xi * h.mxi + yi * h.myi];
          const int pbase = xi * h.pxi + yi * h.pyi + b * h.pb;
#pragma unroll
          for (int o = 0; o < kTileD; ++o)
            if (o < D) acc[o] += m * P[pbase + o * h.po];
        }
      }
#pragma unroll
      for (int o = 0; o < kTileD; ++o)
        if (o < D) out[obase + o * h.oo] = (px ^ par(b, eD)) == (py ^ par(o, eD)) ? acc[o] : 0.0;
    } else {
      for (int o = 0; o < D; ++o) {
        double acc = 0.0;
        if ((px ^ par(b, eD)) == (py ^ par(o, eD)))
          for (int xi = 0; xi < h.ni; ++xi) {
            const int pyi = px ^ py ^ par(xi, h.ei);
            const int y0 = pyi == 0 ? 0 : h.ei, y1 = pyi == 0 ? h.ei : h.ni;
            for (int yi = y0; yi < y1; ++yi)
              acc += M[x * h.mx + y * h.my + xi * h.mxi + yi * h.myi] * P[xi * h.pxi + yi * h.pyi + b * h.pb + o * h.po];
          }
        out[obase + o * h.oo] = acc;
      }
    }
  }
  t.sync();
}

// Doubles of workspace fu_bond needs for these dimensions.
FU_HD inline long long workspace_doubles(const Dims& g) {
  const long long n = (long long)g.na * g.nb;
  const long long t = (long long)g.na * g.d * g.d * g.nb;
  const long long ca = (long long)g.na * g.d * g.D, cb = (long long)g.D * g.d * g.nb;
  const long long nl = (long long)g.na * g.D, nr = (long long)g.D * g.nb;
  const long long nm = nl > nr ? nl : nr;
  long long big = 9 * n * n + 6 * n + 4 * (long long)g.na * g.na + 4 * (long long)g.nb * g.nb;
  big += 8 * t + 16 * (ca > cb ? ca : cb) + 4 * (long long)g.D * g.D;
  big += 4 * nm * nm + 12 * nm * g.d;
  big += 4 * (n * (g.na > g.nb ? g.na : g.nb)) + 4 * t;
  return big + 1024;
}

// ---------------------------------------------------------------------------
// the bond solve
// ---------------------------------------------------------------------------

// N4: the bond metric N[rb,sb,rk,sk]; target[r,p,q,s]; outputs the new cores
// in original coordinates, bond zero-padded to D. Returns the record (the same
// on every lane).
FU_HD inline Result fu_bond(const Team& t, const Dims& g, const Opts& o, View N4, View target, View Aout, View Bout,
                            Work& w) {
  Result res{};
  res.stop = o.max_iterations == 0 ? kDisabled : kIterationLimit;
  const int na = g.na, nb = g.nb, d = g.d, D = g.D;
  const int n = (int)na * nb;
  const int tn = (int)na * d * d * nb;
  const int L = t.lane, NT = t.size;
  // indices below are 32-bit: every array lies inside a workspace of < 2^31 doubles
  if (workspace_doubles(g) > w.cap || workspace_doubles(g) > 0x7fffffff) {
    res.status = kWorkspace;
    return res;
  }
  auto pa = [&](int i) { return par(i, g.ea); };
  auto pb = [&](int i) { return par(i, g.eb); };
  auto pd = [&](int i) { return par(i, g.ed); };
  // ---- environment gauge (full_update.jl:64-104) ----
  double scale = 0.0;
  {
    double sl = 0.0, bad = 0.0;
    for (int i = L; i < n * n; i += NT) {
      const double v = N4[i];
      if (!finite(v)) bad = 1.0;
      else sl += v * v;
    }
    double z = 0.0;
    t.sum3(sl, bad, z);
    if (bad != 0.0) {
      res.status = kNonfiniteMetric;
      return res;
    }
    scale = sqrt(sl);
  }
  if (!(scale > 0.0)) {
    res.status = kZeroMetric;
    return res;
  }
  View Z = w.take(n * n);       // rows: eigenvectors (k), columns (rk, sk)
  View zpar = w.take(n);        // parity of each row of Z
  {
    const long long mark = w.top;
    View bmap = w.take(n);      // block-major position -> (rk, sk) index
    View binv = w.take(n);      // (rk, sk) index -> position inside its block
    int cnt[2] = {0, 0};
    for (int I = 0; I < n; ++I) ++cnt[pa((int)(I / nb)) ^ pb((int)(I % nb))];
    if (L == 0) {
      int fill[2] = {0, 0};
      for (int I = 0; I < n; ++I) {
        const int q = pa((int)(I / nb)) ^ pb((int)(I % nb));
        bmap[(q == 0 ? 0 : cnt[0]) + fill[q]] = (double)I;
        binv[I] = fill[q];
        ++fill[q];
      }
    }
    View Hq = w.take(n * n), cs = w.take(3 * (n / 2 + 1));
    View evals = w.take(n), evecs = w.take(n * n);  // block-major storage
    t.sync();
    int eoff = 0, voff = 0;
    for (int q = 0; q < 2; ++q) {
      const int m = cnt[q];
      if (m == 0) continue;
      const int boff = q == 0 ? 0 : cnt[0];
      for (int idx = L; idx < (int)m * m; idx += NT) {
        const int I = (int)bmap[boff + idx / m], K = (int)bmap[boff + idx % m];
        Hq[idx] = 0.5 * (N4[I * n + K] + N4[K * n + I]) / scale;
      }
      t.sync();
      View wq{&evals[eoff], evals.s};
      View vq{&evecs[voff], evecs.s};
      const int mm = (int)m * m;
      if (2 * mm + 3 * (m / 2 + 1) <= t.lds_cap) {
        // the block and its eigenvectors in LDS
        const View As{t.lds, 1}, Vs{t.lds + mm, 1}, cl{t.lds + 2 * mm, 1};
        for (int i = L; i < mm; i += NT) As[i] = Hq[i];
        t.sync();
        if (!jacobi_eigh(t, As, m, wq, Vs, cl)) ++res.nonconverged;
        for (int i = L; i < mm; i += NT) vq[i] = Vs[i];
        t.sync();
      } else {
        if (!jacobi_eigh(t, Hq, m, wq, vq, cs)) ++res.nonconverged;
      }
      eoff += m;
      voff += (int)m * m;
    }
    // eigenvalues of both blocks set the spectral scale
    double spectral = 0.0;
    for (int k = 0; k < eoff; ++k)
      if (fabs(evals[k]) > spectral) spectral = fabs(evals[k]);
    if (spectral == 0.0) spectral = 1.0;
    double rel = sqrt(kEps);
    if ((double)n * kEps > rel) rel = (double)n * kEps;
    const double floorv = rel * spectral;
    for (int k = 0; k < eoff; ++k)
      if (evals[k] < floorv) ++res.floored;
    // Z row (block q, eigenvector k) = sqrt(max(lambda_k, floor)) v_k on the
    // block's (rk, sk) entries, zero elsewhere
    for (int idx = L; idx < n * n; idx += NT) {
      const int row = idx / n, K = idx % n;
      const int qr = row < cnt[0] ? 0 : 1;
      const int k = (int)(qr == 0 ? row : row - cnt[0]);
      const int m = cnt[qr];
      const int ev0 = qr == 0 ? 0 : cnt[0], vv0 = qr == 0 ? 0 : (int)cnt[0] * cnt[0];
      const int qk = pa((int)(K / nb)) ^ pb((int)(K % nb));
      double v = 0.0;
      if (qk == qr) {
        const double lam = evals[ev0 + k];
        const double root = sqrt(lam > floorv ? lam : floorv);
        v = root * evecs[vv0 + (int)binv[K] * m + k];
      }
      Z[idx] = v;
    }
    for (int row = L; row < n; row += NT) zpar[row] = row < cnt[0] ? 0 : 1;
    t.sync();
    w.top = mark;
  }
  // left = R of ((k, sk) | rk), right = R of ((k, rk) | sk)
  View left = w.take((int)na * na), right = w.take((int)nb * nb);
  View Li = w.take((int)na * na), Ri = w.take((int)nb * nb);
  {
    const long long mark = w.top;
    const View none{nullptr, 0};
    View Ml = w.take(n * nb * na);
    for (int idx = L; idx < n * nb * na; idx += NT) {
      const int kl = idx / na;
      const int a = (int)(idx % na);
      Ml[idx] = Z[(kl / nb) * n + (int)a * nb + kl % nb];
    }
    t.sync();
    int kc[2] = {g.ea, na - g.ea};
    block_qr(t, Ml, (int)(n * nb), na, [&](int i) { return ((int)zpar[i / nb]) ^ pb(i % nb); }, pa, kc, none, left, w);
    w.top = mark;
    View Mr = w.take(n * na * nb);
    for (int idx = L; idx < n * na * nb; idx += NT) {
      const int ka = idx / nb;
      const int l = (int)(idx % nb);
      Mr[idx] = Z[(ka / na) * n + (ka % na) * nb + l];
    }
    t.sync();
    int kd[2] = {g.eb, nb - g.eb};
    block_qr(t, Mr, (int)(n * na), nb, [&](int i) { return ((int)zpar[i / na]) ^ pa(i % na); }, pb, kd, none, right, w);
    w.top = mark;
  }
  const bool li_ok = triangular_inverse(t, left, na, g.ea, Li);
  const bool ri_ok = triangular_inverse(t, right, nb, g.eb, Ri);
  if (!li_ok || !ri_ok) {
    res.status = kSingularGauge;
    return res;
  }
  // metric[x,y,z,w] = sum N[i,j,k,l] Li[i,x] Ri[j,y] Li[k,z] Ri[l,w]
  View metric = w.take(n * n);
  {
    const long long mark = w.top;
    View T1 = w.take(n * n);
    for (int idx = L; idx < n * n; idx += NT) {
      const int x = (int)(idx / (nb * n));
      const int rest = idx % (nb * n);
      double acc = 0.0;
      for (int i = 0; i < na; ++i) acc += Li[(int)i * na + x] * N4[(int)i * nb * n + rest];
      T1[idx] = acc;
    }
    t.sync();
    for (int idx = L; idx < n * n; idx += NT) {
      const int xy = idx / n, rest = idx % n;
      const int x = (int)(xy / nb), y = (int)(xy % nb);
      double acc = 0.0;
      for (int j = 0; j < nb; ++j) acc += Ri[(int)j * nb + y] * T1[((int)x * nb + j) * n + rest];
      metric[idx] = acc;
    }
    t.sync();
    for (int idx = L; idx < n * n; idx += NT) {
      const int xy = idx / n;
      const int z = (int)((idx % n) / nb), l = (int)(idx % nb);
      double acc = 0.0;
      for (int k = 0; k < na; ++k) acc += Li[(int)k * na + z] * metric[xy * n + (int)k * nb + l];
      T1[idx] = acc;
    }
    t.sync();
    for (int idx = L; idx < n * n; idx += NT) {
      const int xyz = idx / nb;
      const int ww = (int)(idx % nb);
      double acc = 0.0;
      for (int l = 0; l < nb; ++l) acc += Ri[(int)l * nb + ww] * T1[xyz * nb + l];
      metric[idx] = acc;
    }
    t.sync();
    w.top = mark;
  }
  // working target wt[x,p,q,y] = sum target[r,p,q,s] left[x,r] right[y,s]
  View wt = w.take(tn);
  for (int idx = L; idx < tn; idx += NT) {
    const int y = (int)(idx % nb);
    const int xpq = idx / nb;
    const int q = (int)(xpq % d), p = (int)((xpq / d) % d), x = (int)(xpq / ((int)d * d));
    double acc = 0.0;
    for (int r = 0; r < na; ++r) {
      const double lx = left[(int)x * na + r];
      if (lx == 0.0) continue;
      for (int s = 0; s < nb; ++s)
        acc += lx * right[(int)y * nb + s] * target[(((int)r * d + p) * d + q) * nb + s];
    }
    wt[idx] = acc;
  }
  t.sync();
  // ---- gauged truncated SVD start (full_update.jl:31-42) ----
  // The cores live in the kept bond (ke even + ko odd, ke + ko <= D) from here
  // on, as the reference's bond space does; the output is expanded to the
  // device's padded capacity (even slots [0, eD), odd slots [eD, D)) at the end.
  const int ca0 = (int)na * d * D, cb0 = (int)D * d * nb;
  View A = w.take(ca0), B = w.take(cb0);
  int ke = 0, ko = 0;
  {
    const long long mark = w.top;
    for (int i = L; i < ca0; i += NT) A[i] = 0.0;
    for (int i = L; i < cb0; i += NT) B[i] = 0.0;
    const int mr = na * d, mc = d * nb;
    const double wn = sqrt(dot(t, wt, wt, tn));
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
      Uv[q] = w.take((int)mq[q] * (kq[q] > 0 ? kq[q] : 1));
      Vv[q] = w.take((int)(kq[q] > 0 ? kq[q] : 1) * nq[q]);
      if (kq[q] == 0) continue;
      View Bm = w.take((int)mq[q] * nq[q]);
      if (L == 0) {
        int bi = 0;
        for (int i = 0; i < mr; ++i) {
          if ((pa(i / d) ^ pd(i % d)) != q) continue;
          int bj = 0;
          for (int j = 0; j < mc; ++j) {
            if ((pd(j / nb) ^ pb(j % nb)) != q) continue;
            Bm[(int)bi * nq[q] + bj] = wt[(int)i * mc + j];
            ++bj;
          }
          ++bi;
        }
      }
      t.sync();
      if (!svd_compact(t, Bm, mq[q], nq[q], Uv[q], Sv[q], Vv[q], w)) ++res.nonconverged;
    }
    // truncrank(D) over both sectors (ties: sector 0 first, earlier first),
    // truncspace caps (the bond capacity), trunctol(atol)
    const int rank = o.rank > 0 && o.rank < D ? o.rank : D;
    int take_q[2] = {0, 0};
    for (int r = 0; r < rank; ++r) {
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
    if (L == 0)
      for (int q = 0; q < 2; ++q) {
        const int off = q == 0 ? 0 : ke;
        int bi = 0;
        for (int i = 0; i < mr; ++i) {
          if ((pa(i / d) ^ pd(i % d)) != q) continue;
          for (int k = 0; k < take_q[q]; ++k)
            A[(int)i * Dk + off + k] = Uv[q][(int)bi * kq[q] + k] * sqrt(Sv[q][k]);
          ++bi;
        }
        for (int k = 0; k < take_q[q]; ++k) {
          int bj = 0;
          for (int j = 0; j < mc; ++j) {
            if ((pd(j / nb) ^ pb(j % nb)) != q) continue;
            B[(int)(off + k) * mc + j] = sqrt(Sv[q][k]) * Vv[q][(int)k * nq[q] + bj];
            ++bj;
          }
        }
      }
    t.sync();
    w.top = mark;
  }
  // ALS dimensions: the kept bond
  Dims gk = g;
  gk.D = ke + ko;
  gk.eD = ke;
  // ---- ALS (full_update.jl:418-482) ----
  {
  const int D = gk.D;
  const int ca = (int)na * d * D, cb = (int)D * d * nb;
  auto pD = [&](int i) { return par(i, gk.eD); };
  View Nt = w.take(tn);
  apply_metric(t, N4, target, Nt, g);
  const double target_norm = dot(t, target, Nt, tn);
  res.target_norm = target_norm;
  View oA = w.take(ca), oB = w.take(cb), bestA = w.take(ca), bestB = w.take(cb);
  original_cores(t, A, B, Li, Ri, oA, oB, gk);
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
  auto keep = [&](View dst, View src, int cnt) {
    for (int i = L; i < cnt; i += NT) dst[i] = src[i];
  };
  double best = 0.0;
  bool fb = false;
  if (!evaluate(residual_cost(t, N4, oA, oB, target, gk, w), best, fb)) return res;
  res.initial_cost = best;
  keep(bestA, oA, ca);
  keep(bestB, oB, cb);
  View fQA = w.take(ca), fC = w.take((int)D * D), fQB = w.take(cb);
  bool have = false;
  if (fb) {
    res.stop = kResidualFallback;
  } else if (o.max_iterations > 0 && best <= o.error_tolerance) {
    res.stop = kErrorTolerance;
  } else if (o.max_iterations > 0) {
    View wtw = w.take(tn);
    apply_metric(t, metric, wt, wtw, g);
    const int nl = na * D, nr = D * nb;
    View NL = w.take((int)nl * nl), yl = w.take((int)nl * d), xl = w.take((int)nl * d);
    View NR = w.take((int)nr * nr), yr = w.take((int)nr * d), xr = w.take((int)nr * d);
    View BB = w.take((int)nr * nr), AA = w.take((int)nl * nl);
    // BB and AA are never live together: one LDS region holds either
    const bool stage = (long long)nr * nr <= t.lds_cap && (long long)nl * nl <= t.lds_cap;
    const View BBs = stage ? View{t.lds, 1} : BB, AAs = stage ? View{t.lds, 1} : AA;
    View QA = w.take(ca), Rm = w.take((int)D * D), rinit = w.take(cb), C = w.take((int)D * D);
    View QB = w.take(cb), solm = w.take(ca), solt = w.take(cb), Qt = w.take(cb);
    const int kcD[2] = {gk.eD, D - gk.eD};
    for (int it = 1; it <= o.max_iterations; ++it) {
      // left normal equation (full_update.jl:402-408):
      // NL[(f,b),(g,o)] = sum_{s,n} metric[f,s,g,n] BB[(b,s),(o,n)], BB = sum_p B B
      for (int idx = L; idx < (int)nr * nr; idx += NT) {
        const int bs = idx / nr, on = idx % nr;
        const int b = (int)(bs / nb), s = (int)(bs % nb), oo = (int)(on / nb), nn = (int)(on % nb);
        double bb = 0.0;
        for (int p = 0; p < d; ++p) bb += B[((int)b * d + p) * nb + s] * B[((int)oo * d + p) * nb + nn];
        BBs[idx] = bb;
      }
      t.sync();
      {
        const NormalShape h{na, g.ea, nb, g.eb,
                            (int)nb * na * nb, nb, (int)na * nb, 1,
                            nr, 1, (int)nb * nr, nb,
                            (int)D * nl, D, nl, 1};
        normal_matrix(t, metric, BBs, NL, h, D, gk.eD);
      }
      for (int idx = L; idx < (int)nl * d; idx += NT) {
        const int p = (int)(idx % d);
        const int fb_ = idx / d;
        const int f = (int)(fb_ / D), b = (int)(fb_ % D);
        double acc = 0.0;
        for (int q = 0; q < d; ++q)
          for (int s = 0; s < nb; ++s)
            acc += wtw[(((int)f * d + p) * d + q) * nb + s] * B[((int)b * d + q) * nb + s];
        yl[idx] = acc;
        xl[idx] = A[((int)f * d + p) * D + b];
      }
      t.sync();
      const double ra = solve_core(t, NL, nl, d, yl, xl, o, w);
      // QA, R = left_orth(sol as ((r, p) | k))
      for (int idx = L; idx < ca; idx += NT) {
        const int k = (int)(idx % D);
        const int rp = idx / D;
        const int r = (int)(rp / d), p = (int)(rp % d);
        solm[idx] = xl[((int)r * D + k) * d + p];
      }
      t.sync();
      block_qr(t, solm, na * d, D, [&](int i) { return pa(i / d) ^ pd(i % d); }, pD, kcD, QA, Rm, w);
      // right_initial = R B
      for (int idx = L; idx < cb; idx += NT) {
        const int k = (int)(idx / ((int)d * nb));
        const int j = idx % ((int)d * nb);
        double acc = 0.0;
        for (int m = 0; m < D; ++m) acc += Rm[(int)k * D + m] * B[(int)m * d * nb + j];
        rinit[idx] = acc;
      }
      // right normal equation (full_update.jl:410-416) with A = QA:
      // NR[(b,s),(o,n)] = sum_{w,e} metric[w,s,e,n] AA[(b,w),(o,e)], AA = sum_p QA QA
      for (int idx = L; idx < (int)nl * nl; idx += NT) {
        const int bw = idx / nl, oe = idx % nl;
        const int b = (int)(bw / na), ww = (int)(bw % na), oo = (int)(oe / na), e = (int)(oe % na);
        double aa = 0.0;
        for (int p = 0; p < d; ++p) aa += QA[((int)ww * d + p) * D + b] * QA[((int)e * d + p) * D + oo];
        AAs[idx] = aa;
      }
      t.sync();
      {
        const NormalShape h{nb, g.eb, na, g.ea,
                            (int)na * nb, 1, (int)nb * na * nb, nb,
                            nl, 1, (int)na * nl, na,
                            nr, 1, (int)nb * nr, nb};
        normal_matrix(t, metric, AAs, NR, h, D, gk.eD);
      }
      for (int idx = L; idx < (int)nr * d; idx += NT) {
        const int p = (int)(idx % d);
        const int bs = idx / d;
        const int b = (int)(bs / nb), s = (int)(bs % nb);
        double acc = 0.0;
        for (int ww = 0; ww < na; ++ww)
          for (int q = 0; q < d; ++q)
            acc += wtw[(((int)ww * d + q) * d + p) * nb + s] * QA[((int)ww * d + q) * D + b];
        yr[idx] = acc;
        xr[idx] = rinit[((int)b * d + p) * nb + s];
      }
      t.sync();
      const double rb = solve_core(t, NR, nr, d, yr, xr, o, w);
      // C, QB = right_orth(sol as (k | (q, s))) via QR of its transpose
      for (int idx = L; idx < cb; idx += NT) {
        const int k = (int)(idx % D);
        const int qs = idx / D;
        const int q = (int)(qs / nb), s = (int)(qs % nb);
        solt[idx] = xr[((int)k * nb + s) * d + q];
      }
      t.sync();
      block_qr(t, solt, d * nb, D, [&](int i) { return pd(i / nb) ^ pb(i % nb); }, pD, kcD, Qt, Rm, w);
      for (int idx = L; idx < (int)D * D; idx += NT) {
        const int k = (int)(idx / D), m = (int)(idx % D);
        C[idx] = Rm[(int)m * D + k];
      }
      for (int idx = L; idx < cb; idx += NT) {
        const int m = (int)(idx / ((int)d * nb));
        const int j = idx % ((int)d * nb);
        QB[idx] = Qt[j * D + m];
      }
      t.sync();
      // A = QA C, B = QB
      for (int idx = L; idx < ca; idx += NT) {
        const int rp = idx / D;
        const int m = (int)(idx % D);
        double acc = 0.0;
        for (int k = 0; k < D; ++k) acc += QA[rp * D + k] * C[(int)k * D + m];
        A[idx] = acc;
      }
      keep(B, QB, cb);
      t.sync();
      original_cores(t, A, B, Li, Ri, oA, oB, gk);
      double cost = 0.0;
      bool fb2 = false;
      if (!evaluate(residual_cost(t, N4, oA, oB, target, gk, w), cost, fb2)) return res;
      res.residual = ra > rb ? ra : rb;
      res.iterations = it;
      if (fb2) {
        res.stop = kResidualFallback;
        break;
      }
      if (cost <= best) {
        best = cost;
        keep(bestA, oA, ca);
        keep(bestB, oB, cb);
        keep(fQA, QA, ca);
        keep(fC, C, (int)D * D);
        keep(fQB, QB, cb);
        have = true;
      }
      t.sync();
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
    View Us = w.take((int)D * D), Vs = w.take((int)D * D);
    for (int i = L; i < (int)D * D; i += NT) Us[i] = Vs[i] = 0.0;
    for (int q = 0; q < 2; ++q) {
      const int lo = q == 0 ? 0 : gk.eD, m = q == 0 ? gk.eD : D - gk.eD;
      if (m == 0) continue;
      View Cm = w.take((int)m * m), U = w.take((int)m * m), S = w.take(m), Vt = w.take((int)m * m);
      for (int idx = L; idx < (int)m * m; idx += NT)
        Cm[idx] = fC[(int)(lo + idx / m) * D + lo + idx % m];
      t.sync();
      if (!svd_compact(t, Cm, m, m, U, S, Vt, w)) ++res.nonconverged;
      for (int idx = L; idx < (int)m * m; idx += NT) {
        const int i = (int)(idx / m), k = (int)(idx % m);
        Us[(int)(lo + i) * D + lo + k] = U[(int)i * m + k] * sqrt(S[k]);
        Vs[(int)(lo + k) * D + lo + i] = sqrt(S[k]) * Vt[(int)k * m + i];
      }
    }
    t.sync();
    for (int idx = L; idx < ca; idx += NT) {
      const int rp = idx / D;
      const int m = (int)(idx % D);
      double acc = 0.0;
      for (int k = 0; k < D; ++k) acc += fQA[rp * D + k] * Us[(int)k * D + m];
      A[idx] = acc;
    }
    for (int idx = L; idx < cb; idx += NT) {
      const int m = (int)(idx / ((int)d * nb));
      const int j = idx % ((int)d * nb);
      double acc = 0.0;
      for (int k = 0; k < D; ++k) acc += Vs[(int)m * D + k] * fQB[(int)k * d * nb + j];
      B[idx] = acc;
    }
    t.sync();
    original_cores(t, A, B, Li, Ri, bestA, bestB, gk);
    w.top = mark;
  }
  // expand the kept bond into the padded capacity
  for (int i = L; i < (int)na * d * g.D; i += NT) Aout[i] = 0.0;
  for (int i = L; i < (int)g.D * d * nb; i += NT) Bout[i] = 0.0;
  t.sync();
  const int Dk = ke + ko;
  for (int idx = L; idx < (int)na * d * Dk; idx += NT) {
    const int rp = idx / Dk;
    const int k = (int)(idx % Dk);
    Aout[rp * g.D + (k < ke ? k : g.eD + k - ke)] = bestA[idx];
  }
  for (int idx = L; idx < (int)Dk * d * nb; idx += NT) {
    const int k = (int)(idx / ((int)d * nb));
    const int j = idx % ((int)d * nb);
    Bout[(int)(k < ke ? k : g.eD + k - ke) * d * nb + j] = bestB[idx];
  }
  t.sync();
  }
  if (res.status == kOk && res.nonconverged > 0) res.status = kNonconverged;
  return res;
}

}  // namespace fu
}  // namespace spp
