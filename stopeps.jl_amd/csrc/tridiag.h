// Householder tridiagonalisation of a symmetric matrix by one workgroup, and
// the Sturm count of a symmetric tridiagonal matrix (shared by the boundary
// entropy's eigenvalue solve and the physical-kind site metric).
#pragma once
#include <hip/hip_runtime.h>

namespace spp {

constexpr int kTriThreads = 512;
constexpr int kTriWaves = kTriThreads / 64;

__device__ inline double tri_wave_sum(double v) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum over the workgroup; every thread returns the same value.
__device__ inline double tri_block_sum(double v, double* red) {
  v = tri_wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kTriWaves; ++i) t += red[i];
  return t;
}

__device__ inline double tri_block_max(double v, double* red) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = red[0];
  for (int i = 1; i < kTriWaves; ++i) t = fmax(t, red[i]);
  return t;
}

// Householder vector of row k (entries k+1..n-1, current values): writes
// vout[k+1] = 1, vout[j > k+1], d[k] and e[k]; returns tau (block-uniform).
// With taus, the vector's entries j > k+1 also replace row k's (which no later
// step reads) and tau goes to taus[k].
__device__ inline double tri_reflector(double* A, int ld, int n, int k, double* vout, double* d, double* e,
                                       double* red, double* taus) {
  double* rk = A + (long long)k * ld;
  const double alpha = rk[k + 1];
  double s = 0.0;
  for (int j = k + 2 + (int)threadIdx.x; j < n; j += kTriThreads) s += rk[j] * rk[j];
  const double sigma = tri_block_sum(s, red);
  double tau = 0.0, scale = 0.0, off = alpha;
  if (sigma > 0.0) {
    const double beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
    tau = (beta - alpha) / beta;
    scale = 1.0 / (alpha - beta);
    off = beta;
  }
  for (int j = k + 2 + (int)threadIdx.x; j < n; j += kTriThreads) {
    const double v = rk[j] * scale;
    vout[j] = v;
    if (taus) rk[j] = v;
  }
  if (threadIdx.x == 0) {
    vout[k + 1] = 1.0;
    d[k] = rk[k];
    e[k] = off;
    if (taus) taus[k] = tau;
  }
  __syncthreads();
  return tau;
}

// Symmetrise the leading n x n block of A (average of the two triangles) and
// reduce it to tridiagonal form T = Q^T A Q in place, Q = H_0 H_1 ... H_{n-3}:
// diagonal d[0..n-1], off-diagonal e[0..n-2]. The rank-two update of step k is
// fused with the matrix-vector product of step k+1, so every step streams the
// trailing block once. With taus, reflector H_k = I - taus[k] u u^T keeps
// u[k+1] = 1 and u[j] = A[k][j] (j > k+1). Scratch (LDS): red [kTriWaves],
// v, v2, p, w [ld] each. Called by all kTriThreads threads of the workgroup.
__device__ inline void tridiagonalize(double* A, int ld, int n, double* red, double* v, double* v2, double* p,
                                      double* w, double* d, double* e, double* taus) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (long long i = tid; i < (long long)n * n; i += kTriThreads) {
    const int r = (int)(i / n), q = (int)(i % n);
    if (r < q) {
      const double s = 0.5 * (A[(long long)r * ld + q] + A[(long long)q * ld + r]);
      A[(long long)r * ld + q] = s;
      A[(long long)q * ld + r] = s;
    }
  }
  __syncthreads();
  if (n <= 2) {
    if (tid == 0) {
      d[0] = A[0];
      if (n == 2) {
        d[1] = A[(long long)ld + 1];
        e[0] = A[1];
      }
    }
    __syncthreads();
    return;
  }
  double tau = tri_reflector(A, ld, n, 0, v, d, e, red, taus);
  for (int i = 1 + wave; i < n; i += kTriWaves) {
    const double* ri = A + (long long)i * ld;
    double acc = 0.0;
    for (int j = 1 + lane; j < n; j += 64) acc += ri[j] * v[j];
    acc = tri_wave_sum(acc);
    if (lane == 0) p[i] = tau * acc;
  }
  __syncthreads();
  for (int k = 0; k <= n - 3; ++k) {
    // w = p + K v with K = -tau/2 (p . v): the step-k update is S -= v w^T + w v^T
    double s = 0.0;
    for (int i = k + 1 + tid; i < n; i += kTriThreads) s += p[i] * v[i];
    const double K = -0.5 * tau * tri_block_sum(s, red);
    for (int i = k + 1 + tid; i < n; i += kTriThreads) w[i] = p[i] + K * v[i];
    __syncthreads();
    if (k == n - 3) {
      if (tid == 0) {
        const int a = n - 2, b = n - 1;
        d[a] = A[(long long)a * ld + a] - 2.0 * v[a] * w[a];
        d[b] = A[(long long)b * ld + b] - 2.0 * v[b] * w[b];
        e[a] = A[(long long)a * ld + b] - (v[a] * w[b] + w[a] * v[b]);
      }
      __syncthreads();
      break;
    }
    {
      const int r = k + 1;
      double* rr = A + (long long)r * ld;
      for (int j = r + tid; j < n; j += kTriThreads) rr[j] -= v[r] * w[j] + w[r] * v[j];
    }
    __syncthreads();
    const double tau2 = tri_reflector(A, ld, n, k + 1, v2, d, e, red, taus);
    // rows k+2.. : apply the step-k update (columns k+2..; column k+1 is not
    // read again) and form the next product p = tau2 * S v2. Each wave takes
    // four rows at a time so that four rows' loads are in flight together.
    for (int i = k + 2 + wave; i < n; i += 4 * kTriWaves) {
      int r[4];
      bool h[4];
      double* rr[4];
      double vi[4], wi[4], acc[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        r[q] = i + q * kTriWaves;
        h[q] = r[q] < n;
        rr[q] = A + (long long)(h[q] ? r[q] : i) * ld;
        vi[q] = h[q] ? v[r[q]] : 0.0;
        wi[q] = h[q] ? w[r[q]] : 0.0;
        acc[q] = 0.0;
      }
      for (int j = k + 2 + lane; j < n; j += 64) {
        const double wj = w[j], vj = v[j], v2j = v2[j];
        double a[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = h[q] ? rr[q][j] : 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (!h[q]) continue;
          a[q] = a[q] - (vi[q] * wj + wi[q] * vj);
          rr[q][j] = a[q];
          acc[q] += a[q] * v2j;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = tri_wave_sum(acc[q]);
      if (lane == 0)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (h[q]) p[r[q]] = tau2 * acc[q];
    }
    __syncthreads();
    double* t = v;
    v = v2;
    v2 = t;
    tau = tau2;
  }
}

// Number of eigenvalues of the tridiagonal (d, off-diagonal squares ee) below x.
__device__ inline int tri_sturm_count(const double* d, const double* ee, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int j = 1; j < n; ++j) {
    q = d[j] - x - ee[j - 1] / q;
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
}

}  // namespace spp
