// Physical-kind conditionals of perfect Born sampling (the resampling
// proposals of src/sampling/born.jl:116-134): at each site the single-site
// metric `_site_hole` (born.jl:75-83) is made positive semidefinite by
// `_positive_metric(metric; square_root=true)` (src/states/tensors.jl:65-94),
// and label k gets the weight ||root v_k||^2 = sum_j max(lambda_j, 0) (u_j^T v_k)^2
// of the site tensor projected on the label's physical basis vector.
//
// Per site and sample the hole H[w,n,e,s] (pair legs, from the contraction
// kernel) is gathered into the two parity blocks of the matrix
// M[(bra w,n,e,s), (ket w,n,e,s)] (FP64). One workgroup per block symmetrises
// it, reduces it to tridiagonal form T = Q^T M Q (Householder, reflectors
// kept) and forms Q^T v_k for both labels; one thread per block then runs the
// implicit QL iteration with the rotations applied to those vectors, which
// gives every eigenvalue lambda_j of the block and u_j^T v_k without the
// eigenvectors. The eigenvalues are those of M, not of the reference's
// M / ||M||: every reported quantity and the normalised probabilities are
// invariant under that scale.
#include "sampler.h"
#include "psd.h"
#include "tridiag.h"
#include <algorithm>
#include <cfloat>

namespace spp {

namespace {

// index of the (ket i, bra j) pair in the parity-sorted pair order of the
// double layers (even-even, odd-odd, even-odd, odd-even blocks, bra fastest)
__device__ inline int pair_of(int i, int j, int dim, int e) {
  const int o = dim - e;
  const bool ie = i < e, je = j < e;
  if (ie && je) return i * e + j;
  if (!ie && !je) return e * e + (i - e) * o + (j - e);
  if (ie) return e * e + o * o + i * o + (j - e);
  return e * e + o * o + e * o + (i - e) * e + j;
}

struct TupleDims {
  int d[4];  // W, N, E, S virtual dimensions
  int e[4];  // their even counts
};

__device__ inline void tuple_of(int t, const TupleDims& td, int* idx) {
  for (int k = 3; k >= 0; --k) {
    idx[k] = t % td.d[k];
    t /= td.d[k];
  }
}

// M[s][q][r][c] = H[s][pair(ket c, bra r)] for the ket / bra tuples of block q.
template <typename T>
__global__ void hole_gather_kernel(const T* H, long long hs, const int* inv, int ld, int n0, int n1, TupleDims td,
                                   double* M, long long ms) {
  const int s = blockIdx.y, q = blockIdx.z;
  const int n = q == 0 ? n0 : n1;
  const long long tot = (long long)n * n;
  const int DN = td.d[1] * td.d[1], DE = td.d[2] * td.d[2], DS = td.d[3] * td.d[3];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / n), c = (int)(i % n);
    int kt[4], bt[4];
    tuple_of(inv[q * ld + c], td, kt);
    tuple_of(inv[q * ld + r], td, bt);
    const long long h = (((long long)pair_of(kt[0], bt[0], td.d[0], td.e[0]) * DN +
                          pair_of(kt[1], bt[1], td.d[1], td.e[1])) * DE +
                         pair_of(kt[2], bt[2], td.d[2], td.e[2])) * DS +
                        pair_of(kt[3], bt[3], td.d[3], td.e[3]);
    M[(long long)s * ms + (long long)q * ld * ld + (long long)r * ld + c] = (double)H[(long long)s * hs + h];
  }
}

// One workgroup per (sample, block): Frobenius norms of the block and of its
// antisymmetric part, then T = Q^T M Q and y_k = Q^T v_k. v_k[c] = sum_p
// U[p][k] A[w,n,p,0,e,s] for the ket tuple c (the site tensor projected on
// label k, `_project(A, B, k, :physical)`, basis.jl:87).
__global__ __launch_bounds__(kTriThreads) void psd_tridiag_kernel(double* M, long long ms, int ld, int n0, int n1,
                                                                  const int* inv, TupleDims td, const double* site,
                                                                  int dP, double u00, double u01, double u10,
                                                                  double u11, double* taus, double* dv, double* ev,
                                                                  double* yv, double* norms) {
  const int prob = blockIdx.x;
  const int s = prob >> 1, q = prob & 1;
  const int n = q == 0 ? n0 : n1;
  const int tid = threadIdx.x;
  extern __shared__ double lds[];
  double* red = lds;
  double* v = red + kTriWaves;
  double* v2 = v + ld;
  double* p = v2 + ld;
  double* w = p + ld;
  double* d = w + ld;
  double* e = d + ld;
  double* A = M + (long long)s * ms + (long long)q * ld * ld;
  // norms of the block as given (before symmetrisation)
  double ss = 0.0, df = 0.0;
  for (long long i = tid; i < (long long)n * n; i += kTriThreads) {
    const int r = (int)(i / n), c = (int)(i % n);
    const double a = A[(long long)r * ld + c];
    ss += a * a;
    if (r < c) {
      const double b = A[(long long)c * ld + r];
      df += 2.0 * (a - b) * (a - b);
    }
  }
  ss = tri_block_sum(ss, red);
  df = tri_block_sum(df, red);
  if (tid == 0) {
    norms[2 * prob] = ss;
    norms[2 * prob + 1] = df;
  }
  if (n <= 0) return;
  double* tau = taus + (long long)prob * ld;
  tridiagonalize(A, ld, n, red, v, v2, p, w, d, e, tau);
  // label vectors of this block (v, v2 are free again); the site is
  // (W,N,P,A,E,S) with an auxiliary leg of one index
  for (int c = tid; c < n; c += kTriThreads) {
    int t[4];
    tuple_of(inv[q * ld + c], td, t);
    const long long base = (((long long)t[0] * td.d[1] + t[1]) * dP) * td.d[2] * td.d[3];
    const long long tail = (long long)t[2] * td.d[3] + t[3];
    const long long pstride = (long long)td.d[2] * td.d[3];
    double x0 = 0.0, x1 = 0.0;
    for (int pp = 0; pp < dP; ++pp) {
      const double a = site[base + pp * pstride + tail];
      x0 += (pp == 0 ? u00 : u10) * a;
      x1 += (pp == 0 ? u01 : u11) * a;
    }
    v[c] = x0;
    v2[c] = x1;
  }
  __syncthreads();
  // y = H_{n-3} ... H_0 v with H_k = I - tau_k u u^T, u[k+1] = 1, u[j > k+1] = A[k][j]
  for (int k = 0; k + 2 < n; ++k) {
    const double* rk = A + (long long)k * ld;
    double a0 = (tid == 0) ? v[k + 1] : 0.0, a1 = (tid == 0) ? v2[k + 1] : 0.0;
    for (int j = k + 2 + tid; j < n; j += kTriThreads) {
      a0 += rk[j] * v[j];
      a1 += rk[j] * v2[j];
    }
    const double t0 = tau[k] * tri_block_sum(a0, red);
    const double t1 = tau[k] * tri_block_sum(a1, red);
    if (tid == 0) {
      v[k + 1] -= t0;
      v2[k + 1] -= t1;
    }
    for (int j = k + 2 + tid; j < n; j += kTriThreads) {
      v[j] -= t0 * rk[j];
      v2[j] -= t1 * rk[j];
    }
    __syncthreads();
  }
  double* dg = dv + (long long)prob * ld;
  double* eg = ev + (long long)prob * ld;
  double* yg = yv + (long long)prob * 2 * ld;
  for (int c = tid; c < n; c += kTriThreads) {
    dg[c] = d[c];
    eg[c] = c < n - 1 ? e[c] : 0.0;
    yg[c] = v[c];
    yg[ld + c] = v2[c];
  }
}

// One wave per (sample, block): the tridiagonal form and the label vectors are
// staged in LDS, lane 0 runs the implicit QL iteration with its rotations
// applied to the vectors that are nonzero in this block (in a Z2 metric each
// label's vector lives in one parity block), then the wave sums per block
// {min lambda, max |lambda|, negatives, sum of negative squares, sum of
// squares, weight of label 0, weight of label 1, failure flag}.
__global__ __launch_bounds__(64) void psd_ql_kernel(double* dv, const double* ev, double* yv, int ld, int n0, int n1,
                                                    int nprob, double* res, int writeback) {
  const int prob = blockIdx.x;
  if (prob >= nprob) return;
  const int n = (prob & 1) == 0 ? n0 : n1;
  const int lane = threadIdx.x;
  extern __shared__ double sh[];
  double* d = sh;
  double* e = d + ld;
  double* y = e + ld;  // [2][ld]
  __shared__ int nz[2];
  __shared__ int fail;
  double* dg = dv + (long long)prob * ld;
  const double* eg = ev + (long long)prob * ld;
  double* yg = yv + (long long)prob * 2 * ld;
  bool any0 = false, any1 = false;
  for (int j = lane; j < n; j += 64) {
    d[j] = dg[j];
    e[j] = eg[j];
    y[j] = yg[j];
    y[ld + j] = yg[ld + j];
    any0 = any0 || yg[j] != 0.0;
    any1 = any1 || yg[ld + j] != 0.0;
  }
  any0 = __any(any0);
  any1 = __any(any1);
  if (lane == 0) {
    nz[0] = any0;
    nz[1] = any1;
  }
  __syncthreads();
  if (lane == 0) {
    // rotate only the nonzero vectors: both, the first, or the second alone
    const int ny = nz[0] + nz[1];
    double* yy = nz[0] ? y : y + ld;
    fail = ql_project(n, d, e, yy, ld, ny, 60);
  }
  __syncthreads();
  double mn = 0.0, mx = 0.0, neg = 0.0, neg2 = 0.0, sum2 = 0.0, w0 = 0.0, w1 = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double l = d[j];
    mn = fmin(mn, l);
    mx = fmax(mx, fabs(l));
    if (l < 0.0) {
      neg += 1.0;
      neg2 += l * l;
    }
    sum2 += l * l;
    if (l > 0.0) {
      if (nz[0]) w0 += l * y[j] * y[j];
      if (nz[1]) w1 += l * y[ld + j] * y[ld + j];
    }
  }
  for (int m = 32; m >= 1; m >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, m, 64));
    mx = fmax(mx, __shfl_xor(mx, m, 64));
    neg += __shfl_xor(neg, m, 64);
    neg2 += __shfl_xor(neg2, m, 64);
    sum2 += __shfl_xor(sum2, m, 64);
    w0 += __shfl_xor(w0, m, 64);
    w1 += __shfl_xor(w1, m, 64);
  }
  if (writeback)  // self-tests read the eigenvalues and projections back
    for (int j = lane; j < n; j += 64) {
      dg[j] = d[j];
      yg[j] = y[j];
      yg[ld + j] = y[ld + j];
    }
  if (lane == 0) {
    double* r = res + (long long)prob * 8;
    r[0] = mn;
    r[1] = mx;
    r[2] = neg;
    r[3] = neg2;
    r[4] = sum2;
    r[5] = w0;
    r[6] = w1;
    r[7] = fail;
  }
}

// Per sample: the site's report and its Born conditional from the block
// results; the allowed labels follow the total charge (born.jl:112-116), a
// zero total falls back to the uniform distribution over the allowed labels
// (born.jl:126-130); then the draw (or the given label) as for the auxiliary
// kind.
__global__ void physical_select_kernel(const double* res, const double* norms, const double* uniforms,
                                       const int* labels_in, long long sample0, int nsite, int site, int c0, int c1,
                                       int target, int reach_next, int* prefix, int* choice, int* labels,
                                       double* logp, double* cond, int* status, int* fail_site, int* done,
                                       double* repairs, int nb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const long long gs = sample0 + s;
  const double* r0 = res + (long long)(2 * s) * 8;
  const double* r1 = r0 + 8;
  const double* nm = norms + (long long)(2 * s) * 2;
  int k = labels_in ? labels_in[gs * nsite + site] : 0;
  double* rep = repairs ? repairs + ((long long)s * nsite + site) * 6 : nullptr;
  if (status[s] == 0 && !done[s]) {
    const double tiny = DBL_MIN;
    const double sumsq = nm[0] + nm[2], defect = nm[1] + nm[3];
    if (!isfinite(sumsq)) {
      status[s] = 4;
      fail_site[s] = site;
    } else if (r0[7] != 0.0 || r1[7] != 0.0) {
      status[s] = 6;
      fail_site[s] = site;
    } else {
      const double mn = fmin(fmin(r0[0], r1[0]), 0.0);
      const double spectral = fmax(fmax(r0[1], r1[1]), tiny);
      const double nneg = r0[2] + r1[2];
      const double neg2 = r0[3] + r1[3], sum2 = r0[4] + r1[4];
      bool allowed[2] = {true, true};
      if (target >= 0) {
        const int cs[2] = {c0, c1};
        for (int j = 0; j < 2; ++j) {
          const int need = target ^ prefix[s] ^ cs[j];
          allowed[j] = (reach_next >> need) & 1;
        }
      }
      if (!allowed[0] && !allowed[1]) {
        status[s] = 5;
        fail_site[s] = site;
      } else {
        double val[2] = {allowed[0] ? r0[5] + r1[5] : 0.0, allowed[1] ? r0[6] + r1[6] : 0.0};
        const double total = val[0] + val[1];
        double pr[2];
        int zero_allowed = 0;
        if (total == 0.0) {
          zero_allowed = (int)allowed[0] + (int)allowed[1];
          pr[0] = allowed[0] ? 1.0 / zero_allowed : 0.0;
          pr[1] = allowed[1] ? 1.0 / zero_allowed : 0.0;
        } else {
          pr[0] = val[0] / total;
          pr[1] = val[1] / total;
        }
        if (rep) {
          rep[0] = mn / spectral;
          rep[1] = nneg;
          rep[2] = nneg;
          rep[3] = sqrt(neg2) / fmax(sqrt(sum2), tiny);
          rep[4] = sqrt(defect) / fmax(sqrt(sumsq), tiny);
          rep[5] = zero_allowed;
        }
        if (!labels_in) k = categorical_index(pr, 2, uniforms[gs * nsite + site]);
        const double pk = pr[k];
        cond[(long long)s * nsite + site] = pk;
        if (pk == 0.0) {
          logp[s] = -INFINITY;
          done[s] = 1;
        } else {
          logp[s] += log(pk);
        }
        prefix[s] ^= k == 0 ? c0 : c1;
      }
    }
  }
  choice[s] = k;
  labels[(long long)s * nsite + site] = k;
}

}  // namespace

void launch_physical_select(Context& ctx, const PsdScratch& w, const double* uniforms, const int* labels_in,
                            long long sample0, int nsite, int site, const int* charges, int target, int reach_next,
                            int* prefix, int* choice, int* labels, double* logp, double* cond, int* status,
                            int* fail_site, int* done, double* repairs, int nb) {
  hipLaunchKernelGGL(physical_select_kernel, dim3((nb + 255) / 256), dim3(256), 0, ctx.stream, w.res, w.norms,
                     uniforms, labels_in, sample0, nsite, site, charges[0], charges[1], target, reach_next, prefix,
                     choice, labels, logp, cond, status, fail_site, done, repairs, nb);
  SPP_CHECK(hipGetLastError());
}

void physical_conditionals(Context& ctx, const void* H, bool h_fp32, long long hs, const PhysSite& ps,
                           const double* U, int nb, const PsdScratch& w) {
  hipStream_t st = ctx.stream;
  const int ld = w.ld;
  if (ps.nq[0] > ld || ps.nq[1] > ld) throw HipError("physical metric block exceeds the workspace");
  TupleDims td;
  for (int k = 0; k < 4; ++k) {
    const int slot = k < 2 ? k : k + 2;  // W, N, E, S of (W,N,P,A,E,S)
    td.d[k] = ps.dim[slot];
    td.e[k] = ps.even[slot];
  }
  const long long ms = 2LL * ld * ld;
  const int nmax = std::max(ps.nq[0], ps.nq[1]);
  if (nmax > 0) {
    const long long tot = (long long)nmax * nmax;
    dim3 grid((unsigned)std::min<long long>((tot + 255) / 256, 1024), (unsigned)nb, 2);
    if (h_fp32)
      hipLaunchKernelGGL(hole_gather_kernel<float>, grid, dim3(256), 0, st, (const float*)H, hs, ps.inv, ld,
                         ps.nq[0], ps.nq[1], td, w.M, ms);
    else
      hipLaunchKernelGGL(hole_gather_kernel<double>, grid, dim3(256), 0, st, (const double*)H, hs, ps.inv, ld,
                         ps.nq[0], ps.nq[1], td, w.M, ms);
    SPP_CHECK(hipGetLastError());
  }
  const size_t lds = sizeof(double) * (kTriWaves + 6 * (size_t)ld);
  hipLaunchKernelGGL(psd_tridiag_kernel, dim3(2 * nb), dim3(kTriThreads), lds, st, w.M, ms, ld, ps.nq[0], ps.nq[1],
                     ps.inv, td, ps.a, ps.dim[2], U[0], U[1], U[2], U[3], w.taus, w.d, w.e, w.y, w.norms);
  SPP_CHECK(hipGetLastError());
  const int nprob = 2 * nb;
  hipLaunchKernelGGL(psd_ql_kernel, dim3(nprob), dim3(64), sizeof(double) * 4 * (size_t)ld, st, w.d, w.e, w.y, ld,
                     ps.nq[0], ps.nq[1], nprob, w.res, 0);
  SPP_CHECK(hipGetLastError());
}

// Self-test: eigenvalues and label weights of nprob symmetric blocks through
// the same tridiagonalisation and QL kernels (no hole, no gather).
void psd_selftest(Context& ctx, int nprob, int ld, const int* n_h, const double* mats_h, const double* vecs_h,
                  double* eig_h, double* proj_h, int* fail_h) {
  hipStream_t st = ctx.stream;
  for (int i = 0; i < nprob; ++i)
    if (n_h[i] < 0 || n_h[i] > ld) throw HipError("block size out of range");
  double *M = nullptr, *taus = nullptr, *d = nullptr, *e = nullptr, *y = nullptr, *nrm = nullptr, *res = nullptr;
  double* site = nullptr;
  int* inv = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)M, (void*)taus, (void*)d, (void*)e, (void*)y, (void*)nrm, (void*)res, (void*)site,
                    (void*)inv})
      if (p) (void)hipFree(p);
  };
  try {
    // one problem per launch slot pair: run each block as block 0 of a "sample"
    SPP_CHECK(hipMalloc(&M, sizeof(double) * 2 * (size_t)ld * ld));
    SPP_CHECK(hipMalloc(&taus, sizeof(double) * 2 * ld));
    SPP_CHECK(hipMalloc(&d, sizeof(double) * 2 * ld));
    SPP_CHECK(hipMalloc(&e, sizeof(double) * 2 * ld));
    SPP_CHECK(hipMalloc(&y, sizeof(double) * 4 * ld));
    SPP_CHECK(hipMalloc(&nrm, sizeof(double) * 4));
    SPP_CHECK(hipMalloc(&res, sizeof(double) * 16));
    SPP_CHECK(hipMalloc(&site, sizeof(double) * 2 * ld));
    SPP_CHECK(hipMalloc(&inv, sizeof(int) * 2 * ld));
    // the "site" is a one-leg tensor (W = ld, P = 2) holding the two vectors:
    // v_k = site[c][k] with the identity basis
    std::vector<int> hinv(2 * ld);
    for (int c = 0; c < ld; ++c) hinv[c] = hinv[ld + c] = c;
    upload(inv, hinv.data(), sizeof(int) * 2 * ld, st);
    TupleDims td;
    td.d[0] = ld;
    td.d[1] = td.d[2] = td.d[3] = 1;
    td.e[0] = ld;
    td.e[1] = td.e[2] = td.e[3] = 1;
    std::vector<double> hsite(2 * ld), hd(ld), hy(2 * ld), hres(16);
    for (int i = 0; i < nprob; ++i) {
      const int n = n_h[i];
      SPP_CHECK(hipMemcpy2DAsync(M, sizeof(double) * ld, mats_h + (size_t)i * ld * ld, sizeof(double) * ld,
                                 sizeof(double) * ld, ld, hipMemcpyHostToDevice, st));
      for (int c = 0; c < ld; ++c) {
        hsite[2 * c] = vecs_h[((size_t)i * 2) * ld + c];
        hsite[2 * c + 1] = vecs_h[((size_t)i * 2 + 1) * ld + c];
      }
      upload(site, hsite.data(), sizeof(double) * 2 * ld, st);
      const size_t lds = sizeof(double) * (kTriWaves + 6 * (size_t)ld);
      hipLaunchKernelGGL(psd_tridiag_kernel, dim3(1), dim3(kTriThreads), lds, st, M, 2LL * ld * ld, ld, n, 0, inv, td,
                         site, 2, 1.0, 0.0, 0.0, 1.0, taus, d, e, y, nrm);
      SPP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(psd_ql_kernel, dim3(1), dim3(64), sizeof(double) * 4 * (size_t)ld, st, d, e, y, ld, n, 0, 1,
                         res, 1);
      SPP_CHECK(hipGetLastError());
      SPP_CHECK(hipMemcpyAsync(hd.data(), d, sizeof(double) * ld, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(hy.data(), y, sizeof(double) * 2 * ld, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(hres.data(), res, sizeof(double) * 8, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
      for (int j = 0; j < ld; ++j) {
        eig_h[(size_t)i * ld + j] = j < n ? hd[j] : 0.0;
        proj_h[((size_t)i * 2) * ld + j] = j < n ? hy[j] : 0.0;
        proj_h[((size_t)i * 2 + 1) * ld + j] = j < n ? hy[ld + j] : 0.0;
      }
      fail_h[i] = (int)hres[7];
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace spp
