// Pre-truncation Schmidt entropy of boundary cuts (_schmidt_entropy and
// _boundary_svd, src/contraction/boundary_mps.jl:20-43).
//
// At the cut after site x the truncation problem is the compact parity-block
// matrix rho_x = Y G Y^T (one block per sample and bond parity) whose
// eigenvalues are the squared singular values of the reference's
// pre-truncation compact SVD. The entropy of a sample's cut is
//   S = -sum_i p_i log p_i,  p_i = lambda_i / sum_j lambda_j  (both parities;
// Z2 sectors have dimension one), NaN for a zero spectrum, at least 0.
//
// Blocks that the direct Jacobi solve of the truncation already diagonalised
// give their eigenvalues straight away. Larger blocks are copied aside during
// the sweeps and solved together when the caller flushes (once per batch of
// samples, so that thousands of blocks fill the device), one workgroup per
// block: Householder tridiagonalisation in place (the rank-two update of one
// step fused with the matrix-vector product of the next, so every step streams
// the trailing block once) followed by Sturm-count bisection for every
// eigenvalue.
#include "engine.h"
#include "tridiag.h"
#include <cfloat>

namespace spp {

namespace {

constexpr int kSpecThreads = kTriThreads;
constexpr int kSpecWaves = kTriWaves;

// All eigenvalues (ascending) of each symmetric block: blockIdx.x = block
// (sample * 2 + parity), blockIdx.y = table entry. The block is symmetrised and
// then overwritten.
__global__ __launch_bounds__(kSpecThreads) void spectrum_kernel(EntropyTable tab) {
  const int c = blockIdx.y;
  const int prob = blockIdx.x;
  if (prob >= 2 * tab.nb[c]) return;
  const int ld = tab.ld[c];
  const int n = tab.cnt[c][prob];
  double* A = tab.mat[c] + (long long)prob * ld * ld;
  double* wout = tab.w[c] + (long long)prob * tab.wstride[c];
  extern __shared__ double lds[];
  double* red = lds;              // [kSpecWaves]
  double* v = red + kSpecWaves;   // [ld] reflector of the current step
  double* v2 = v + ld;            // [ld] reflector of the next step
  double* p = v2 + ld;            // [ld] tau * S v
  double* w = p + ld;             // [ld] p + K v, then squared off-diagonal
  double* d = w + ld;             // [ld] diagonal of the tridiagonal form
  double* e = d + ld;             // [ld] off-diagonal
  const int tid = threadIdx.x;
  if (n <= 0) return;
  tridiagonalize(A, ld, n, red, v, v2, p, w, d, e, nullptr);
  // bisection on the tridiagonal form (Gershgorin interval, Sturm counts)
  double* ee = w;
  for (int j = tid; j < n - 1; j += kSpecThreads) ee[j] = e[j] * e[j];
  double lo = INFINITY, hi = -INFINITY, emax = 0.0;
  for (int i = tid; i < n; i += kSpecThreads) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i < n - 1) emax = fmax(emax, e[i] * e[i]);
  }
  lo = -tri_block_max(-lo, red);
  hi = tri_block_max(hi, red);
  emax = tri_block_max(emax, red);
  const double pivmin = DBL_MIN * fmax(1.0, emax);
  const double tol = 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin;
  for (int i = tid; i < n; i += kSpecThreads) {
    double a = lo, b = hi;
    for (int it = 0; it < 128 && b - a > tol; ++it) {
      const double mid = 0.5 * (a + b);
      if (tri_sturm_count(d, ee, n, mid, pivmin) > i)
        b = mid;
      else
        a = mid;
    }
    wout[i] = 0.5 * (a + b);
  }
}

__device__ inline double nan_max(double m, double v) { return (isnan(m) || isnan(v)) ? NAN : fmax(m, v); }

// Entropy of every sample (blockIdx.x) at every cut of the table, from the
// spectra of its two parity blocks; out[cut * nb + s] and the running maximum
// (one block per sample updates it entry after entry).
__global__ __launch_bounds__(256) void entropy_kernel(EntropyTable tab) {
  const int s = blockIdx.x;
  __shared__ double part[256];
  for (int c = 0; c < tab.count; ++c) {
    if (s >= tab.nb[c]) continue;
    double tot = 0.0;
    for (int q = 0; q < 2; ++q) {
      const double* wq = tab.w[c] + (long long)(2 * s + q) * tab.wstride[c];
      const int n = tab.cnt[c][2 * s + q];
      for (int i = threadIdx.x; i < n; i += 256) tot += fmax(wq[i], 0.0);
    }
    part[threadIdx.x] = tot;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
      __syncthreads();
    }
    const double A = part[0];
    __syncthreads();
    double h = 0.0;
    if (A > 0.0) {
      for (int q = 0; q < 2; ++q) {
        const double* wq = tab.w[c] + (long long)(2 * s + q) * tab.wstride[c];
        const int n = tab.cnt[c][2 * s + q];
        for (int i = threadIdx.x; i < n; i += 256) {
          const double pr = fmax(wq[i], 0.0) / A;
          if (pr > 0.0) h -= pr * log(pr);
        }
      }
    }
    part[threadIdx.x] = h;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
      __syncthreads();
    }
    const double S = A > 0.0 ? fmax(0.0, part[0]) : NAN;
    __syncthreads();
    if (threadIdx.x == 0) {
      if (tab.out[c]) tab.out[c][(long long)tab.cut[c] * tab.nb[c] + s] = S;
      if (tab.run_max[c]) tab.run_max[c][s] = nan_max(tab.run_max[c][s], S);
    }
  }
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

EntropyWork::~EntropyWork() {
  if (buf) (void)hipFree(buf);
}

void entropy_cut(Context& ctx, const double* w, long long wstride, const int* cnt, int nb, int cut, double* out,
                 double* run_max) {
  EntropyTable tab{};
  tab.w[0] = const_cast<double*>(w);
  tab.wstride[0] = wstride;
  tab.cnt[0] = cnt;
  tab.cut[0] = cut;
  tab.nb[0] = nb;
  tab.out[0] = out;
  tab.run_max[0] = run_max;
  tab.count = 1;
  hipLaunchKernelGGL(entropy_kernel, dim3(nb), dim3(256), 0, ctx.stream, tab);
  SPP_CHECK(hipGetLastError());
}

void entropy_flush(Context& ctx) {
  EntropyWork& ew = ctx.entropy;
  for (size_t g = 0; g < ew.pending.size(); g += kEntropyCuts) {
    EntropyTable tab{};
    int ldmax = 1, nbmax = 1;
    for (size_t i = g; i < ew.pending.size() && tab.count < kEntropyCuts; ++i) {
      const EntropyWork::Cut& c = ew.pending[i];
      const int k = tab.count++;
      tab.mat[k] = (double*)(ew.buf + c.mat);
      tab.w[k] = (double*)(ew.buf + c.w);
      tab.cnt[k] = (const int*)(ew.buf + c.cnt);
      tab.wstride[k] = c.ld;
      tab.ld[k] = c.ld;
      tab.cut[k] = c.cut;
      tab.nb[k] = c.nb;
      tab.out[k] = c.out;
      tab.run_max[k] = c.run_max;
      ldmax = std::max(ldmax, c.ld);
      nbmax = std::max(nbmax, c.nb);
    }
    const size_t lds = sizeof(double) * (kSpecWaves + 6 * (size_t)ldmax);
    if (lds > 64 * 1024) {
      int dev = 0, lmax = 0;
      SPP_CHECK(hipGetDevice(&dev));
      SPP_CHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeSharedMemPerBlockOptin, dev));
      if (lmax <= 0) SPP_CHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
      if (lds > (size_t)lmax)
        throw HipError("boundary entropy: cut block of dimension " + std::to_string(ldmax) +
                       " exceeds the local memory of one workgroup");
      SPP_CHECK(hipFuncSetAttribute((const void*)spectrum_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    }
    // blocks with rows past a block's dimension idle; blockIdx.x covers the largest batch
    hipLaunchKernelGGL(spectrum_kernel, dim3(2 * nbmax, tab.count), dim3(kSpecThreads), lds, ctx.stream, tab);
    SPP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(entropy_kernel, dim3(nbmax), dim3(256), 0, ctx.stream, tab);
    SPP_CHECK(hipGetLastError());
  }
  ew.pending.clear();
  ew.used = 0;
}

void entropy_reset(Context& ctx) {
  EntropyWork& ew = ctx.entropy;
  ew.pending.clear();
  ew.used = 0;
  if (ew.buf) {
    (void)hipStreamSynchronize(ctx.stream);
    (void)hipFree(ew.buf);
  }
  ew.buf = nullptr;
  ew.cap = 0;
}

void entropy_defer(Context& ctx, const double* rho, long long rps, int ld, const int* cnt, int nb, int cut,
                   double* out, double* run_max) {
  EntropyWork& ew = ctx.entropy;
  const size_t bm = align256(sizeof(double) * 2 * (size_t)nb * (size_t)ld * ld);
  const size_t bw = align256(sizeof(double) * 2 * (size_t)nb * ld);
  const size_t bc = align256(sizeof(int) * 2 * (size_t)nb);
  const size_t need = bm + bw + bc;
  if (ew.used + need > ew.cap) {
    entropy_flush(ctx);
    if (need > ew.cap) {
      // room for a few hundred cuts of this size, within a quarter of the free memory
      size_t freeb = 0, totalb = 0;
      SPP_CHECK(hipStreamSynchronize(ctx.stream));
      if (ew.buf) SPP_CHECK(hipFree(ew.buf));
      ew.buf = nullptr;
      ew.cap = 0;
      SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
      const size_t want = std::max(need, std::min(need * 256, freeb / 4));
      SPP_CHECK(hipMalloc((void**)&ew.buf, want));
      ew.cap = want;
    }
  }
  EntropyWork::Cut c{ew.used, ew.used + bm, ew.used + bm + bw, ld, cut, nb, out, run_max};
  SPP_CHECK(hipMemcpyAsync(ew.buf + c.mat, rho, sizeof(double) * 2 * (size_t)nb * (size_t)rps,
                           hipMemcpyDeviceToDevice, ctx.stream));
  SPP_CHECK(hipMemcpyAsync(ew.buf + c.cnt, cnt, sizeof(int) * 2 * (size_t)nb, hipMemcpyDeviceToDevice, ctx.stream));
  ew.used += need;
  ew.pending.push_back(c);
}

void spectrum_selftest(Context& ctx, int nprob, int ld, const int* n_h, const double* mats_h, double* eig_h) {
  if (nprob <= 0 || ld <= 0) throw HipError("nprob and ld must be positive");
  for (int i = 0; i < nprob; ++i)
    if (n_h[i] < 0 || n_h[i] > ld) throw HipError("block dimension outside [0, ld]");
  hipStream_t st = ctx.stream;
  const size_t mb = sizeof(double) * (size_t)nprob * ld * ld, wb = sizeof(double) * (size_t)nprob * ld;
  double *m = nullptr, *w = nullptr;
  int* c = nullptr;
  SPP_CHECK(hipMalloc((void**)&m, mb));
  try {
    SPP_CHECK(hipMalloc((void**)&w, wb));
    SPP_CHECK(hipMalloc((void**)&c, sizeof(int) * nprob));
    upload(m, mats_h, mb, st);
    upload(c, n_h, sizeof(int) * nprob, st);
    SPP_CHECK(launch_fill<double>(w, (long long)nprob * ld, 0.0, st));
    EntropyTable tab{};
    tab.mat[0] = m;
    tab.w[0] = w;
    tab.cnt[0] = c;
    tab.wstride[0] = ld;
    tab.ld[0] = ld;
    tab.nb[0] = (nprob + 1) / 2;
    tab.count = 1;
    const size_t lds = sizeof(double) * (kSpecWaves + 6 * (size_t)ld);
    if (lds > 64 * 1024) {
      int lmax = 0;
      SPP_CHECK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeSharedMemPerBlockOptin, ctx.device));
      if (lds > (size_t)lmax) throw HipError("ld exceeds the local memory of one workgroup");
      SPP_CHECK(hipFuncSetAttribute((const void*)spectrum_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    }
    hipLaunchKernelGGL(spectrum_kernel, dim3(nprob, 1), dim3(kSpecThreads), lds, st, tab);
    SPP_CHECK(hipGetLastError());
    SPP_CHECK(hipMemcpyAsync(eig_h, w, wb, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(m);
    if (w) (void)hipFree(w);
    if (c) (void)hipFree(c);
    throw;
  }
  (void)hipFree(m);
  (void)hipFree(w);
  (void)hipFree(c);
}

}  // namespace spp
