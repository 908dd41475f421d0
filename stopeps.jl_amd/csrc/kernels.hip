// Auxiliary kernels: neutral-entry reductions and updates, double layers,
// batched Jacobi eigensolver. All are memory- or latency-bound helpers around
// the MFMA contraction kernel.
#include "kernels.h"

namespace spp {

// ---------------------------------------------------------------------------
// Neutral-entry kernels. One block group per sample; tuples of parity 0 over
// all legs are enumerated with the shared group decoder.

template <typename T>
__device__ inline T* fv_base(const FullView& v, int s) {
  return v.ptrs ? reinterpret_cast<T*>(v.ptrs[s]) : reinterpret_cast<T*>(v.base) + (long long)s * v.sstride;
}

__device__ inline long long fv_off(const FullView& v, const GroupView& g, const int* legs) {
  long long o = 0;
  for (int j = 0; j < g.n; ++j) o += (long long)legs[j] * v.stride[j];
  return o;
}

template <typename T>
__global__ void dot_kernel(FullView x, FullView y, double* out) {
  const int s = blockIdx.x;
  __shared__ GroupView g;
  __shared__ double red[256];
  if (threadIdx.x == 0) resolve_group(x.g, s, g);
  __syncthreads();
  const long long n = group_count(g, 0);
  const T* xb = fv_base<T>(x, s);
  const T* yb = fv_base<T>(y, s);
  double acc = 0.0;
  int legs[kMaxLegs];
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    decode(g, 0, i, legs);
    acc += (double)xb[fv_off(x, g, legs)] * (double)yb[fv_off(y, g, legs)];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[s] = red[0];
}

// Partial sums of chunk blockIdx.x of sample blockIdx.y (entries i with
// i / 256 = chunk mod kDotChunks) into part[s * kDotChunks + chunk].
template <typename T>
__global__ void dot_chunk_kernel(FullView x, FullView y, double* part) {
  const int s = blockIdx.y, c = blockIdx.x;
  __shared__ GroupView g;
  __shared__ double red[256];
  if (threadIdx.x == 0) resolve_group(x.g, s, g);
  __syncthreads();
  const long long n = group_count(g, 0);
  const T* xb = fv_base<T>(x, s);
  const T* yb = fv_base<T>(y, s);
  double acc = 0.0;
  int legs[kMaxLegs];
  for (long long i = (long long)c * blockDim.x + threadIdx.x; i < n; i += (long long)kDotChunks * blockDim.x) {
    decode(g, 0, i, legs);
    acc += (double)xb[fv_off(x, g, legs)] * (double)yb[fv_off(y, g, legs)];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long long)s * kDotChunks + c] = red[0];
}

__global__ void dot_finish_kernel(const double* part, double* out, int nsamples) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsamples) return;
  double acc = 0.0;
  for (int c = 0; c < kDotChunks; ++c) acc += part[(long long)s * kDotChunks + c];
  out[s] = acc;
}

template <typename T>
hipError_t launch_dot(const FullView& x, const FullView& y, double* out, int nsamples, hipStream_t st,
                      double* partials) {
  if (nsamples <= 0) return hipSuccess;
  if (!partials) {
    hipLaunchKernelGGL(dot_kernel<T>, dim3(nsamples), dim3(256), 0, st, x, y, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(dot_chunk_kernel<T>, dim3(kDotChunks, nsamples), dim3(256), 0, st, x, y, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(dot_finish_kernel, dim3((nsamples + 255) / 256), dim3(256), 0, st, (const double*)partials,
                     out, nsamples);
  return hipGetLastError();
}

template <typename T>
__global__ void scale_kernel(FullView x, const double* f) {
  const int s = blockIdx.y;
  const double fs = f[s];
  if (fs == 0.0 || fs == 1.0) return;
  __shared__ GroupView g;
  if (threadIdx.x == 0) resolve_group(x.g, s, g);
  __syncthreads();
  const long long n = group_count(g, 0);
  T* xb = fv_base<T>(x, s);
  int legs[kMaxLegs];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    decode(g, 0, i, legs);
    long long o = fv_off(x, g, legs);
    xb[o] = (T)((double)xb[o] * fs);
  }
}

template <typename T>
hipError_t launch_scale(const FullView& x, const double* f, int nsamples, hipStream_t st) {
  if (nsamples <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_kernel<T>, dim3(8, nsamples), dim3(256), 0, st, x, f);
  return hipGetLastError();
}

template <typename T>
__global__ void select_copy_kernel(FullView out, FullView a, FullView b, const int* choice, const double* f) {
  const int s = blockIdx.y;
  __shared__ GroupView g;
  if (threadIdx.x == 0) resolve_group(out.g, s, g);
  __syncthreads();
  const long long n = group_count(g, 0);
  const int c = choice ? choice[s] : 0;
  const FullView& src = c == 0 ? a : b;
  const T* sb = fv_base<T>(src, s);
  T* ob = fv_base<T>(out, s);
  const double fs = f ? (f[s] == 0.0 ? 1.0 : f[s]) : 1.0;
  int legs[kMaxLegs];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    decode(g, 0, i, legs);
    ob[fv_off(out, g, legs)] = (T)((double)sb[fv_off(src, g, legs)] * fs);
  }
}

template <typename T>
hipError_t launch_select_copy(const FullView& out, const FullView& a, const FullView& b, const int* choice,
                              const double* f, int nsamples, hipStream_t st) {
  if (nsamples <= 0) return hipSuccess;
  hipLaunchKernelGGL(select_copy_kernel<T>, dim3(8, nsamples), dim3(256), 0, st, out, a, b, choice, f);
  return hipGetLastError();
}

template <typename T>
__global__ void fill_kernel(T* p, long long n, T v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
hipError_t launch_fill(T* p, long long n, T v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

// _safe_scale (boundary_mps.jl:257-260): factor = 1/norm (0 keeps the tensor),
// nonfinite norm -> status 3 at the given site.
__global__ void safe_scale_kernel(const double* nrm2, double* factor, int* status, int* fail_site, int site,
                                  int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  double n = sqrt(nrm2[s]);
  if (!isfinite(n)) {
    if (status[s] == 0) {
      status[s] = 3;
      fail_site[s] = site;
    }
    factor[s] = 0.0;
    return;
  }
  factor[s] = n == 0.0 ? 0.0 : 1.0 / n;
}

hipError_t launch_safe_factor(const double* nrm2, double* factor, int* status, int* fail_site, int site, int nb,
                              hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(safe_scale_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, nrm2, factor, status, fail_site,
                     site, nb);
  return hipGetLastError();
}

// out[s] = sum_i cand_lognorm[2 i + labels[s][i]] (log norms of the chosen projected sites).
__global__ void gather_lognorm_kernel(const double* cand_lognorm, const int* labels, int nsite, double* out, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  double acc = 0.0;
  for (int i = 0; i < nsite; ++i) acc += cand_lognorm[2 * i + labels[(long long)s * nsite + i]];
  out[s] = acc;
}

// Mantissa of a closed boundary whose bonds all have dimension one
// (_close_boundary, boundary_mps.jl:153-160).
template <typename T>
__global__ void close_unit_kernel(T* const* tensors, const long long* sstrides, int L, double* out, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  double v = 1.0;
  for (int x = 0; x < L; ++x) v *= (double)tensors[x][(long long)s * sstrides[x]];
  out[s] = v;
}

hipError_t launch_gather_lognorm(const double* cand_lognorm, const int* labels, int nsite, double* out, int nb,
                                 hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_lognorm_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, cand_lognorm, labels, nsite,
                     out, nb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_close_unit(T* const* tensors, const long long* sstrides, int L, double* out, int nb,
                             hipStream_t st) {
  if (nb <= 0) return hipSuccess;
  hipLaunchKernelGGL(close_unit_kernel<T>, dim3((nb + 255) / 256), dim3(256), 0, st, tensors, sstrides, L, out, nb);
  return hipGetLastError();
}

template <typename TO, typename TI>
__global__ void convert_kernel(TO* out, const TI* in, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = (TO)in[i];
}

template <typename TO, typename TI>
hipError_t launch_convert(TO* out, const TI* in, long long n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((convert_kernel<TO, TI>), dim3((unsigned)blocks), dim3(256), 0, st, out, in, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Double layers. Delta index of a (ket i, bra j) pair in parity-sorted pair
// order: even pairs (ee then oo), then odd pairs (eo then oe), bra fastest.

__device__ inline void split_pair(int idx, int dk, int ek, int db, int eb, int& i, int& j) {
  const int ok = dk - ek, ob = db - eb;
  const int ee = ek * eb, oo = ok * ob, eo = ek * ob;
  if (idx < ee) { i = idx / eb; j = idx % eb; return; }
  idx -= ee;
  if (idx < oo) { i = ek + idx / ob; j = eb + idx % ob; return; }
  idx -= oo;
  if (idx < eo) { i = idx / ob; j = eb + idx % ob; return; }
  idx -= eo;
  i = ek + idx / eb;
  j = idx % eb;
}

__device__ inline int par_of(int i, int e) { return i >= e ? 1 : 0; }

template <typename T>
__global__ void double_layer_kernel(const LayerJob* jobs) {
  const LayerJob& J = jobs[blockIdx.y];
  const int DW = J.dk[0] * J.db[0], DN = J.dk[1] * J.db[1], DE = J.dk[4] * J.db[4], DS = J.dk[5] * J.db[5];
  const long long total = (long long)DW * DN * DE * DS;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  long long r = t;
  const int is = (int)(r % DS); r /= DS;
  const int ie = (int)(r % DE); r /= DE;
  const int in_ = (int)(r % DN); r /= DN;
  const int iw = (int)r;
  // logical (parity-sorted) ket even counts; flipped legs list the odd indices first
  int lek[6];
  for (int k = 0; k < 6; ++k) lek[k] = J.flipk[k] ? J.dk[k] - J.ek[k] : J.ek[k];
  int wk, wb, nk, nb, ek, eb, sk, sb;
  split_pair(iw, J.dk[0], lek[0], J.db[0], J.eb[0], wk, wb);
  split_pair(in_, J.dk[1], lek[1], J.db[1], J.eb[1], nk, nb);
  split_pair(ie, J.dk[4], lek[4], J.db[4], J.eb[4], ek, eb);
  split_pair(is, J.dk[5], lek[5], J.db[5], J.eb[5], sk, sb);
  T* out = reinterpret_cast<T*>(J.out);
  const int pk = par_of(wk, lek[0]) + par_of(nk, lek[1]) + par_of(ek, lek[4]) + par_of(sk, lek[5]);
  const int pb = par_of(wb, J.eb[0]) + par_of(nb, J.eb[1]) + par_of(eb, J.eb[4]) + par_of(sb, J.eb[5]);
  if (((pk + pb) & 1) != 0) {
    out[t] = T(0);
    return;
  }
  // logical -> storage ket index of flipped legs
  auto unflip = [&](int i, int k) { return J.flipk[k] ? (i < lek[k] ? J.ek[k] + i : i - lek[k]) : i; };
  wk = unflip(wk, 0);
  nk = unflip(nk, 1);
  ek = unflip(ek, 4);
  sk = unflip(sk, 5);
  // strides of the (W,N,P,A,E,S) row-major site arrays
  long long skt[6], sbt[6];
  skt[5] = 1;
  sbt[5] = 1;
  for (int k = 4; k >= 0; --k) {
    skt[k] = skt[k + 1] * J.dk[k + 1];
    sbt[k] = sbt[k + 1] * J.db[k + 1];
  }
  const long long bk = wk * skt[0] + nk * skt[1] + ek * skt[4] + sk * skt[5];
  const long long bb = wb * sbt[0] + nb * sbt[1] + eb * sbt[4] + sb * sbt[5];
  const int dp = J.dk[2], da_k = J.dk[3], da_b = J.db[3];
  double acc = 0.0;
  for (int p = 0; p < dp; ++p) {
    if (J.use_aux) {
      double kv = 0.0, bv = 0.0;
      for (int a = 0; a < da_k; ++a) {
        double v = 0.0;
        if (J.use_op) {
          for (int pp = 0; pp < dp; ++pp) v += J.op[p * 2 + pp] * J.ket[bk + pp * skt[2] + a * skt[3]];
        } else {
          v = J.ket[bk + p * skt[2] + a * skt[3]];
        }
        kv += J.aux[a] * v;
      }
      for (int a = 0; a < da_b; ++a) bv += J.aux[a] * J.bra[bb + p * sbt[2] + a * sbt[3]];
      acc += kv * bv;
    } else {
      for (int a = 0; a < da_k; ++a) {
        double v = 0.0;
        if (J.use_op) {
          for (int pp = 0; pp < dp; ++pp) v += J.op[p * 2 + pp] * J.ket[bk + pp * skt[2] + a * skt[3]];
        } else {
          v = J.ket[bk + p * skt[2] + a * skt[3]];
        }
        acc += v * J.bra[bb + p * sbt[2] + a * sbt[3]];
      }
    }
  }
  out[t] = (T)acc;
}

template <typename T>
hipError_t launch_double_layers(const LayerJob* jobs_dev, int njobs, long long maxout, hipStream_t st) {
  if (njobs <= 0 || maxout <= 0) return hipSuccess;
  dim3 grid((unsigned)((maxout + 255) / 256), (unsigned)njobs);
  hipLaunchKernelGGL(double_layer_kernel<T>, grid, dim3(256), 0, st, jobs_dev);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cyclic Jacobi eigensolver with round-robin parallel ordering, n <= 64, FP64.
//
// Each round rotates m/2 disjoint index pairs at once. With G the block
// diagonal rotation of the round, A <- G A G^T changes a 2x2 block
// A[{p1,q1},{p2,q2}] only through the rotations of its own two pairs, so one
// pass over all (pair, pair) blocks applies the whole round with a single
// barrier. A pair is rotated when |a_pq| exceeds both eps sqrt(|a_pp a_qq|)
// (relative accuracy of the small eigenvalues) and 1e-17 ||A||_F (below that
// the rotation changes no eigenvalue in FP64); the sweeps stop after one
// without rotations.

constexpr int JMAX = 64;
constexpr int JLD = JMAX + 1;
constexpr int JT = 256;                            // threads
constexpr int JB = (JMAX / 2) * (JMAX / 2) / JT;   // 2x2 blocks per thread
constexpr int JV = JMAX * (JMAX / 2) / JT;         // V row updates per thread

__global__ __launch_bounds__(JT) void jacobi_kernel(const double* a, long long astride, int lda, double* w,
                                                    long long wstride, double* v, long long vstride, int ldv,
                                                    const int* n_of, const int* pmask) {
  const int prob = blockIdx.x;
  if (pmask && !pmask[prob]) return;
  const int n = n_of[prob];
  if (n <= 0) return;
  __shared__ double A[JMAX * JLD];
  __shared__ double V[JMAX * JLD];
  __shared__ double cs[JMAX / 2], sn[JMAX / 2];
  __shared__ int pp_[JMAX / 2], qq_[JMAX / 2];
  __shared__ int order[JMAX + 1];
  __shared__ double red[JT];
  __shared__ int rotated;
  const int tid = threadIdx.x;
  const int m = (n & 1) ? n + 1 : n;  // padded size (dummy index n when odd)
  const int half = m / 2;
  const double* ap = a + (long long)prob * astride;
  for (int i = tid; i < m * m; i += JT) {
    const int r = i / m, c = i - (i / m) * m;
    double val = 0.0;
    if (r < n && c < n) val = 0.5 * (ap[(long long)r * lda + c] + ap[(long long)c * lda + r]);
    A[r * JLD + c] = val;
    V[r * JLD + c] = (r == c) ? 1.0 : 0.0;
  }
  if (tid <= m) order[tid] = tid;
  // this thread's (pair, pair) blocks and (row, pair) V updates
  int bk1[JB], bk2[JB], vj[JV], vk[JV];
#pragma unroll
  for (int i = 0; i < JB; ++i) {
    const int b = tid + JT * i;
    bk1[i] = half > 0 ? b / half : 0;
    bk2[i] = half > 0 ? b - bk1[i] * half : 0;
  }
#pragma unroll
  for (int i = 0; i < JV; ++i) {
    const int e = tid + JT * i;
    vk[i] = half > 0 ? e / m : 0;
    vj[i] = e - vk[i] * m;
  }
  __syncthreads();
  double loc = 0.0;
  for (int i = tid; i < m * m; i += JT) {
    const double x = A[(i / m) * JLD + (i - (i / m) * m)];
    loc += x * x;
  }
  red[tid] = loc;
  __syncthreads();
  for (int s = JT / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double floor_abs = 1e-17 * sqrt(red[0]);
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < m - 1; ++round) {
      if (tid < half) {
        // round-robin pairing: position 0 fixed, the others rotate
        const int i1 = tid, i2 = m - 1 - tid;
        int p = 0;
        if (i1 != 0) {
          p = i1 - 1 + round;
          if (p >= m - 1) p -= m - 1;
          p += 1;
        }
        int q = i2 - 1 + round;
        if (q >= m - 1) q -= m - 1;
        q += 1;
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        pp_[tid] = p;
        qq_[tid] = q;
        const double apq = A[p * JLD + q];
        const double app = A[p * JLD + p], aqq = A[q * JLD + q];
        double c = 1.0, sv = 0.0;
        if (fabs(apq) > 2.220446049250313e-16 * sqrt(fabs(app * aqq)) && fabs(apq) > floor_abs) {
          const double theta = (aqq - app) / (2.0 * apq);
          const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(t * t + 1.0);
          sv = t * c;
          rotated = 1;
        }
        cs[tid] = c;
        sn[tid] = sv;
      }
      __syncthreads();
      // A <- G A G^T block by block; G rows: p' = c p - s q, q' = s p + c q
#pragma unroll
      for (int i = 0; i < JB; ++i) {
        const int k1 = bk1[i], k2 = bk2[i];
        if (k1 >= half) continue;
        const double s1 = sn[k1], s2 = sn[k2];
        if (s1 == 0.0 && s2 == 0.0) continue;
        const double c1 = cs[k1], c2 = cs[k2];
        const int p1 = pp_[k1], q1 = qq_[k1], p2 = pp_[k2], q2 = qq_[k2];
        const double a00 = A[p1 * JLD + p2], a01 = A[p1 * JLD + q2];
        const double a10 = A[q1 * JLD + p2], a11 = A[q1 * JLD + q2];
        // left: rows (p1, q1)
        const double l00 = c1 * a00 - s1 * a10, l01 = c1 * a01 - s1 * a11;
        const double l10 = s1 * a00 + c1 * a10, l11 = s1 * a01 + c1 * a11;
        // right: columns (p2, q2)
        double r00 = c2 * l00 - s2 * l01, r01 = s2 * l00 + c2 * l01;
        double r10 = c2 * l10 - s2 * l11, r11 = s2 * l10 + c2 * l11;
        if (k1 == k2) {  // the rotated pair's own block: exact zero off the diagonal
          r01 = 0.0;
          r10 = 0.0;
        }
        A[p1 * JLD + p2] = r00;
        A[p1 * JLD + q2] = r01;
        A[q1 * JLD + p2] = r10;
        A[q1 * JLD + q2] = r11;
      }
#pragma unroll
      for (int i = 0; i < JV; ++i) {
        const int k = vk[i], j = vj[i];
        if (k >= half) continue;
        const double sv = sn[k];
        if (sv == 0.0) continue;
        const double c = cs[k];
        const int p = pp_[k], q = qq_[k];
        const double vp = V[j * JLD + p], vq = V[j * JLD + q];
        V[j * JLD + p] = c * vp - sv * vq;
        V[j * JLD + q] = sv * vp + c * vq;
      }
      __syncthreads();
    }
    if (!rotated) break;
  }
  // sort eigenvalues (descending) over the n real indices
  if (tid == 0) {
    int cnt = 0;
    for (int i = 0; i < m; ++i)
      if (i < n) order[cnt++] = i;
    for (int i = 1; i < n; ++i) {
      const int key = order[i];
      const double kv = A[key * JLD + key];
      int j = i - 1;
      while (j >= 0 && A[order[j] * JLD + order[j]] < kv) {
        order[j + 1] = order[j];
        --j;
      }
      order[j + 1] = key;
    }
  }
  __syncthreads();
  double* wp = w + (long long)prob * wstride;
  double* vp = v + (long long)prob * vstride;
  for (int i = tid; i < n; i += JT) wp[i] = A[order[i] * JLD + order[i]];
  for (int i = tid; i < n * n; i += JT) {
    const int r = i / n, c = i - (i / n) * n;
    vp[(long long)r * ldv + c] = V[r * JLD + order[c]];
  }
}

hipError_t launch_jacobi(const double* a, long long astride, int lda, double* w, long long wstride, double* v,
                         long long vstride, int ldv, const int* n_of, const int* pmask, int nprob,
                         hipStream_t st) {
  if (nprob <= 0) return hipSuccess;
  hipLaunchKernelGGL(jacobi_kernel, dim3(nprob), dim3(JT), 0, st, a, astride, lda, w, wstride, v, vstride, ldv,
                     n_of, pmask);
  return hipGetLastError();
}

template hipError_t launch_dot<float>(const FullView&, const FullView&, double*, int, hipStream_t, double*);
template hipError_t launch_dot<double>(const FullView&, const FullView&, double*, int, hipStream_t, double*);
template hipError_t launch_scale<float>(const FullView&, const double*, int, hipStream_t);
template hipError_t launch_scale<double>(const FullView&, const double*, int, hipStream_t);
template hipError_t launch_select_copy<float>(const FullView&, const FullView&, const FullView&, const int*,
                                              const double*, int, hipStream_t);
template hipError_t launch_select_copy<double>(const FullView&, const FullView&, const FullView&, const int*,
                                               const double*, int, hipStream_t);
template hipError_t launch_fill<float>(float*, long long, float, hipStream_t);
template hipError_t launch_fill<double>(double*, long long, double, hipStream_t);
template hipError_t launch_fill<int>(int*, long long, int, hipStream_t);
template hipError_t launch_convert<float, double>(float*, const double*, long long, hipStream_t);
template hipError_t launch_convert<double, float>(double*, const float*, long long, hipStream_t);
template hipError_t launch_convert<double, double>(double*, const double*, long long, hipStream_t);
template hipError_t launch_convert<float, float>(float*, const float*, long long, hipStream_t);
template hipError_t launch_close_unit<float>(float* const*, const long long*, int, double*, int, hipStream_t);
template hipError_t launch_close_unit<double>(double* const*, const long long*, int, double*, int, hipStream_t);
template hipError_t launch_double_layers<float>(const LayerJob*, int, long long, hipStream_t);
template hipError_t launch_double_layers<double>(const LayerJob*, int, long long, hipStream_t);

}  // namespace spp
