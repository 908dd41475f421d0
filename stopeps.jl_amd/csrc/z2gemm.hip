// Batched Z2 block-sparse GEMM on CDNA4 matrix cores (see z2gemm.h).
//
// Block tile 128x128, K tile 16, 256 threads = 4 waves in a 2x2 grid, each
// wave owning a 64x64 sub-tile. FP32 uses v_mfma_f32_32x32x2f32 (2x2 tiles per
// wave), FP64 uses v_mfma_f64_16x16x4f64 (4x4 tiles per wave). Operand tiles
// are gathered from global memory through the decoded row/column offsets and
// staged k-major in LDS so that each MFMA operand read is a contiguous run of
// lanes.
#include "z2gemm.h"

namespace spp {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int kThreads = 256;

template <typename T>
struct WaveAcc;

template <>
struct WaveAcc<float> {
  floatx16 acc[2][2];
  __device__ void zero() {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  }
  // As/Bs are k-major [BK][BM] / [BK][BN]; wm/wn are the wave's origin.
  __device__ void step(const float* As, const float* Bs, int wm, int wn, int lane) {
    const int li = lane & 31, lk = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      float a0 = As[(kk + lk) * BM + wm + li];
      float a1 = As[(kk + lk) * BM + wm + 32 + li];
      float b0 = Bs[(kk + lk) * BN + wn + li];
      float b1 = Bs[(kk + lk) * BN + wn + 32 + li];
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  // Row/column (within the wave's 64x64 tile) of accumulator element (ta,tb,v).
  static __device__ void coord(int ta, int tb, int v, int lane, int& i, int& j) {
    i = ta * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
    j = tb * 32 + (lane & 31);
  }
  __device__ float get(int ta, int tb, int v) const { return acc[ta][tb][v]; }
  static constexpr int TA = 2, TB = 2, NV = 16;
};

template <>
struct WaveAcc<double> {
  doublex4 acc[4][4];
  __device__ void zero() {
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
  }
  __device__ void step(const double* As, const double* Bs, int wm, int wn, int lane) {
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = As[(kk + lk) * BM + wm + 16 * t + li];
        b[t] = Bs[(kk + lk) * BN + wn + 16 * t + li];
      }
#pragma unroll
      for (int ta = 0; ta < 4; ++ta)
#pragma unroll
        for (int tb = 0; tb < 4; ++tb)
          acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
    }
  }
  // v_mfma_f64_16x16x4f64 output layout on gfx950 (checked with a one-wave
  // probe): accumulator v of lane l holds row (l / 16) + 4 v, column l % 16.
  static __device__ void coord(int ta, int tb, int v, int lane, int& i, int& j) {
    i = ta * 16 + (lane >> 4) + 4 * v;
    j = tb * 16 + (lane & 15);
  }
  __device__ double get(int ta, int tb, int v) const { return acc[ta][tb][v]; }
  static constexpr int TA = 4, TB = 4, NV = 4;
};

template <typename T>
__device__ inline T* operand_base(const Operand& op, int s, int q) {
  T* p = op.ptrs ? reinterpret_cast<T*>(op.ptrs[s]) : reinterpret_cast<T*>(op.base) + (long long)s * op.sstride;
  return p + (long long)q * op.pstride;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void z2gemm_kernel(GemmArgs args) {
  __shared__ T As[BK * BM];
  __shared__ T Bs[BK * BN];
  __shared__ long long offAm[BM], offCm[BM], offBn[BN], offCn[BN];
  __shared__ long long offAk[BK], offBk[BK];
  __shared__ GroupView gvM, gvN, gvK;

  const int tid = threadIdx.x;
  const int prob = blockIdx.z;
  const int sp = prob >> 1;
  const int q = prob & 1;
  if (args.pmask && !args.pmask[prob]) return;
  const int s = args.sample_map ? args.sample_map[sp] : sp;

  if (tid == 0) resolve_group(args.gm, s, gvM);
  if (tid == 1) resolve_group(args.gn, s, gvN);
  if (tid == 2) resolve_group(args.gk, s, gvK);
  __syncthreads();
  const long long cM = group_count(gvM, q);
  const long long cN = group_count(gvN, q);
  const long long cK = group_count(gvK, q);
  const long long m0 = (long long)blockIdx.y * BM;
  const long long n0 = (long long)blockIdx.x * BN;
  if (m0 >= cM || n0 >= cN) return;

  int legs[kMaxLegs];
  if (tid < BM) {
    long long idx = m0 + tid;
    if (idx < cM) {
      decode(gvM, q, idx, legs);
      offAm[tid] = operand_offset(args.a.g[0], gvM, idx, legs);
      offCm[tid] = operand_offset(args.c.g[0], gvM, idx, legs);
    } else {
      offAm[tid] = -1;
      offCm[tid] = -1;
    }
  } else {
    int t = tid - BM;
    long long idx = n0 + t;
    if (idx < cN) {
      decode(gvN, q, idx, legs);
      offBn[t] = operand_offset(args.b.g[1], gvN, idx, legs);
      offCn[t] = operand_offset(args.c.g[1], gvN, idx, legs);
    } else {
      offBn[t] = -1;
      offCn[t] = -1;
    }
  }

  // The offset tables above are read by every thread in the epilogue. When the
  // K block of this parity is empty (cK == 0) the K loop below, whose barriers
  // would otherwise order those writes, does not run: barrier here.
  __syncthreads();

  const T* Ab = operand_base<T>(args.a, s, q);
  const T* Bb = operand_base<T>(args.b, s, q);
  T* Cb = operand_base<T>(args.c, s, q);

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 64;
  const int wn = (wave & 1) * 64;
  WaveAcc<T> acc;
  acc.zero();

  for (long long k0 = 0; k0 < cK; k0 += BK) {
    if (tid < BK) {
      long long idx = k0 + tid;
      if (idx < cK) {
        decode(gvK, q, idx, legs);
        offAk[tid] = operand_offset(args.a.g[1], gvK, idx, legs);
        offBk[tid] = operand_offset(args.b.g[0], gvK, idx, legs);
      } else {
        offAk[tid] = -1;
        offBk[tid] = -1;
      }
    }
    __syncthreads();
    {
      const int r = tid >> 1;
      const int kb = (tid & 1) * 8;
      const long long om = offAm[r];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const long long ok = offAk[kb + t];
        T v = (om >= 0 && ok >= 0) ? Ab[om + ok] : T(0);
        As[(kb + t) * BM + r] = v;
      }
    }
    {
      const int k = tid >> 4;
      const int nb = (tid & 15) * 8;
      const long long ok = offBk[k];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const long long on = offBn[nb + t];
        T v = (on >= 0 && ok >= 0) ? Bb[ok + on] : T(0);
        Bs[k * BN + nb + t] = v;
      }
    }
    __syncthreads();
    acc.step(As, Bs, wm, wn, lane);
    __syncthreads();
  }

  const T alpha = (T)args.alpha;
#pragma unroll
  for (int ta = 0; ta < WaveAcc<T>::TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < WaveAcc<T>::TB; ++tb)
#pragma unroll
      for (int v = 0; v < WaveAcc<T>::NV; ++v) {
        int i, j;
        WaveAcc<T>::coord(ta, tb, v, lane, i, j);
        const long long om = offCm[wm + i];
        const long long on = offCn[wn + j];
        if (om >= 0 && on >= 0) {
          T val = alpha * acc.get(ta, tb, v);
          if (args.accumulate) val += Cb[om + on];
          Cb[om + on] = val;
        }
      }
}

template <typename T>
hipError_t launch_z2gemm(const GemmArgs& args, long long mmax, long long nmax, hipStream_t stream) {
  if (mmax <= 0 || nmax <= 0 || args.nsamples <= 0) return hipSuccess;
  long long gy = (mmax + BM - 1) / BM;
  long long gx = (nmax + BN - 1) / BN;
  long long gz = 2LL * args.nsamples;
  if (gy > 65535 || gz > 65535 || gx > 2147483647LL) return hipErrorInvalidConfiguration;
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)gz);
  hipLaunchKernelGGL(z2gemm_kernel<T>, grid, dim3(kThreads), 0, stream, args);
  return hipGetLastError();
}

template hipError_t launch_z2gemm<float>(const GemmArgs&, long long, long long, hipStream_t);
template hipError_t launch_z2gemm<double>(const GemmArgs&, long long, long long, hipStream_t);

}  // namespace spp
