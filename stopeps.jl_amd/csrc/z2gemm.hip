// Batched Z2 block-sparse GEMM on CDNA4 matrix cores (see z2gemm.h).
//
// Block tile 128x128, K tile 32 (FP32) / 16 (FP64), 256 threads = 4 waves in a
// 2x2 grid, each wave owning a 64x64 sub-tile. FP32 uses v_mfma_f32_32x32x2f32
// (2x2 tiles per wave), FP64 uses v_mfma_f64_16x16x4f64 (4x4 tiles per wave).
// Operand tiles are gathered from global memory through the decoded row/column
// offsets (consecutive lanes on consecutive k of an A row / n of a B row), held
// in registers one tile ahead of the MFMA work, and staged k-major in padded
// LDS so that each MFMA operand read is a contiguous run of lanes.
#include "z2gemm.h"

namespace spp {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int kThreads = 256;
constexpr int PAD = 1;  // LDS row padding: k-major stores from 16/32 distinct k hit distinct banks

template <typename T>
struct TileCfg;
template <>
struct TileCfg<float> {
  static constexpr int BK = 32;
};
template <>
struct TileCfg<double> {
  static constexpr int BK = 16;
};

template <typename T>
struct WaveAcc;

template <>
struct WaveAcc<float> {
  static constexpr int BK = TileCfg<float>::BK;
  floatx16 acc[2][2];
  __device__ void zero() {
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  }
  // As/Bs are k-major [BK][BM+PAD] / [BK][BN+PAD]; wm/wn are the wave's origin.
  __device__ void step(const float* As, const float* Bs, int wm, int wn, int lane) {
    const int li = lane & 31, lk = lane >> 5;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const float* ar = As + (kk + lk) * (BM + PAD) + wm + li;
      const float* br = Bs + (kk + lk) * (BN + PAD) + wn + li;
      float a0 = ar[0], a1 = ar[32], b0 = br[0], b1 = br[32];
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
    }
  }
  // Row/column (within the wave's 64x64 tile) of accumulator element (ta,tb,v):
  // v_mfma_f32_32x32x2f32 lane l, element v -> row (v%4) + 4(l/32) + 8(v/4), column l%32.
  static __device__ void coord(int ta, int tb, int v, int lane, int& i, int& j) {
    i = ta * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
    j = tb * 32 + (lane & 31);
  }
  __device__ float get(int ta, int tb, int v) const { return acc[ta][tb][v]; }
  static constexpr int TA = 2, TB = 2, NV = 16;
};

template <>
struct WaveAcc<double> {
  static constexpr int BK = TileCfg<double>::BK;
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
        a[t] = As[(kk + lk) * (BM + PAD) + wm + 16 * t + li];
        b[t] = Bs[(kk + lk) * (BN + PAD) + wn + 16 * t + li];
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
  constexpr int BK = TileCfg<T>::BK;
  constexpr int AL = BM * BK / kThreads;  // A elements per thread per tile
  constexpr int BL = BN * BK / kThreads;  // B elements per thread per tile
  __shared__ T As[BK * (BM + PAD)];
  __shared__ T Bs[BK * (BN + PAD)];
  __shared__ long long offAm[BM], offCm[BM], offBn[BN], offCn[BN];
  __shared__ long long offAk[2][BK], offBk[2][BK];
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
  auto decode_k = [&](long long k0, int buf) {
    if (tid < BK) {
      long long idx = k0 + tid;
      if (idx < cK) {
        decode(gvK, q, idx, legs);
        offAk[buf][tid] = operand_offset(args.a.g[1], gvK, idx, legs);
        offBk[buf][tid] = operand_offset(args.b.g[0], gvK, idx, legs);
      } else {
        offAk[buf][tid] = -1;
        offBk[buf][tid] = -1;
      }
    }
  };
  if (cK > 0) decode_k(0, 0);
  // Orders the offset tables before their first use (also when cK == 0: the
  // epilogue reads offCm/offCn and no K-loop barrier runs).
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

  // Tile loads: consecutive lanes read consecutive k of one A row and
  // consecutive n of one B row, so each wave touches few cache lines.
  const int a_k = tid % BK, a_r0 = tid / BK;            // rows a_r0 + j * (kThreads / BK)
  const int b_n = tid % BN, b_k0 = tid / BN;            // k = b_k0 + j * (kThreads / BN)
  constexpr int AR = kThreads / BK, BR = kThreads / BN;
  T ra[AL], rb[BL];
  auto load_tile = [&](int buf) {
    const long long ok = offAk[buf][a_k];
#pragma unroll
    for (int j = 0; j < AL; ++j) {
      const long long om = offAm[a_r0 + j * AR];
      ra[j] = (om >= 0 && ok >= 0) ? Ab[om + ok] : T(0);
    }
    const long long on = offBn[b_n];
#pragma unroll
    for (int j = 0; j < BL; ++j) {
      const long long okb = offBk[buf][b_k0 + j * BR];
      rb[j] = (on >= 0 && okb >= 0) ? Bb[okb + on] : T(0);
    }
  };
  const long long ntiles = (cK + BK - 1) / BK;
  if (ntiles > 0) load_tile(0);
  for (long long t = 0; t < ntiles; ++t) {
#pragma unroll
    for (int j = 0; j < AL; ++j) As[a_k * (BM + PAD) + a_r0 + j * AR] = ra[j];
#pragma unroll
    for (int j = 0; j < BL; ++j) Bs[(b_k0 + j * BR) * (BN + PAD) + b_n] = rb[j];
    if (t + 1 < ntiles) decode_k((t + 1) * BK, (int)((t + 1) & 1));
    __syncthreads();
    if (t + 1 < ntiles) load_tile((int)((t + 1) & 1));
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
