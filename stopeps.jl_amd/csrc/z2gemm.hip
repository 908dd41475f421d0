// Batched Z2 block-sparse GEMM on CDNA4 matrix cores (see z2gemm.h).
//
// Block tile 128x128, K tile 32 (FP32) / 16 (FP64), 256 threads = 4 waves in a
// 2x2 grid, each wave owning a 64x64 sub-tile. FP32 uses v_mfma_f32_32x32x2f32
// (2x2 tiles per wave), FP64 uses v_mfma_f64_16x16x4f64 (4x4 tiles per wave).
// Operand tiles are gathered from global memory through the decoded row/column
// offsets, with consecutive lanes along each operand's unit-stride index, held
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

// Raw buffer loads: offsets are bytes from the (sample, parity) base; an
// offset at or beyond the descriptor's extent reads as zero.
constexpr int kBufferFlags = 0x00020000;
constexpr unsigned kInvalidK = 0x80000000u;  // K index past the end: a zero operand
__device__ inline float buffer_load(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ inline double buffer_load(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

constexpr int KLDS = 2048;  // K offsets held in LDS per pass over K

template <typename T, int AMF, int BKF>
__global__ __launch_bounds__(kThreads) void z2gemm_kernel(GemmArgs args) {
  constexpr int BK = TileCfg<T>::BK;
  constexpr int AL = BM * BK / kThreads;  // A elements per thread per tile
  constexpr int BL = BN * BK / kThreads;  // B elements per thread per tile
  __shared__ T As[BK * (BM + PAD)];
  __shared__ T Bs[BK * (BN + PAD)];
  // byte offsets of A rows / B columns (invalid: 0, any readable element) and
  // element offsets of C rows / columns (invalid: -1)
  __shared__ unsigned offAm[BM], offBn[BN];
  __shared__ int offCm[BM], offCn[BN];
  __shared__ unsigned offAk[KLDS], offBk[KLDS];  // byte offsets along K (invalid: kInvalidK)
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
  const int cM = (int)group_count(gvM, q);
  const int cN = (int)group_count(gvN, q);
  const int cK = (int)group_count(gvK, q);
  const int m0 = blockIdx.y * BM;
  const int n0 = blockIdx.x * BN;
  if (m0 >= cM || n0 >= cN) return;

  int legs[kMaxLegs];
  if (tid < BM) {
    const int idx = m0 + tid;
    if (idx < cM) {
      decode32(gvM, q, idx, legs);
      offAm[tid] = (unsigned)operand_offset(args.a.g[0], gvM, idx, legs) * (unsigned)sizeof(T);
      offCm[tid] = (int)operand_offset(args.c.g[0], gvM, idx, legs);
    } else {
      offAm[tid] = 0;
      offCm[tid] = -1;
    }
  } else {
    const int t = tid - BM;
    const int idx = n0 + t;
    if (idx < cN) {
      decode32(gvN, q, idx, legs);
      offBn[t] = (unsigned)operand_offset(args.b.g[1], gvN, idx, legs) * (unsigned)sizeof(T);
      offCn[t] = (int)operand_offset(args.c.g[1], gvN, idx, legs);
    } else {
      offBn[t] = 0;
      offCn[t] = -1;
    }
  }

  const __amdgpu_buffer_rsrc_t ra_rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)operand_base<T>(args.a, s, q), 0, (int)args.a_bytes, kBufferFlags);
  const __amdgpu_buffer_rsrc_t rb_rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)operand_base<T>(args.b, s, q), 0, (int)args.b_bytes, kBufferFlags);
  T* Cb = operand_base<T>(args.c, s, q);

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave >> 1) * 64;
  const int wn = (wave & 1) * 64;
  WaveAcc<T> acc;
  acc.zero();

  // Tile loads. Consecutive lanes follow the operand's unit-stride direction
  // (AMF / BKF, chosen on the host): along K (k = tid % BK, rows tid / BK +
  // j AR) or along the row index (row = tid % 128, k = tid / 128 + j KR), so
  // that a wave's load touches few cache lines in either layout.
  constexpr int AR = kThreads / BK;    // row step of a K-fast assignment
  constexpr int KR = kThreads / BM;    // k step of a row-fast assignment (BM == BN)
  const int a_r = AMF ? tid % BM : tid / BK, a_k = AMF ? tid / BM : tid % BK;
  const int b_r = BKF ? tid / BK : tid % BN, b_k = BKF ? tid % BK : tid / BN;
  constexpr int a_dr = AMF ? 0 : AR, a_dk = AMF ? KR : 0;
  constexpr int b_dr = BKF ? AR : 0, b_dk = BKF ? 0 : KR;
  T ra[AL], rb[BL];
  auto load_tile = [&](int kt) {  // kt: first k of the tile within the LDS offset table
#pragma unroll
    for (int j = 0; j < AL; ++j)
      ra[j] = buffer_load(ra_rsrc, offAm[a_r + j * a_dr] + offAk[kt + a_k + j * a_dk], T(0));
#pragma unroll
    for (int j = 0; j < BL; ++j)
      rb[j] = buffer_load(rb_rsrc, offBn[b_r + j * b_dr] + offBk[kt + b_k + j * b_dk], T(0));
  };

  for (int kc = 0; kc < cK; kc += KLDS) {
    const int klen = min(KLDS, cK - kc);
    const int nt = (klen + BK - 1) / BK;
    // offsets of this pass's K range (the previous pass's tiles are done:
    // every K loop iteration ends with a barrier)
    for (int k = tid; k < nt * BK; k += kThreads) {
      if (k < klen) {
        decode32(gvK, q, kc + k, legs);
        offAk[k] = (unsigned)operand_offset(args.a.g[1], gvK, kc + k, legs) * (unsigned)sizeof(T);
        offBk[k] = (unsigned)operand_offset(args.b.g[0], gvK, kc + k, legs) * (unsigned)sizeof(T);
      } else {
        offAk[k] = kInvalidK;
        offBk[k] = kInvalidK;
      }
    }
    __syncthreads();
    load_tile(0);
    for (int t = 0; t < nt; ++t) {
#pragma unroll
      for (int j = 0; j < AL; ++j) As[(a_k + j * a_dk) * (BM + PAD) + a_r + j * a_dr] = ra[j];
#pragma unroll
      for (int j = 0; j < BL; ++j) Bs[(b_k + j * b_dk) * (BN + PAD) + b_r + j * b_dr] = rb[j];
      __syncthreads();
      if (t + 1 < nt) load_tile((t + 1) * BK);
      acc.step(As, Bs, wm, wn, lane);
      __syncthreads();
    }
  }
  // Orders the C offset tables before the epilogue when cK == 0 (no K loop).
  if (cK == 0) __syncthreads();

  const T alpha = (T)args.alpha;
#pragma unroll
  for (int ta = 0; ta < WaveAcc<T>::TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < WaveAcc<T>::TB; ++tb)
#pragma unroll
      for (int v = 0; v < WaveAcc<T>::NV; ++v) {
        int i, j;
        WaveAcc<T>::coord(ta, tb, v, lane, i, j);
        const int om = offCm[wm + i];
        const int on = offCn[wn + j];
        if (om >= 0 && on >= 0) {
          T val = alpha * acc.get(ta, tb, v);
          T* c = Cb + (unsigned)(om + on);
          if (args.accumulate) val += *c;
          *c = val;
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
  const int mode = (args.a_mfast ? 1 : 0) + (args.b_kfast ? 2 : 0);
  if (mode == 0) hipLaunchKernelGGL((z2gemm_kernel<T, 0, 0>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 1) hipLaunchKernelGGL((z2gemm_kernel<T, 1, 0>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 2) hipLaunchKernelGGL((z2gemm_kernel<T, 0, 1>), grid, dim3(kThreads), 0, stream, args);
  else hipLaunchKernelGGL((z2gemm_kernel<T, 1, 1>), grid, dim3(kThreads), 0, stream, args);
  return hipGetLastError();
}

template hipError_t launch_z2gemm<float>(const GemmArgs&, long long, long long, hipStream_t);
template hipError_t launch_z2gemm<double>(const GemmArgs&, long long, long long, hipStream_t);

}  // namespace spp
