// Batched Z2 block-sparse GEMM on CDNA4 matrix cores (see z2gemm.h).
//
// 256 threads = 4 waves per block; the block tile BM x BN is split into a
// WGM x (4 / WGM) grid of wave tiles. FP32 uses v_mfma_f32_32x32x2f32, FP64
// v_mfma_f64_16x16x4f64, TA x TB MFMA tiles per wave. K advances in tiles of
// BK. Operand tiles are gathered from global memory through decoded
// row/column offsets (the K offsets of up to 1024 indices decoded once per
// block into LDS), with consecutive lanes along each operand's unit-stride
// index, held in registers one tile ahead of the MFMA work, and staged k-major
// in padded LDS so that each MFMA operand read is a contiguous run of lanes.
//
// Tile shapes (chosen per launch by the padded area of the problem; see
// Shapes below for the K tile and occupancy target of each):
//   128 x 128, 2 x 2 waves           general
//   FP32 256 x 96 / FP64 128 x 96    N a few multiples of 96 (e.g. 648 = 6.75 x 96)
//   256 x 32, 4 x 1 waves            N <= 32
#include "z2gemm.h"

namespace spp {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int PAD = 1;  // LDS row padding: k-major stores from distinct k hit distinct banks

template <typename T>
struct KTile;
template <>
struct KTile<float> {
  static constexpr int MT = 32;  // MFMA tile edge
};
template <>
struct KTile<double> {
  static constexpr int MT = 16;
};

template <typename T, int TA, int TB, int BK>
struct WaveAcc;

template <int TA, int TB, int BK>
struct WaveAcc<float, TA, TB, BK> {
  static constexpr int NV = 16;
  floatx16 acc[TA][TB];
  __device__ void zero() {
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
      for (int b = 0; b < TB; ++b)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
  }
  // As/Bs are k-major [BK][LDA] / [BK][LDB]; wm/wn are the wave's origin.
  // The operands of k step kk + 2 are read from LDS before the MFMAs of step
  // kk are issued, so the LDS latency overlaps the matrix work.
  template <int LDA, int LDB>
  __device__ void step(const float* As, const float* Bs, int wm, int wn, int lane) {
    const int li = lane & 31, lk = lane >> 5;
    float a[2][TA], b[2][TB];
#pragma unroll
    for (int t = 0; t < TA; ++t) a[0][t] = As[lk * LDA + wm + 32 * t + li];
#pragma unroll
    for (int t = 0; t < TB; ++t) b[0][t] = Bs[lk * LDB + wn + 32 * t + li];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      const int c = (kk >> 1) & 1;
      if (kk + 2 < BK) {
#pragma unroll
        for (int t = 0; t < TA; ++t) a[c ^ 1][t] = As[(kk + 2 + lk) * LDA + wm + 32 * t + li];
#pragma unroll
        for (int t = 0; t < TB; ++t) b[c ^ 1][t] = Bs[(kk + 2 + lk) * LDB + wn + 32 * t + li];
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this step's MFMAs
#pragma unroll
      for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
          acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[c][ta], b[c][tb], acc[ta][tb], 0, 0, 0);
    }
  }
  // v_mfma_f32_32x32x2f32 lane l, element v -> row (v%4) + 4(l/32) + 8(v/4), column l%32.
  static __device__ void coord(int ta, int tb, int v, int lane, int& i, int& j) {
    i = ta * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
    j = tb * 32 + (lane & 31);
  }
  __device__ float get(int ta, int tb, int v) const { return acc[ta][tb][v]; }
};

template <int TA, int TB, int BK>
struct WaveAcc<double, TA, TB, BK> {
  static constexpr int NV = 4;
  doublex4 acc[TA][TB];
  __device__ void zero() {
#pragma unroll
    for (int a = 0; a < TA; ++a)
#pragma unroll
      for (int b = 0; b < TB; ++b)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[a][b][v] = 0.0;
  }
  template <int LDA, int LDB>
  __device__ void step(const double* As, const double* Bs, int wm, int wn, int lane) {
    const int li = lane & 15, lk = lane >> 4;
    double a[2][TA], b[2][TB];
#pragma unroll
    for (int t = 0; t < TA; ++t) a[0][t] = As[lk * LDA + wm + 16 * t + li];
#pragma unroll
    for (int t = 0; t < TB; ++t) b[0][t] = Bs[lk * LDB + wn + 16 * t + li];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int c = (kk >> 2) & 1;
      if (kk + 4 < BK) {
#pragma unroll
        for (int t = 0; t < TA; ++t) a[c ^ 1][t] = As[(kk + 4 + lk) * LDA + wm + 16 * t + li];
#pragma unroll
        for (int t = 0; t < TB; ++t) b[c ^ 1][t] = Bs[(kk + 4 + lk) * LDB + wn + 16 * t + li];
      }
      __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of this step's MFMAs
#pragma unroll
      for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
          acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c][ta], b[c][tb], acc[ta][tb], 0, 0, 0);
    }
  }
  // v_mfma_f64_16x16x4f64 output layout on gfx950 (checked with a one-wave
  // probe): accumulator v of lane l holds row (l / 16) + 4 v, column l % 16.
  static __device__ void coord(int ta, int tb, int v, int lane, int& i, int& j) {
    i = ta * 16 + (lane >> 4) + 4 * v;
    j = tb * 16 + (lane & 15);
  }
  __device__ double get(int ta, int tb, int v) const { return acc[ta][tb][v]; }
};

template <typename T>
__device__ inline T* operand_base(const Operand& op, int s, int q) {
  T* p = op.ptrs ? reinterpret_cast<T*>(op.ptrs[s]) : reinterpret_cast<T*>(op.base) + (long long)s * op.sstride;
  return p + (long long)q * op.pstride + op.offset;
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

constexpr int KLDS = 1024;  // K offsets held in LDS per pass over K

// Thread -> (row, k) assignment of an R x BK operand tile, j = 0 .. R*BK/256-1.
// Row-fast: runs of RUN consecutive rows per k (RUN = min(R, 256)); K-fast:
// consecutive k of one row. Both are affine in j with compile-time steps.
template <int R, int BK, int ROWFAST>
struct TileMap {
  static constexpr int RUN = R < kThreads ? (R % 32 == 0 && R % 64 != 0 ? 32 : R) : kThreads;
  static constexpr int GROUPS = R / RUN;        // row runs per k (row-fast)
  static constexpr int KSTEP = kThreads / RUN;  // k rows covered by one j group (row-fast)
  static constexpr int N = R * BK / kThreads;   // elements per thread
  static constexpr int RSTEP = kThreads / BK;   // rows covered per j (K-fast)
  static_assert(kThreads % RUN == 0 && R % RUN == 0 && BK % KSTEP == 0 && R % RSTEP == 0, "tile map");
  __device__ static int row(int tid, int j) {
    return ROWFAST ? tid % RUN + RUN * (j % GROUPS) : tid / BK + RSTEP * j;
  }
  __device__ static int k(int tid, int j) { return ROWFAST ? tid / RUN + KSTEP * (j / GROUPS) : tid % BK; }
};

// Tile order. Workgroups are dealt to the 8 XCDs round-robin by linear id,
// and each XCD has its own L2. The linear id is remapped so that each XCD
// receives a contiguous range of tiles in the order [sample][M tile][parity]
// [N tile]: the N tiles of one A row panel run together on one XCD, and so do
// the panels of both parity problems of a sample, whose rows interleave in
// memory (a parity-sorted last leg puts both parities in each cache line).
// The map is a bijection of [0, total).
constexpr int kXcds = 8;
__device__ inline void tile_order(int& bn, int& bm, int& prob) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned total = gx * gy * gridDim.z;
  const unsigned bid = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned per = (total + kXcds - 1) / kXcds;
  const unsigned tall = total % kXcds == 0 ? kXcds : total % kXcds;  // XCDs that receive `per` tiles
  const unsigned xcd = bid % kXcds, local = bid / kXcds;
  unsigned t = xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
  bn = (int)(t % gx);
  t /= gx;
  const int q = (int)(t & 1);
  t >>= 1;
  bm = (int)(t % gy);
  prob = 2 * (int)(t / gy) + q;
}

// MINW: waves per SIMD the register allocation must allow (occupancy target);
// DB: double-buffered LDS tiles.
template <typename T, int BM, int BN, int WGM, int BK, int MINW, int DB, int AMF, int BKF>
__global__ __launch_bounds__(kThreads, MINW) void z2gemm_kernel(GemmArgs args) {
  constexpr int MT = KTile<T>::MT;
  constexpr int WGN = 4 / WGM;
  constexpr int WTM = BM / WGM, WTN = BN / WGN;  // wave tile
  constexpr int TA = WTM / MT, TB = WTN / MT;
  static_assert(TA * MT == WTM && TB * MT == WTN, "wave tile must be a multiple of the MFMA tile");
  constexpr int LDA = BM + PAD, LDB = BN + PAD;
  using MA = TileMap<BM, BK, AMF>;
  using MB = TileMap<BN, BK, 1 - BKF>;
  constexpr int NBUF = DB ? 2 : 1;  // LDS tile buffers
  __shared__ T As[NBUF][BK * LDA];
  __shared__ T Bs[NBUF][BK * LDB];
  // byte offsets of A rows / B columns (invalid: 0, any readable element) and
  // element offsets of C rows / columns (invalid: -1)
  __shared__ unsigned offAm[BM], offBn[BN];
  __shared__ int offCm[BM], offCn[BN];
  __shared__ unsigned offAk[KLDS], offBk[KLDS];  // byte offsets along K (invalid: kInvalidK)
  __shared__ GroupView gvM, gvN, gvK;

  const int tid = threadIdx.x;
  int bx, by, prob;
  tile_order(bx, by, prob);
  // wave-uniform: keeps the operand bases and buffer descriptors in SGPRs
  bx = __builtin_amdgcn_readfirstlane(bx);
  by = __builtin_amdgcn_readfirstlane(by);
  prob = __builtin_amdgcn_readfirstlane(prob);
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
  const int m0 = by * BM;
  const int n0 = bx * BN;
  if (m0 >= cM || n0 >= cN) return;

  int legs[kMaxLegs];
  for (int r = tid; r < BM + BN; r += kThreads) {
    if (r < BM) {
      const int idx = m0 + r;
      if (idx < cM) {
        decode32(gvM, q, idx, legs);
        offAm[r] = (unsigned)operand_offset(args.a.g[0], gvM, idx, legs) * (unsigned)sizeof(T);
        offCm[r] = (int)operand_offset(args.c.g[0], gvM, idx, legs);
      } else {
        offAm[r] = 0;
        offCm[r] = -1;
      }
    } else {
      const int t = r - BM;
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
  }

  const __amdgpu_buffer_rsrc_t ra_rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)operand_base<T>(args.a, s, q), 0, (int)args.a_bytes, kBufferFlags);
  const __amdgpu_buffer_rsrc_t rb_rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)operand_base<T>(args.b, s, q), 0, (int)args.b_bytes, kBufferFlags);
  T* Cb = operand_base<T>(args.c, s, q);

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave / WGN) * WTM;
  const int wn = (wave % WGN) * WTN;
  WaveAcc<T, TA, TB, BK> acc;
  acc.zero();

  T ra[MA::N], rb[MB::N];
  auto load_tile = [&](int kt) {  // kt: first k of the tile within the LDS offset table
#pragma unroll
    for (int j = 0; j < MA::N; ++j)
      ra[j] = buffer_load(ra_rsrc, offAm[MA::row(tid, j)] + offAk[kt + MA::k(tid, j)], T(0));
#pragma unroll
    for (int j = 0; j < MB::N; ++j)
      rb[j] = buffer_load(rb_rsrc, offBn[MB::row(tid, j)] + offBk[kt + MB::k(tid, j)], T(0));
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
    auto store_tile = [&](int b) {
#pragma unroll
      for (int j = 0; j < MA::N; ++j) As[b][MA::k(tid, j) * LDA + MA::row(tid, j)] = ra[j];
#pragma unroll
      for (int j = 0; j < MB::N; ++j) Bs[b][MB::k(tid, j) * LDB + MB::row(tid, j)] = rb[j];
    };
    if constexpr (DB) {
      // two LDS buffers: tile t + 1 is stored while tile t is multiplied, one
      // barrier per tile (buffer (t + 1) & 1 was last read before the
      // previous barrier)
      store_tile(0);
      __syncthreads();
      for (int t = 0; t < nt; ++t) {
        if (t + 1 < nt) load_tile((t + 1) * BK);
        acc.template step<LDA, LDB>(As[t & 1], Bs[t & 1], wm, wn, lane);
        if (t + 1 < nt) store_tile((t + 1) & 1);
        __syncthreads();
      }
    } else {
      for (int t = 0; t < nt; ++t) {
        store_tile(0);
        __syncthreads();
        if (t + 1 < nt) load_tile((t + 1) * BK);
        acc.template step<LDA, LDB>(As[0], Bs[0], wm, wn, lane);
        __syncthreads();
      }
    }
  }
  // Orders the C offset tables before the epilogue when cK == 0 (no K loop).
  if (cK == 0) __syncthreads();

  const T alpha = (T)args.alpha;
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb)
#pragma unroll
      for (int v = 0; v < WaveAcc<T, TA, TB, BK>::NV; ++v) {
        int i, j;
        WaveAcc<T, TA, TB, BK>::coord(ta, tb, v, lane, i, j);
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

// ---------------------------------------------------------------------------
// BF16 split products (the "bf16x6" / "bf16x3" precision modes).
//
// FP32 operands are split when staged into LDS: hi = bf16(x), mid =
// bf16(x - hi), lo = bf16(x - hi - mid) (round to nearest even, two values per
// conversion), so that x = hi + mid + lo to 2^-24 relative (finite operands
// below the BF16 overflow threshold, about 3.39e38). The product sum_k a_k b_k
// is formed on the BF16 matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulation) from the
// piece products whose orders add up to at most NP - 1: NP = 3 pieces take
// hh, hm, mh, hl, lh, mm (six products; the omitted ones are below 2^-24
// relative, so the result matches an FP32 product with FP32 accumulation),
// NP = 2 pieces take hh, hm, mh (three products, about 2^-16 relative).
// The LDS tiles are row-major with the K index contiguous, one array per piece,
// so each MFMA operand (eight consecutive K values of one row) is one 16-byte read.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two FP32 values -> two BF16 (round to nearest even, v_cvt_pk_bf16_f32),
// element 0 in the low half.
__device__ inline unsigned bf16_pair(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

constexpr int SBK = 16;       // K tile of the split kernel: one MFMA step
constexpr int SLD = SBK + 8;  // LDS row stride in BF16 elements (48 bytes)

// Thread -> (row, k pair) assignment of an R x 16 operand tile: 8 R pairs of
// consecutive k, R / 32 per thread. Row-fast: consecutive lanes along rows;
// K-fast: the 8 pairs of a row on consecutive lanes.
template <int R, int ROWFAST>
struct PairMap {
  static constexpr int N = R / 32;
  static_assert(N * 32 == R, "pair map rows");
  __device__ static int row(int tid, int j) {
    const int t = tid + kThreads * j;
    return ROWFAST ? t % R : t >> 3;
  }
  __device__ static int pair(int tid, int j) {
    const int t = tid + kThreads * j;
    return ROWFAST ? t / R : t & 7;
  }
};

template <int BM, int BN, int WGM, int MINW, int AMF, int BKF, int NP>
__global__ __launch_bounds__(kThreads, MINW) void z2gemm_split_kernel(GemmArgs args) {
  constexpr int BK = SBK;
  constexpr int WGN = 4 / WGM;
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TA = WTM / 32, TB = WTN / 32;
  static_assert(TA * 32 == WTM && TB * 32 == WTN, "wave tile must be a multiple of 32");
  using MA = PairMap<BM, AMF>;
  using MB = PairMap<BN, 1 - BKF>;
  __shared__ __align__(16) unsigned short As[NP][BM * SLD];
  __shared__ __align__(16) unsigned short Bs[NP][BN * SLD];
  __shared__ unsigned offAm[BM], offBn[BN];
  __shared__ int offCm[BM], offCn[BN];
  __shared__ unsigned offAk[KLDS], offBk[KLDS];
  __shared__ GroupView gvM, gvN, gvK;

  const int tid = threadIdx.x;
  int bx, by, prob;
  tile_order(bx, by, prob);
  bx = __builtin_amdgcn_readfirstlane(bx);
  by = __builtin_amdgcn_readfirstlane(by);
  prob = __builtin_amdgcn_readfirstlane(prob);
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
  const int m0 = by * BM;
  const int n0 = bx * BN;
  if (m0 >= cM || n0 >= cN) return;

  int legs[kMaxLegs];
  for (int r = tid; r < BM + BN; r += kThreads) {
    if (r < BM) {
      const int idx = m0 + r;
      if (idx < cM) {
        decode32(gvM, q, idx, legs);
        offAm[r] = (unsigned)operand_offset(args.a.g[0], gvM, idx, legs) * (unsigned)sizeof(float);
        offCm[r] = (int)operand_offset(args.c.g[0], gvM, idx, legs);
      } else {
        offAm[r] = 0;
        offCm[r] = -1;
      }
    } else {
      const int t = r - BM;
      const int idx = n0 + t;
      if (idx < cN) {
        decode32(gvN, q, idx, legs);
        offBn[t] = (unsigned)operand_offset(args.b.g[1], gvN, idx, legs) * (unsigned)sizeof(float);
        offCn[t] = (int)operand_offset(args.c.g[1], gvN, idx, legs);
      } else {
        offBn[t] = 0;
        offCn[t] = -1;
      }
    }
  }

  const __amdgpu_buffer_rsrc_t ra_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)operand_base<float>(args.a, s, q), 0, (int)args.a_bytes, kBufferFlags);
  const __amdgpu_buffer_rsrc_t rb_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)operand_base<float>(args.b, s, q), 0, (int)args.b_bytes, kBufferFlags);
  float* Cb = operand_base<float>(args.c, s, q);

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave / WGN) * WTM;
  const int wn = (wave % WGN) * WTN;
  const int li = lane & 31, lk = lane >> 5;
  floatx16 acc[TA][TB];
#pragma unroll
  for (int a = 0; a < TA; ++a)
#pragma unroll
    for (int b = 0; b < TB; ++b)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

  float ra[MA::N][2], rb[MB::N][2];
  auto load_tile = [&](int kt) {
#pragma unroll
    for (int j = 0; j < MA::N; ++j) {
      const unsigned o = offAm[MA::row(tid, j)];
      const int k = kt + 2 * MA::pair(tid, j);
      ra[j][0] = buffer_load(ra_rsrc, o + offAk[k], 0.f);
      ra[j][1] = buffer_load(ra_rsrc, o + offAk[k + 1], 0.f);
    }
#pragma unroll
    for (int j = 0; j < MB::N; ++j) {
      const unsigned o = offBn[MB::row(tid, j)];
      const int k = kt + 2 * MB::pair(tid, j);
      rb[j][0] = buffer_load(rb_rsrc, o + offBk[k], 0.f);
      rb[j][1] = buffer_load(rb_rsrc, o + offBk[k + 1], 0.f);
    }
  };
  // pieces of a pair of consecutive k values: one 32-bit LDS word per piece;
  // each remainder x - piece is exact in FP32
  auto split_pair = [&](unsigned short* dst, int stride, float x0, float x1) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const unsigned w = bf16_pair(x0, x1);
      *reinterpret_cast<unsigned*>(dst + p * stride) = w;
      if (p + 1 < NP) {
        x0 -= __uint_as_float(w << 16);
        x1 -= __uint_as_float(w & 0xFFFF0000u);
      }
    }
  };
  auto store_tile = [&]() {
#pragma unroll
    for (int j = 0; j < MA::N; ++j)
      split_pair(&As[0][MA::row(tid, j) * SLD + 2 * MA::pair(tid, j)], BM * SLD, ra[j][0], ra[j][1]);
#pragma unroll
    for (int j = 0; j < MB::N; ++j)
      split_pair(&Bs[0][MB::row(tid, j) * SLD + 2 * MB::pair(tid, j)], BN * SLD, rb[j][0], rb[j][1]);
  };

  for (int kc = 0; kc < cK; kc += KLDS) {
    const int klen = min(KLDS, cK - kc);
    const int nt = (klen + BK - 1) / BK;
    for (int k = tid; k < nt * BK; k += kThreads) {
      if (k < klen) {
        decode32(gvK, q, kc + k, legs);
        offAk[k] = (unsigned)operand_offset(args.a.g[1], gvK, kc + k, legs) * (unsigned)sizeof(float);
        offBk[k] = (unsigned)operand_offset(args.b.g[0], gvK, kc + k, legs) * (unsigned)sizeof(float);
      } else {
        offAk[k] = kInvalidK;
        offBk[k] = kInvalidK;
      }
    }
    __syncthreads();
    load_tile(0);
    for (int t = 0; t < nt; ++t) {
      store_tile();
      __syncthreads();
      if (t + 1 < nt) load_tile((t + 1) * BK);
      // operands of this K step: eight consecutive K values per lane and piece
      bf16x8 av[NP][TA], bv[NP][TB];
#pragma unroll
      for (int p = 0; p < NP; ++p) {
#pragma unroll
        for (int ta = 0; ta < TA; ++ta)
          av[p][ta] = *reinterpret_cast<const bf16x8*>(&As[p][(wm + 32 * ta + li) * SLD + 8 * lk]);
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
          bv[p][tb] = *reinterpret_cast<const bf16x8*>(&Bs[p][(wn + 32 * tb + li) * SLD + 8 * lk]);
      }
#pragma unroll
      for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb) {
          // smallest products first, the leading one last
#pragma unroll
          for (int o = NP - 1; o >= 0; --o)
#pragma unroll
            for (int pa = 0; pa <= o; ++pa)
              acc[ta][tb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[pa][ta], bv[o - pa][tb], acc[ta][tb], 0, 0,
                                                                    0);
        }
      __syncthreads();
    }
  }
  if (cK == 0) __syncthreads();

  const float alpha = (float)args.alpha;
#pragma unroll
  for (int ta = 0; ta < TA; ++ta)
#pragma unroll
    for (int tb = 0; tb < TB; ++tb)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int i = ta * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
        const int j = tb * 32 + (lane & 31);
        const int om = offCm[wm + i];
        const int on = offCn[wn + j];
        if (om >= 0 && on >= 0) {
          float val = alpha * acc[ta][tb][v];
          float* c = Cb + (unsigned)(om + on);
          if (args.accumulate) val += *c;
          *c = val;
        }
      }
}

template <int BM, int BN, int WGM, int MINW, int NP>
static hipError_t launch_split_shape(const GemmArgs& args, long long mmax, long long nmax, hipStream_t stream) {
  const long long gy = (mmax + BM - 1) / BM;
  const long long gx = (nmax + BN - 1) / BN;
  const long long gz = 2LL * args.nsamples;
  if (gy > 65535 || gz > 65535 || gx > 2147483647LL || gx * gy * gz >= (1LL << 32))
    return hipErrorInvalidConfiguration;
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)gz);
  const int mode = (args.a_mfast ? 1 : 0) + (args.b_kfast ? 2 : 0);
  if (mode == 0)
    hipLaunchKernelGGL((z2gemm_split_kernel<BM, BN, WGM, MINW, 0, 0, NP>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 1)
    hipLaunchKernelGGL((z2gemm_split_kernel<BM, BN, WGM, MINW, 1, 0, NP>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 2)
    hipLaunchKernelGGL((z2gemm_split_kernel<BM, BN, WGM, MINW, 0, 1, NP>), grid, dim3(kThreads), 0, stream, args);
  else
    hipLaunchKernelGGL((z2gemm_split_kernel<BM, BN, WGM, MINW, 1, 1, NP>), grid, dim3(kThreads), 0, stream, args);
  return hipGetLastError();
}

// Split-kernel shapes: 0 = 128x128, 1 = 256x96, 2 = 256x32, 3 = 128x64.
template <int NP>
static hipError_t launch_split_tile(const GemmArgs& args, int tile, long long mmax, long long nmax,
                                    hipStream_t stream) {
  switch (tile) {
    case 0: return launch_split_shape<128, 128, 2, 2, NP>(args, mmax, nmax, stream);
    case 1: return launch_split_shape<256, 96, 4, 2, NP>(args, mmax, nmax, stream);
    case 2: return launch_split_shape<256, 32, 4, 2, NP>(args, mmax, nmax, stream);
    case 3: return launch_split_shape<128, 64, 2, 2, NP>(args, mmax, nmax, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_z2gemm_split(const GemmArgs& args, int pieces, long long mmax, long long nmax,
                               hipStream_t stream) {
  return launch_z2gemm_tile(args, 2, pieces, z2gemm_default_tile(2, mmax, nmax), mmax, nmax, stream);
}

// Block tile shapes per element type: (BM, BN, wave rows, K tile, waves per
// SIMD, double-buffered LDS).
template <typename T>
struct Shapes;
template <>
struct Shapes<float> {
  static constexpr int M0 = 128, N0 = 128, W0 = 2, K0 = 16, O0 = 3, D0 = 0;  // wave tile 64 x 64
  static constexpr int M1 = 256, N1 = 96, W1 = 4, K1 = 8, O1 = 3, D1 = 0;    // wave tile 64 x 96
  static constexpr int M2 = 256, N2 = 32, W2 = 4, K2 = 32, O2 = 2, D2 = 0;   // wave tile 64 x 32
};
template <>
struct Shapes<double> {
  static constexpr int M0 = 128, N0 = 128, W0 = 2, K0 = 16, O0 = 1, D0 = 0;  // wave tile 64 x 64
  static constexpr int M1 = 128, N1 = 96, W1 = 2, K1 = 8, O1 = 2, D1 = 0;    // wave tile 64 x 48
  static constexpr int M2 = 256, N2 = 32, W2 = 4, K2 = 16, O2 = 2, D2 = 0;   // wave tile 64 x 32
};

template <typename T, int BM, int BN, int WGM, int BK, int MINW, int DB>
static hipError_t launch_shape(const GemmArgs& args, long long mmax, long long nmax, hipStream_t stream) {
  const long long gy = (mmax + BM - 1) / BM;
  const long long gx = (nmax + BN - 1) / BN;
  const long long gz = 2LL * args.nsamples;
  if (gy > 65535 || gz > 65535 || gx > 2147483647LL || gx * gy * gz >= (1LL << 32))
    return hipErrorInvalidConfiguration;
  dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)gz);
  const int mode = (args.a_mfast ? 1 : 0) + (args.b_kfast ? 2 : 0);
  if (mode == 0)
    hipLaunchKernelGGL((z2gemm_kernel<T, BM, BN, WGM, BK, MINW, DB, 0, 0>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 1)
    hipLaunchKernelGGL((z2gemm_kernel<T, BM, BN, WGM, BK, MINW, DB, 1, 0>), grid, dim3(kThreads), 0, stream, args);
  else if (mode == 2)
    hipLaunchKernelGGL((z2gemm_kernel<T, BM, BN, WGM, BK, MINW, DB, 0, 1>), grid, dim3(kThreads), 0, stream, args);
  else
    hipLaunchKernelGGL((z2gemm_kernel<T, BM, BN, WGM, BK, MINW, DB, 1, 1>), grid, dim3(kThreads), 0, stream, args);
  return hipGetLastError();
}

// Shapes 0-2 of a family are the Shapes<T> above; 3 and up are the extra
// candidates of the autotuner: FP32 3 = 128x64 (2x2 waves, wave tile 64x32),
// 4 = 256x64 (4x1, 64x64), 5 = 64x128 (1x4, 64x32); FP64 3 = 64x64 (2x2,
// 32x32), 4 = 128x64 (2x2, 64x32).
template <typename T>
static hipError_t launch_tile_of(const GemmArgs& args, int tile, long long mmax, long long nmax, hipStream_t stream);
template <>
hipError_t launch_tile_of<float>(const GemmArgs& args, int tile, long long mmax, long long nmax, hipStream_t stream) {
  using S = Shapes<float>;
  switch (tile) {
    case 0: return launch_shape<float, S::M0, S::N0, S::W0, S::K0, S::O0, S::D0>(args, mmax, nmax, stream);
    case 1: return launch_shape<float, S::M1, S::N1, S::W1, S::K1, S::O1, S::D1>(args, mmax, nmax, stream);
    case 2: return launch_shape<float, S::M2, S::N2, S::W2, S::K2, S::O2, S::D2>(args, mmax, nmax, stream);
    case 3: return launch_shape<float, 128, 64, 2, 16, 3, 0>(args, mmax, nmax, stream);
    case 4: return launch_shape<float, 256, 64, 4, 16, 3, 0>(args, mmax, nmax, stream);
    case 5: return launch_shape<float, 64, 128, 1, 16, 3, 0>(args, mmax, nmax, stream);
    default: return hipErrorInvalidValue;
  }
}
template <>
hipError_t launch_tile_of<double>(const GemmArgs& args, int tile, long long mmax, long long nmax,
                                  hipStream_t stream) {
  using S = Shapes<double>;
  switch (tile) {
    case 0: return launch_shape<double, S::M0, S::N0, S::W0, S::K0, S::O0, S::D0>(args, mmax, nmax, stream);
    case 1: return launch_shape<double, S::M1, S::N1, S::W1, S::K1, S::O1, S::D1>(args, mmax, nmax, stream);
    case 2: return launch_shape<double, S::M2, S::N2, S::W2, S::K2, S::O2, S::D2>(args, mmax, nmax, stream);
    case 3: return launch_shape<double, 64, 64, 2, 16, 2, 0>(args, mmax, nmax, stream);
    case 4: return launch_shape<double, 128, 64, 2, 16, 2, 0>(args, mmax, nmax, stream);
    default: return hipErrorInvalidValue;
  }
}

int z2gemm_tile_count(int family) { return family == 1 ? 6 : family == 0 ? 5 : 4; }

int z2gemm_default_tile(int family, long long mmax, long long nmax) {
  // (BM, BN) of shapes 0, 1 and the N <= 32 shape 2 per family
  const long long m0 = 128, n0 = 128, m1 = family == 0 ? 128 : 256, n1 = 96;
  if (nmax <= 32) return 2;
  auto padded = [&](long long bm, long long bn) {
    return ((mmax + bm - 1) / bm) * bm * (((nmax + bn - 1) / bn) * bn);
  };
  return padded(m1, n1) < padded(m0, n0) ? 1 : 0;
}

const char* z2gemm_tile_name(int family, int tile) {
  static const char* f64[] = {"128x128", "128x96", "256x32", "64x64", "128x64"};
  static const char* f32[] = {"128x128", "256x96", "256x32", "128x64", "256x64", "64x128"};
  static const char* spl[] = {"128x128", "256x96", "256x32", "128x64"};
  if (tile < 0 || tile >= z2gemm_tile_count(family)) return "?";
  return family == 0 ? f64[tile] : family == 1 ? f32[tile] : spl[tile];
}

hipError_t launch_z2gemm_tile(const GemmArgs& args, int family, int pieces, int tile, long long mmax,
                              long long nmax, hipStream_t stream) {
  if (mmax <= 0 || nmax <= 0 || args.nsamples <= 0) return hipSuccess;
  if (family == 0) return launch_tile_of<double>(args, tile, mmax, nmax, stream);
  if (family == 1) return launch_tile_of<float>(args, tile, mmax, nmax, stream);
  if (family == 2 && pieces == 2) return launch_split_tile<2>(args, tile, mmax, nmax, stream);
  if (family == 2 && pieces == 3) return launch_split_tile<3>(args, tile, mmax, nmax, stream);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_z2gemm(const GemmArgs& args, long long mmax, long long nmax, hipStream_t stream) {
  const int family = sizeof(T) == 4 ? 1 : 0;
  return launch_z2gemm_tile(args, family, 0, z2gemm_default_tile(family, mmax, nmax), mmax, nmax, stream);
}

template hipError_t launch_z2gemm<float>(const GemmArgs&, long long, long long, hipStream_t);
template hipError_t launch_z2gemm<double>(const GemmArgs&, long long, long long, hipStream_t);

}  // namespace spp
