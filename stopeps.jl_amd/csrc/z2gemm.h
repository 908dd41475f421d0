// Batched Z2 block-sparse contraction on CDNA4 MFMA.
//
// Every tensor is stored dense per sample, each leg sorted by parity (the even
// indices first). A contraction C[M,N] = sum_K A[M,K] B[K,N] over leg groups
// M, N, K splits into one independent GEMM per (sample, parity pi): because
// every operand is Z2-neutral, an entry is nonzero only when the M, K and N
// tuples all carry parity pi. Tuples of a group with parity pi are enumerated
// prefix-parity-major with the last leg fastest, so the last leg's contiguous
// parity range maps to contiguous memory. Each (operand, group) pair is
// addressed either through per-leg strides (dense tensor) or linearly
// (idx * stride, for compact per-parity block matrices).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace spp {

constexpr int kMaxLegs = 6;

struct Group {
  int n;            // number of legs
  int graded;       // 1: parity-enumerated tuples; 0: flat (same count for both parities)
  int dim[kMaxLegs];
  int even[kMaxLegs];            // fixed even counts (used when split == nullptr)
  const int* split[kMaxLegs];    // per-sample even counts, indexed by sample, or nullptr
};

struct OperandGroup {
  int linear;                    // 1: offset = idx * lstride; 0: sum of leg index * leg stride
  long long lstride;
  long long stride[kMaxLegs];
  int fold_j;                    // leg of this group stored folded (its index within its parity half), or -1
  int fold_even;                 // even count of that leg
};

struct Operand {
  void* base;                    // element pointer
  long long sstride;             // elements between samples (0 = shared)
  void* const* ptrs;             // per-sample base pointers (overrides base/sstride) or nullptr
  long long pstride;             // elements between the two parity blocks (compact operands)
  long long offset;              // elements added to every (sample, parity) base (sub-range views)
  OperandGroup g[2];             // A: (M,K)  B: (K,N)  C: (M,N)
};

struct GemmArgs {
  Group gm, gn, gk;
  Operand a, b, c;
  int nsamples;
  const int* sample_map;         // optional: problem sample index -> storage sample index
  const int* pmask;              // optional per (sample*2+parity) flag: 0 skips the problem
  double alpha;
  int accumulate;                // C = alpha*AB + (accumulate ? C : 0)
  int conj_b;                    // reserved (real arithmetic only)
  int a_mfast;                   // A tiles are read with consecutive lanes along M (else along K)
  int b_kfast;                   // B tiles are read with consecutive lanes along K (else along N)
  unsigned a_bytes, b_bytes;     // extent of A / B from a (sample, parity) base, bytes (< 2^31)
};

// Host-side launcher. T is float or double. Grid bounds mmax/nmax are upper
// bounds of the per-problem M and N counts.
template <typename T>
hipError_t launch_z2gemm(const GemmArgs& args, long long mmax, long long nmax, hipStream_t stream);
// FP32 operands and result with the products formed from BF16 pieces on the
// BF16 matrix cores (pieces = 3: six piece products, FP32-equivalent;
// pieces = 2: three piece products, about 2^-16 relative).
hipError_t launch_z2gemm_split(const GemmArgs& args, int pieces, long long mmax, long long nmax,
                               hipStream_t stream);

// Block tile shapes by kernel family: 0 = FP64 kernel, 1 = FP32 kernel,
// 2 = split-product kernel. Every shape of a family computes bit-identical
// results (the same matrix-core instruction walks K in the same order); they
// differ in speed only. The default is the padded-area heuristic; the
// autotuner of the contraction layer may pick another shape per launch class.
constexpr int kMaxTileShapes = 6;
int z2gemm_tile_count(int family);
int z2gemm_default_tile(int family, long long mmax, long long nmax);
const char* z2gemm_tile_name(int family, int tile);
// Launch with an explicit tile shape (pieces: the split kernel's BF16 pieces).
hipError_t launch_z2gemm_tile(const GemmArgs& args, int family, int pieces, int tile, long long mmax,
                              long long nmax, hipStream_t stream);

// ---------------------------------------------------------------------------
// Group decoding, shared with the other kernels.
struct GroupView {
  int n, graded;
  int dim[kMaxLegs];
  int ev[kMaxLegs];
  long long pc[kMaxLegs + 1][2];   // prefix counts per parity
};

__host__ __device__ inline int leg_count(const GroupView& g, int j, int q) {
  return q == 0 ? g.ev[j] : g.dim[j] - g.ev[j];
}

__device__ inline void resolve_group(const Group& gr, int s, GroupView& g) {
  g.n = gr.n;
  g.graded = gr.graded;
  g.pc[0][0] = 1;
  g.pc[0][1] = 0;
  for (int j = 0; j < gr.n; ++j) {
    g.dim[j] = gr.dim[j];
    g.ev[j] = gr.split[j] ? gr.split[j][s] : gr.even[j];
    if (gr.graded) {
      long long c0 = g.ev[j], c1 = g.dim[j] - g.ev[j];
      g.pc[j + 1][0] = g.pc[j][0] * c0 + g.pc[j][1] * c1;
      g.pc[j + 1][1] = g.pc[j][0] * c1 + g.pc[j][1] * c0;
    } else {
      g.pc[j + 1][0] = g.pc[j][0] * g.dim[j];
      g.pc[j + 1][1] = g.pc[j][0] * g.dim[j];
    }
  }
  if (!gr.graded && gr.n == 0) {
    g.pc[0][1] = 1;
  }
}

__device__ inline long long group_count(const GroupView& g, int q) { return g.pc[g.n][q]; }

// Decode tuple index ``idx`` of parity q into leg indices.
__device__ inline void decode(const GroupView& g, int q, long long idx, int* legs) {
  if (!g.graded) {
    for (int j = g.n - 1; j >= 0; --j) {
      legs[j] = (int)(idx % g.dim[j]);
      idx /= g.dim[j];
    }
    return;
  }
  for (int j = g.n - 1; j >= 0; --j) {
    long long c_q = (q == 0) ? g.ev[j] : g.dim[j] - g.ev[j];
    long long part0 = g.pc[j][0] * c_q;
    int lastpar, qp;
    long long cl;
    if (idx < part0) {
      qp = 0;
      lastpar = q;
      cl = c_q;
    } else {
      idx -= part0;
      qp = 1;
      lastpar = q ^ 1;
      cl = (lastpar == 0) ? g.ev[j] : g.dim[j] - g.ev[j];
    }
    long long pr = idx / cl;
    int lr = (int)(idx - pr * cl);
    legs[j] = lr + (lastpar ? g.ev[j] : 0);
    idx = pr;
    q = qp;
  }
}

// decode() for tuple indices below 2^31 (32-bit division).
__device__ inline void decode32(const GroupView& g, int q, int idx, int* legs) {
  if (!g.graded) {
    for (int j = g.n - 1; j >= 0; --j) {
      legs[j] = idx % g.dim[j];
      idx /= g.dim[j];
    }
    return;
  }
  for (int j = g.n - 1; j >= 0; --j) {
    const int c_q = (q == 0) ? g.ev[j] : g.dim[j] - g.ev[j];
    const int part0 = (int)g.pc[j][0] * c_q;
    int lastpar, qp, cl;
    if (idx < part0) {
      qp = 0;
      lastpar = q;
      cl = c_q;
    } else {
      idx -= part0;
      qp = 1;
      lastpar = q ^ 1;
      cl = (lastpar == 0) ? g.ev[j] : g.dim[j] - g.ev[j];
    }
    const int pr = idx / cl;
    legs[j] = (idx - pr * cl) + (lastpar ? g.ev[j] : 0);
    idx = pr;
    q = qp;
  }
}

__device__ inline long long operand_offset(const OperandGroup& og, const GroupView& g, long long idx,
                                           const int* legs) {
  if (og.linear) return idx * og.lstride;
  long long off = 0;
  for (int j = 0; j < g.n; ++j) {
    int l = legs[j];
    if (j == og.fold_j && l >= og.fold_even) l -= og.fold_even;
    off += (long long)l * og.stride[j];
  }
  return off;
}

}  // namespace spp
