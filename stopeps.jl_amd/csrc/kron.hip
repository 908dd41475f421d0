// Row Kronecker apply of a double layer through its single-layer factors
// (declarations and the operation in kron.h).
//
// Per row r of parity rho and for the single layer S' of parity q:
//   stage 1, per bra tuple (c1b, c2b):
//     T[c1b, c2b; phi, o1k, o2k] = sum_{c1k, c2k} X[r, (c1k, c1b), (c2k, c2b)] S'(c1k, c2k, phi, o1k, o2k)
//   stage 2, per ket tuple (o1k, o2k):
//     Y[r, (o1k, o1b), (o2k, o2b)] = sum_{c1b, c2b, phi} T[c1b, c2b; phi, o1k, o2k] S'(c1b, c2b, phi, o1b, o2b)
// Z2 fixes which terms are nonzero: a stage-1 lane of class c = par(c1b, c2b) ^ rho sums over the
// ket tuples (c1k, c2k) of parity c into the outputs (phi, o1k, o2k) of parity q ^ c; a stage-2
// lane of class c2 = par(o1k, o2k) ^ rho sums over (c1b, c2b, phi) of parity q ^ c2 into the
// bra tuples (o1b, o2b) of parity c2. Lanes are grouped by class so that every coefficient
// a wave reads is the same for all its lanes (scalar loads); the coefficient blocks K1, K2
// are tabulated per (site, label) on the host.
//
// One block holds `rows` rows: the X rows are staged in LDS, stage 1 writes T (rows x
// tuples x no1) to LDS, stage 2 reads it, writes the Y rows over the X area, and the block
// stores them contiguously.
#include "kron.h"
#include "engine.h"
#include <algorithm>
#include <cstdint>
#include <cstring>

namespace spp {

namespace {

constexpr int kKronLdsBytes = 160 * 1024;

struct Inst {
  int no1, no2;
};
// Register tiles (stage-1 outputs, stage-2 outputs per lane) with a kernel instantiation.
constexpr Inst kInst[] = {{8, 8}, {16, 16}, {18, 18}, {32, 32}, {64, 64}};

// Roles (c1, c2, o1, o2) of the layer legs W=0, N=1, E=2, S=3 per stage.
constexpr int kRole[kKronStages][4] = {{1, 2, 0, 3}, {2, 3, 0, 1}, {0, 1, 2, 3}, {3, 0, 2, 1}};

inline int parl(int i, int e) { return i >= e ? 1 : 0; }

struct RoleLegs {
  int d[4], e[4];
};
RoleLegs roles(const KronSite& ks, int stage) {
  RoleLegs r;
  for (int k = 0; k < 4; ++k) {
    r.d[k] = ks.d[kRole[stage][k]];
    r.e[k] = ks.e[kRole[stage][k]];
  }
  return r;
}

// tuples (a, b) of two single legs with par(a) + par(b) = pi
int tuple_count(int da, int ea, int db, int eb, int pi) {
  int n = 0;
  for (int a = 0; a < da; ++a)
    for (int b = 0; b < db; ++b)
      if (((parl(a, ea) + parl(b, eb)) & 1) == pi) ++n;
  return n;
}

// stage-1 outputs (phi, o1k, o2k) of parity pi
int out1_count(const RoleLegs& r, int dphi, int ephi, int pi) {
  int n = 0;
  for (int f = 0; f < dphi; ++f)
    for (int a = 0; a < r.d[2]; ++a)
      for (int b = 0; b < r.d[3]; ++b)
        if (((parl(f, ephi) + parl(a, r.e[2]) + parl(b, r.e[3])) & 1) == pi) ++n;
  return n;
}

// stage-2 inner tuples (c1b, c2b, phi) of parity pi
int inner2_count(const RoleLegs& r, int dphi, int ephi, int pi) {
  int n = 0;
  for (int a = 0; a < r.d[0]; ++a)
    for (int b = 0; b < r.d[1]; ++b)
      for (int f = 0; f < dphi; ++f)
        if (((parl(a, r.e[0]) + parl(b, r.e[1]) + parl(f, ephi)) & 1) == pi) ++n;
  return n;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

double layer_at(const KronSite& ks, const double* S, int w, int n, int f, int e, int s) {
  return S[((((long long)w * ks.d[1] + n) * ks.dphi + f) * ks.d[2] + e) * ks.d[3] + s];
}

// S'(c1k, c2k, phi, o1k, o2k) of a stage
double role_at(const KronSite& ks, const double* S, int stage, int c1, int c2, int f, int o1, int o2) {
  int leg[4];
  leg[kRole[stage][0]] = c1;
  leg[kRole[stage][1]] = c2;
  leg[kRole[stage][2]] = o1;
  leg[kRole[stage][3]] = o2;
  return layer_at(ks, S, leg[0], leg[1], f, leg[2], leg[3]);
}

}  // namespace

bool kron_site_layout(const int d[4], const int e[4], int dphi, int ephi, KronSite& ks) {
  for (int k = 0; k < 4; ++k) {
    if (d[k] < 1 || d[k] > 8 || e[k] < 0 || e[k] > d[k]) return false;
    ks.d[k] = d[k];
    ks.e[k] = e[k];
  }
  if (dphi < 1 || dphi > 2 || ephi < 0 || ephi > dphi) return false;
  ks.dphi = dphi;
  ks.ephi = ephi;
  // per pass the stage-1 outputs (o1k, o2k) and the stage-2 bra outputs
  // (o1b, o2b) are tuples of one parity; the ket / bra inner lists too
  int need = 0, j1 = 0;
  for (int stage = 0; stage < kKronStages; ++stage) {
    const RoleLegs r = roles(ks, stage);
    for (int pi = 0; pi < 2; ++pi) {
      need = std::max(need, tuple_count(r.d[2], r.e[2], r.d[3], r.e[3], pi));
      j1 = std::max(j1, tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], pi));
    }
  }
  int best = -1;
  for (int i = 0; i < (int)(sizeof(kInst) / sizeof(kInst[0])); ++i)
    if (kInst[i].no1 >= need && kInst[i].no2 >= need &&
        (best < 0 || kInst[i].no1 + kInst[i].no2 < kInst[best].no1 + kInst[best].no2))
      best = i;
  if (best < 0 || j1 > kKronMaxTuple) return false;
  ks.no1 = kInst[best].no1;
  ks.no2 = kInst[best].no2;
  ks.j1cap = j1;
  ks.j2cap = j1;
  ks.stage_bytes = align256(sizeof(KronTab) + sizeof(float) * 2 * (size_t)dphi *
                                                  ((size_t)j1 * ks.no1 + (size_t)j1 * ks.no2));
  return true;
}

double kron_row_macs(const KronSite& ks, int stage) {
  const RoleLegs r = roles(ks, stage);
  double macs = 0;
  for (int rho = 0; rho < 2; ++rho)
    for (int c = 0; c < 2; ++c) {
      // per pass phi: stage 1 over the ket tuples of parity c, stage 2 over
      // the bra tuples of parity c ^ par(phi) (layer parity 0)
      macs += (double)tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], c ^ rho) *
              tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], c) * ks.no1 * ks.dphi;
      for (int f = 0; f < ks.dphi; ++f)
        macs += (double)tuple_count(r.d[2], r.e[2], r.d[3], r.e[3], c ^ rho) *
                tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], c ^ parl(f, ks.ephi)) * ks.no2;
    }
  return 0.5 * macs;
}

int kron_layer_parity(const int d[4], const int e[4], int dphi, int ephi, const double* S) {
  int seen[2] = {0, 0};
  long long i = 0;
  for (int w = 0; w < d[0]; ++w)
    for (int n = 0; n < d[1]; ++n)
      for (int f = 0; f < dphi; ++f)
        for (int x = 0; x < d[2]; ++x)
          for (int s = 0; s < d[3]; ++s, ++i)
            if (S[i] != 0.0)
              seen[(parl(w, e[0]) + parl(n, e[1]) + parl(f, ephi) + parl(x, e[2]) + parl(s, e[3])) & 1] = 1;
  if (seen[0] && seen[1]) return -1;
  return seen[1] ? 1 : 0;
}

void kron_build_tables(const KronSite& ks, const double* S, int q, std::vector<char>& out) {
  out.assign(kKronStages * ks.stage_bytes, 0);
  for (int stage = 0; stage < kKronStages; ++stage) {
    const RoleLegs r = roles(ks, stage);
    char* base = out.data() + stage * ks.stage_bytes;
    KronTab* tab = reinterpret_cast<KronTab*>(base);
    tab->q = q;
    float* K1 = reinterpret_cast<float*>(base + sizeof(KronTab));
    float* K2 = K1 + 2 * (size_t)ks.dphi * ks.j1cap * ks.no1;
    for (int pi = 0; pi < 2; ++pi) {
      int n = 0;
      for (int a = 0; a < r.d[0]; ++a)
        for (int b = 0; b < r.d[1]; ++b)
          if (((parl(a, r.e[0]) + parl(b, r.e[1])) & 1) == pi) tab->p12[pi][n++] = a | b << 8;
      tab->n12[pi] = n;
      n = 0;
      for (int a = 0; a < r.d[2]; ++a)
        for (int b = 0; b < r.d[3]; ++b)
          if (((parl(a, r.e[2]) + parl(b, r.e[3])) & 1) == pi) tab->po[pi][n++] = a | b << 8;
      tab->no[pi] = n;
    }
    for (int c = 0; c < 2; ++c)
      for (int f = 0; f < ks.dphi; ++f) {
        const int pf = parl(f, ks.ephi);
        // stage 1, class c, pass f: ket tuples of parity c into the outputs
        // (o1k, o2k) of parity q ^ c ^ par(f), lexicographic
        const int po = q ^ c ^ pf;
        for (int j = 0; j < tab->n12[c]; ++j) {
          const int c1 = tab->p12[c][j] & 255, c2 = tab->p12[c][j] >> 8;
          for (int o = 0; o < tab->no[po]; ++o) {
            const int o1 = tab->po[po][o] & 255, o2 = tab->po[po][o] >> 8;
            K1[(((size_t)c * ks.dphi + f) * ks.j1cap + j) * ks.no1 + o] =
                (float)role_at(ks, S, stage, c1, c2, f, o1, o2);
          }
        }
        // stage 2, class c, pass f: bra tuples (c1b, c2b) of parity q ^ c ^ par(f)
        // into the bra outputs (o1b, o2b) of parity c
        int n = 0;
        for (int j = 0; j < tab->n12[po]; ++j) {
          const int a = tab->p12[po][j] & 255, b = tab->p12[po][j] >> 8;
          tab->j2[c][f][n] = a * r.d[1] + b;
          for (int o = 0; o < tab->no[c]; ++o) {
            const int o1 = tab->po[c][o] & 255, o2 = tab->po[c][o] >> 8;
            K2[(((size_t)c * ks.dphi + f) * ks.j2cap + n) * ks.no2 + o] =
                (float)role_at(ks, S, stage, a, b, f, o1, o2);
          }
          ++n;
        }
        tab->nj2[c][f] = n;
      }
  }
}

// ---------------------------------------------------------------------------
struct KronArgs {
  int d[4], e[4];          // single legs of the roles c1, c2, o1, o2
  int pe[4], pf[4], pd[4]; // their pair legs: even count, folded length, dimension
  int dphi, ephi;
  long long nrows;
  int nf, ef, rc;
  const int* fsplit;
  const int* lsplit;
  int leven;
  const int* rsplit;
  int reven;
  const float* x;
  long long xs, sx_slab, sxf, sx1, sx2;
  float* y;
  long long ys, rowY;
  void* const* tabs;
  const void* tab_shared;  // tables of every sample when tabs is null
  long long stage_off;
  int rows, groups, rowmax, w1, w2, j1cap, j2cap, ntup1;
  float inv_rowx, inv_pf1;  // 1 / (pd[0] * pf[1]), 1 / pf[1]
  int sx1i, sx2i;                     // sx1, sx2 (a sample's X has fewer than 2^31 elements)
  int xflat;                          // X rows are packed and consecutive in global memory
  float inv_pd0;                      // 1 / pd[0]
  // matrix-core variant
  int tno1, tno2;           // row widths of the K1 / K2 tables
  int nofmax, ncol1, ts;    // outputs per pass, stage-1 columns (dphi * nofmax), T row stride (odd)
  int wc;                   // waves per class
  int n1max, n2max, m1cap, m2cap;
};

// n / m for 0 <= n < 2^22 and 1 <= m <= 4096 from inv = 1 / m: (n + 1/2) / m
// stays at least 1/(2m) away from an integer, far beyond the FP32 error.
__device__ __forceinline__ int small_div(int n, float inv) { return (int)(((float)n + 0.5f) * inv); }

__device__ __forceinline__ int pair_base(int pi, int pj, int e, int o) {
  // parity-sorted pair blocks: ee, oo, eo, oe
  return pi == pj ? (pi ? e * e : 0) : (e * e + o * o + (pi ? e * o : 0));
}

constexpr int kKronPrefetch = 16;  // X loads in flight per thread

// X rows of a block into LDS (packed, row stride pd[0] * pf[1]), PF independent
// loads in flight per thread.
// Fields of KronArgs that the X staging reads, copied field by field in the
// kernel (taking a reference to the kernel argument block would move it to
// private memory and turn every table load into a vector load).
struct StageX {
  int xflat, pd0, pf1, sx1i, sx2i;
  long long sxf, sx_slab;
  float inv_rowx, inv_pf1, inv_pd0;
};

#define KRON_STAGE_ARGS(a) \
  StageX { a.xflat, a.pd[0], a.pf[1], a.sx1i, a.sx2i, a.sxf, a.sx_slab, a.inv_rowx, a.inv_pf1, a.inv_pd0 }

// LDS of both kernels: X rows from offset 0 (then kernel-specific areas), the
// row parities and X row offsets of the block's rows. Declared here so that the
// staging below stores to LDS through the array itself (a pointer argument
// would make the stores generic, and generic stores keep the table loads that
// follow them off the scalar unit).
extern __shared__ float kron_lds[];
__shared__ int kron_rho[32];
__shared__ int kron_xrow0[32];

template <int PF>
__device__ __forceinline__ void kron_stage_x(const StageX a, const float* xs, int nr, int tid, int nthr) {
  float* Xl = kron_lds;
  const int* xrow0 = kron_xrow0;
  const int rowX = a.pd0 * a.pf1;
  if (a.xflat) {
    // consecutive rows are contiguous in global memory: 16-byte copies where aligned
    const float* src = xs + xrow0[0];
    const int total = nr * rowX;
    if ((reinterpret_cast<uintptr_t>(src) & 15) == 0 && (total & 3) == 0) {
      constexpr int PF4 = PF >= 4 ? PF / 4 : 1;
      const int t4 = total >> 2;
      const float4* s4 = reinterpret_cast<const float4*>(src);
      float4* d4 = reinterpret_cast<float4*>(Xl);
      for (int base = 0; base < t4; base += PF4 * nthr) {
        float4 v[PF4];
#pragma unroll
        for (int u = 0; u < PF4; ++u) v[u] = s4[min(base + u * nthr + tid, t4 - 1)];  // clamped: stored only in range
#pragma unroll
        for (int u = 0; u < PF4; ++u) {
          const int i = base + u * nthr + tid;
          if (i < t4) d4[i] = v[u];
        }
      }
      return;
    }
    for (int base = 0; base < total; base += PF * nthr) {
      float v[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int i = base + u * nthr + tid;
        if (i < total) v[u] = src[i];
      }
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int i = base + u * nthr + tid;
        if (i < total) Xl[i] = v[u];
      }
    }
  } else if (a.sx2i == 1 && (a.pf1 & 1) == 0 && (a.sx1i & 1) == 0 && (a.sxf & 1) == 0 && (a.sx_slab & 1) == 0) {
    // rows of pd[0] chunks of pf[1] contiguous values (even lengths and offsets): one
    // chunk per thread, copied as 8-byte pairs
    const int nch = nr * a.pd0, half = a.pf1 >> 1;
    for (int ch = tid; ch < nch; ch += nthr) {
      const int rr = small_div(ch, a.inv_pd0), c1 = ch - rr * a.pd0;
      const float2* src = reinterpret_cast<const float2*>(xs + xrow0[rr] + c1 * a.sx1i);
      float2* dst = reinterpret_cast<float2*>(Xl + rr * rowX + c1 * a.pf1);
      for (int u = 0; u < half; ++u) dst[u] = src[u];
    }
  } else {
    const bool contig = a.sx2i == 1 && a.sx1i == a.pf1;
    const int total = nr * rowX;
    for (int base = 0; base < total; base += PF * nthr) {
      float v[PF];
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int i = base + u * nthr + tid;
        if (i < total) {
          const int rr = small_div(i, a.inv_rowx), k = i - rr * rowX;
          int off = xrow0[rr];
          if (contig) {
            off += k;
          } else {
            const int c1 = small_div(k, a.inv_pf1);
            off += c1 * a.sx1i + (k - c1 * a.pf1) * a.sx2i;
          }
          v[u] = xs[off];
        }
      }
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int i = base + u * nthr + tid;
        if (i < total) Xl[i] = v[u];
      }
    }
  }
}

// The block's Y rows from LDS to global memory, 16 bytes at a time where the
// destination is aligned.
__device__ __forceinline__ void kron_copy_out(float* ys, const float* Xl, int totY, int tid, int nthr) {
  if ((reinterpret_cast<uintptr_t>(ys) & 15) == 0 && (totY & 3) == 0) {
    const float4* s4 = reinterpret_cast<const float4*>(Xl);
    float4* d4 = reinterpret_cast<float4*>(ys);
    for (int i = tid; i < (totY >> 2); i += nthr) d4[i] = s4[i];
    return;
  }
  for (int i = tid; i < totY; i += nthr) ys[i] = Xl[i];
}

// Threads per block: the wide register tiles keep 512 threads so that their
// accumulators (both stages live across the passes) fit without spilling.
constexpr int kron_max_threads(int no1) { return no1 > 24 ? 512 : 1024; }

template <int NO1, int NO2>
__global__ void __launch_bounds__(kron_max_threads(NO1)) kron_rows_kernel(KronArgs a) {
  constexpr int kNO1P = NO1 | 1;  // odd stride of a tuple's stage-1 outputs in LDS (no bank conflicts)
  int* rho = kron_rho;
  int* xrow0 = kron_xrow0;
  const int s = blockIdx.y;
  const char* tb = reinterpret_cast<const char*>(a.tabs ? a.tabs[s] : a.tab_shared) + a.stage_off;
  const KronTab* tab = reinterpret_cast<const KronTab*>(tb);
  const float* K1 = reinterpret_cast<const float*>(tb + sizeof(KronTab));
  const float* K2 = K1 + 2 * a.dphi * a.j1cap * NO1;
  const long long row0 = (long long)blockIdx.x * a.rows;
  const int nr = (int)min((long long)a.rows, a.nrows - row0);
  float* Xl = kron_lds;
  float* Tl = kron_lds + a.rows * a.rowmax;
  const float* xs = a.x + (long long)s * a.xs;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (tid < nr) {
    const int el = a.lsplit ? a.lsplit[s] : a.leven;
    const int er = a.rsplit ? a.rsplit[s] : a.reven;
    const long long r = row0 + tid;
    const long long slab = r / a.nf;
    const int f = (int)(r - slab * a.nf);
    const int l = (int)(slab / a.rc), R = (int)(slab - (long long)l * a.rc);
    rho[tid] = ((l >= el) + (R >= er) + (f >= (a.fsplit ? a.fsplit[s] : a.ef))) & 1;
    xrow0[tid] = (int)(slab * a.sx_slab + f * a.sxf);
  }
  __syncthreads();
  const int rowX = a.pd[0] * a.pf[1];
  kron_stage_x<kKronPrefetch>(KRON_STAGE_ARGS(a), xs, nr, tid, nthr);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  // stage-1 lane: a bra tuple (c1b, c2b) of a row, class c1 = par(c1b, c2b) ^ rho
  const bool s1 = wave < 2 * a.w1;
  const int c1 = wave >= a.w1 ? 1 : 0;
  int r1 = 0, tup1 = 0;
  bool act1 = false;
  if (s1) {
    int t = (wave - c1 * a.w1) * 64 + lane;
    const int n0 = tab->n12[0], n1 = tab->n12[1];
    if (n0 == n1) {
      // every row holds n0 tuples of the class: the lane's row is a quotient
      const int rr = n0 > 0 ? t / n0 : nr;
      if (rr < nr) {
        r1 = rr;
        tup1 = tab->p12[c1 ^ rho[rr]][t - rr * n0];
        act1 = true;
      }
    } else {
      for (int rr = 0; rr < nr; ++rr) {
        const int pi = c1 ^ rho[rr];
        const int n = pi ? n1 : n0;
        if (t < n) {
          r1 = rr;
          tup1 = tab->p12[pi][t];
          act1 = true;
          break;
        }
        t -= n;
      }
    }
  }
  // stage-2 lane: a ket tuple (o1k, o2k) of a row, class c2 = par(o1k, o2k) ^ rho;
  // its index t2 among its row's tuples of that parity is its rank in the stored outputs
  const bool s2 = wave < 2 * a.w2;
  const int c2 = wave >= a.w2 ? 1 : 0;
  int r2 = 0, tup2 = 0, t2 = 0;
  bool act2 = false;
  if (s2) {
    int t = (wave - c2 * a.w2) * 64 + lane;
    const int n0 = tab->no[0], n1 = tab->no[1];
    if (n0 == n1) {
      const int rr = n0 > 0 ? t / n0 : nr;
      if (rr < nr) {
        r2 = rr;
        t2 = t - rr * n0;
        tup2 = tab->po[c2 ^ rho[rr]][t2];
        act2 = true;
      }
    } else {
      for (int rr = 0; rr < nr; ++rr) {
        const int pi = c2 ^ rho[rr];
        const int n = pi ? n1 : n0;
        if (t < n) {
          r2 = rr;
          tup2 = tab->po[pi][t];
          t2 = t;
          act2 = true;
          break;
        }
        t -= n;
      }
    }
  }
  float acc2[NO2];
#pragma unroll
  for (int o = 0; o < NO2; ++o) acc2[o] = 0.f;
  for (int f = 0; f < a.dphi; ++f) {
    // ---------------- stage 1, pass f ----------------
    if (s1) {
      const int e1 = a.e[0], o1 = a.d[0] - a.e[0], e2 = a.e[1], o2 = a.d[1] - a.e[1];
      const int c1b = tup1 & 255, c2b = tup1 >> 8;
      const int p1b = c1b >= e1 ? 1 : 0, p2b = c2b >= e2 ? 1 : 0;
      const int r1b = c1b - (p1b ? e1 : 0), r2b = c2b - (p2b ? e2 : 0);
      const int mul1 = p1b ? o1 : e1, mul2 = p2b ? o2 : e2;
      const int b10 = pair_base(0, p1b, e1, o1) + r1b, b11 = pair_base(1, p1b, e1, o1) + r1b;
      const int b20 = pair_base(0, p2b, e2, o2) + r2b - (p2b ? a.pe[1] : 0);
      const int b21 = pair_base(1, p2b, e2, o2) + r2b - (p2b ? 0 : a.pe[1]);
      const float* xrow = Xl + r1 * rowX;
      float acc[NO1];
#pragma unroll
      for (int o = 0; o < NO1; ++o) acc[o] = 0.f;
      // ket tuples (c1k, c2k) of parity c1 in the tables' order: the even c1k with the c2k of
      // parity c1, then the odd c1k with the c2k of parity c1 ^ 1; in each group the X offset
      // advances by mul1 * pf per c1k and by mul2 per c2k
      const float* kj = K1 + (c1 * a.dphi + f) * a.j1cap * NO1;
      const int M1 = mul1 * a.pf[1];
      for (int g = 0; g < 2; ++g) {
        const int p2k = c1 ^ g;
        const int n1 = g ? o1 : e1, n2 = p2k ? o2 : e2;
        const float* xg = xrow + (g ? b11 : b10) * a.pf[1] + (p2k ? b21 : b20);
        for (int i1 = 0; i1 < n1; ++i1) {
          const float* xp = xg + i1 * M1;
#pragma unroll 2  // two tuples' coefficient loads per wait
          for (int i2 = 0; i2 < n2; ++i2) {
            const float xv = xp[i2 * mul2];  // lanes without a tuple read row 0, never stored
#pragma unroll
            for (int o = 0; o < NO1; ++o) acc[o] = fmaf(xv, kj[o], acc[o]);
            kj += NO1;
          }
        }
      }
      if (act1) {
        float* tp = Tl + (r1 * a.ntup1 + c1b * a.d[1] + c2b) * kNO1P;
#pragma unroll
        for (int o = 0; o < NO1; ++o) tp[o] = acc[o];
      }
    }
    __syncthreads();
    // ---------------- stage 2, pass f ----------------
    if (s2) {
      const float* trow = Tl + r2 * a.ntup1 * kNO1P + t2;
      // bra tuples (c1b, c2b) of parity q ^ c2 ^ par(f), one flat pass over the table's list
      const float* kj = K2 + (c2 * a.dphi + f) * a.j2cap * NO2;
      const int nj = tab->nj2[c2][f];
      const int* jl = tab->j2[c2][f];
      for (int j = 0; j < nj; ++j) {
        const float tv = trow[jl[j] * kNO1P];  // lanes without a tuple read row 0, never stored
#pragma unroll
        for (int o = 0; o < NO2; ++o) acc2[o] = fmaf(tv, kj[o], acc2[o]);
        kj += NO2;
      }
    }
    if (f + 1 < a.dphi) __syncthreads();  // T of the next pass overwrites this one
  }
  if (act2) {
    // Y rows (packed, row stride rowY) overwrite the X rows (X was last read
    // before the barrier above); the pair index of (o1k, o1b) is
    // A1[par o1b] + rank of o1b, that of (o2k, o2b) folded A2[par o2b] + rank of o2b
    const int eo1 = a.e[2], oo1 = a.d[2] - a.e[2], eo2 = a.e[3], oo2 = a.d[3] - a.e[3];
    const int o1k = tup2 & 255, o2k = tup2 >> 8;
    const int p1k = o1k >= eo1 ? 1 : 0, p2k = o2k >= eo2 ? 1 : 0;
    const int r1k = o1k - (p1k ? eo1 : 0), r2k = o2k - (p2k ? eo2 : 0);
    float* yrow = Xl + r2 * (int)a.rowY;
    const int A10 = pair_base(p1k, 0, eo1, oo1) + r1k * eo1, A11 = pair_base(p1k, 1, eo1, oo1) + r1k * oo1;
    const int A20 = pair_base(p2k, 0, eo2, oo2) + r2k * eo2 - (p2k ? a.pe[3] : 0);
    const int A21 = pair_base(p2k, 1, eo2, oo2) + r2k * oo2 - (p2k ? 0 : a.pe[3]);
    const int no = tab->no[c2];  // bra tuples (o1b, o2b) of parity c2, wave-uniform
#pragma unroll
    for (int o = 0; o < NO2; ++o) {
      if (o < no) {
        const int ot = tab->po[c2][o];
        const int o1b = ot & 255, o2b = ot >> 8;
        const int q1b = o1b >= eo1 ? 1 : 0, q2b = o2b >= eo2 ? 1 : 0;
        const int po1 = (q1b ? A11 : A10) + (o1b - (q1b ? eo1 : 0));
        const int po2 = (q2b ? A21 : A20) + (o2b - (q2b ? eo2 : 0));
        yrow[po1 * a.pf[3] + po2] = acc2[o];
      }
    }
  }
  __syncthreads();
  float* ys = a.y + (long long)s * a.ys + row0 * a.rowY;
  const int totY = nr * (int)a.rowY;
  kron_copy_out(ys, Xl, totY, tid, nthr);
}

// ---------------------------------------------------------------------------
// Matrix-core variant of the same two products. Per class c (the first wc waves
// take class 0, the next wc class 1) both are small dense products on the
// 16x16x4 FP32 MFMA:
//   stage 1: C1[m, n] = sum_j X[m; j] K1[c, phi, j, o], rows m = (row, bra tuple (c1b, c2b) of
//            parity c ^ rho), K = ket tuples j of parity c, columns n = phi * nofmax + o;
//   stage 2: C2[m, o] = sum_k T[m; k] K2[c, k, o], rows m = (row, ket tuple (o1k, o2k) of parity
//            c ^ rho), K = (phi, bra tuples of parity q ^ c ^ par(phi)), columns the bra outputs
//            (o1b, o2b) of parity c.
// v_mfma_f32_16x16x4f32: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; accumulator
// v of lane l is C[4 (l / 16) + v][l % 16]. The wave's coefficient tiles stay in registers
// for the whole block (row groups included); per row group X is staged in LDS, stage 1 writes
// T (rows x tuples x columns, odd row stride) to LDS, stage 2 scatters the Y rows over the X
// rows, and the block stores them contiguously. Row lists and offsets are tabulated in LDS
// per row group.
typedef float kfloatx4 __attribute__((ext_vector_type(4)));

struct MfmaInst {
  int kt1, nt1, kt2, nt2;
};
// every tile of an instantiation is issued (padding multiplies zeros): the D=6 shapes are exact
constexpr MfmaInst kMfmaInst[] = {{2, 1, 3, 1}, {5, 2, 5, 2}, {5, 3, 9, 2}, {8, 4, 16, 2}};
constexpr int kMfmaMaxThreads = 512;

template <int KT1, int NT1, int KT2, int NT2>
__global__ void __launch_bounds__(kMfmaMaxThreads) kron_mfma_kernel(KronArgs a) {
  int* rho = kron_rho;
  int* xrow0 = kron_xrow0;
  __shared__ int pre[2][2][33];  // [stage][class][row]: first list entry of the row
  const int s = blockIdx.y;
  const char* tb = reinterpret_cast<const char*>(a.tabs ? a.tabs[s] : a.tab_shared) + a.stage_off;
  const KronTab* tab = reinterpret_cast<const KronTab*>(tb);
  const float* K1 = reinterpret_cast<const float*>(tb + sizeof(KronTab));
  const float* K2 = K1 + 2 * a.dphi * a.j1cap * a.tno1;
  const int R = a.rows;
  float* Xl = kron_lds;
  float* Tl = Xl + R * a.rowmax;
  int* G1 = reinterpret_cast<int*>(Tl + R * a.ntup1 * a.ts);  // [2][4][KT1 * 4]
  int* E1 = G1 + 2 * 4 * KT1 * 4;                               // [2][m1cap]: X base | bra parities << 24
  int* E1t = E1 + 2 * a.m1cap;                                  // [2][m1cap]: T row offset
  int* E2t = E1t + 2 * a.m1cap;                                 // [2][m2cap]: T offset of the ket tuple
  int* E2y = E2t + 2 * a.m2cap;                                 // [2][m2cap][4]: Y base per bra parities
  const float* xs = a.x + (long long)s * a.xs;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int c = wave >= a.wc ? 1 : 0;  // class of the wave
  const int wcl = wave - c * a.wc;     // its index among the class's waves
  const int e1 = a.e[0], o1 = a.d[0] - a.e[0], e2 = a.e[1], o2 = a.d[1] - a.e[1];
  const int eo1 = a.e[2], oo1 = a.d[2] - a.e[2], eo2 = a.e[3], oo2 = a.d[3] - a.e[3];
  const int rowX = a.pd[0] * a.pf[1];
  const int rowY = (int)a.rowY;

  // X offset of ket tuple j of class cc relative to the bra tuple's base, per bra parities
  for (int i = tid; i < 2 * 4 * KT1 * 4; i += nthr) {
    const int j = i % (KT1 * 4), cb = (i / (KT1 * 4)) & 3, cc = i / (16 * KT1);
    int g = 0;
    if (j < tab->n12[cc]) {
      const int t = tab->p12[cc][j];
      const int c1k = t & 255, c2k = t >> 8;
      const int p1b = cb & 1, p2b = cb >> 1;
      const int p1k = c1k >= e1 ? 1 : 0, p2k = c2k >= e2 ? 1 : 0;
      const int r1k = c1k - (p1k ? e1 : 0), r2k = c2k - (p2k ? e2 : 0);
      g = (pair_base(p1k, p1b, e1, o1) + r1k * (p1b ? o1 : e1)) * a.pf[1] + pair_base(p2k, p2b, e2, o2) +
          r2k * (p2b ? o2 : e2) - ((p2k ^ p2b) ? a.pe[1] : 0);
    }
    G1[i] = g;
  }
  // coefficient tiles of the wave's class (zero outside the class's tuples)
  const int q = tab->q;
  float b1[KT1][NT1], b2[KT2][NT2];
  int k2o[KT2], yc[NT2];
#pragma unroll
  for (int ks = 0; ks < KT1; ++ks)
#pragma unroll
    for (int nt = 0; nt < NT1; ++nt) {
      const int j = 4 * ks + lk, n = 16 * nt + li;
      const int f = n / a.nofmax, o = n - f * a.nofmax;
      const bool v = j < tab->n12[c] && f < a.dphi && o < tab->no[q ^ c ^ (f >= a.ephi ? 1 : 0)];
      b1[ks][nt] = v ? K1[((c * a.dphi + f) * a.j1cap + j) * a.tno1 + o] : 0.f;
    }
#pragma unroll
  for (int ks = 0; ks < KT2; ++ks) {
    int j = 4 * ks + lk, f = 0;
    if (j >= tab->nj2[c][0]) {
      j -= tab->nj2[c][0];
      f = 1;
    }
    const bool v = f < a.dphi && j < tab->nj2[c][f];
    k2o[ks] = v ? tab->j2[c][f][j] * a.ts + f * a.nofmax : -1;
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
      const int n = 16 * nt + li;
      b2[ks][nt] = (v && n < tab->no[c]) ? K2[((c * a.dphi + f) * a.j2cap + j) * a.tno2 + n] : 0.f;
    }
  }
#pragma unroll
  for (int nt = 0; nt < NT2; ++nt) {
    const int n = 16 * nt + li;
    yc[nt] = -1;
    if (n < tab->no[c]) {
      const int t = tab->po[c][n];
      const int o1b = t & 255, o2b = t >> 8;
      const int q1b = o1b >= eo1 ? 1 : 0, q2b = o2b >= eo2 ? 1 : 0;
      yc[nt] = (q1b | q2b << 1) | ((o1b - (q1b ? eo1 : 0)) * a.pf[3] + (o2b - (q2b ? eo2 : 0))) << 2;
    }
  }

  for (int g = 0; g < a.groups; ++g) {
    const long long row0 = ((long long)blockIdx.x * a.groups + g) * R;
    if (row0 >= a.nrows) break;
    const int nr = (int)min((long long)R, a.nrows - row0);
    if (tid < nr) {
      const int el = a.lsplit ? a.lsplit[s] : a.leven;
      const int er = a.rsplit ? a.rsplit[s] : a.reven;
      const long long r = row0 + tid;
      const long long slab = r / a.nf;
      const int f = (int)(r - slab * a.nf);
      const int l = (int)(slab / a.rc), Rr = (int)(slab - (long long)l * a.rc);
      rho[tid] = ((l >= el) + (Rr >= er) + (f >= (a.fsplit ? a.fsplit[s] : a.ef))) & 1;
      xrow0[tid] = (int)(slab * a.sx_slab + f * a.sxf);
    }
    __syncthreads();
    if (tid < 4) {
      const int st = tid >> 1, cc = tid & 1;
      int acc = 0;
      for (int r = 0; r < nr; ++r) {
        pre[st][cc][r] = acc;
        const int pi = cc ^ rho[r];
        acc += st ? tab->no[pi] : tab->n12[pi];
      }
      pre[st][cc][nr] = acc;
    }
    kron_stage_x<4>(KRON_STAGE_ARGS(a), xs, nr, tid, nthr);  // 4 loads in flight: 3 waves per SIMD
    __syncthreads();
    // row lists of both stages
    for (int i = tid; i < 2 * nr * a.n1max; i += nthr) {
      const int cc = i / (nr * a.n1max), rem = i - cc * nr * a.n1max;
      const int r = rem / a.n1max, t = rem - r * a.n1max;
      const int pi = cc ^ rho[r];
      if (t < tab->n12[pi]) {
        const int bt = tab->p12[pi][t], c1b = bt & 255, c2b = bt >> 8;
        const int p1b = c1b >= e1 ? 1 : 0, p2b = c2b >= e2 ? 1 : 0;
        const int m = cc * a.m1cap + pre[0][cc][r] + t;
        E1[m] = (r * rowX + (c1b - (p1b ? e1 : 0)) * a.pf[1] + (c2b - (p2b ? e2 : 0))) | (p1b | p2b << 1) << 24;
        E1t[m] = (r * a.ntup1 + c1b * a.d[1] + c2b) * a.ts;
      }
    }
    for (int i = tid; i < 2 * nr * a.n2max; i += nthr) {
      const int cc = i / (nr * a.n2max), rem = i - cc * nr * a.n2max;
      const int r = rem / a.n2max, t = rem - r * a.n2max;
      const int pi = cc ^ rho[r];
      if (t < tab->no[pi]) {
        const int ot = tab->po[pi][t], o1k = ot & 255, o2k = ot >> 8;
        const int p1k = o1k >= eo1 ? 1 : 0, p2k = o2k >= eo2 ? 1 : 0;
        const int r1k = o1k - (p1k ? eo1 : 0), r2k = o2k - (p2k ? eo2 : 0);
        const int A10 = pair_base(p1k, 0, eo1, oo1) + r1k * eo1, A11 = pair_base(p1k, 1, eo1, oo1) + r1k * oo1;
        const int A20 = pair_base(p2k, 0, eo2, oo2) + r2k * eo2 - (p2k ? a.pe[3] : 0);
        const int A21 = pair_base(p2k, 1, eo2, oo2) + r2k * oo2 - (p2k ? 0 : a.pe[3]);
        const int m = cc * a.m2cap + pre[1][cc][r] + t;
        E2t[m] = r * a.ntup1 * a.ts + t;
        int* ey = E2y + 4 * m;
        const int base = r * rowY;
        ey[0] = base + A10 * a.pf[3] + A20;
        ey[1] = base + A11 * a.pf[3] + A20;
        ey[2] = base + A10 * a.pf[3] + A21;
        ey[3] = base + A11 * a.pf[3] + A21;
      }
    }
    __syncthreads();
    // ---------------- stage 1: T = X K1 ----------------
    {
      const int nM = pre[0][c][nr];
      const int ntl = (nM + 15) >> 4;
      const int n12c = tab->n12[c];
      const int* e1c = E1 + c * a.m1cap;
      const int* e1tc = E1t + c * a.m1cap;
      for (int mt = wcl; mt < ntl; mt += a.wc) {
        const int m = (mt << 4) + li;
        const bool mv = m < nM;
        const int pk = mv ? e1c[m] : 0;
        const int xb = pk & 0xffffff;
        const int* gp = G1 + (c * 4 + (pk >> 24)) * (KT1 * 4) + lk;
        float av[KT1];
#pragma unroll
        for (int ks = 0; ks < KT1; ++ks) av[ks] = (mv && 4 * ks + lk < n12c) ? Xl[xb + gp[4 * ks]] : 0.f;
        kfloatx4 acc[NT1];
#pragma unroll
        for (int nt = 0; nt < NT1; ++nt) acc[nt] = kfloatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KT1; ++ks)
#pragma unroll
          for (int nt = 0; nt < NT1; ++nt)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks], b1[ks][nt], acc[nt], 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int mo = (mt << 4) + 4 * lk + v;
          if (mo < nM) {
            float* tp = Tl + e1tc[mo] + li;
#pragma unroll
            for (int nt = 0; nt < NT1; ++nt)
              if (16 * nt + li < a.ncol1) tp[16 * nt] = acc[nt][v];
          }
        }
      }
    }
    __syncthreads();
    // ---------------- stage 2: Y = T K2, scattered over the X rows ----------------
    {
      const int nM = pre[1][c][nr];
      const int ntl = (nM + 15) >> 4;
      const int* e2tc = E2t + c * a.m2cap;
      const int* e2yc = E2y + 4 * c * a.m2cap;
      for (int mt = wcl; mt < ntl; mt += a.wc) {
        const int m = (mt << 4) + li;
        const bool mv = m < nM;
        const int tb0 = mv ? e2tc[m] : 0;
        float av[KT2];
#pragma unroll
        for (int ks = 0; ks < KT2; ++ks) av[ks] = (mv && k2o[ks] >= 0) ? Tl[tb0 + k2o[ks]] : 0.f;
        kfloatx4 acc[NT2];
#pragma unroll
        for (int nt = 0; nt < NT2; ++nt) acc[nt] = kfloatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KT2; ++ks)
#pragma unroll
          for (int nt = 0; nt < NT2; ++nt)
            acc[nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[ks], b2[ks][nt], acc[nt], 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int mo = (mt << 4) + 4 * lk + v;
          if (mo < nM) {
            const int* ey = e2yc + 4 * mo;
#pragma unroll
            for (int nt = 0; nt < NT2; ++nt)
              if (yc[nt] >= 0) Xl[ey[yc[nt] & 3] + (yc[nt] >> 2)] = acc[nt][v];
          }
        }
      }
    }
    __syncthreads();
    float* ys = a.y + (long long)s * a.ys + row0 * a.rowY;
    const int totY = nr * rowY;
    kron_copy_out(ys, Xl, totY, tid, nthr);
    __syncthreads();
  }
}

template <int KT1, int NT1, int KT2, int NT2>
static hipError_t launch_mfma(const KronArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((kron_mfma_kernel<KT1, NT1, KT2, NT2>), grid, block, lds, st, a);
  return hipGetLastError();
}

template <int NO1, int NO2>
static hipError_t launch_inst(const KronArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((kron_rows_kernel<NO1, NO2>), grid, block, lds, st, a);
  return hipGetLastError();
}

template <int NO1, int NO2>
static void raise_lds() {
  (void)hipFuncSetAttribute((const void*)kron_rows_kernel<NO1, NO2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kKronLdsBytes - 1024);
}

void kron_prepare() {
  static bool done = false;
  if (done) return;
  raise_lds<8, 8>();
  raise_lds<16, 16>();
  raise_lds<18, 18>();
  raise_lds<32, 32>();
  raise_lds<64, 64>();
  (void)hipFuncSetAttribute((const void*)kron_mfma_kernel<2, 1, 3, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kKronLdsBytes - 4096);
  (void)hipFuncSetAttribute((const void*)kron_mfma_kernel<5, 2, 5, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kKronLdsBytes - 4096);
  (void)hipFuncSetAttribute((const void*)kron_mfma_kernel<5, 3, 9, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kKronLdsBytes - 4096);
  (void)hipFuncSetAttribute((const void*)kron_mfma_kernel<8, 4, 16, 2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            kKronLdsBytes - 4096);
  done = true;
}

namespace {

// Tuple counts and the matrix-core tile counts of one stage of a site.
struct MfmaShape {
  int n12[2], no[2];
  int n1max, n2max, kt1, nt1, kt2, nt2, inst;
};

MfmaShape mfma_shape(const KronSite& ks, int stage) {
  const RoleLegs r = roles(ks, stage);
  MfmaShape m{};
  for (int pi = 0; pi < 2; ++pi) {
    m.n12[pi] = tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], pi);
    m.no[pi] = tuple_count(r.d[2], r.e[2], r.d[3], r.e[3], pi);
  }
  m.n1max = std::max(m.n12[0], m.n12[1]);
  m.n2max = std::max(m.no[0], m.no[1]);
  // stage-2 inner length of class c for either layer parity: sum over phi of n12[c ^ q ^ par(phi)]
  int k2 = 0;
  for (int cq = 0; cq < 2; ++cq) {
    int k = 0;
    for (int f = 0; f < ks.dphi; ++f) k += m.n12[cq ^ parl(f, ks.ephi)];
    k2 = std::max(k2, k);
  }
  m.kt1 = (m.n1max + 3) / 4;
  m.nt1 = (ks.dphi * m.n2max + 15) / 16;
  m.kt2 = (k2 + 3) / 4;
  m.nt2 = (m.n2max + 15) / 16;
  // the covering instantiation that issues the fewest tile steps
  m.inst = -1;
  for (int i = 0; i < (int)(sizeof(kMfmaInst) / sizeof(kMfmaInst[0])); ++i) {
    const MfmaInst& in = kMfmaInst[i];
    if (in.kt1 < m.kt1 || in.nt1 < m.nt1 || in.kt2 < m.kt2 || in.nt2 < m.nt2) continue;
    const MfmaInst* b = m.inst >= 0 ? &kMfmaInst[m.inst] : nullptr;
    if (!b || in.kt1 * in.nt1 + in.kt2 * in.nt2 < b->kt1 * b->nt1 + b->kt2 * b->nt2) m.inst = i;
  }
  if (m.n1max < 1 || m.n2max < 1) m.inst = -1;
  return m;
}

size_t mfma_lds_bytes(const KronArgs& a, int R, int KT1) {
  return sizeof(float) * ((size_t)R * a.rowmax + (size_t)R * a.ntup1 * a.ts + 2 * 4 * (size_t)KT1 * 4 +
                          4 * (size_t)R * a.n1max + 10 * (size_t)R * a.n2max);
}

}  // namespace

double kron_mfma_row_macs(const KronSite& ks, int stage) {
  const MfmaShape m = mfma_shape(ks, stage);
  if (m.inst < 0) return kron_row_macs(ks, stage);
  // per row of parity rho: n12[c ^ rho] stage-1 rows and no[c ^ rho] stage-2 rows per class,
  // each 16 x 4 multiply-adds per issued tile step and column tile (row tiles of 16 not rounded)
  double macs = 0;
  const MfmaInst& in = kMfmaInst[m.inst];
  for (int rho = 0; rho < 2; ++rho)
    for (int c = 0; c < 2; ++c)
      macs += 64.0 * ((double)m.n12[c ^ rho] * in.kt1 * in.nt1 + (double)m.no[c ^ rho] * in.kt2 * in.nt2);
  return 0.5 * macs;
}

hipError_t launch_kron_rows(const KronSite& ks, int stage, const float* x, float* y, const KronRows& rw,
                            void* const* tabs, int nb, int rows_per_block, hipStream_t st, int groups,
                            const void* tab_shared, bool mfma) {
  if (nb <= 0 || rw.nrows <= 0) return hipSuccess;
  KronArgs a{};
  const RoleLegs r = roles(ks, stage);
  for (int k = 0; k < 4; ++k) {
    a.d[k] = r.d[k];
    a.e[k] = r.e[k];
    const int o = r.d[k] - r.e[k];
    a.pd[k] = r.d[k] * r.d[k];
    a.pe[k] = r.e[k] * r.e[k] + o * o;
    a.pf[k] = std::max(a.pe[k], a.pd[k] - a.pe[k]);
  }
  a.dphi = ks.dphi;
  a.ephi = ks.ephi;
  a.nrows = rw.nrows;
  a.nf = rw.nf;
  a.ef = rw.ef;
  a.rc = rw.rc;
  a.fsplit = rw.fsplit;
  a.lsplit = rw.lsplit;
  a.leven = rw.leven;
  a.rsplit = rw.rsplit;
  a.reven = rw.reven;
  a.x = x;
  a.xs = rw.xs;
  a.sx_slab = rw.sx_slab;
  a.sxf = rw.sxf;
  a.sx1 = rw.sx1;
  a.sx2 = rw.sx2;
  a.y = y;
  a.ys = rw.ys;
  a.rowY = (long long)a.pd[2] * a.pf[3];
  a.tabs = tabs;
  a.tab_shared = tab_shared;
  if (!tabs && !tab_shared) return hipErrorInvalidValue;
  a.stage_off = (long long)stage * (long long)ks.stage_bytes;
  a.j1cap = ks.j1cap;
  a.j2cap = ks.j2cap;
  a.ntup1 = r.d[0] * r.d[1];
  a.rowmax = std::max(a.pd[0] * a.pf[1], a.pd[2] * a.pf[3]);
  int n1 = 0, n2 = 0;
  for (int pi = 0; pi < 2; ++pi) {
    n1 = std::max(n1, tuple_count(r.d[0], r.e[0], r.d[1], r.e[1], pi));
    n2 = std::max(n2, tuple_count(r.d[2], r.e[2], r.d[3], r.e[3], pi));
  }
  int rows = rows_per_block > 0 ? std::min(rows_per_block, 32) : 7;
  auto fits = [&](int R) {
    const int w1 = (R * n1 + 63) / 64, w2 = (R * n2 + 63) / 64;
    const size_t lds = sizeof(float) * ((size_t)R * a.rowmax + (size_t)R * a.ntup1 * (ks.no1 | 1));
    return 64 * 2 * std::max(w1, w2) <= kron_max_threads(ks.no1) && lds <= (size_t)kKronLdsBytes - 1024;
  };
  a.inv_rowx = 1.0f / (float)(a.pd[0] * a.pf[1]);
  a.inv_pf1 = 1.0f / (float)a.pf[1];
  // 32-bit element offsets within a sample's X
  if (rw.sx_slab * ((rw.nrows + rw.nf - 1) / rw.nf) >= (1LL << 31) || rw.sx1 >= (1LL << 31) || rw.sx2 >= (1LL << 31))
    return hipErrorInvalidValue;
  a.sx1i = (int)rw.sx1;
  a.sx2i = (int)rw.sx2;
  a.inv_pd0 = 1.0f / (float)a.pd[0];
  a.xflat = rw.sx2 == 1 && rw.sx1 == a.pf[1] && rw.sxf == (long long)a.pd[0] * a.pf[1] &&
            rw.sx_slab == (long long)rw.nf * rw.sxf;
  if (32LL * a.rowmax >= (1 << 22) || a.rowmax > 4096) return hipErrorInvalidValue;
  if (mfma) {
    const MfmaShape m = mfma_shape(ks, stage);
    if (m.inst >= 0 && m.n2max <= ks.no1 && m.n2max <= ks.no2) {
      const MfmaInst in = kMfmaInst[m.inst];
      a.tno1 = ks.no1;
      a.tno2 = ks.no2;
      a.nofmax = m.n2max;
      a.ncol1 = ks.dphi * m.n2max;
      a.ts = a.ncol1 | 1;
      a.n1max = m.n1max;
      a.n2max = m.n2max;
      int R = rows_per_block > 0 ? std::min(rows_per_block, 32) : 8;
      while (R > 1 && mfma_lds_bytes(a, R, in.kt1) > (size_t)kKronLdsBytes - 4096) --R;
      if (mfma_lds_bytes(a, R, in.kt1) <= (size_t)kKronLdsBytes - 4096) {
        a.rows = R;
        a.m1cap = R * m.n1max;
        a.m2cap = R * m.n2max;
        const int tiles = (R * std::max(m.n1max, m.n2max) + 15) / 16;
        a.wc = std::min(4, std::max(1, (tiles + 2) / 3));
        a.groups = std::max(1, std::min(groups, 64));
        const long long nblk = (rw.nrows + (long long)R * a.groups - 1) / ((long long)R * a.groups);
        if (nblk > 0x7fffffffLL || nb > 65535) return hipErrorInvalidValue;
        const size_t lds = mfma_lds_bytes(a, R, in.kt1);
        dim3 grid((unsigned)nblk, (unsigned)nb), block((unsigned)(128 * a.wc));
        switch (m.inst) {
          case 0: return launch_mfma<2, 1, 3, 1>(a, grid, block, lds, st);
          case 1: return launch_mfma<5, 2, 5, 2>(a, grid, block, lds, st);
          case 2: return launch_mfma<5, 3, 9, 2>(a, grid, block, lds, st);
          default: return launch_mfma<8, 4, 16, 2>(a, grid, block, lds, st);
        }
      }
    }
  }
  while (rows > 1 && !fits(rows)) --rows;
  if (!fits(rows)) return hipErrorInvalidValue;
  a.rows = rows;
  a.groups = 1;
  a.w1 = (rows * n1 + 63) / 64;
  a.w2 = (rows * n2 + 63) / 64;
  const size_t lds = sizeof(float) * ((size_t)rows * a.rowmax + (size_t)rows * a.ntup1 * (ks.no1 | 1));
  const long long nblk = (rw.nrows + (long long)rows * a.groups - 1) / ((long long)rows * a.groups);
  if (nblk > 0x7fffffffLL || nb > 65535) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk, (unsigned)nb), block((unsigned)(64 * 2 * std::max(a.w1, a.w2)));
  switch (ks.no1 * 100 + ks.no2) {
    case 808: return launch_inst<8, 8>(a, grid, block, lds, st);
    case 1616: return launch_inst<16, 16>(a, grid, block, lds, st);
    case 1818: return launch_inst<18, 18>(a, grid, block, lds, st);
    case 3232: return launch_inst<32, 32>(a, grid, block, lds, st);
    case 6464: return launch_inst<64, 64>(a, grid, block, lds, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace spp
