// Full-update evolution of a batch of states (src/evolution/full_update.jl):
// `evolution_step!` with FullUpdate and BoundaryMPS(chi) on the device.
//
// Every sample keeps the bond spaces it was uploaded with (the capacity); the
// sectors a sample actually uses on each bond (its "live" even/odd counts, the
// reference's truncated bond space) are tracked per sample, and every QR,
// SVD and ALS step works in the live dimensions only, so that padded slots
// stay exactly zero. Per bond update (all samples of a chunk at once):
//   fu_pair_kernel     QR of both sites into Q (exterior legs | r) and cores,
//                      the gated target (_reduced_pair, full_update.jl:21-29)
//   double layers      of the Q's (reduced layers, full_update.jl:142-148)
//   contractions       bond metric N (_reduced_metric, full_update.jl:150-172)
//   fu_core_kernel     gauge + SVD start + ALS + balance (fu_core.h)
//   fu_restore_kernel  new sites Q * core (_restore_pair, full_update.jl:128-133)
// then the sites' double layers and the rolling row message; rows are
// re-absorbed into the incremental boundary (_sweep!, full_update.jl:510-561).
#include "fu.h"
#include "../../include/stopeps_hip.h"
#include "boundary_ops.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace spp {

fu::Result fu_bond_host(const fu::Dims& g, const fu::Opts& o, const double* N4, const double* target, double* A,
                        double* B) {
  const long long cap = fu::workspace_doubles(g);
  std::vector<double> ws((size_t)cap);
  fu::Work w{ws.data(), 1, 0, cap};
  fu::View n{const_cast<double*>(N4), 1}, t{const_cast<double*>(target), 1}, a{A, 1}, b{B, 1};
  const fu::Team one{0, 1, nullptr, nullptr, 0};
  return fu::fu_bond(one, g, o, n, t, a, b, w);
}

bool jacobi_host(int kind, int m, int n, const double* A, int max_sweeps, double* values) {
  const fu::Team one{0, 1, nullptr, nullptr, 0};
  std::vector<double> a(A, A + (size_t)m * n);
  fu::View av{a.data(), 1}, w{values, 1};
  if (kind == 0) {
    std::vector<double> V((size_t)n * n), cs(3 * (size_t)(n / 2 + 1));
    return fu::jacobi_eigh(one, av, n, w, fu::View{V.data(), 1}, fu::View{cs.data(), 1}, max_sweeps);
  }
  std::vector<double> U((size_t)m * n), Vt((size_t)n * n), Vw((size_t)n * n);
  return fu::jacobi_svd_tall(one, av, m, n, fu::View{U.data(), 1}, w, fu::View{Vt.data(), 1},
                             fu::View{Vw.data(), 1}, max_sweeps);
}

namespace {

// (ket i, bra j) pair index in the parity-sorted pair order of the double layers
__host__ __device__ inline int pair_index(int i, int j, int dim, int e) {
  const int o = dim - e;
  const bool ie = i < e, je = j < e;
  if (ie && je) return i * e + j;
  if (!ie && !je) return e * e + (i - e) * o + (j - e);
  if (ie) return e * e + o * o + i * o + (j - e);
  return e * e + o * o + e * o + (i - e) * e + j;
}

FU_HD inline bool live_slot(int i, int ecap, int le, int lo) {
  return i < ecap ? i < le : (i - ecap) < lo;
}

// Site-side description of one reduced pair factor.
// rows: the three exterior virtual legs (site leg indices) and Aux; core: the
// bond leg; pseudo-site slots of the Q tensor (W,N,P,A,E,S): the core leg's
// slot carries r, P carries Aux, A has dimension 1.
struct FuSide {
  int leg[6], even[6];   // site dims / even counts (W,N,P,A,E,S)
  int rowv[3];           // exterior virtual legs
  int core;              // bond leg (1 of W,N,E,S)
  int rcap, recap;       // r capacity and its even count
  long long rows;        // capacity rows = prod of exterior dims * Aux
};

struct PairJob {
  const double* site[2];   // per sample: A, B
  const int* live[2];      // per sample: live counts [4 legs][2] of A and B
  int aux_even[2];         // per sample: even count of A's and B's auxiliary leg
};

FU_HD inline int vidx(int leg) { return leg == 0 ? 0 : leg == 1 ? 1 : leg == 4 ? 2 : 3; }

// Householder QR of both sites of one sample into pseudo-site Q tensors and
// cores, then the gated target; solved by a Team of lanes (fu_core.h). The
// sample's workspace is contiguous (work_per doubles); the cores and the
// target are interleaved by sample (stride nbtot).
FU_HD inline void fu_pair_sample(const fu::Team& t, const PairJob& J, int s, const FuSide& sa, const FuSide& sb,
                                  const double* gate, int d, int ed, double* qa, long long qa_stride, double* qb,
                                  long long qb_stride, double* ca, double* cb, double* target, int* rlive,
                                  double* work, long long work_per, int nbtot) {
  const int L = t.lane, NT = t.size;
  fu::Work w{work + (long long)s * work_per, 1, 0, work_per};
  const FuSide* sides[2] = {&sa, &sb};
  double* qout[2] = {qa + (long long)s * qa_stride, qb + (long long)s * qb_stride};
  double* cout[2] = {ca, cb};
  for (int side = 0; side < 2; ++side) {
    const FuSide& F = *sides[side];
    const double* A = J.site[side];
    const int* lv = J.live[side];
    // strides of the (W,N,P,A,E,S) site array
    long long st[6];
    st[5] = 1;
    for (int k = 4; k >= 0; --k) st[k] = st[k + 1] * F.leg[k + 1];
    // pseudo-site dims / strides
    int pd[6];
    for (int k = 0; k < 6; ++k) pd[k] = F.leg[k];
    pd[2] = F.leg[3];
    pd[3] = 1;
    pd[F.core] = F.rcap;
    long long ps[6];
    ps[5] = 1;
    for (int k = 4; k >= 0; --k) ps[k] = ps[k + 1] * pd[k + 1];
    double* Q = qout[side];
    const long long qsize = ps[0] * pd[0];
    for (long long i = L; i < qsize; i += NT) Q[i] = 0.0;
    const int ncore = F.leg[F.core];
    // cores C[r][p][k] (interleaved)
    fu::View C{cout[side] + s, nbtot};
    for (long long i = L; i < (long long)F.rcap * d * ncore; i += NT) C[i] = 0.0;
    t.sync();
    const int dv0 = F.leg[F.rowv[0]], dv1 = F.leg[F.rowv[1]], dv2 = F.leg[F.rowv[2]], dax = F.leg[3];
    const int l0 = F.rowv[0], l1 = F.rowv[1], l2 = F.rowv[2];
    auto row_live = [&](int i0, int i1, int i2) {
      return live_slot(i0, F.even[l0], lv[2 * vidx(l0)], lv[2 * vidx(l0) + 1]) &&
             live_slot(i1, F.even[l1], lv[2 * vidx(l1)], lv[2 * vidx(l1) + 1]) &&
             live_slot(i2, F.even[l2], lv[2 * vidx(l2)], lv[2 * vidx(l2) + 1]);
    };
    auto row_par = [&](int i0, int i1, int i2, int ax) {
      return fu::par(i0, F.even[l0]) ^ fu::par(i1, F.even[l1]) ^ fu::par(i2, F.even[l2]) ^ fu::par(ax, J.aux_even[side]);
    };
    auto col_live = [&](int k) {
      return live_slot(k, F.even[F.core], lv[2 * vidx(F.core)], lv[2 * vidx(F.core) + 1]);
    };
    int rl[2] = {0, 0};
    for (int q = 0; q < 2; ++q) {
      const long long mark = w.top;
      // live rows / columns of parity q (every lane counts)
      int mq = 0, nq = 0;
      for (int i0 = 0; i0 < dv0; ++i0)
        for (int i1 = 0; i1 < dv1; ++i1)
          for (int i2 = 0; i2 < dv2; ++i2)
            for (int ax = 0; ax < dax; ++ax)
              if (row_live(i0, i1, i2)) mq += row_par(i0, i1, i2, ax) == q;
      for (int p = 0; p < d; ++p)
        for (int k = 0; k < ncore; ++k)
          if (col_live(k)) nq += (fu::par(p, ed) ^ fu::par(k, F.even[F.core])) == q;
      if (mq == 0 || nq == 0) continue;
      fu::View M = w.take((long long)mq * nq);
      fu::View rowid = w.take(mq), colid = w.take(nq);
      if (L == 0) {
        int bi = 0;
        for (int i0 = 0; i0 < dv0; ++i0)
          for (int i1 = 0; i1 < dv1; ++i1)
            for (int i2 = 0; i2 < dv2; ++i2)
              for (int ax = 0; ax < dax; ++ax) {
                if (!row_live(i0, i1, i2) || row_par(i0, i1, i2, ax) != q) continue;
                rowid[bi] = (double)(((i0 * dv1 + i1) * dv2 + i2) * dax + ax);
                ++bi;
              }
        int bj = 0;
        for (int p = 0; p < d; ++p)
          for (int k = 0; k < ncore; ++k) {
            if (!col_live(k) || (fu::par(p, ed) ^ fu::par(k, F.even[F.core])) != q) continue;
            colid[bj] = (double)(p * ncore + k);
            ++bj;
          }
      }
      t.sync();
      for (int bi = L; bi < mq; bi += NT) {
        int rid = (int)rowid[bi];
        const int ax = rid % dax;
        rid /= dax;
        const int i2 = rid % dv2;
        rid /= dv2;
        const int i1 = rid % dv1;
        const int i0 = rid / dv1;
        long long base = (long long)ax * st[3] + (long long)i0 * st[F.rowv[0]] + (long long)i1 * st[F.rowv[1]] +
                         (long long)i2 * st[F.rowv[2]];
        for (int bj = 0; bj < nq; ++bj) {
          const int cid = (int)colid[bj];
          const int p = cid / ncore, k = cid % ncore;
          M[(long long)bi * nq + bj] = A[base + (long long)p * st[2] + (long long)k * st[F.core]];
        }
      }
      const int kq = mq < nq ? mq : nq;
      fu::View Qb = w.take((long long)mq * kq), Rb = w.take((long long)kq * nq);
      fu::View vs = w.take((long long)mq * kq), tau = w.take(kq);
      t.sync();
      fu::householder_qr(t, M, mq, nq, Qb, Rb, vs, tau);
      const int off = q == 0 ? 0 : F.recap;
      rl[q] = kq;
      for (int bi = L; bi < mq; bi += NT) {
        int rid = (int)rowid[bi];
        const int ax = rid % dax;
        rid /= dax;
        const int i2 = rid % dv2;
        rid /= dv2;
        const int i1 = rid % dv1;
        const int i0 = rid / dv1;
        for (int c = 0; c < kq; ++c) {
          int idx[6] = {0, 0, 0, 0, 0, 0};
          idx[F.rowv[0]] = i0;
          idx[F.rowv[1]] = i1;
          idx[F.rowv[2]] = i2;
          idx[2] = ax;
          idx[F.core] = off + c;
          long long o = 0;
          for (int k = 0; k < 6; ++k) o += (long long)idx[k] * ps[k];
          Q[o] = Qb[(long long)bi * kq + c];
        }
      }
      for (long long idx = L; idx < (long long)kq * nq; idx += NT) {
        const int c = (int)(idx / nq), bj = (int)(idx % nq);
        const int cid = (int)colid[bj];
        const int p = cid / ncore, k = cid % ncore;
        C[((long long)(off + c) * d + p) * ncore + k] = Rb[idx];
      }
      t.sync();
      w.top = mark;
    }
    if (L == 0) {
      rlive[(long long)s * 4 + side * 2] = rl[0];
      rlive[(long long)s * 4 + side * 2 + 1] = rl[1];
    }
  }
  t.sync();
  // target[r,p,q,s] = sum_{u,k,v} Ca[r,u,k] Cb[s,v,k] gate[p,q,u,v]
  const int ra = sa.rcap, rb = sb.rcap, nk = sa.leg[sa.core];
  fu::View Ca{ca + s, nbtot}, Cb{cb + s, nbtot}, Tg{target + s, nbtot};
  for (long long idx = L; idx < (long long)ra * d * d * rb; idx += NT) {
    const int tt = (int)(idx % rb);
    const long long rpq = idx / rb;
    const int q = (int)(rpq % d), p = (int)((rpq / d) % d), r = (int)(rpq / ((long long)d * d));
    double acc = 0.0;
    for (int u = 0; u < d; ++u)
      for (int v = 0; v < d; ++v) {
        const double gv = gate[((p * d + q) * d + u) * d + v];
        if (gv == 0.0) continue;
        double ab = 0.0;
        for (int k = 0; k < nk; ++k) ab += Ca[((long long)r * d + u) * nk + k] * Cb[((long long)tt * d + v) * nk + k];
        acc += gv * ab;
      }
    Tg[idx] = acc;
  }
  t.sync();
}

__global__ void __launch_bounds__(64)
fu_pair_kernel(const PairJob* jobs, int nb, FuSide sa, FuSide sb, const double* gate, int d, int ed, double* qa,
               long long qa_stride, double* qb, long long qb_stride, double* ca, double* cb, double* target,
               int* rlive, double* work, long long work_per, int nbtot) {
  const int s = blockIdx.x;
  if (s >= nb) return;
  __shared__ double red[3];
  const fu::Team tm{(int)threadIdx.x, 64, red, nullptr, 0};
  fu_pair_sample(tm, jobs[s], s, sa, sb, gate, d, ed, qa, qa_stride, qb, qb_stride, ca, cb, target, rlive, work,
                 work_per, nbtot);
}

// Bond solve of every sample, one 512-lane workgroup per sample (fu_core.h's
// Team): compact the metric and the target to the live r dimensions, solve,
// expand the cores. Each sample's workspace is contiguous (work_per doubles).
constexpr int kCoreLanes = 512;

template <typename T>
__global__ void __launch_bounds__(kCoreLanes)
fu_core_kernel(const T* nf, long long nf_stride, const double* target, const int* rlive, int ra, int era, int rb,
               int erb, int d, int ed, int D, int eD, fu::Opts o, double* aout, double* bout, double* rec,
               int* bond_live, double* work, long long work_per, int nb, int nbtot, long long lds_doubles,
               int s_first) {
  const int s = blockIdx.x + s_first;  // s_first > 0: the solve of one sample again (PSD repair)
  if (s >= nb) return;
  __shared__ double red[3 * (kCoreLanes / 64)];
  extern __shared__ double fu_lds[];
  const fu::Team tm{(int)threadIdx.x, kCoreLanes, red, fu_lds, lds_doubles};
  const int L = tm.lane;
  const int la[2] = {rlive[(long long)s * 4], rlive[(long long)s * 4 + 1]};
  const int lb[2] = {rlive[(long long)s * 4 + 2], rlive[(long long)s * 4 + 3]};
  fu::Dims g{la[0] + la[1], la[0], lb[0] + lb[1], lb[0], d, ed, D, eD};
  fu::Work w{work + (long long)s * work_per, 1, 0, work_per};
  const long long na = g.na, nbk = g.nb;
  fu::View N4 = w.take(na * nbk * na * nbk), tg = w.take(na * d * d * nbk);
  fu::View A = w.take(na * d * D), B = w.take((long long)D * d * nbk);
  auto slot_a = [&](int i) { return i < la[0] ? i : era + (i - la[0]); };
  auto slot_b = [&](int i) { return i < lb[0] ? i : erb + (i - lb[0]); };
  const T* N = nf + (long long)s * nf_stride;
  const int nbr = rb * rb;
  // N[rb,sb,rk,sk] = Nf[pair(rk, rb)][pair(sk, sb)]. The contraction kernel
  // writes only the parity-allowed blocks of its output (pair parities equal);
  // the others hold whatever the buffer held before and are zero by symmetry.
  const long long nn = na * nbk * na * nbk;
  for (long long idx = L; idx < nn; idx += kCoreLanes) {
    const int l = (int)(idx % nbk), k = (int)((idx / nbk) % na);
    const int j = (int)((idx / (nbk * na)) % nbk), i = (int)(idx / (nbk * na * nbk));
    const int pR = fu::par(i, la[0]) ^ fu::par(k, la[0]);
    const int pS = fu::par(j, lb[0]) ^ fu::par(l, lb[0]);
    double v = 0.0;
    if (pR == pS) {
      const int R = pair_index(slot_a(k), slot_a(i), ra, era);
      const int S = pair_index(slot_b(l), slot_b(j), rb, erb);
      v = (double)N[(long long)R * nbr + S];
    }
    N4[idx] = v;
  }
  fu::View Tg{const_cast<double*>(target) + s, nbtot};
  const long long nt = na * d * d * nbk;
  for (long long idx = L; idx < nt; idx += kCoreLanes) {
    const int t = (int)(idx % nbk);
    const long long rpq = idx / nbk;
    const int q = (int)(rpq % d), p = (int)((rpq / d) % d), r = (int)(rpq / ((long long)d * d));
    tg[idx] = Tg[(((long long)slot_a(r) * d + p) * d + q) * rb + slot_b(t)];
  }
  tm.sync();
  fu::Work inner{w.base, 1, w.top, work_per - w.top};
  const fu::Result res = fu::fu_bond(tm, g, o, N4, tg, A, B, inner);
  fu::View Ao{aout + s, nbtot}, Bo{bout + s, nbtot};
  for (long long i = L; i < (long long)ra * d * D; i += kCoreLanes) Ao[i] = 0.0;
  for (long long i = L; i < (long long)D * d * rb; i += kCoreLanes) Bo[i] = 0.0;
  tm.sync();
  if (res.status == 0) {
    for (long long idx = L; idx < na * d * D; idx += kCoreLanes) {
      const int r = (int)(idx / ((long long)d * D));
      Ao[(long long)slot_a(r) * d * D + idx % ((long long)d * D)] = A[idx];
    }
    for (long long idx = L; idx < (long long)D * d * nbk; idx += kCoreLanes) {
      const long long kq = idx / nbk;
      Bo[kq * rb + slot_b((int)(idx % nbk))] = B[idx];
    }
  }
  if (L == 0) {
    double* rr = rec + (long long)s * 12;
    rr[0] = res.initial_cost;
    rr[1] = res.final_cost;
    rr[2] = res.residual;
    rr[3] = res.target_norm;
    rr[4] = res.iterations;
    rr[5] = res.limit_reached;
    rr[6] = res.stop;
    rr[7] = res.status;
    rr[8] = res.kept_even;
    rr[9] = res.kept_odd;
    rr[10] = res.floored;
    rr[11] = res.nonconverged;
    bond_live[(long long)s * 2] = res.kept_even;
    bond_live[(long long)s * 2 + 1] = res.kept_odd;
  }
}

// New site = Q * core, written into the state's site array (thread per entry).
struct RestoreJob {
  double* site[2];
};

__global__ void fu_restore_kernel(const RestoreJob* jobs, int nb, FuSide sa, FuSide sb, int d, const double* qa,
                                  long long qa_stride, const double* qb, long long qb_stride, const double* aout,
                                  const double* bout, int D, int nbtot, const int* failed) {
  const int s = blockIdx.y;
  const int side = blockIdx.z;
  if (s >= nb || failed[s]) return;  // a failed sample keeps its sites
  const FuSide& F = side == 0 ? sa : sb;
  long long total = 1;
  for (int k = 0; k < 6; ++k) total *= F.leg[k];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int idx[6];
  long long r = t;
  for (int k = 5; k >= 0; --k) {
    idx[k] = (int)(r % F.leg[k]);
    r /= F.leg[k];
  }
  int pd[6];
  for (int k = 0; k < 6; ++k) pd[k] = F.leg[k];
  pd[2] = F.leg[3];
  pd[3] = 1;
  pd[F.core] = F.rcap;
  long long ps[6];
  ps[5] = 1;
  for (int k = 4; k >= 0; --k) ps[k] = ps[k + 1] * pd[k + 1];
  int qi[6] = {idx[0], idx[1], idx[3], 0, idx[4], idx[5]};
  qi[F.core] = 0;
  long long qbase = 0;
  for (int k = 0; k < 6; ++k) qbase += (long long)qi[k] * ps[k];
  const double* Q = side == 0 ? qa + (long long)s * qa_stride : qb + (long long)s * qb_stride;
  const int p = idx[2], k = idx[F.core];
  double acc = 0.0;
  if (side == 0) {
    fu::View Ao{const_cast<double*>(aout) + s, nbtot};
    for (int c = 0; c < F.rcap; ++c) acc += Q[qbase + (long long)c * ps[F.core]] * Ao[((long long)c * d + p) * D + k];
  } else {
    fu::View Bo{const_cast<double*>(bout) + s, nbtot};
    for (int c = 0; c < F.rcap; ++c) acc += Q[qbase + (long long)c * ps[F.core]] * Bo[((long long)k * d + p) * F.rcap + c];
  }
  jobs[s].site[side][t] = acc;
}

// live[(s*ns + site)*8 + v*2 + q] <- bond_live[s*2 + q] for the bond legs of A and B
__global__ void fu_set_live_kernel(int* live, int ns, int site_a, int va, int site_b, int vb, const int* bond_live,
                                   int nb, const int* failed) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb || failed[s]) return;
  for (int q = 0; q < 2; ++q) {
    live[((long long)s * ns + site_a) * 8 + va * 2 + q] = bond_live[(long long)s * 2 + q];
    live[((long long)s * ns + site_b) * 8 + vb * 2 + q] = bond_live[(long long)s * 2 + q];
  }
}

// sum of squares of each sample's site array (block per sample)
// A[.., p, ..] <- sum_u G[p][u] A[.., u, ..] on the physical leg (leg 2) of one
// site per sample (blockIdx.x); outer = dims of legs 0-1, inner = legs 3-5.
__global__ void onsite_gate_kernel(double* const* p, long long outer, int d, long long inner, const double* G) {
  double* a = p[blockIdx.x];
  for (long long t = threadIdx.x; t < outer * inner; t += blockDim.x) {
    const long long o = t / inner, i = t % inner;
    double v[8];
    for (int u = 0; u < d; ++u) v[u] = a[(o * d + u) * inner + i];
    for (int q = 0; q < d; ++q) {
      double acc = 0.0;
      for (int u = 0; u < d; ++u) acc += G[q * d + u] * v[u];
      a[(o * d + q) * inner + i] = acc;
    }
  }
}

__global__ void site_sumsq_kernel(double* const* p, long long n, double* out) {
  const double* a = p[blockIdx.x];
  __shared__ double part[256];
  double acc = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) acc += a[i] * a[i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0];
}

__global__ void site_scale_kernel(double* const* p, long long n, const double* f) {
  double* a = p[blockIdx.x];
  const double v = f[blockIdx.x];
  for (long long i = (long long)blockIdx.y * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.y * blockDim.x)
    a[i] *= v;
}

long long count_legs(const std::vector<Leg>& l) {
  long long n = 1;
  for (auto& g : l) n *= g.dim;
  return n;
}

// any_aux: size the r capacity for either parity split of the auxiliary leg
// (a batch whose samples carry different auxiliary charges).
FuSide make_side(const SiteDev& sd, int rowv0, int rowv1, int rowv2, int core, bool any_aux = false) {
  FuSide F{};
  for (int k = 0; k < 6; ++k) {
    F.leg[k] = sd.dim[k];
    F.even[k] = sd.even[k];
  }
  F.rowv[0] = rowv0;
  F.rowv[1] = rowv1;
  F.rowv[2] = rowv2;
  F.core = core;
  // r capacity per parity: min(rows of parity q, columns of parity q)
  long long cq[2] = {0, 0};
  for (int p = 0; p < sd.dim[2]; ++p)
    for (int k = 0; k < sd.dim[core]; ++k) ++cq[fu::par(p, sd.even[2]) ^ fu::par(k, sd.even[core])];
  long long k0 = 0, k1 = 0;
  for (int flip = 0; flip < (any_aux ? 2 : 1); ++flip) {
    const int ea = flip ? sd.dim[3] - sd.even[3] : sd.even[3];
    long long rq[2] = {0, 0};
    for (int a = 0; a < sd.dim[rowv0]; ++a)
      for (int b = 0; b < sd.dim[rowv1]; ++b)
        for (int c = 0; c < sd.dim[rowv2]; ++c)
          for (int x = 0; x < sd.dim[3]; ++x)
            ++rq[fu::par(a, sd.even[rowv0]) ^ fu::par(b, sd.even[rowv1]) ^ fu::par(c, sd.even[rowv2]) ^
                 fu::par(x, ea)];
    k0 = std::max(k0, std::min(rq[0], cq[0]));
    k1 = std::max(k1, std::min(rq[1], cq[1]));
  }
  F.recap = (int)k0;
  F.rcap = (int)(k0 + k1);
  F.rows = (long long)sd.dim[rowv0] * sd.dim[rowv1] * sd.dim[rowv2] * sd.dim[3];
  return F;
}

}  // namespace

// Host run of the reduced-pair stage for one sample (selftest).
int fu_pair_host(int horizontal, const int* dims_a, const int* evens_a, const int* dims_b, const int* evens_b,
                 const double* site_a, const double* site_b, const int* live_a, const int* live_b, const double* gate,
                 double* qa, double* qb, double* ca, double* cb, double* target, int* geometry) {
  SiteDev A, B;
  for (int k = 0; k < 6; ++k) {
    A.dim[k] = dims_a[k];
    A.even[k] = evens_a[k];
    B.dim[k] = dims_b[k];
    B.even[k] = evens_b[k];
  }
  FuSide sa = horizontal ? make_side(A, 0, 1, 5, 4) : make_side(A, 0, 1, 4, 5);
  FuSide sb = horizontal ? make_side(B, 1, 4, 5, 0) : make_side(B, 0, 4, 5, 1);
  A.d = nullptr;
  B.d = nullptr;
  const int d = A.dim[2], ed = A.even[2];
  long long work = 64;
  for (const FuSide* F : {&sa, &sb}) {
    const long long cols = (long long)d * F->leg[F->core];
    work = std::max(work, 3 * F->rows * cols + 2 * F->rows + 2 * cols + 64);
  }
  std::vector<double> ws((size_t)work);
  PairJob J{};
  J.site[0] = site_a;
  J.site[1] = site_b;
  J.live[0] = live_a;
  J.live[1] = live_b;
  J.aux_even[0] = evens_a[3];
  J.aux_even[1] = evens_b[3];
  int rl[4] = {0, 0, 0, 0};
  const fu::Team one{0, 1, nullptr, nullptr, 0};
  fu_pair_sample(one, J, 0, sa, sb, gate, d, ed, qa, 0, qb, 0, ca, cb, target, rl, ws.data(), work, 1);
  geometry[0] = sa.rcap;
  geometry[1] = sa.recap;
  geometry[2] = sb.rcap;
  geometry[3] = sb.recap;
  for (int k = 0; k < 4; ++k) geometry[4 + k] = rl[k];
  return 0;
}

// ---------------------------------------------------------------------------
// the evolution step
// ---------------------------------------------------------------------------

// One-site lattice (full_update.jl:568-570): A <- exp(-delta h_onsite) on the
// physical leg, then normalize_state!. The gate is P.gates_h [d][d].
static void onsite_step(Context& ctx, const std::vector<StateDev*>& states, const FuParams& P, FuOutputs& out) {
  hipStream_t st = ctx.stream;
  const int n = (int)states.size();
  const SiteDev& s0 = states[0]->sites[0];
  const int d = s0.dim[2], ed = s0.even[2];
  if (!P.gates_h) throw HipError("full update: the onsite gate is required");
  if (d > 8) throw HipError("full update: physical dimension above 8");
  for (int q = 0; q < d; ++q)
    for (int u = 0; u < d; ++u)
      if ((q >= ed) != (u >= ed) && P.gates_h[q * d + u] != 0.0)
        throw HipError("full update: the onsite gate does not respect the Z2 grading of the physical space");
  const long long outer = (long long)s0.dim[0] * s0.dim[1];
  const long long inner = (long long)s0.dim[3] * s0.dim[4] * s0.dim[5];
  const long long size = s0.size();
  std::vector<double*> hp(n);
  for (int s = 0; s < n; ++s) hp[s] = states[s]->sites[0].d;
  double** dp = nullptr;
  double *dg = nullptr, *dn = nullptr;
  auto release = [&]() {
    if (dp) (void)hipFree(dp);
    if (dg) (void)hipFree(dg);
    if (dn) (void)hipFree(dn);
  };
  try {
    SPP_CHECK(hipMalloc((void**)&dp, sizeof(double*) * n));
    SPP_CHECK(hipMalloc((void**)&dg, sizeof(double) * d * d));
    SPP_CHECK(hipMalloc((void**)&dn, sizeof(double) * n));
    upload(dp, hp.data(), sizeof(double*) * n, st);
    upload(dg, P.gates_h, sizeof(double) * d * d, st);
    hipLaunchKernelGGL(onsite_gate_kernel, dim3(n), dim3(256), 0, st, dp, outer, d, inner, (const double*)dg);
    SPP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(site_sumsq_kernel, dim3(n), dim3(256), 0, st, dp, size, dn);
    SPP_CHECK(hipGetLastError());
    std::vector<double> hn(n);
    SPP_CHECK(hipMemcpyAsync(hn.data(), dn, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
    for (int s = 0; s < n; ++s) {
      const double v = std::sqrt(hn[s]);
      if (!(v > 0.0 && std::isfinite(v))) throw HipError("zero or nonfinite tensor after the full update");
      states[s]->log_amplitude += std::log(v);
      if (out.normalizations) out.normalizations[s] = std::log(v);
      if (out.max_env_trunc) out.max_env_trunc[s] = 0.0;
      if (out.max_entropy) out.max_entropy[s] = 0.0;  // no boundary cuts
      states[s]->beta += 2.0 * P.delta;
      hn[s] = 1.0 / v;
    }
    upload(dn, hn.data(), sizeof(double) * n, st);
    hipLaunchKernelGGL(site_scale_kernel, dim3(n, (unsigned)std::min<long long>((size + 255) / 256, 64)), dim3(256),
                       0, st, dp, size, (const double*)dn);
    SPP_CHECK(hipGetLastError());
    SPP_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamSynchronize(st);
    release();
    throw;
  }
  release();
}

template <typename T>
void fu_evolution_step(Context& ctx, const std::vector<StateDev*>& states, int chi, const FuParams& P,
                       FuOutputs& out) {
  hipStream_t st = ctx.stream;
  const int n = (int)states.size();
  if (n == 0) return;
  if (chi < 1) throw HipError("chi must be positive");
  const StateDev& s0 = *states[0];
  const int Lx = s0.Lx, Ly = s0.Ly, ns = Lx * Ly;
  for (const StateDev* sd : states) {
    if (!sd || sd->Lx != Lx || sd->Ly != Ly || (int)sd->sites.size() != ns)
      throw HipError("full update: states must share the lattice");
    for (int i = 0; i < ns; ++i)
      if (!batchable_sites(sd->sites[i], s0.sites[i]))
        throw HipError("full update: states must share every leg space (up to the auxiliary charge)");
    if ((int)sd->live.size() != ns * 8) throw HipError("full update: state without bond bookkeeping");
  }
  const int d = s0.sites[0].dim[2], ed = s0.sites[0].even[2];
  for (int i = 0; i < ns; ++i)
    if (s0.sites[i].dim[2] != d || s0.sites[i].even[2] != ed) throw HipError("full update: physical legs differ");
  if (ns == 1) {
    onsite_step(ctx, states, P, out);
    if (out.status)
      for (int s = 0; s < n; ++s) out.status[s] = 0;
    return;
  }
  const int nh = Ly * (Lx - 1), nv = (Ly - 1) * Lx;
  if (!P.gates_h || !P.gates_v) throw HipError("full update: gates are required");
  // leg specs of the double layers (fixed capacities)
  std::vector<LayerSpec> spec(ns);
  std::vector<long long> lsize(ns);
  long long D = 1, maxout = 1, layer_elems = 0;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = s0.sites[i];
    spec[i] = LayerSpec{pair_leg(s.dim[0], s.even[0]), pair_leg(s.dim[1], s.even[1]), pair_leg(s.dim[4], s.even[4]),
                        pair_leg(s.dim[5], s.even[5])};
    lsize[i] = (long long)spec[i].w.dim * spec[i].n.dim * spec[i].e.dim * spec[i].s.dim;
    layer_elems += lsize[i];
    maxout = std::max(maxout, lsize[i]);
    D = std::max<long long>(D, std::max(std::max(spec[i].w.dim, spec[i].n.dim), std::max(spec[i].e.dim, spec[i].s.dim)));
  }
  long long pmax = 1;
  for (const LayerSpec& q : spec) pmax = std::max<long long>(pmax, std::max(q.n.dim, q.s.dim));
  const long long cap = (long long)chi * pmax * chi;
  // the sides of every bond: capacities of r, Q sizes, reduced layer sizes
  struct BondGeo {
    int ia, ib;
    FuSide sa, sb;
    int Dcap, eDcap;
    long long qa_size, qb_size, la_size, lb_size;
    LayerSpec la, lb;
  };
  // doubles of LDS the bond solve stages: the larger gauge eigenproblem block
  // with its eigenvectors and rotations, or one normal-matrix operand
  auto core_lds_doubles = [](const BondGeo& G) -> long long {
    const long long ra = G.sa.rcap, era = G.sa.recap, rb = G.sb.rcap, erb = G.sb.recap;
    const long long c0 = era * erb + (ra - era) * (rb - erb), c1 = era * (rb - erb) + (ra - era) * erb;
    const long long m = std::max(c0, c1);
    const long long nl = ra * G.Dcap, nr = G.Dcap * rb;
    return std::max(2 * m * m + 3 * (m / 2 + 1), std::max(nl * nl, nr * nr));
  };
  auto geo = [&](int ia, int ib, bool horizontal) {
    BondGeo g{};
    g.ia = ia;
    g.ib = ib;
    const SiteDev& A = s0.sites[ia];
    const SiteDev& B = s0.sites[ib];
    bool mixed_a = false, mixed_b = false;
    for (const StateDev* sd : states) {
      mixed_a = mixed_a || sd->sites[ia].even[3] != A.even[3];
      mixed_b = mixed_b || sd->sites[ib].even[3] != B.even[3];
    }
    if (horizontal) {
      g.sa = make_side(A, 0, 1, 5, 4, mixed_a);
      g.sb = make_side(B, 1, 4, 5, 0, mixed_b);
      g.Dcap = A.dim[4];
      g.eDcap = A.even[4];
    } else {
      g.sa = make_side(A, 0, 1, 4, 5, mixed_a);
      g.sb = make_side(B, 0, 4, 5, 1, mixed_b);
      g.Dcap = A.dim[5];
      g.eDcap = A.even[5];
    }
    auto qsize = [&](const FuSide& F) {
      long long v = 1;
      for (int k = 0; k < 6; ++k) v *= (k == F.core ? F.rcap : k == 3 ? 1 : k == 2 ? F.leg[3] : F.leg[k]);
      return v;
    };
    g.qa_size = qsize(g.sa);
    g.qb_size = qsize(g.sb);
    auto lspec = [&](const FuSide& F) {
      int dm[4] = {F.leg[0], F.leg[1], F.leg[4], F.leg[5]}, ev[4] = {F.even[0], F.even[1], F.even[4], F.even[5]};
      const int ci = F.core == 0 ? 0 : F.core == 1 ? 1 : F.core == 4 ? 2 : 3;
      dm[ci] = F.rcap;
      ev[ci] = F.recap;
      return LayerSpec{pair_leg(dm[0], ev[0]), pair_leg(dm[1], ev[1]), pair_leg(dm[2], ev[2]), pair_leg(dm[3], ev[3])};
    };
    g.la = lspec(g.sa);
    g.lb = lspec(g.sb);
    g.la_size = (long long)g.la.w.dim * g.la.n.dim * g.la.e.dim * g.la.s.dim;
    g.lb_size = (long long)g.lb.w.dim * g.lb.n.dim * g.lb.e.dim * g.lb.s.dim;
    return g;
  };
  std::vector<BondGeo> hb(nh), vb(nv);
  for (int y = 0; y < Ly; ++y)
    for (int x = 0; x + 1 < Lx; ++x) hb[y * (Lx - 1) + x] = geo(x + y * Lx, x + 1 + y * Lx, true);
  for (int y = 0; y + 1 < Ly; ++y)
    for (int x = 0; x < Lx; ++x) vb[y * Lx + x] = geo(x + y * Lx, x + (y + 1) * Lx, false);
  long long qmax = 1, lredmax = 1, rmax = 1, coremax = 1, cmax = 1;
  for (const auto* set : {&hb, &vb})
    for (const BondGeo& g : *set) {
      qmax = std::max(qmax, std::max(g.qa_size, g.qb_size));
      lredmax = std::max(lredmax, std::max(g.la_size, g.lb_size));
      rmax = std::max<long long>(rmax, std::max(g.sa.rcap, g.sb.rcap));
      fu::Dims cg{g.sa.rcap, g.sa.recap, g.sb.rcap, g.sb.recap, d, ed, g.Dcap, g.eDcap};
      coremax = std::max(coremax, fu::workspace_doubles(cg) + 2 * (long long)cg.na * cg.nb * cg.na * cg.nb +
                                      4 * (long long)cg.na * d * d * cg.nb + 4 * (long long)cg.na * d * cg.D +
                                      4 * (long long)cg.D * d * cg.nb);
      cmax = std::max<long long>(cmax, (long long)std::max(g.sa.rcap, g.sb.rcap) * d * g.Dcap);
    }
  long long pairwork = 64;
  for (const auto* set : {&hb, &vb})
    for (const BondGeo& g : *set)
      for (const FuSide* F : {&g.sa, &g.sb}) {
        const long long cols = (long long)d * F->leg[F->core];
        pairwork = std::max(pairwork, 3 * F->rows * cols + 2 * F->rows + 2 * cols + 64);
      }
  const long long tsz = rmax * d * d * rmax;
  const long long nfs = rmax * rmax * rmax * rmax;
  // ---- chunk size from memory ----
  const double ts = sizeof(T);
  const double small = (double)chi * D * chi, mid = small * D;
  const double slab = mid * (double)rmax * rmax * D / std::max<long long>(D, 1);
  const double row_b = ts * ((Lx + 6) * mid + 2.0 * small + 4.0 * mid * D + 4.0 * chi * chi * D * rmax * rmax +
                             slab * 2.0);
  const size_t per_arena =
      (size_t)(1.2 * std::max(row_b, absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub))) + 16384;
  const size_t per_persist = (size_t)(ts * (layer_elems + 2.0 * (Ly + 1) * Lx * cap + 2.0 * lredmax + nfs)) +
                             sizeof(double) * (2 * qmax + 3 * tsz + 4 * cmax + pairwork + coremax + 64) +
                             sizeof(int) * (ns * 8 + 16) + sizeof(LayerJob) * ns + 1024;
  SPP_CHECK(hipStreamSynchronize(st));
  size_t freeb = 0, totalb = 0;
  SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
  const double avail = 0.6 * (double)freeb + (double)ctx.arena.capacity();
  long long chunk = std::max<long long>(1, (long long)(avail / (double)(per_arena + per_persist)));
  if (ctx.chunk > 0) chunk = std::min<long long>(chunk, ctx.chunk);
  chunk = std::min<long long>(chunk, n);
  chunk = std::min<long long>(chunk, std::max(1, 65535 / ns));
  const int nb = (int)chunk;
  ctx.arena.reserve(per_arena * (size_t)nb);

  std::vector<Boundary<T>> north(Ly + 1), south(Ly + 1);
  std::vector<void*> scratch;
  auto dev = [&](size_t bytes) {
    void* p = nullptr;
    SPP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    scratch.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    entropy_reset(ctx);  // pending cuts write into buffers freed below
    for (auto& b : north) b.release();
    for (auto& b : south) b.release();
    for (void* p : scratch) (void)hipFree(p);
    scratch.clear();
  };
  const int nrec_bonds = P.order == 2 ? 2 * (nh + nv) : nh + nv;
  int first_failure = 0;  // first per-sample failure code of the call (reported when no status array)
  try {
    T* layers = (T*)dev(sizeof(T) * (size_t)layer_elems * nb);
    std::vector<T*> site_base(ns);
    {
      long long off = 0;
      for (int i = 0; i < ns; ++i) {
        site_base[i] = layers + off * nb;
        off += lsize[i];
      }
    }
    void** ptrs = (void**)dev(sizeof(void*) * (size_t)ns * nb);
    {
      std::vector<void*> hptr((size_t)ns * nb);
      for (int i = 0; i < ns; ++i)
        for (int s = 0; s < nb; ++s) hptr[(size_t)i * nb + s] = site_base[i] + (long long)s * lsize[i];
      upload(ptrs, hptr.data(), sizeof(void*) * hptr.size(), st);
      SPP_CHECK(hipStreamSynchronize(st));  // hptr is released at the end of this block
    }
    auto table = [&](int site) { return (void* const*)(ptrs + (size_t)site * nb); };
    LayerJob* djobs = (LayerJob*)dev(sizeof(LayerJob) * (size_t)std::max(ns, 2) * nb);
    for (int y = 0; y <= Ly; ++y) {
      prepare_boundary(north[y], Lx, nb, cap);
      prepare_boundary(south[y], Lx, nb, cap);
    }
    double* qa = (double*)dev(sizeof(double) * (size_t)qmax * nb);
    double* qbuf = (double*)dev(sizeof(double) * (size_t)qmax * nb);
    T* lred_a = (T*)dev(sizeof(T) * (size_t)lredmax * nb);
    T* lred_b = (T*)dev(sizeof(T) * (size_t)lredmax * nb);
    T* nfused = (T*)dev(sizeof(T) * (size_t)nfs * nb);
    double* ca = (double*)dev(sizeof(double) * (size_t)cmax * nb);
    double* cb = (double*)dev(sizeof(double) * (size_t)cmax * nb);
    double* tgt = (double*)dev(sizeof(double) * (size_t)tsz * nb);
    double* aout = (double*)dev(sizeof(double) * (size_t)cmax * nb);
    double* bout = (double*)dev(sizeof(double) * (size_t)cmax * nb);
    double* pwork = (double*)dev(sizeof(double) * (size_t)pairwork * nb);
    double* cwork = (double*)dev(sizeof(double) * (size_t)coremax * nb);
    // LDS pool of the bond solve: what the largest bond needs, within the
    // device's per-workgroup limit (the kernel stages only what fits)
    long long lds_limit = 0;
    {
      int maxlds = 0, optin = 0;
      SPP_CHECK(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx.device));
      if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, ctx.device) == hipSuccess)
        maxlds = std::max(maxlds, optin);
      (void)hipGetLastError();
      long long need = 0;
      for (const auto* set : {&hb, &vb})
        for (const BondGeo& g : *set) need = std::max(need, core_lds_doubles(g));
      lds_limit = std::min<long long>(need, (maxlds - 256) / (long long)sizeof(double));
      if (lds_limit < 0) lds_limit = 0;
      if (lds_limit * (long long)sizeof(double) > 65536)
        SPP_CHECK(hipFuncSetAttribute((const void*)fu_core_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(lds_limit * sizeof(double))));
    }
    double* drec = (double*)dev(sizeof(double) * 12 * (size_t)nb);
    int* rlive = (int*)dev(sizeof(int) * 4 * (size_t)nb);
    int* blive = (int*)dev(sizeof(int) * 2 * (size_t)nb);
    int* live = (int*)dev(sizeof(int) * (size_t)ns * 8 * nb);
    double* gate_dev = (double*)dev(sizeof(double) * (size_t)(nh + nv) * d * d * d * d);
    upload(gate_dev, P.gates_h, sizeof(double) * (size_t)nh * d * d * d * d, st);
    upload(gate_dev + (size_t)nh * d * d * d * d, P.gates_v, sizeof(double) * (size_t)nv * d * d * d * d, st);
    PairJob* pjobs = (PairJob*)dev(sizeof(PairJob) * (size_t)nb);
    RestoreJob* rjobs = (RestoreJob*)dev(sizeof(RestoreJob) * (size_t)nb);
    int* status = (int*)dev(sizeof(int) * nb);
    int* fail = (int*)dev(sizeof(int) * nb);
    int* dfailed = (int*)dev(sizeof(int) * nb);  // per sample: first failure code of this step, 0 ok
    std::vector<int> hfailed(nb, 0);
    double* nrm = (double*)dev(sizeof(double) * nb);
    double* fac = (double*)dev(sizeof(double) * nb);
    std::vector<double> hrec((size_t)12 * nb);
    std::vector<int> hstatus(nb);
    std::vector<double> hmt(nb);
    // boundary entropy: every absorb of the step raises a per-sample maximum
    double* emax = out.max_entropy ? (double*)dev(sizeof(double) * (size_t)nb) : nullptr;
    for (auto* side : {&north, &south})
      for (Boundary<T>& b : *side) b.entropy_max = emax;
    for (long long c0 = 0; c0 < n; c0 += nb) {
      const int cur = (int)std::min<long long>(nb, n - c0);
      // per-sample bookkeeping
      {
        std::vector<int> hl((size_t)ns * 8 * nb, 0);
        for (int s = 0; s < cur; ++s)
          std::memcpy(hl.data() + (size_t)s * ns * 8, states[c0 + s]->live.data(), sizeof(int) * ns * 8);
        upload(live, hl.data(), sizeof(int) * hl.size(), st);
        SPP_CHECK(hipStreamSynchronize(st));  // hl is released at the end of this block
      }
      SPP_CHECK(launch_fill<int>(status, nb, 0, st));
      SPP_CHECK(launch_fill<int>(fail, nb, -1, st));
      SPP_CHECK(launch_fill<int>(dfailed, nb, 0, st));
      std::fill(hfailed.begin(), hfailed.end(), 0);
      if (emax) SPP_CHECK(launch_fill<double>(emax, nb, 0.0, st));
      std::vector<double> env_err(cur, 0.0);
      auto relayer = [&](const std::vector<int>& sites) {
        std::vector<LayerJob> jobs;
        for (int i : sites)
          for (int s = 0; s < cur; ++s) {
            const SiteDev& sd = states[c0 + s]->sites[i];
            LayerJob J{};
            J.ket = sd.d;
            J.bra = sd.d;
            for (int k = 0; k < 6; ++k) {
              J.dk[k] = J.db[k] = sd.dim[k];
              J.ek[k] = J.eb[k] = sd.even[k];
            }
            J.out = site_base[i] + (long long)s * lsize[i];
            jobs.push_back(J);
          }
        // the job table may still be read by queued kernels: wait before rewriting it
        SPP_CHECK(hipStreamSynchronize(st));
        upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
        SPP_CHECK(hipStreamSynchronize(st));  // jobs is released when this call returns
        SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
      };
      {
        std::vector<int> all(ns);
        for (int i = 0; i < ns; ++i) all[i] = i;
        relayer(all);
      }
      const int nsweep = P.order == 2 ? 4 : 2;
      const bool sweep_h[4] = {true, false, false, true};
      const bool sweep_rev[4] = {false, false, true, true};
      int rec_bond = 0;
      for (int sw = 0; sw < nsweep; ++sw) {
        const bool horizontal = sweep_h[sw], reverse = sweep_rev[sw];
        if (horizontal ? nh == 0 : nv == 0) continue;
        for (auto& b : north) b.nb = cur;
        for (auto& b : south) b.nb = cur;
        // boundary environments of the far side (boundary_environments, side = reverse ? north : south)
        std::vector<Leg> physN(Lx), physS(Lx);
        for (int x = 0; x < Lx; ++x) physN[x] = spec[x].n;
        for (int xr = 0; xr < Lx; ++xr) physS[xr] = spec[(Lx - 1 - xr) + (Ly - 1) * Lx].s;
        auto absorb_row = [&](int y, bool into_south) {
          if (into_south) {
            RowLayers<T> rr;
            rr.L = Lx;
            rr.rotated = true;
            rr.spec.resize(Lx);
            rr.ptrs.resize(Lx);
            for (int xr = 0; xr < Lx; ++xr) {
              const int x = Lx - 1 - xr;
              const LayerSpec& sp = spec[x + y * Lx];
              rr.spec[xr] = LayerSpec{sp.e, sp.s, sp.w, sp.n};
              rr.ptrs[xr] = table(x + y * Lx);
            }
            absorb_forward<T>(ctx, south[y + 1], rr, chi, south[y]);
          } else {
            RowLayers<T> row;
            row.L = Lx;
            row.spec.assign(spec.begin() + y * Lx, spec.begin() + (y + 1) * Lx);
            row.ptrs.resize(Lx);
            for (int x = 0; x < Lx; ++x) row.ptrs[x] = table(x + y * Lx);
            absorb_forward<T>(ctx, north[y], row, chi, north[y + 1]);
          }
        };
        set_trivial(ctx, north[0], physN);
        set_trivial(ctx, south[Ly], physS);
        if (reverse)
          for (int y = 0; y < Ly; ++y) absorb_row(y, false);
        else
          for (int y = Ly - 1; y >= 0; --y) absorb_row(y, true);
        const int ymax = horizontal ? Ly : Ly - 1;
        const int xmax = horizontal ? Lx - 1 : Lx;
        for (int yi = 0; yi < ymax; ++yi) {
          const int y = reverse ? ymax - 1 - yi : yi;
          const int lastrow = y + (horizontal ? 0 : 1);
          const Boundary<T>& N = north[y];
          const Boundary<T>& S = south[lastrow + 1];
          const size_t mark = ctx.arena.mark();
          auto take = [&](long long elems) {
            return (T*)ctx.arena.take(sizeof(T) * (size_t)cur * (size_t)std::max(1LL, elems));
          };
          auto normalize = [&](const TRef& tr, int site) {
            FullView fv = full_view(tr);
            SPP_CHECK(launch_dot<T>(fv, fv, nrm, cur, st, dot_partials(ctx, cur)));
            SPP_CHECK(launch_safe_factor(nrm, fac, status, fail, site, cur, st));
            SPP_CHECK(launch_scale<T>(fv, fac, cur, st));
          };
          auto lay = [&](int site) {
            const LayerSpec& sp = spec[site];
            TRef r = dense_ref(nullptr, 0, "wnes", {sp.w, sp.n, sp.e, sp.s});
            r.ptrs = table(site);
            return r;
          };
          auto lay_l = [&](int site, const char* letters) { return with_letters(lay(site), letters); };
          auto s1 = [&](int x) { return x + y * Lx; };
          auto s2 = [&](int x) { return x + (y + 1) * Lx; };
          // message storage: messages[x] for x = 0..Lx (the precomputed side)
          std::vector<T*> Mb(Lx + 1);
          std::vector<std::vector<Leg>> Ml(Lx + 1);
          for (int x = 0; x <= Lx; ++x) {
            if (horizontal) {
              Leg wl = x < Lx ? spec[s1(x)].w : spec[s1(Lx - 1)].e;
              Ml[x] = {N.bond(x), wl, south_bond(S, Lx, x)};
            } else {
              Leg w1 = x < Lx ? spec[s1(x)].w : spec[s1(Lx - 1)].e;
              Leg w2 = x < Lx ? spec[s2(x)].w : spec[s2(Lx - 1)].e;
              Ml[x] = {N.bond(x), w1, w2, south_bond(S, Lx, x)};
            }
            Mb[x] = take(count_legs(Ml[x]));
          }
          const long long bigm = std::max((long long)chi * D * D * chi * (horizontal ? 1 : D), 1LL);
          T* BX = take(bigm);
          T* BY = take(bigm);
          T* Roll = take(horizontal ? (long long)chi * D * chi : (long long)chi * D * D * chi);
          T* Rnew = take(horizontal ? (long long)chi * D * chi : (long long)chi * D * D * chi);
          const char* ml = horizontal ? "awb" : "awvb";
          // one right step (messages covering columns >= x) from Rin (covering > x)
          auto right_step = [&](int x, const TRef& Rin, T* outbuf, std::vector<Leg>& ol) -> TRef {
            if (horizontal) {
              const LayerSpec& sp = spec[s1(x)];
              std::vector<Leg> xl = {N.bond(x), Rin.leg[2], N.phys(x), Rin.leg[1]};
              TRef Xr = folded_ref(BX, count_legs(xl), "adne", xl);
              contract<T>(ctx, Xr, N.ref(x, "anc"), with_letters(Rin, "ced"), cur);
              std::vector<Leg> yl = {N.bond(x), Rin.leg[2], sp.w, sp.s};
              TRef Yr = folded_ref(BY, count_legs(yl), "adws", yl);
              contract<T>(ctx, Yr, Xr, lay(s1(x)), cur);
              ol = {N.bond(x), sp.w, south_bond(S, Lx, x)};
              TRef Or = dense_ref(outbuf, count_legs(ol), "awb", ol);
              contract<T>(ctx, Or, Yr, bottom_ref(S, Lx, x, "bsd", true), cur);
              return Or;
            }
            const LayerSpec& u = spec[s1(x)];
            const LayerSpec& l = spec[s2(x)];
            const Leg dd = Rin.leg[3];
            std::vector<Leg> xl = {N.bond(x), u.n, dd, Rin.leg[1], Rin.leg[2]};
            TRef Xr = folded_ref(BX, count_legs(xl), "andef", xl);
            contract<T>(ctx, Xr, N.ref(x, "anc"), with_letters(Rin, "cefd"), cur);
            std::vector<Leg> yl = {N.bond(x), u.w, dd, u.s, Rin.leg[2]};
            TRef Yr = folded_ref(BY, count_legs(yl), "awdkf", yl);
            contract<T>(ctx, Yr, Xr, lay_l(s1(x), "wnek"), cur);
            std::vector<Leg> zl = {N.bond(x), u.w, dd, l.w, l.s};
            TRef Zr = folded_ref(BX, count_legs(zl), "awdvs", zl);
            contract<T>(ctx, Zr, Yr, lay_l(s2(x), "vkfs"), cur);
            ol = {N.bond(x), u.w, l.w, south_bond(S, Lx, x)};
            TRef Or = dense_ref(outbuf, count_legs(ol), "awvb", ol);
            contract<T>(ctx, Or, Zr, bottom_ref(S, Lx, x, "bsd", true), cur);
            return Or;
          };
          // one left step (message covering columns <= x) from Lin (covering < x)
          auto left_step = [&](const TRef& Lin, int x, T* outbuf, std::vector<Leg>& ol) -> TRef {
            if (horizontal) {
              const LayerSpec& sp = spec[s1(x)];
              std::vector<Leg> xl = {Lin.leg[1], Lin.leg[2], N.bond(x + 1), N.phys(x)};
              TRef Xr = folded_ref(BX, count_legs(xl), "wbcn", xl);
              contract<T>(ctx, Xr, with_letters(Lin, "awb"), N.ref(x, "anc"), cur);
              std::vector<Leg> yl = {Lin.leg[2], N.bond(x + 1), sp.e, sp.s};
              TRef Yr = folded_ref(BY, count_legs(yl), "bces", yl);
              contract<T>(ctx, Yr, Xr, lay(s1(x)), cur);
              ol = {N.bond(x + 1), sp.e, south_bond(S, Lx, x + 1)};
              TRef Or = dense_ref(outbuf, count_legs(ol), "ced", ol);
              contract<T>(ctx, Or, Yr, bottom_ref(S, Lx, x, "bsd", true), cur);
              return with_letters(Or, "awb");
            }
            const LayerSpec& u = spec[s1(x)];
            const LayerSpec& l = spec[s2(x)];
            const Leg b = Lin.leg[3], c = N.bond(x + 1);
            std::vector<Leg> xl = {Lin.leg[1], Lin.leg[2], b, c, N.phys(x)};
            TRef Xr = folded_ref(BX, count_legs(xl), "wvbcn", xl);
            contract<T>(ctx, Xr, with_letters(Lin, "awvb"), N.ref(x, "anc"), cur);
            std::vector<Leg> yl = {Lin.leg[2], b, c, u.e, u.s};
            TRef Yr = folded_ref(BY, count_legs(yl), "vbcek", yl);
            contract<T>(ctx, Yr, Xr, lay_l(s1(x), "wnek"), cur);
            std::vector<Leg> zl = {b, c, u.e, l.e, l.s};
            TRef Zr = folded_ref(BX, count_legs(zl), "bcefs", zl);
            contract<T>(ctx, Zr, Yr, lay_l(s2(x), "vkfs"), cur);
            ol = {c, u.e, l.e, south_bond(S, Lx, x + 1)};
            TRef Or = dense_ref(outbuf, count_legs(ol), "cefd", ol);
            contract<T>(ctx, Or, Zr, bottom_ref(S, Lx, x, "bsd", true), cur);
            return with_letters(Or, "awvb");
          };
          auto Mref = [&](int x) { return dense_ref(Mb[x], count_legs(Ml[x]), ml, Ml[x]); };
          // precomputed messages: right (forward) or left (reverse)
          if (!reverse) {
            SPP_CHECK(launch_fill<T>(Mb[Lx], cur, T(1), st));
            for (int x = Lx - 1; x >= 0; --x) {
              std::vector<Leg> ol;
              TRef Or = right_step(x, Mref(x + 1), Mb[x], ol);
              normalize(Or, s1(x));
            }
          } else {
            SPP_CHECK(launch_fill<T>(Mb[0], cur, T(1), st));
            for (int x = 0; x < Lx; ++x) {
              std::vector<Leg> ol;
              TRef Or = left_step(Mref(x), x, Mb[x + 1], ol);
              normalize(Or, s1(x));
            }
          }
          // rolling message starts at the trivial edge
          std::vector<Leg> rl = reverse ? Ml[Lx] : Ml[0];
          SPP_CHECK(launch_fill<T>(Roll, cur, T(1), st));
          for (int xi = 0; xi < xmax; ++xi) {
            const int x = reverse ? xmax - 1 - xi : xi;
            const int ia = s1(x), ib = horizontal ? s1(x + 1) : s2(x);
            const int finalx = horizontal ? x + 1 : x;
            const BondGeo& G = horizontal ? hb[y * (Lx - 1) + x] : vb[y * Lx + x];
            const double* gate = gate_dev + (size_t)(horizontal ? y * (Lx - 1) + x : nh + y * Lx + x) * d * d * d * d;
            TRef leftm = reverse ? Mref(x) : dense_ref(Roll, count_legs(rl), ml, rl);
            TRef rightm = reverse ? dense_ref(Roll, count_legs(rl), ml, rl) : Mref(finalx + 1);
            // --- reduced pair ---
            {
              std::vector<PairJob> pj(cur);
              std::vector<RestoreJob> rj(cur);
              for (int s = 0; s < cur; ++s) {
                pj[s].site[0] = states[c0 + s]->sites[ia].d;
                pj[s].site[1] = states[c0 + s]->sites[ib].d;
                pj[s].aux_even[0] = states[c0 + s]->sites[ia].even[3];
                pj[s].aux_even[1] = states[c0 + s]->sites[ib].even[3];
                rj[s].site[0] = states[c0 + s]->sites[ia].d;
                rj[s].site[1] = states[c0 + s]->sites[ib].d;
              }
              // live pointers resolved on the device side: per-sample offsets into `live`
              std::vector<PairJob> pjd(cur);
              for (int s = 0; s < cur; ++s) {
                pjd[s] = pj[s];
                pjd[s].live[0] = live + ((size_t)s * ns + ia) * 8;
                pjd[s].live[1] = live + ((size_t)s * ns + ib) * 8;
              }
              // the previous bond's restore kernel reads rjobs: wait before rewriting the tables
              SPP_CHECK(hipStreamSynchronize(st));
              upload(pjobs, pjd.data(), sizeof(PairJob) * cur, st);
              upload(rjobs, rj.data(), sizeof(RestoreJob) * cur, st);
              SPP_CHECK(hipStreamSynchronize(st));  // pjd and rj are released at the end of this block
            }
            hipLaunchKernelGGL(fu_pair_kernel, dim3(cur), dim3(64), 0, st, pjobs, cur, G.sa, G.sb, gate, d,
                               ed, qa, qmax, qbuf, qmax, ca, cb, tgt, rlive, pwork, pairwork, nb);
            SPP_CHECK(hipGetLastError());
            // --- reduced layers of the Q pseudo-sites ---
            {
              std::vector<LayerJob> jobs;
              const FuSide* sides[2] = {&G.sa, &G.sb};
              double* qs[2] = {qa, qbuf};
              T* outs[2] = {lred_a, lred_b};
              const long long lsz[2] = {G.la_size, G.lb_size};
              for (int side = 0; side < 2; ++side) {
                const FuSide& F = *sides[side];
                int pdm[6], pev[6];
                for (int k = 0; k < 6; ++k) {
                  pdm[k] = F.leg[k];
                  pev[k] = F.even[k];
                }
                pdm[2] = F.leg[3];
                pev[2] = F.even[3];
                pdm[3] = 1;
                pev[3] = 1;
                pdm[F.core] = F.rcap;
                pev[F.core] = F.recap;
                for (int s = 0; s < cur; ++s) {
                  LayerJob J{};
                  J.ket = qs[side] + (long long)s * qmax;
                  J.bra = J.ket;
                  for (int k = 0; k < 6; ++k) {
                    J.dk[k] = J.db[k] = pdm[k];
                    J.ek[k] = J.eb[k] = pev[k];
                  }
                  J.out = outs[side] + (long long)s * lsz[side];
                  jobs.push_back(J);
                }
              }
              SPP_CHECK(hipStreamSynchronize(st));
              upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
              SPP_CHECK(hipStreamSynchronize(st));  // jobs is released at the end of this block
              SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), std::max(G.la_size, G.lb_size), st));
            }
            // --- bond metric N[R, S] (R = (rk, rb) pairs of A, S of B) ---
            const Leg Ra = pair_leg(G.sa.rcap, G.sa.recap), Sb = pair_leg(G.sb.rcap, G.sb.recap);
            const long long nfsz = (long long)Ra.dim * Sb.dim;
            {
              const size_t m2 = ctx.arena.mark();
              TRef Nr = dense_ref(nfused, nfsz, "RS", {Ra, Sb});
              if (horizontal) {
                const LayerSpec& A_ = G.la;  // (W2, N2, R2, S2)
                const LayerSpec& B_ = G.lb;  // (R2', N2, E2, S2)
                TRef Ar = dense_ref(lred_a, G.la_size, "wnRs", {A_.w, A_.n, A_.e, A_.s});
                TRef Br = dense_ref(lred_b, G.lb_size, "SveT", {B_.w, B_.n, B_.e, B_.s});
                const Leg lb = leftm.leg[2], la = leftm.leg[0];
                std::vector<Leg> xl = {lb, N.phys(x), N.bond(x + 1), leftm.leg[1]};
                T* X = take(count_legs(xl));
                TRef Xr = dense_ref(X, count_legs(xl), "bncw", xl);
                contract<T>(ctx, Xr, with_letters(leftm, "awb"), N.ref(x, "anc"), cur);
                (void)la;
                std::vector<Leg> yl = {lb, N.bond(x + 1), Ra, A_.s};
                T* Y = take(count_legs(yl));
                TRef Yr = dense_ref(Y, count_legs(yl), "bcRs", yl);
                contract<T>(ctx, Yr, Xr, Ar, cur);
                std::vector<Leg> lm = {N.bond(x + 1), south_bond(S, Lx, x + 1), Ra};
                T* Lm = take(count_legs(lm));
                TRef Lr = dense_ref(Lm, count_legs(lm), "cfR", lm);
                contract<T>(ctx, Lr, Yr, bottom_ref(S, Lx, x, "bsf", true), cur);
                std::vector<Leg> ul = {N.bond(x + 1), N.phys(x + 1), rightm.leg[1], rightm.leg[2]};
                T* U = take(count_legs(ul));
                TRef Ur = dense_ref(U, count_legs(ul), "cveg", ul);
                contract<T>(ctx, Ur, with_letters(rightm, "deg"), N.ref(x + 1, "cvd"), cur);
                std::vector<Leg> vl = {N.bond(x + 1), rightm.leg[2], Sb, B_.s};
                T* V = take(count_legs(vl));
                TRef Vr = dense_ref(V, count_legs(vl), "cgST", vl);
                contract<T>(ctx, Vr, Ur, Br, cur);
                std::vector<Leg> rm = {N.bond(x + 1), south_bond(S, Lx, x + 1), Sb};
                T* Rm = take(count_legs(rm));
                TRef Rr = dense_ref(Rm, count_legs(rm), "cfS", rm);
                contract<T>(ctx, Rr, Vr, bottom_ref(S, Lx, x + 1, "fTg", true), cur);
                contract<T>(ctx, Nr, Lr, Rr, cur);
              } else {
                const LayerSpec& A_ = G.la;  // (W2, N2, E2, R2)
                const LayerSpec& B_ = G.lb;  // (W2, R2', E2, S2)
                TRef Ar = dense_ref(lred_a, G.la_size, "wneR", {A_.w, A_.n, A_.e, A_.s});
                TRef Br = dense_ref(lred_b, G.lb_size, "vSfs", {B_.w, B_.n, B_.e, B_.s});
                // Rb[c,e,b,v,S] = sum_{f,s} (right[c,e,f,d] bottom[b,s,d]) Bred[v,S,f,s]
                std::vector<Leg> rbl = {rightm.leg[0], rightm.leg[1], leftm.leg[3], B_.w, Sb};
                T* Rb = take(count_legs(rbl));
                const size_t mg = ctx.arena.mark();
                std::vector<Leg> g1 = {rightm.leg[0], rightm.leg[1], rightm.leg[2], leftm.leg[3], B_.s};
                T* G1 = take(count_legs(g1));
                TRef G1r = dense_ref(G1, count_legs(g1), "cefbs", g1);
                contract<T>(ctx, G1r, with_letters(rightm, "cefd"), bottom_ref(S, Lx, x, "bsd", true), cur);
                TRef Rbr = dense_ref(Rb, count_legs(rbl), "cebvS", rbl);
                contract<T>(ctx, Rbr, G1r, Br, cur);
                ctx.arena.release(mg);  // G1 is no longer needed
                std::vector<Leg> xl = {leftm.leg[1], leftm.leg[2], leftm.leg[3], N.phys(x), N.bond(x + 1)};
                T* X = take(count_legs(xl));
                TRef Xr = dense_ref(X, count_legs(xl), "wvbnc", xl);
                contract<T>(ctx, Xr, with_letters(leftm, "awvb"), N.ref(x, "anc"), cur);
                std::vector<Leg> ql = {A_.w, A_.n, A_.e, Sb};
                T* Qt = take(count_legs(ql));
                TRef Qr = dense_ref(Qt, count_legs(ql), "wneS", ql);
                contract<T>(ctx, Qr, Xr, Rbr, cur);
                contract<T>(ctx, Nr, Ar, Qr, cur);
              }
              ctx.arena.release(m2);
            }
            // --- the bond solve ---
            const fu::Opts o = P.opts;
            const long long lds = std::min(core_lds_doubles(G), lds_limit);
            hipLaunchKernelGGL(fu_core_kernel<T>, dim3(cur), dim3(kCoreLanes), (unsigned)(lds * sizeof(double)), st,
                               nfused, nfsz, tgt, rlive, G.sa.rcap, G.sa.recap, G.sb.rcap, G.sb.recap, d, ed, G.Dcap,
                               G.eDcap, o, aout, bout, drec, blive, cwork, coremax, cur, nb, lds, 0);
            SPP_CHECK(hipGetLastError());
            SPP_CHECK(hipMemcpyAsync(hrec.data(), drec, sizeof(double) * 12 * cur, hipMemcpyDeviceToHost, st));
            SPP_CHECK(hipStreamSynchronize(st));
            // PSD repair (update_bond!, full_update.jl:491-501): the complete bond
            // hole of one sample, its PSD projection by the host callback, and the
            // bond solve of that sample again on the repaired reduced metric
            auto repair = [&](int s) -> int {
              const int rec_before = (int)hrec[(size_t)s * 12 + 7];
              std::vector<void*> bufs;
              auto dmal = [&](long long elems) {
                void* p = nullptr;
                SPP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)std::max(elems, 1LL)));
                bufs.push_back(p);
                return (T*)p;
              };
              auto free_all = [&]() {
                (void)hipStreamSynchronize(st);
                for (void* p : bufs) (void)hipFree(p);
                bufs.clear();
              };
              try {
                auto view = [&](const TRef& t) {
                  if (t.ptrs || t.compact || t.fold) throw HipError("PSD repair: unsupported operand layout");
                  TRef v = t;
                  v.base = (T*)t.base + (long long)s * t.sstride;
                  for (Leg& l : v.leg)
                    if (l.split) l.split += s;
                  return v;
                };
                auto leg = [](const TRef& t, char c) {
                  const size_t k = t.letters.find(c);
                  if (k == std::string::npos) throw HipError("PSD repair: missing leg");
                  return t.leg[k];
                };
                auto make = [&](const std::string& letters, const std::vector<Leg>& legs) {
                  const long long n = count_legs(legs);
                  return dense_ref(dmal(n), n, letters, legs);
                };
                TRef H;
                if (horizontal) {
                  TRef L = view(with_letters(leftm, "awb")), T1 = view(N.ref(x, "anc")),
                       B1 = view(bottom_ref(S, Lx, x, "bsf", true)), T2 = view(N.ref(x + 1, "cvd")),
                       B2 = view(bottom_ref(S, Lx, x + 1, "ftg", true)), R = view(with_letters(rightm, "deg"));
                  TRef X = make("wbnc", {leg(L, 'w'), leg(L, 'b'), leg(T1, 'n'), leg(T1, 'c')});
                  contract<T>(ctx, X, L, T1, 1);
                  TRef Lh = make("wncsf", {leg(L, 'w'), leg(T1, 'n'), leg(T1, 'c'), leg(B1, 's'), leg(B1, 'f')});
                  contract<T>(ctx, Lh, X, B1, 1);
                  TRef U = make("cveg", {leg(T2, 'c'), leg(T2, 'v'), leg(R, 'e'), leg(R, 'g')});
                  contract<T>(ctx, U, R, T2, 1);
                  TRef Rh = make("cveft", {leg(T2, 'c'), leg(T2, 'v'), leg(R, 'e'), leg(B2, 'f'), leg(B2, 't')});
                  contract<T>(ctx, Rh, U, B2, 1);
                  H = make("wnsvet", {leg(L, 'w'), leg(T1, 'n'), leg(B1, 's'), leg(T2, 'v'), leg(R, 'e'),
                                      leg(B2, 't')});
                  contract<T>(ctx, H, Lh, Rh, 1);
                } else {
                  TRef L = view(with_letters(leftm, "awvb")), T1 = view(N.ref(x, "anc")),
                       B1 = view(bottom_ref(S, Lx, x, "bsd", true)), R = view(with_letters(rightm, "cefd"));
                  TRef Lh = make("wvbnc", {leg(L, 'w'), leg(L, 'v'), leg(L, 'b'), leg(T1, 'n'), leg(T1, 'c')});
                  contract<T>(ctx, Lh, L, T1, 1);
                  TRef Rh = make("bscef", {leg(B1, 'b'), leg(B1, 's'), leg(R, 'c'), leg(R, 'e'), leg(R, 'f')});
                  contract<T>(ctx, Rh, B1, R, 1);
                  H = make("wnevfs", {leg(L, 'w'), leg(T1, 'n'), leg(R, 'e'), leg(L, 'v'), leg(R, 'f'),
                                      leg(B1, 's')});
                  contract<T>(ctx, H, Lh, Rh, 1);
                }
                const long long hn = count_legs(H.leg);
                std::vector<T> ht((size_t)hn);
                SPP_CHECK(hipMemcpyAsync(ht.data(), H.base, sizeof(T) * hn, hipMemcpyDeviceToHost, st));
                // Q factors of the reduced pair (pseudo sites, FP64)
                const FuSide* F2[2] = {&G.sa, &G.sb};
                const double* qsrc[2] = {qa + (long long)s * qmax, qbuf + (long long)s * qmax};
                spp_fu_repair_problem pr{};
                std::vector<double> hq[2];
                for (int side = 0; side < 2; ++side) {
                  const FuSide& F = *F2[side];
                  int* qd = side == 0 ? pr.qa_dims : pr.qb_dims;
                  int* qe = side == 0 ? pr.qa_evens : pr.qb_evens;
                  for (int k = 0; k < 6; ++k) {
                    qd[k] = F.leg[k];
                    qe[k] = F.even[k];
                  }
                  qd[2] = F.leg[3];
                  qe[2] = F.even[3];
                  qd[3] = 1;
                  qe[3] = 1;
                  qd[F.core] = F.rcap;
                  qe[F.core] = F.recap;
                  (side == 0 ? pr.qa_core : pr.qb_core) = F.core;
                  long long qn = 1;
                  for (int k = 0; k < 6; ++k) qn *= qd[k];
                  hq[side].resize((size_t)qn);
                  SPP_CHECK(hipMemcpyAsync(hq[side].data(), qsrc[side], sizeof(double) * qn, hipMemcpyDeviceToHost,
                                           st));
                }
                SPP_CHECK(hipStreamSynchronize(st));
                std::vector<double> hd(ht.begin(), ht.end());
                // exterior legs of the hole: (owner Q, pseudo-site slot)
                static const int hq_h[6] = {0, 0, 0, 1, 1, 1}, hs_h[6] = {0, 1, 5, 1, 4, 5};
                static const int hs_v[6] = {0, 1, 4, 0, 4, 5};
                for (int k = 0; k < 6; ++k) {
                  pr.hole_q[k] = hq_h[k];
                  pr.hole_slot[k] = horizontal ? hs_h[k] : hs_v[k];
                  const int* qd = pr.hole_q[k] == 0 ? pr.qa_dims : pr.qb_dims;
                  const int* qe = pr.hole_q[k] == 0 ? pr.qa_evens : pr.qb_evens;
                  pr.hole_dims[k] = qd[pr.hole_slot[k]];
                  pr.hole_evens[k] = qe[pr.hole_slot[k]];
                }
                pr.sample = (int)(c0 + s);
                pr.bond = rec_bond;
                pr.horizontal = horizontal ? 1 : 0;
                pr.hole = hd.data();
                pr.qa = hq[0].data();
                pr.qb = hq[1].data();
                pr.target_norm = hrec[(size_t)s * 12 + 3];
                std::vector<double> nm((size_t)nfsz, 0.0);
                const int cb = ((spp_fu_repair_fn)ctx.fu_repair_fn)(ctx.fu_repair_user, &pr, nm.data());
                if (cb != 0) throw HipError("PSD repair: the host callback failed");
                std::vector<T> nt(nm.begin(), nm.end());
                upload(nfused + (long long)s * nfsz, nt.data(), sizeof(T) * nfsz, st);
                hipLaunchKernelGGL(fu_core_kernel<T>, dim3(1), dim3(kCoreLanes), (unsigned)(lds * sizeof(double)), st,
                                   nfused, nfsz, tgt, rlive, G.sa.rcap, G.sa.recap, G.sb.rcap, G.sb.recap, d, ed,
                                   G.Dcap, G.eDcap, o, aout, bout, drec, blive, cwork, coremax, cur, nb, lds, s);
                SPP_CHECK(hipGetLastError());
                SPP_CHECK(hipMemcpyAsync(hrec.data() + (size_t)s * 12, drec + (size_t)s * 12, sizeof(double) * 12,
                                         hipMemcpyDeviceToHost, st));
                SPP_CHECK(hipStreamSynchronize(st));
                free_all();
              } catch (...) {
                free_all();
                throw;
              }
              ctx.counters.fu_repairs += 1;
              (void)rec_before;
              const int after = (int)hrec[(size_t)s * 12 + 7];
              return after == fu::kRepairNeeded ? kFuRepairFailed : after;
            };
            bool newly_failed = false;
            for (int s = 0; s < cur; ++s) {
              int code = (int)hrec[(size_t)s * 12 + 7];
              if (ctx.fu_fail_sample == c0 + s && ctx.fu_fail_bond == rec_bond) code = ctx.fu_fail_code;
              if (code == fu::kRepairNeeded && ctx.fu_repair_fn && hfailed[s] == 0) code = repair(s);
              if (code != 0 && hfailed[s] == 0) {
                // the sample keeps its sites from here on (restores are skipped)
                // and reports the first failure; the others go on
                hfailed[s] = code;
                newly_failed = true;
              }
              if (hfailed[s] != 0) continue;
              if (out.records) {
                double* rr = out.records + ((size_t)(c0 + s) * nrec_bonds + rec_bond) * 8;
                rr[0] = hrec[(size_t)s * 12 + 0];
                rr[1] = hrec[(size_t)s * 12 + 1];
                rr[2] = hrec[(size_t)s * 12 + 2];
                rr[3] = hrec[(size_t)s * 12 + 4];
                rr[4] = hrec[(size_t)s * 12 + 5];
                rr[5] = hrec[(size_t)s * 12 + 6];
                rr[6] = hrec[(size_t)s * 12 + 8];
                rr[7] = hrec[(size_t)s * 12 + 9];
              }
            }
            if (newly_failed) upload(dfailed, hfailed.data(), sizeof(int) * cur, st);
            ++rec_bond;
            // --- restore the sites, their bond bookkeeping and layers ---
            {
              long long tmax = std::max(s0.sites[ia].size(), s0.sites[ib].size());
              dim3 grid((unsigned)((tmax + 255) / 256), (unsigned)cur, 2);
              hipLaunchKernelGGL(fu_restore_kernel, grid, dim3(256), 0, st, rjobs, cur, G.sa, G.sb, d, qa, qmax, qbuf,
                                 qmax, aout, bout, G.Dcap, nb, dfailed);
              SPP_CHECK(hipGetLastError());
              const int va = horizontal ? 2 : 3, vbl = horizontal ? 0 : 1;
              hipLaunchKernelGGL(fu_set_live_kernel, dim3((cur + 63) / 64), dim3(64), 0, st, live, ns, ia, va, ib, vbl,
                                 blive, cur, dfailed);
              SPP_CHECK(hipGetLastError());
              relayer({ia, ib});
            }
            // --- rolling message ---
            const int column = reverse ? finalx : x;
            if (x != (reverse ? 0 : xmax - 1)) {
              std::vector<Leg> ol;
              TRef Rin = dense_ref(Roll, count_legs(rl), ml, rl);
              TRef Or = reverse ? right_step(column, Rin, Rnew, ol) : left_step(Rin, column, Rnew, ol);
              normalize(Or, s1(column));
              std::swap(Roll, Rnew);
              rl = ol;
            }
          }
          ctx.arena.release(mark);
          if (reverse) absorb_row(lastrow, true);
          else absorb_row(y, false);
        }
        // environment errors: the far side's chain and the incremental one
        const double* ends[2] = {reverse ? north[Ly].max_trunc : south[0].max_trunc,
                                 reverse ? south[horizontal ? 0 : 1].max_trunc : north[ymax].max_trunc};
        for (const double* e : ends) {
          SPP_CHECK(hipMemcpyAsync(hmt.data(), e, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
          SPP_CHECK(hipStreamSynchronize(st));
          for (int s = 0; s < cur; ++s) env_err[s] = std::max(env_err[s], hmt[s]);
        }
      }
      SPP_CHECK(hipMemcpyAsync(hstatus.data(), status, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s)
        if (hstatus[s] != 0 && hfailed[s] == 0) hfailed[s] = kFuNonfiniteEnvironment;
      // normalize_state! (finite.jl:89-100): site norms into log_amplitude, row-major sites
      {
        std::vector<double*> hp((size_t)ns * cur);
        for (int i = 0; i < ns; ++i)
          for (int s = 0; s < cur; ++s) hp[(size_t)i * cur + s] = states[c0 + s]->sites[i].d;
        double** dp = (double**)ctx.arena.take(sizeof(double*) * hp.size());
        double* dn = (double*)ctx.arena.take(sizeof(double) * hp.size());
        upload(dp, hp.data(), sizeof(double*) * hp.size(), st);
        std::vector<long long> sizes(ns);
        for (int i = 0; i < ns; ++i) sizes[i] = s0.sites[i].size();
        for (int i = 0; i < ns; ++i) {
          hipLaunchKernelGGL(site_sumsq_kernel, dim3(cur), dim3(256), 0, st, dp + (size_t)i * cur, sizes[i],
                             dn + (size_t)i * cur);
          SPP_CHECK(hipGetLastError());
        }
        std::vector<double> hn(hp.size());
        SPP_CHECK(hipMemcpyAsync(hn.data(), dn, sizeof(double) * hn.size(), hipMemcpyDeviceToHost, st));
        SPP_CHECK(hipStreamSynchronize(st));
        for (size_t k = 0; k < hn.size(); ++k) {
          const double v = std::sqrt(hn[k]);
          const int s = (int)(k % cur);
          if (!(v > 0.0 && std::isfinite(v)) && hfailed[s] == 0) hfailed[s] = kFuZeroTensor;
          hn[k] = v > 0.0 && std::isfinite(v) ? v : 1.0;  // a failed sample is left unscaled
        }
        for (int i = 0; i < ns; ++i)
          for (int s = 0; s < cur; ++s) {
            const double v = hn[(size_t)i * cur + s];
            states[c0 + s]->log_amplitude += std::log(v);
            if (out.normalizations) out.normalizations[(size_t)(c0 + s) * ns + i] = std::log(v);
            hn[(size_t)i * cur + s] = 1.0 / v;
          }
        upload(dn, hn.data(), sizeof(double) * hn.size(), st);
        for (int i = 0; i < ns; ++i) {
          hipLaunchKernelGGL(site_scale_kernel, dim3(cur, (unsigned)std::min<long long>((sizes[i] + 255) / 256, 64)),
                             dim3(256), 0, st, dp + (size_t)i * cur, sizes[i], dn + (size_t)i * cur);
          SPP_CHECK(hipGetLastError());
        }
        SPP_CHECK(hipStreamSynchronize(st));
        ctx.arena.release(0);
      }
      // bond bookkeeping back to the states
      {
        std::vector<int> hl((size_t)ns * 8 * nb);
        SPP_CHECK(hipMemcpyAsync(hl.data(), live, sizeof(int) * hl.size(), hipMemcpyDeviceToHost, st));
        SPP_CHECK(hipStreamSynchronize(st));
        for (int s = 0; s < cur; ++s)
          std::memcpy(states[c0 + s]->live.data(), hl.data() + (size_t)s * ns * 8, sizeof(int) * ns * 8);
      }
      if (emax) {
        entropy_flush(ctx);
        SPP_CHECK(hipMemcpyAsync(hmt.data(), emax, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
        SPP_CHECK(hipStreamSynchronize(st));
      }
      for (int s = 0; s < cur; ++s) {
        states[c0 + s]->beta += 2.0 * P.delta;
        if (out.max_env_trunc) out.max_env_trunc[c0 + s] = env_err[s];
        if (emax) out.max_entropy[c0 + s] = hmt[s];
        if (out.status) out.status[c0 + s] = hfailed[s];
        if (hfailed[s] != 0) {
          states[c0 + s]->failed = hfailed[s];
          if (first_failure == 0) first_failure = hfailed[s];
        }
      }
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    ctx.arena.release(0);
    cleanup();
    throw;
  }
  SPP_CHECK(hipStreamSynchronize(st));
  cleanup();
  if (first_failure != 0 && !out.status) throw HipError(std::string("full update: ") + fu_failure_message(first_failure));
}

template void fu_evolution_step<float>(Context&, const std::vector<StateDev*>&, int, const FuParams&, FuOutputs&);
template void fu_evolution_step<double>(Context&, const std::vector<StateDev*>&, int, const FuParams&, FuOutputs&);

}  // namespace spp
