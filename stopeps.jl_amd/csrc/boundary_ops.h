// Boundary-MPS helpers shared by the sampler and the energy measurement.
#pragma once
#include "engine.h"
#include <algorithm>
#include <cmath>

namespace spp {

template <typename T>
inline void prepare_boundary(Boundary<T>& b, int L, int nb, long long cap_elems) {
  b.L = L;
  b.nb = nb;
  b.bdim.assign(L + 1, 1);
  b.beven.assign(L + 1, 1);
  b.bsplit.assign(L + 1, nullptr);
  b.pdim.assign(L, 1);
  b.peven.assign(L, 1);
  b.t.assign(L, nullptr);
  for (int x = 0; x < L; ++x) SPP_CHECK(hipMalloc(&b.t[x], sizeof(T) * nb * cap_elems));
  for (int k = 1; k < L; ++k) SPP_CHECK(hipMalloc(&b.bsplit[k], sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&b.log_scale, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&b.max_trunc, sizeof(double) * nb));
  b.owns = true;
}

template <typename T>
inline void set_trivial(Context& ctx, Boundary<T>& b, const std::vector<Leg>& phys) {
  for (int x = 0; x <= b.L; ++x) {
    b.bdim[x] = 1;
    b.beven[x] = 1;
  }
  for (int k = 1; k < b.L; ++k) SPP_CHECK(launch_fill<int>(b.bsplit[k], b.nb, 1, ctx.stream));
  for (int x = 0; x < b.L; ++x) {
    if (phys[x].dim != 1) throw HipError("nontrivial OBC leg");
    b.pdim[x] = 1;
    b.peven[x] = phys[x].even;
    SPP_CHECK(launch_fill<T>(b.t[x], b.nb, T(1), ctx.stream));
  }
  SPP_CHECK(launch_fill<double>(b.log_scale, b.nb, 0.0, ctx.stream));
  SPP_CHECK(launch_fill<double>(b.max_trunc, b.nb, 0.0, ctx.stream));
}

// Replace per-sample splits of a batch-1 boundary by fixed even counts.
template <typename T>
inline void freeze_splits(Context& ctx, Boundary<T>& b) {
  for (int k = 0; k <= b.L; ++k) {
    if (!b.bsplit[k]) continue;
    int v = 0;
    SPP_CHECK(hipMemcpyAsync(&v, b.bsplit[k], sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
    SPP_CHECK(hipStreamSynchronize(ctx.stream));
    SPP_CHECK(hipFree(b.bsplit[k]));
    b.bsplit[k] = nullptr;
    b.beven[k] = v;
  }
}

// Elements per sample and per index of the bra bond R of the three Gram
// transfer intermediates of absorb_forward at one site (environment bond l,
// layer legs w, n, e, s): the largest of lneF, lFwp and lwWN.
inline long long gram_unit_elems(long long l, long long w, long long n, long long e, long long s) {
  return std::max(l * n * e * e, std::max(l * e * w * s, l * w * w * n));
}
// Length of the R ranges of a site: all of R when that fits, else the largest
// range that keeps one intermediate within `limit` bytes per sample (at least 1).
inline int gram_rchunk(long long unit, long long rdim, size_t elem, double limit = kGramChunkBytes) {
  const double fit = limit / ((double)elem * (double)std::max(unit, 1LL));
  return (int)std::max(1.0, std::min((double)rdim, std::floor(fit)));
}

// Workspace bytes per sample of absorb_forward on a row of Lx layers with
// double-layer legs of dimension at most D: the Gram transfers (over R ranges),
// the right Grams and the truncation sweep (compact rho, subspace blocks, kept
// tensors).
inline double absorb_bytes_per_sample(long long Lx, long long chi, long long D, size_t elem, int ksub) {
  const double ts = (double)elem;
  // rows of one compact rho block: chi * D / 2 when the double-layer leg splits
  // evenly by parity, up to chi * D when it does not (e.g. D = 2 + 1)
  const double half = (double)chi * D + 1;
  const long long unit = chi * D * D * D;
  double z = ts * 2.0 * (double)unit * gram_rchunk(unit, chi, elem);
  double g = ts * (Lx + 1) * (double)(chi * D) * (chi * D);
  double sweep = ts * 5.0 * chi * chi * D * D + (8.0 + ts) * 2.0 * half * half +
                 8.0 * 2.0 * (4.0 * half * ksub + 3.0 * ksub * ksub) + ts * Lx * chi * chi * D;
  return z + g + sweep;
}

// Reference to the south tensor below original site x with letters (b, s, d).
template <typename T>
inline TRef bottom_ref(const Boundary<T>& sb, int Lx, int x, const std::string& bsd, bool per_sample = false) {
  int xr = Lx - 1 - x;
  std::string stl = {bsd[2], bsd[1], bsd[0]};
  return dense_ref(sb.t[xr], per_sample ? sb.tsize(xr) : 0, stl, {sb.bond(xr), sb.phys(xr), sb.bond(xr + 1)});
}
template <typename T>
inline Leg south_bond(const Boundary<T>& sb, int Lx, int b) {
  return sb.bond(Lx - b);
}


}  // namespace spp
