// Batched energy measurement (measure_energy, src/measurement/energy.jl:31-78).
//
// All samples of a chunk are projected PEPS that share the PEPO and differ in
// the auxiliary label of every site, so every double layer a sample needs is
// one of two per (site, kind) and is addressed through per-sample pointer
// tables. Kinds: the plain layer (the sampling cache's candidate), the onsite
// impurity (operator X on the ket) and the four halves of the two-site term
// (_operator_pair, energy.jl:13-23): F on the ket with the factor leg fused
// into E (horizontal, left site) or S (vertical, upper site), G with the
// factor leg fused into W or N.
//
// Per chunk: the south boundaries of every row (boundary_environments,
// boundary_mps.jl:193-209, in the rotated frame), then a north sweep that at
// row y evaluates the row terms against north[y] and south[y+1] (energy.jl:
// 44-60) and the vertical terms with two-row slab messages against north[y]
// and south[y+2] (energy.jl:61-72), and absorbs row y.
//
// Each left step from the normalised message left[x] is shared by the norm
// (den), the onsite numerator and the horizontal numerator; the horizontal
// denominator of bond (x, x+1) equals den[x+1] / f[x], f[x] being the factor
// that normalised the plain step into left[x+1]. The ratios are therefore the
// reference's num/den up to rounding.
#include "energy.h"
#include "boundary_ops.h"
#include "kron.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace spp {
namespace {

enum Kind { kPlain = 0, kOnsite, kHorzC, kHorzD, kVertC, kVertD, kKinds };

// tables[(kind * nsite + site) * ld + s] = bases[(kind * nsite + site) * 2 + labels[s][site]]
__global__ void layer_tables_kernel(const int* labels, int nsite, void* const* bases, int nkinds, void** tables,
                                    int nb, int ld) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nkinds * nsite * nb;
  if (t >= total) return;
  const int s = (int)(t % nb);
  const long long ks = t / nb;
  const int site = (int)(ks % nsite);
  const int lab = labels[(long long)s * nsite + site];
  tables[ks * ld + s] = bases[ks * 2 + lab];
}

// _real_ratio (energy.jl:1-6) for real scalars: acc += num * f / den, where f
// rescales a numerator taken from an unnormalised message (0 means 1).
__global__ void ratio_kernel(const double* num, const double* den, const double* fac, double* acc, int* status,
                             int* fail_site, int site, int nb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb || status[s] != 0) return;
  const double d = den[s];
  if (d == 0.0) {
    status[s] = 4;
    fail_site[s] = site;
    return;
  }
  double f = fac ? fac[s] : 1.0;
  if (f == 0.0) f = 1.0;
  const double z = num[s] * f / d;
  if (!isfinite(z)) {
    status[s] = 5;
    fail_site[s] = site;
    return;
  }
  acc[s] += z;
}

// Pair leg of a double layer whose ket leg carries an extra odd unit charge.
inline Leg flipped_pair_leg(int dim, int even) {
  const int odd = dim - even;
  return fixed_leg(dim * dim, 2 * odd * even);
}

// Physical-operator entry allowed by the Z2 grading of a leg with `even` of 2 indices.
bool op_respects_grading(const double* op, int even, int odd_op) {
  if (even == 2) return true;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      const int pa = a >= even, pb = b >= even;
      if (((pa ^ pb) != odd_op) && op[a * 2 + b] != 0.0) return false;
    }
  return true;
}

// One d = 2 physical space on every site, operators that respect its grading.
void check_energy_ops(const StateDev& st, const EnergyOps& ops) {
  const int pe = st.sites[0].even[2];
  for (const SiteDev& sd : st.sites)
    if (sd.dim[2] != 2 || sd.even[2] != pe) throw HipError("energy needs one d = 2 physical space on every site");
  if (pe == 2 && ops.odd) throw HipError("an odd factor leg needs a Z2-graded physical space");
  if (!op_respects_grading(ops.X, pe, 0) || !op_respects_grading(ops.F, pe, ops.odd) ||
      !op_respects_grading(ops.G, pe, ops.odd))
    throw HipError("operator does not respect the Z2 grading of the physical space");
}

// Layer specs of every kind at one site (plain spec p; the factor leg of the
// two-site halves fused into E, W, S or N).
void impurity_specs(const SiteDev& sd, const LayerSpec& p, int odd, LayerSpec out[kKinds]) {
  auto flip = [&](int k) { return odd ? flipped_pair_leg(sd.dim[k], sd.even[k]) : pair_leg(sd.dim[k], sd.even[k]); };
  out[kPlain] = p;
  out[kOnsite] = p;
  out[kHorzC] = LayerSpec{p.w, p.n, flip(4), p.s};
  out[kHorzD] = LayerSpec{flip(0), p.n, p.e, p.s};
  out[kVertC] = LayerSpec{p.w, p.n, p.e, flip(5)};
  out[kVertD] = LayerSpec{p.w, flip(1), p.e, p.s};
}

// Whether a kind's layer is read at site (x, y).
bool kind_used(int kind, int x, int y, int Lx, int Ly) {
  return kind == kPlain || kind == kOnsite || (kind == kHorzC && x < Lx - 1) || (kind == kHorzD && x > 0) ||
         (kind == kVertC && y < Ly - 1) || (kind == kVertD && y > 0);
}

// Impurity layer job: the site with the kind's operator on the ket's physical
// leg (and its factor leg fused into a virtual leg); aux traced (use_aux 0) or
// projected on a basis vector.
LayerJob impurity_job(const SiteDev& sd, int kind, const EnergyOps& ops, void* out) {
  LayerJob J{};
  J.ket = sd.d;
  J.bra = sd.d;
  for (int l = 0; l < 6; ++l) {
    J.dk[l] = J.db[l] = sd.dim[l];
    J.ek[l] = J.eb[l] = sd.even[l];
  }
  J.out = out;
  if (kind != kPlain) {
    J.use_op = 1;
    const double* op = kind == kOnsite ? ops.X : (kind == kHorzC || kind == kVertC) ? ops.F : ops.G;
    std::memcpy(J.op, op, sizeof(double) * 4);
    const int fl = kind == kHorzC ? 4 : kind == kHorzD ? 0 : kind == kVertC ? 5 : kind == kVertD ? 1 : -1;
    if (fl >= 0) J.flipk[fl] = ops.odd;
  }
  return J;
}

// Boundary-entropy observer of a measurement chunk (boundary_environments
// with an entropy argument, boundary_mps.jl:193-209, as measure_energy runs it,
// energy.jl:31-78): the per-cut entropies of every south and north absorb and
// their running maximum. On the host the cuts map to Se.north / Se.south
// [Ly-1][Lx-1]: the north absorb of row y (0-based) is row_cut y+1, its cut
// after site x is bond x+1; the south absorb of row y is row_cut y, and its
// rotated cut after site xr is bond Lx-1-xr. Rows outside 1..Ly-1 only feed
// the maximum.
struct EntropyObserver {
  int Lx = 0, Ly = 0, nb = 0, cur = 0;
  double* cuts = nullptr;  // device [2][Ly][Lx-1][nb]: south absorbs, then north absorbs, by row
  double* emax = nullptr;  // device [nb]
  std::vector<double> hc, hm;
  bool on() const { return emax != nullptr; }
  template <typename Dev>
  void init(int lx, int ly, int nbb, Dev&& dev) {
    Lx = lx;
    Ly = ly;
    nb = nbb;
    if (Lx > 1) cuts = (double*)dev(sizeof(double) * 2 * (size_t)Ly * (Lx - 1) * nb);
    emax = (double*)dev(sizeof(double) * (size_t)nb);
    hc.assign(Lx > 1 ? 2 * (size_t)Ly * (Lx - 1) * nb : 0, 0.0);
    hm.assign(nb, 0.0);
  }
  // the absorb into b of row y on side 0 (south) or 1 (north)
  template <typename B>
  void attach(B& b, int side, int y) const {
    if (!on()) return;
    b.entropy = cuts ? cuts + ((size_t)side * Ly + y) * (size_t)(Lx - 1) * nb : nullptr;
    b.entropy_max = emax;
  }
  void start(int c, hipStream_t st) {
    cur = c;
    if (on()) SPP_CHECK(launch_fill<double>(emax, nb, 0.0, st));
  }
  void fetch(hipStream_t st) {
    if (!on()) return;
    if (cuts) SPP_CHECK(hipMemcpyAsync(hc.data(), cuts, sizeof(double) * hc.size(), hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipMemcpyAsync(hm.data(), emax, sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, st));
  }
  // after fetch and a stream synchronisation: sample s of the chunk is output g
  void write(int s, long long g, double* profile, double* maximum) const {
    if (!on()) return;
    const int C = Lx - 1, R = Ly - 1;
    if (profile && C > 0 && R > 0) {
      double* P = profile + (size_t)g * 2 * R * C;
      for (int i = 0; i < 2 * R * C; ++i) P[i] = NAN;
      auto at = [&](int side, int y, int cut) {
        return hc[((size_t)side * Ly + y) * (size_t)C * nb + (size_t)cut * cur + s];
      };
      for (int y = 0; y + 1 <= R; ++y)
        for (int x = 0; x < C; ++x) P[(size_t)y * C + x] = at(1, y, x);
      for (int y = 1; y <= R; ++y)
        for (int xr = 0; xr < C; ++xr) P[(size_t)R * C + (size_t)(y - 1) * C + (C - 1 - xr)] = at(0, y, xr);
    }
    if (maximum) maximum[g] = hm[s];
  }
};

template <typename T>
struct EnergyRun {
  Context& ctx;
  int Lx, Ly, ns, chi, nb = 0, cur = 0;
  long long D = 1;                          // largest double-layer leg
  std::vector<LayerSpec> spec[kKinds];      // [kind][site]
  void** tables = nullptr;                  // [kind][site][nb]
  int *status = nullptr, *fail = nullptr;
  double *xsum = nullptr, *zzsum = nullptr, *den = nullptr, *num = nullptr, *numh = nullptr;
  double *fac = nullptr, *facp = nullptr, *nrm = nullptr;

  EnergyRun(Context& cx, int lx, int ly, int chi_) : ctx(cx), Lx(lx), Ly(ly), ns(lx * ly), chi(chi_) {}

  // Single-layer tables of the plain (projected) layers, FP32 caches that carry
  // them (kron.h): the layout per site and the table pointer of every sample,
  // ktab[site * nb + s]. The plain-layer products of the row and slab terms run
  // through them; impurity layers (an operator on the ket) keep the GEMM.
  const KronSite* ksite = nullptr;
  void** ktab = nullptr;
  bool kron_plain(int kind) const { return kind == kPlain && ksite != nullptr && ktab != nullptr; }
  static long long fold(const Leg& l) { return std::max(l.even, l.dim - l.even); }
  // rows of (l, R, f) with the contracted pair leg c1 at stride sx1 and the folded c2 contiguous
  static KronRows rows3(const Leg& l, const Leg& R, const Leg* f, long long sx1) {
    KronRows kr;
    const int nf = f ? f->dim : 1;
    kr.nrows = (long long)l.dim * R.dim * nf;
    kr.nf = nf;
    kr.ef = f ? f->even : 1;
    kr.fsplit = f ? f->split : nullptr;
    kr.rc = R.dim;
    kr.lsplit = l.split;
    kr.leven = l.even;
    kr.rsplit = R.split;
    kr.reven = R.even;
    kr.sx1 = sx1;
    kr.sx2 = 1;
    return kr;
  }
  void kron_apply(int stage, int site, const TRef& X, const TRef& Y, KronRows kr) {
    if constexpr (sizeof(T) != sizeof(float)) {
      throw HipError("factorised layer products need an FP32 cache");
    } else {
      kr.xs = X.sstride;
      kr.ys = Y.sstride;
      const KronSite& ks = ksite[site];
      const double fl =
          2.0 * cur * (ctx.kron == 2 ? kron_mfma_row_macs(ks, stage) : kron_row_macs(ks, stage)) * kr.nrows;
      timed_kron(ctx, fl, [&] {
        return launch_kron_rows(ks, stage, reinterpret_cast<const float*>(X.base) + X.offset,
                                reinterpret_cast<float*>(Y.base) + Y.offset, kr, ktab + (size_t)site * nb, cur,
                                ctx.kron_rows, ctx.stream, ctx.kron_groups, nullptr, ctx.kron == 2);
      });
      ctx.counters.kron_launches += 1;
      ctx.counters.gemm_flops += fl;
    }
  }

  void* const* table(int kind, int site) const { return tables + ((long long)kind * ns + site) * nb; }
  TRef layer(int kind, int site, const std::string& letters) const {
    const LayerSpec& sp = spec[kind][site];
    TRef r = dense_ref(nullptr, 0, letters, {sp.w, sp.n, sp.e, sp.s});
    r.ptrs = table(kind, site);
    return r;
  }
  T* take(long long elems) { return (T*)ctx.arena.take(sizeof(T) * (size_t)cur * (size_t)std::max(1LL, elems)); }
  void dot(const TRef& a, const TRef& b, double* out) {
    SPP_CHECK(launch_dot<T>(full_view(a), full_view(b), out, cur, ctx.stream, dot_partials(ctx, cur)));
  }
  // _safe_scale in place; the factor stays in `fac`
  void normalize(const TRef& t, int site) {
    FullView fv = full_view(t);
    SPP_CHECK(launch_dot<T>(fv, fv, nrm, cur, ctx.stream, dot_partials(ctx, cur)));
    SPP_CHECK(launch_safe_factor(nrm, fac, status, fail, site, cur, ctx.stream));
    SPP_CHECK(launch_scale<T>(fv, fac, cur, ctx.stream));
  }
  void ratio(const double* nu, const double* de, const double* f, double* acc, int site) {
    hipLaunchKernelGGL(ratio_kernel, dim3((cur + 255) / 256), dim3(256), 0, ctx.stream, nu, de, f, acc, status, fail,
                       site, cur);
    SPP_CHECK(hipGetLastError());
  }
  static long long count(const std::vector<Leg>& l) {
    long long n = 1;
    for (auto& g : l) n *= g.dim;
    return n;
  }

  // ---------------- one-row terms (energy.jl:44-60) ----------------
  void row_terms(const Boundary<T>& N, const Boundary<T>& S, int y) {
    const size_t mark = ctx.arena.mark();
    const std::vector<LayerSpec>& P = spec[kPlain];
    const long long small = (long long)chi * D * chi, big = (long long)chi * D * D * chi;
    std::vector<T*> R(Lx + 1);
    std::vector<std::vector<Leg>> Rl(Lx + 1);
    for (int x = 0; x <= Lx; ++x) {
      Leg wl = x < Lx ? P[x + y * Lx].w : P[(Lx - 1) + y * Lx].e;
      Rl[x] = {N.bond(x), wl, south_bond(S, Lx, x)};
      R[x] = take(count(Rl[x]));
    }
    T* BX = take(big);
    T* BY = take(big);
    SPP_CHECK(launch_fill<T>(R[Lx], cur, T(1), ctx.stream));
    auto Rref = [&](int x) { return dense_ref(R[x], count(Rl[x]), "ced", Rl[x]); };
    // right messages (_right_step, boundary_mps.jl:240-246)
    for (int x = Lx - 1; x >= 0; --x) {
      const int site = x + y * Lx;
      const LayerSpec& sp = P[site];
      std::vector<Leg> xl = {N.bond(x), Rl[x + 1][2], N.phys(x), Rl[x + 1][1]};
      TRef Xr = folded_ref(BX, count(xl), "adne", xl);
      contract<T>(ctx, Xr, N.ref(x, "anc"), Rref(x + 1), cur);
      std::vector<Leg> yl = {N.bond(x), Rl[x + 1][2], sp.w, sp.s};
      TRef Yr = folded_ref(BY, count(yl), "adws", yl);
      if (kron_plain(kPlain)) {
        // rows (a, d), contract (n, e) into (w, s): role assignment 0
        KronRows kr = rows3(xl[0], xl[1], nullptr, fold(xl[3]));
        kr.sx_slab = (long long)xl[2].dim * fold(xl[3]);
        kr.sxf = 0;
        kron_apply(0, site, Xr, Yr, kr);
      } else {
        contract<T>(ctx, Yr, Xr, layer(kPlain, site, "wnes"), cur);
      }
      TRef Rr = dense_ref(R[x], count(Rl[x]), "awb", Rl[x]);
      contract<T>(ctx, Rr, Yr, bottom_ref(S, Lx, x, "bsd", true), cur);
      normalize(Rr, site);
    }
    // left sweep (_left_step, boundary_mps.jl:233-239)
    T* Lb = take(small);
    T* O0 = take(small);
    T* Ot = take(small);
    T* OC = take(small);
    SPP_CHECK(launch_fill<T>(Lb, cur, T(1), ctx.stream));
    std::vector<Leg> Ll = {N.bond(0), P[y * Lx].w, south_bond(S, Lx, 0)};
    // X = L.T over the north bond
    auto make_x = [&](const TRef& L, int x) {
      std::vector<Leg> xl = {L.leg[1], L.leg[2], N.bond(x + 1), N.phys(x)};
      TRef Xr = folded_ref(BX, count(xl), "wbcn", xl);
      contract<T>(ctx, Xr, L, N.ref(x, "anc"), cur);
      return Xr;
    };
    // (X.E).B for the layer of the given kind at column x
    auto finish = [&](const TRef& Xr, int kind, int x, T* outbuf) {
      const LayerSpec& q = spec[kind][x + y * Lx];
      std::vector<Leg> yl = {Xr.leg[1], N.bond(x + 1), q.e, q.s};
      TRef Yr = folded_ref(BY, count(yl), "bces", yl);
      if (kron_plain(kind)) {
        // X[w][b][c][n]: rows (b, c), contract (w, n) into (e, s): role assignment 2
        const long long fn = fold(Xr.leg[3]);
        KronRows kr = rows3(Xr.leg[1], Xr.leg[2], nullptr, (long long)Xr.leg[1].dim * Xr.leg[2].dim * fn);
        kr.sx_slab = fn;
        kr.sxf = 0;
        kron_apply(2, x + y * Lx, Xr, Yr, kr);
      } else {
        contract<T>(ctx, Yr, Xr, layer(kind, x + y * Lx, "wnes"), cur);
      }
      std::vector<Leg> ol = {N.bond(x + 1), q.e, south_bond(S, Lx, x + 1)};
      TRef Or = dense_ref(outbuf, count(ol), "ced", ol);
      contract<T>(ctx, Or, Yr, bottom_ref(S, Lx, x, "bsd", true), cur);
      return Or;
    };
    for (int x = 0; x < Lx; ++x) {
      const int site = x + y * Lx;
      TRef Lr = dense_ref(Lb, count(Ll), "awb", Ll);
      TRef Xr = make_x(Lr, x);
      TRef O0r = finish(Xr, kPlain, x, O0);
      dot(O0r, Rref(x + 1), den);
      TRef Otr = finish(Xr, kOnsite, x, Ot);
      dot(Otr, Rref(x + 1), num);
      ratio(num, den, nullptr, xsum, site);
      if (x > 0) ratio(numh, den, facp, zzsum, site - 1);
      if (x < Lx - 1) {
        TRef OCr = finish(Xr, kHorzC, x, OC);
        TRef XD = make_x(with_letters(OCr, "awb"), x + 1);
        TRef ODr = finish(XD, kHorzD, x + 1, Ot);
        dot(ODr, Rref(x + 2), numh);
        normalize(O0r, site);
        std::swap(fac, facp);
        std::swap(Lb, O0);
        Ll = O0r.leg;
      }
    }
    ctx.arena.release(mark);
  }

  // ---------------- two-row (vertical) terms (energy.jl:61-72) ----------------
  void slab_terms(const Boundary<T>& N, const Boundary<T>& S, int y) {
    const size_t mark = ctx.arena.mark();
    const std::vector<LayerSpec>& P = spec[kPlain];
    const long long mid = (long long)chi * D * D * chi, big = mid * D;
    auto s1 = [&](int x) { return x + y * Lx; };
    auto s2 = [&](int x) { return x + (y + 1) * Lx; };
    std::vector<T*> R(Lx + 1);
    std::vector<std::vector<Leg>> Rl(Lx + 1);
    for (int x = 0; x <= Lx; ++x) {
      Leg w1 = x < Lx ? P[s1(x)].w : P[s1(Lx - 1)].e;
      Leg w2 = x < Lx ? P[s2(x)].w : P[s2(Lx - 1)].e;
      Rl[x] = {N.bond(x), w1, w2, south_bond(S, Lx, x)};
      R[x] = take(count(Rl[x]));
    }
    T* BX = take(big);
    T* BY = take(big);
    SPP_CHECK(launch_fill<T>(R[Lx], cur, T(1), ctx.stream));
    auto Rref = [&](int x) { return dense_ref(R[x], count(Rl[x]), "cefd", Rl[x]); };
    // right slab messages (SlabColumn _right_step, boundary_mps.jl:224-231)
    for (int x = Lx - 1; x >= 0; --x) {
      const LayerSpec& u = P[s1(x)];
      const LayerSpec& l = P[s2(x)];
      const Leg d = Rl[x + 1][3];
      std::vector<Leg> xl = {N.bond(x), u.n, d, Rl[x + 1][1], Rl[x + 1][2]};
      TRef Xr = folded_ref(BX, count(xl), "andef", xl);
      contract<T>(ctx, Xr, N.ref(x, "anc"), Rref(x + 1), cur);
      std::vector<Leg> yl = {N.bond(x), u.w, d, u.s, Rl[x + 1][2]};
      TRef Yr = folded_ref(BY, count(yl), "awdkf", yl);
      contract<T>(ctx, Yr, Xr, layer(kPlain, s1(x), "wnek"), cur);
      std::vector<Leg> zl = {N.bond(x), u.w, d, l.w, l.s};
      TRef Zr = folded_ref(BX, count(zl), "awdvs", zl);
      if (kron_plain(kPlain)) {
        // Y[a][w][d][k][f]: rows (a, w, d), contract (k, f) into (v, s): role assignment 0
        const long long ff = fold(yl[4]);
        KronRows kr = rows3(yl[0], yl[1], &yl[2], ff);
        kr.sxf = (long long)yl[3].dim * ff;
        kr.sx_slab = (long long)yl[2].dim * kr.sxf;
        kron_apply(0, s2(x), Yr, Zr, kr);
      } else {
        contract<T>(ctx, Zr, Yr, layer(kPlain, s2(x), "vkfs"), cur);
      }
      TRef Rr = dense_ref(R[x], count(Rl[x]), "awvb", Rl[x]);
      contract<T>(ctx, Rr, Zr, bottom_ref(S, Lx, x, "bsd", true), cur);
      normalize(Rr, s1(x));
    }
    // left slab sweep (SlabColumn _left_step, boundary_mps.jl:216-223)
    T* Lb = take(mid);
    T* O0 = take(mid);
    T* Oi = take(mid);
    SPP_CHECK(launch_fill<T>(Lb, cur, T(1), ctx.stream));
    std::vector<Leg> Ll = {N.bond(0), P[s1(0)].w, P[s2(0)].w, south_bond(S, Lx, 0)};
    auto step = [&](const TRef& L, int x, int ku, int kl, T* outbuf) {
      const LayerSpec& u = spec[ku][s1(x)];
      const LayerSpec& l = spec[kl][s2(x)];
      const Leg b = L.leg[3], c = N.bond(x + 1);
      std::vector<Leg> xl = {L.leg[1], L.leg[2], b, c, N.phys(x)};
      TRef Xr = folded_ref(BX, count(xl), "wvbcn", xl);
      contract<T>(ctx, Xr, L, N.ref(x, "anc"), cur);
      std::vector<Leg> yl = {L.leg[2], b, c, u.e, u.s};
      TRef Yr = folded_ref(BY, count(yl), "vbcek", yl);
      if (kron_plain(ku)) {
        // X[w][v][b][c][n]: rows (v, b, c), contract (w, n) into (e, k): role assignment 2
        const long long fn = fold(xl[4]);
        KronRows kr = rows3(xl[1], xl[2], &xl[3], (long long)xl[1].dim * xl[2].dim * xl[3].dim * fn);
        kr.sxf = fn;
        kr.sx_slab = (long long)xl[3].dim * fn;
        kron_apply(2, s1(x), Xr, Yr, kr);
      } else {
        contract<T>(ctx, Yr, Xr, layer(ku, s1(x), "wnek"), cur);
      }
      std::vector<Leg> zl = {b, c, u.e, l.e, l.s};
      TRef Zr = folded_ref(BX, count(zl), "bcefs", zl);
      if (kron_plain(kl)) {
        // Y[v][b][c][e][k]: rows (b, c, e), contract (v, k) into (f, s): role assignment 2
        const long long fk = fold(yl[4]);
        KronRows kr = rows3(yl[1], yl[2], &yl[3], (long long)yl[1].dim * yl[2].dim * yl[3].dim * fk);
        kr.sxf = fk;
        kr.sx_slab = (long long)yl[3].dim * fk;
        kron_apply(2, s2(x), Yr, Zr, kr);
      } else {
        contract<T>(ctx, Zr, Yr, layer(kl, s2(x), "vkfs"), cur);
      }
      std::vector<Leg> ol = {c, u.e, l.e, south_bond(S, Lx, x + 1)};
      TRef Or = dense_ref(outbuf, count(ol), "cefd", ol);
      contract<T>(ctx, Or, Zr, bottom_ref(S, Lx, x, "bsd", true), cur);
      return Or;
    };
    for (int x = 0; x < Lx; ++x) {
      TRef Lr = dense_ref(Lb, count(Ll), "awvb", Ll);
      TRef O0r = step(Lr, x, kPlain, kPlain, O0);
      dot(O0r, Rref(x + 1), den);
      TRef Oir = step(Lr, x, kVertC, kVertD, Oi);
      dot(Oir, Rref(x + 1), num);
      ratio(num, den, nullptr, zzsum, s1(x));
      if (x < Lx - 1) {
        normalize(O0r, s1(x));
        std::swap(Lb, O0);
        Ll = O0r.leg;
      }
    }
    ctx.arena.release(mark);
  }
};

}  // namespace

template <typename T>
void measure_energy_batch(Context& ctx, const StateDev& pepo, const CacheDev<T>& c, int chi, int n,
                          const int* labels_h, const EnergyOps& ops, EnergyOutputs& out) {
  hipStream_t st = ctx.stream;
  const int Lx = c.Lx, Ly = c.Ly, ns = Lx * Ly;
  if (pepo.Lx != Lx || pepo.Ly != Ly || (int)pepo.sites.size() != ns)
    throw HipError("the state does not match the sampling cache");
  if (chi < 1) throw HipError("chi must be positive");
  if (n <= 0) return;
  if (!labels_h) throw HipError("labels are required");
  for (long long i = 0; i < (long long)n * ns; ++i)
    if (labels_h[i] != 0 && labels_h[i] != 1) throw HipError("labels must be 0 or 1");
  check_energy_ops(pepo, ops);

  EnergyRun<T> run(ctx, Lx, Ly, chi);
  // ---------------- impurity layers ----------------
  std::vector<void*> owned;
  std::vector<void*> base((size_t)kKinds * ns * 2, nullptr);
  std::vector<LayerJob> jobs;
  long long maxout = 1;
  for (int k = 0; k < kKinds; ++k) run.spec[k].resize(ns);
  for (int i = 0; i < ns; ++i) {
    const SiteDev& sd = pepo.sites[i];
    const int x = i % Lx, y = i / Lx;
    LayerSpec ks[kKinds];
    impurity_specs(sd, c.spec[i], ops.odd, ks);
    for (int k = 0; k < kKinds; ++k) run.spec[k][i] = ks[k];
    for (int k = 0; k < 2; ++k) base[((size_t)kPlain * ns + i) * 2 + k] = c.cand[k][i];
    for (int kind = kOnsite; kind < kKinds; ++kind) {
      const bool used = kind_used(kind, x, y, Lx, Ly);
      for (int k = 0; k < 2; ++k) {
        void*& slot = base[((size_t)kind * ns + i) * 2 + k];
        if (!used) {
          slot = c.cand[k][i];  // never read
          continue;
        }
        const LayerSpec& q = run.spec[kind][i];
        const long long cnt = (long long)q.w.dim * q.n.dim * q.e.dim * q.s.dim;
        maxout = std::max(maxout, cnt);
        void* buf = nullptr;
        SPP_CHECK(hipMalloc(&buf, sizeof(T) * cnt));
        owned.push_back(buf);
        slot = buf;
        LayerJob J = impurity_job(sd, kind, ops, buf);
        J.use_aux = 1;
        J.aux[0] = c.basis[0 * 2 + k];
        J.aux[1] = c.basis[1 * 2 + k];
        jobs.push_back(J);
      }
    }
  }
  for (int k = 0; k < kKinds; ++k)
    for (const LayerSpec& q : run.spec[k])
      run.D = std::max<long long>(run.D, std::max(std::max(q.w.dim, q.n.dim), std::max(q.e.dim, q.s.dim)));

  std::vector<Boundary<T>> south(Ly + 1);
  Boundary<T> NB[2];
  LayerJob* djobs = nullptr;
  void** dbase = nullptr;
  int* dlabels = nullptr;
  double *mant = nullptr, *lnamp = nullptr;
  long long* dsst = nullptr;
  T** dtens = nullptr;
  std::vector<void*> scratch;
  auto dev = [&](size_t bytes) {
    void* p = nullptr;
    SPP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    scratch.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    entropy_reset(ctx);  // pending cuts write into buffers freed below
    for (auto& b : south) b.release();
    NB[0].release();
    NB[1].release();
    for (void* p : scratch) (void)hipFree(p);
    scratch.clear();
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
  };
  try {
    if (!jobs.empty()) {
      djobs = (LayerJob*)dev(sizeof(LayerJob) * jobs.size());
      upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
      SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
    }
    // ---------------- chunk size from the memory budget ----------------
    const long long D = run.D;
    const double ts = sizeof(T);
    const double small = (double)chi * D * chi, mid = small * D, big = mid * D;
    const double row_b = ts * ((Lx + 5) * small + 2.0 * mid);
    const double slab_b = ts * ((Lx + 4) * mid + 2.0 * big);
    const double absorb_b = absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub);
    const size_t per_arena = (size_t)(1.15 * std::max(std::max(row_b, slab_b), absorb_b)) + 8192;
    long long pmax = 1;
    for (const LayerSpec& q : c.spec) pmax = std::max<long long>(pmax, std::max(q.n.dim, q.s.dim));
    const long long capS = (long long)chi * pmax * chi;
    const size_t per_persist = (size_t)(ts * capS * Lx * (Ly + 3)) + sizeof(int) * (Lx + 1) * (Ly + 3) +
                               sizeof(void*) * kKinds * ns + 128;
    SPP_CHECK(hipStreamSynchronize(st));
    size_t freeb = 0, totalb = 0;
    SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
    const double avail = 0.6 * (double)freeb + (double)ctx.arena.capacity();
    long long chunk = std::max<long long>(1, (long long)(avail / (double)(per_arena + per_persist)));
    if (ctx.chunk > 0) chunk = std::min<long long>(chunk, ctx.chunk);
    chunk = std::min<long long>(chunk, n);
    chunk = std::min<long long>(chunk, 16384);
    const int nb = (int)chunk;
    ctx.arena.reserve(per_arena * (size_t)nb);
    run.nb = nb;
    for (int y = 0; y <= Ly; ++y) prepare_boundary(south[y], Lx, nb, capS);
    prepare_boundary(NB[0], Lx, nb, capS);
    prepare_boundary(NB[1], Lx, nb, capS);
    EntropyObserver eobs;
    if (out.entropy_profile || out.entropy_max) eobs.init(Lx, Ly, nb, dev);
    run.tables = (void**)dev(sizeof(void*) * (size_t)kKinds * ns * nb);
    dbase = (void**)dev(sizeof(void*) * base.size());
    dlabels = (int*)dev(sizeof(int) * (size_t)n * ns);
    run.status = (int*)dev(sizeof(int) * nb);
    run.fail = (int*)dev(sizeof(int) * nb);
    run.xsum = (double*)dev(sizeof(double) * nb);
    run.zzsum = (double*)dev(sizeof(double) * nb);
    run.den = (double*)dev(sizeof(double) * nb);
    run.num = (double*)dev(sizeof(double) * nb);
    run.numh = (double*)dev(sizeof(double) * nb);
    run.fac = (double*)dev(sizeof(double) * nb);
    run.facp = (double*)dev(sizeof(double) * nb);
    run.nrm = (double*)dev(sizeof(double) * nb);
    mant = (double*)dev(sizeof(double) * nb);
    lnamp = (double*)dev(sizeof(double) * nb);
    dsst = (long long*)dev(sizeof(long long) * Lx);
    dtens = (T**)dev(sizeof(T*) * Lx);
    upload(dbase, base.data(), sizeof(void*) * base.size(), st);
    upload(dlabels, labels_h, sizeof(int) * (size_t)n * ns, st);
    // The sweeps' absorbs apply each projected layer through its single-layer
    // tables (FP32 caches that carry them, kron.h): per chunk, the (site, label)
    // table pointer of every sample, dktab[site * nb + s].
    const bool kron = sizeof(T) == sizeof(float) && ctx.kron && !c.kron_site.empty();
    std::vector<void*> kbase;
    void** dkbase = nullptr;
    void** dktab = nullptr;
    if (kron) {
      kbase.resize((size_t)ns * 2);
      for (int i = 0; i < ns; ++i)
        for (int k = 0; k < 2; ++k) kbase[(size_t)i * 2 + k] = c.kron_tab[k][i];
      dkbase = (void**)dev(sizeof(void*) * kbase.size());
      dktab = (void**)dev(sizeof(void*) * (size_t)ns * nb);
      upload(dkbase, kbase.data(), sizeof(void*) * kbase.size(), st);
      run.ksite = c.kron_site.data();
      run.ktab = dktab;
    }
    auto kron_row = [&](RowLayers<T>& r, int slot, int site) {
      if (!kron) return;
      r.kron_site.resize(Lx, nullptr);
      r.kron_tabs.resize(Lx, nullptr);
      r.kron_site[slot] = &c.kron_site[site];
      r.kron_tabs[slot] = dktab + (size_t)site * nb;
    };
    std::vector<int> h_status(nb), h_fail(nb);
    std::vector<double> h_x(nb), h_zz(nb), h_mant(nb), h_ls(nb), h_mtn(nb), h_mts(nb), h_amp(nb);
    for (long long s0 = 0; s0 < n; s0 += nb) {
      const int cur = (int)std::min<long long>(nb, n - s0);
      run.cur = cur;
      const int* lab = dlabels + s0 * ns;
      const long long tot = (long long)kKinds * ns * cur;
      hipLaunchKernelGGL(layer_tables_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, lab, ns,
                         (void* const*)dbase, (int)kKinds, run.tables, cur, nb);
      SPP_CHECK(hipGetLastError());
      if (kron) {
        hipLaunchKernelGGL(layer_tables_kernel, dim3((unsigned)(((long long)ns * cur + 255) / 256)), dim3(256), 0,
                           st, lab, ns, (void* const*)dkbase, 1, dktab, cur, nb);
        SPP_CHECK(hipGetLastError());
      }
      SPP_CHECK(launch_fill<int>(run.status, nb, 0, st));
      SPP_CHECK(launch_fill<int>(run.fail, nb, -1, st));
      SPP_CHECK(launch_fill<double>(run.xsum, nb, 0.0, st));
      SPP_CHECK(launch_fill<double>(run.zzsum, nb, 0.0, st));
      eobs.start(cur, st);
      // south boundaries of the projected samples, rotated frame
      for (auto& b : south) b.nb = cur;
      std::vector<Leg> physS(Lx);
      for (int xr = 0; xr < Lx; ++xr) physS[xr] = c.spec[(Lx - 1 - xr) + (Ly - 1) * Lx].s;
      set_trivial(ctx, south[Ly], physS);
      for (int y = Ly - 1; y >= 0; --y) {
        RowLayers<T> rr;
        rr.L = Lx;
        rr.rotated = true;
        rr.spec.resize(Lx);
        rr.ptrs.resize(Lx);
        for (int xr = 0; xr < Lx; ++xr) {
          const int x = Lx - 1 - xr;
          const LayerSpec& sp = c.spec[x + y * Lx];
          rr.spec[xr] = LayerSpec{sp.e, sp.s, sp.w, sp.n};
          rr.ptrs[xr] = run.table(kPlain, x + y * Lx);
          kron_row(rr, xr, x + y * Lx);
        }
        eobs.attach(south[y], 0, y);
        absorb_forward<T>(ctx, south[y + 1], rr, chi, south[y]);
      }
      // north sweep with the row and slab terms
      NB[0].nb = cur;
      NB[1].nb = cur;
      std::vector<Leg> physN(Lx);
      for (int x = 0; x < Lx; ++x) physN[x] = c.spec[x].n;
      set_trivial(ctx, NB[0], physN);
      int which = 0;
      for (int y = 0; y < Ly; ++y) {
        Boundary<T>& Nc = NB[which];
        run.row_terms(Nc, south[y + 1], y);
        if (y < Ly - 1) run.slab_terms(Nc, south[y + 2], y);
        RowLayers<T> row;
        row.L = Lx;
        row.spec.assign(c.spec.begin() + y * Lx, c.spec.begin() + (y + 1) * Lx);
        row.ptrs.resize(Lx);
        for (int x = 0; x < Lx; ++x) {
          row.ptrs[x] = run.table(kPlain, x + y * Lx);
          kron_row(row, x, x + y * Lx);
        }
        eobs.attach(NB[1 - which], 1, y);
        absorb_forward<T>(ctx, Nc, row, chi, NB[1 - which]);
        which = 1 - which;
      }
      // log norm (lognorm, boundary_mps.jl:180-191)
      Boundary<T>& Nf = NB[which];
      for (int x = 0; x <= Lx; ++x)
        if (Nf.bdim[x] != 1) throw HipError("final boundary must have unit bonds");
      std::vector<long long> sst(Lx);
      for (int x = 0; x < Lx; ++x) sst[x] = Nf.tsize(x);
      upload(dsst, sst.data(), sizeof(long long) * Lx, st);
      upload(dtens, Nf.t.data(), sizeof(T*) * Lx, st);
      SPP_CHECK(launch_close_unit<T>(dtens, dsst, Lx, mant, cur, st));
      SPP_CHECK(launch_gather_lognorm(c.cand_lognorm, lab, ns, lnamp, cur, st));
      SPP_CHECK(hipMemcpyAsync(h_status.data(), run.status, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_fail.data(), run.fail, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_x.data(), run.xsum, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_zz.data(), run.zzsum, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mant.data(), mant, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_ls.data(), Nf.log_scale, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mtn.data(), Nf.max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mts.data(), south[0].max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_amp.data(), lnamp, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      if (eobs.on()) entropy_flush(ctx);
      eobs.fetch(st);
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s) {
        const long long g = s0 + s;
        int stt = h_status[s];
        const double m = h_mant[s];
        if (stt == 0 && !(m > 0.0 && std::isfinite(m))) stt = 6;
        const double tln = std::log(m) + h_ls[s] - 2.0 * h_amp[s];
        const double e = stt == 0 ? -ops.J * h_zz[s] - ops.h * h_x[s] : NAN;
        if (out.energy_total) out.energy_total[g] = e;
        if (out.energy_per_site) out.energy_per_site[g] = e / ns;
        if (out.tensor_log_norm) out.tensor_log_norm[g] = stt == 0 ? tln : NAN;
        if (out.log_norm) out.log_norm[g] = stt == 0 ? tln + 2.0 * (c.pepo_log_amplitude + h_amp[s]) : NAN;
        if (out.max_trunc) out.max_trunc[g] = std::max(h_mtn[s], h_mts[s]);
        if (out.status) out.status[g] = stt;
        if (out.fail_site) out.fail_site[g] = stt == 0 ? -1 : h_fail[s];
        eobs.write(s, g, out.entropy_profile, out.entropy_max);
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
}

// Per chunk of states: the plain double layers of every site, the south
// boundaries, then the north sweep; before row y the impurity layers of row y
// (onsite, both horizontal halves, the upper vertical half) and the lower
// vertical half of row y+1 are built into per-(kind, column) buffers, so the
// impurity storage is two rows' worth whatever the lattice height.
template <typename T>
void measure_energy_states(Context& ctx, const std::vector<const StateDev*>& states, int chi, const EnergyOps& ops,
                           EnergyOutputs& out) {
  hipStream_t st = ctx.stream;
  const int n = (int)states.size();
  if (n == 0) return;
  if (chi < 1) throw HipError("chi must be positive");
  const StateDev& s0 = *states[0];
  const int Lx = s0.Lx, Ly = s0.Ly, ns = Lx * Ly;
  for (const StateDev* sd : states) {
    if (!sd || sd->Lx != Lx || sd->Ly != Ly || (int)sd->sites.size() != ns)
      throw HipError("energy batch: states must share the lattice");
    for (int i = 0; i < ns; ++i)
      if (!batchable_sites(sd->sites[i], s0.sites[i]))
        throw HipError("energy batch: states must share every leg space (up to the auxiliary charge)");
  }
  check_energy_ops(s0, ops);

  EnergyRun<T> run(ctx, Lx, Ly, chi);
  for (int k = 0; k < kKinds; ++k) run.spec[k].resize(ns);
  std::vector<long long> lsize(ns);
  std::vector<long long> isize((size_t)kKinds * Lx, 0);  // per (kind, column): largest over rows
  long long maxout = 1;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& sd = s0.sites[i];
    const int x = i % Lx, y = i / Lx;
    const LayerSpec p{pair_leg(sd.dim[0], sd.even[0]), pair_leg(sd.dim[1], sd.even[1]),
                      pair_leg(sd.dim[4], sd.even[4]), pair_leg(sd.dim[5], sd.even[5])};
    LayerSpec ks[kKinds];
    impurity_specs(sd, p, ops.odd, ks);
    for (int k = 0; k < kKinds; ++k) {
      run.spec[k][i] = ks[k];
      const long long cnt = (long long)ks[k].w.dim * ks[k].n.dim * ks[k].e.dim * ks[k].s.dim;
      if (!kind_used(k, x, y, Lx, Ly)) continue;
      maxout = std::max(maxout, cnt);
      if (k == kPlain) lsize[i] = cnt;
      else isize[(size_t)k * Lx + x] = std::max(isize[(size_t)k * Lx + x], cnt);
      run.D = std::max<long long>(run.D, std::max(std::max(ks[k].w.dim, ks[k].n.dim), std::max(ks[k].e.dim, ks[k].s.dim)));
    }
  }
  long long layer_elems = 0, imp_elems = 0;
  for (long long v : lsize) layer_elems += v;
  for (long long v : isize) imp_elems += v;

  std::vector<Boundary<T>> south(Ly + 1);
  Boundary<T> NB[2];
  std::vector<void*> scratch;
  auto dev = [&](size_t bytes) {
    void* p = nullptr;
    SPP_CHECK(hipMalloc(&p, std::max<size_t>(bytes, 16)));
    scratch.push_back(p);
    return p;
  };
  auto cleanup = [&]() {
    entropy_reset(ctx);  // pending cuts write into buffers freed below
    for (auto& b : south) b.release();
    NB[0].release();
    NB[1].release();
    for (void* p : scratch) (void)hipFree(p);
    scratch.clear();
  };
  try {
    // ---------------- chunk size from the memory budget ----------------
    const long long D = run.D;
    const double ts = sizeof(T);
    const double small = (double)chi * D * chi, mid = small * D, big = mid * D;
    const double row_b = ts * ((Lx + 5) * small + 2.0 * mid);
    const double slab_b = ts * ((Lx + 4) * mid + 2.0 * big);
    const double absorb_b = absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub);
    const size_t per_arena = (size_t)(1.15 * std::max(std::max(row_b, slab_b), absorb_b)) + 8192;
    long long pmax = 1;
    for (const LayerSpec& q : run.spec[kPlain]) pmax = std::max<long long>(pmax, std::max(q.n.dim, q.s.dim));
    const long long capS = (long long)chi * pmax * chi;
    const size_t per_persist = (size_t)(ts * (capS * Lx * (Ly + 3) + layer_elems + imp_elems)) +
                               sizeof(int) * (Lx + 1) * (Ly + 3) + sizeof(void*) * kKinds * ns +
                               sizeof(LayerJob) * (ns + kKinds * Lx) + 128;
    SPP_CHECK(hipStreamSynchronize(st));
    size_t freeb = 0, totalb = 0;
    SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
    const double avail = 0.6 * (double)freeb + (double)ctx.arena.capacity();
    long long chunk = std::max<long long>(1, (long long)(avail / (double)(per_arena + per_persist)));
    if (ctx.chunk > 0) chunk = std::min<long long>(chunk, ctx.chunk);
    chunk = std::min<long long>(chunk, n);
    // double-layer grid: one row per job (a site, or a row's impurity layers, of each sample)
    chunk = std::min<long long>(chunk, std::max(1, 65535 / std::max(ns, (int)kKinds * Lx)));
    const int nb = (int)chunk;
    ctx.arena.reserve(per_arena * (size_t)nb);
    run.nb = nb;
    for (int y = 0; y <= Ly; ++y) prepare_boundary(south[y], Lx, nb, capS);
    prepare_boundary(NB[0], Lx, nb, capS);
    prepare_boundary(NB[1], Lx, nb, capS);
    EntropyObserver eobs;
    if (out.entropy_profile || out.entropy_max) eobs.init(Lx, Ly, nb, dev);
    T* layers = (T*)dev(sizeof(T) * (size_t)layer_elems * nb);
    T* imp = (T*)dev(sizeof(T) * (size_t)std::max(imp_elems, 1LL) * nb);
    std::vector<T*> site_base(ns), imp_base((size_t)kKinds * Lx, nullptr);
    {
      long long off = 0;
      for (int i = 0; i < ns; ++i) {
        site_base[i] = layers + off * nb;
        off += lsize[i];
      }
      off = 0;
      for (size_t k = 0; k < isize.size(); ++k) {
        imp_base[k] = imp + off * nb;
        off += isize[k];
      }
    }
    run.tables = (void**)dev(sizeof(void*) * (size_t)kKinds * ns * nb);
    // host mirror of the tables; unused (kind, site) entries point at the plain layer
    std::vector<void*> htab((size_t)kKinds * ns * nb);
    for (int k = 0; k < kKinds; ++k)
      for (int i = 0; i < ns; ++i)
        for (int s = 0; s < nb; ++s) htab[((size_t)k * ns + i) * nb + s] = site_base[i] + (long long)s * lsize[i];
    LayerJob* djobs = (LayerJob*)dev(sizeof(LayerJob) * (size_t)std::max(ns, kKinds * Lx) * nb);
    run.status = (int*)dev(sizeof(int) * nb);
    run.fail = (int*)dev(sizeof(int) * nb);
    run.xsum = (double*)dev(sizeof(double) * nb);
    run.zzsum = (double*)dev(sizeof(double) * nb);
    run.den = (double*)dev(sizeof(double) * nb);
    run.num = (double*)dev(sizeof(double) * nb);
    run.numh = (double*)dev(sizeof(double) * nb);
    run.fac = (double*)dev(sizeof(double) * nb);
    run.facp = (double*)dev(sizeof(double) * nb);
    run.nrm = (double*)dev(sizeof(double) * nb);
    double* mant = (double*)dev(sizeof(double) * nb);
    long long* dsst = (long long*)dev(sizeof(long long) * Lx);
    T** dtens = (T**)dev(sizeof(T*) * Lx);
    std::vector<int> h_status(nb), h_fail(nb);
    std::vector<double> h_x(nb), h_zz(nb), h_mant(nb), h_ls(nb), h_mtn(nb), h_mts(nb);
    std::vector<LayerJob> jobs;
    // build the given jobs (the job table and the pointer tables may still be
    // read by queued kernels: wait before rewriting them)
    auto run_jobs = [&]() {
      SPP_CHECK(hipStreamSynchronize(st));
      upload(run.tables, htab.data(), sizeof(void*) * htab.size(), st);
      if (!jobs.empty()) {
        upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
        SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
      }
    };
    for (long long c0 = 0; c0 < n; c0 += nb) {
      const int cur = (int)std::min<long long>(nb, n - c0);
      run.cur = cur;
      jobs.clear();
      for (int i = 0; i < ns; ++i)
        for (int s = 0; s < cur; ++s)
          jobs.push_back(impurity_job(states[c0 + s]->sites[i], kPlain, ops, site_base[i] + (long long)s * lsize[i]));
      run_jobs();
      SPP_CHECK(launch_fill<int>(run.status, nb, 0, st));
      SPP_CHECK(launch_fill<int>(run.fail, nb, -1, st));
      SPP_CHECK(launch_fill<double>(run.xsum, nb, 0.0, st));
      SPP_CHECK(launch_fill<double>(run.zzsum, nb, 0.0, st));
      eobs.start(cur, st);
      // south boundaries, rotated frame
      for (auto& b : south) b.nb = cur;
      const std::vector<LayerSpec>& P = run.spec[kPlain];
      std::vector<Leg> physS(Lx);
      for (int xr = 0; xr < Lx; ++xr) physS[xr] = P[(Lx - 1 - xr) + (Ly - 1) * Lx].s;
      set_trivial(ctx, south[Ly], physS);
      for (int y = Ly - 1; y >= 0; --y) {
        RowLayers<T> rr;
        rr.L = Lx;
        rr.rotated = true;
        rr.spec.resize(Lx);
        rr.ptrs.resize(Lx);
        for (int xr = 0; xr < Lx; ++xr) {
          const int x = Lx - 1 - xr;
          const LayerSpec& sp = P[x + y * Lx];
          rr.spec[xr] = LayerSpec{sp.e, sp.s, sp.w, sp.n};
          rr.ptrs[xr] = run.table(kPlain, x + y * Lx);
        }
        eobs.attach(south[y], 0, y);
        absorb_forward<T>(ctx, south[y + 1], rr, chi, south[y]);
      }
      // north sweep with the row and slab terms
      NB[0].nb = cur;
      NB[1].nb = cur;
      std::vector<Leg> physN(Lx);
      for (int x = 0; x < Lx; ++x) physN[x] = P[x].n;
      set_trivial(ctx, NB[0], physN);
      int which = 0;
      for (int y = 0; y < Ly; ++y) {
        jobs.clear();
        auto add = [&](int kind, int x, int yy) {
          const int i = x + yy * Lx;
          if (yy >= Ly || !kind_used(kind, x, yy, Lx, Ly)) return;
          const LayerSpec& q = run.spec[kind][i];
          const long long cnt = (long long)q.w.dim * q.n.dim * q.e.dim * q.s.dim;
          T* base = imp_base[(size_t)kind * Lx + x];
          for (int s = 0; s < cur; ++s) {
            T* o = base + (long long)s * cnt;
            htab[((size_t)kind * ns + i) * nb + s] = o;
            jobs.push_back(impurity_job(states[c0 + s]->sites[i], kind, ops, o));
          }
        };
        for (int x = 0; x < Lx; ++x) {
          add(kOnsite, x, y);
          add(kHorzC, x, y);
          add(kHorzD, x, y);
          add(kVertC, x, y);
          add(kVertD, x, y + 1);
        }
        run_jobs();
        Boundary<T>& Nc = NB[which];
        run.row_terms(Nc, south[y + 1], y);
        if (y < Ly - 1) run.slab_terms(Nc, south[y + 2], y);
        RowLayers<T> row;
        row.L = Lx;
        row.spec.assign(P.begin() + y * Lx, P.begin() + (y + 1) * Lx);
        row.ptrs.resize(Lx);
        for (int x = 0; x < Lx; ++x) row.ptrs[x] = run.table(kPlain, x + y * Lx);
        eobs.attach(NB[1 - which], 1, y);
        absorb_forward<T>(ctx, Nc, row, chi, NB[1 - which]);
        which = 1 - which;
      }
      // log norm (lognorm, boundary_mps.jl:180-191)
      Boundary<T>& Nf = NB[which];
      for (int x = 0; x <= Lx; ++x)
        if (Nf.bdim[x] != 1) throw HipError("final boundary must have unit bonds");
      std::vector<long long> sst(Lx);
      for (int x = 0; x < Lx; ++x) sst[x] = Nf.tsize(x);
      upload(dsst, sst.data(), sizeof(long long) * Lx, st);
      upload(dtens, Nf.t.data(), sizeof(T*) * Lx, st);
      SPP_CHECK(launch_close_unit<T>(dtens, dsst, Lx, mant, cur, st));
      SPP_CHECK(hipMemcpyAsync(h_status.data(), run.status, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_fail.data(), run.fail, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_x.data(), run.xsum, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_zz.data(), run.zzsum, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mant.data(), mant, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_ls.data(), Nf.log_scale, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mtn.data(), Nf.max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mts.data(), south[0].max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      if (eobs.on()) entropy_flush(ctx);
      eobs.fetch(st);
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s) {
        const long long g = c0 + s;
        int stt = h_status[s];
        const double m = h_mant[s];
        if (stt == 0 && !(m > 0.0 && std::isfinite(m))) stt = 6;
        const double tln = std::log(m) + h_ls[s];
        const double e = stt == 0 ? -ops.J * h_zz[s] - ops.h * h_x[s] : NAN;
        if (out.energy_total) out.energy_total[g] = e;
        if (out.energy_per_site) out.energy_per_site[g] = e / ns;
        if (out.tensor_log_norm) out.tensor_log_norm[g] = stt == 0 ? tln : NAN;
        if (out.log_norm) out.log_norm[g] = stt == 0 ? tln + 2.0 * states[g]->log_amplitude : NAN;
        if (out.max_trunc) out.max_trunc[g] = std::max(h_mtn[s], h_mts[s]);
        if (out.status) out.status[g] = stt;
        if (out.fail_site) out.fail_site[g] = stt == 0 ? -1 : h_fail[s];
        eobs.write(s, g, out.entropy_profile, out.entropy_max);
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
}

template void measure_energy_states<float>(Context&, const std::vector<const StateDev*>&, int, const EnergyOps&,
                                           EnergyOutputs&);
template void measure_energy_states<double>(Context&, const std::vector<const StateDev*>&, int, const EnergyOps&,
                                            EnergyOutputs&);

template void measure_energy_batch<float>(Context&, const StateDev&, const CacheDev<float>&, int, int, const int*,
                                          const EnergyOps&, EnergyOutputs&);
template void measure_energy_batch<double>(Context&, const StateDev&, const CacheDev<double>&, int, int, const int*,
                                           const EnergyOps&, EnergyOutputs&);

}  // namespace spp
