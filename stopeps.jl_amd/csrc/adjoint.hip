// Environment adjoint of the log norm. For real tensors
//   d log <psi|psi> / d A_i = (N_i A_i + N_i^T A_i) / (A_i N_i A_i),
// where N_i is the single-site norm matrix: the site hole
//   H[w,n,e,s] = (L[a,w,b] T[a,n,c]) (B[b,s,d] R[c,e,d])
// of _site_hole (src/sampling/born.jl:75-83) over the fused (ket, bra) pair
// legs, built from the north boundary above row y, the south boundary below it
// (boundary_environments, boundary_mps.jl:193-209) and the row messages
// (row_environments, boundary_mps.jl:261-273). The scale of the normalised
// messages cancels in the ratio. Exact when chi reaches the exact bond.
#include "adjoint.h"
#include "boundary_ops.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace spp {
namespace {

// index of the (ket i, bra j) pair in the parity-sorted pair order of
// split_pair (double_layer_kernel): ee, oo, eo, oe blocks
__device__ inline int pair_index(int i, int dk, int ek, int j, int db, int eb) {
  const int ok = dk - ek, ob = db - eb;
  const int ee = ek * eb, oo = ok * ob, eo = ek * ob;
  const bool ie = i < ek, je = j < eb;
  if (ie && je) return i * eb + j;
  if (!ie && !je) return ee + (i - ek) * ob + (j - eb);
  if (ie) return ee + oo + i * ob + (j - eb);
  return ee + oo + eo + (i - ek) * eb + j;
}

// G[k] = sum_b H[pair(k, b)] A[b] + sum_b H[pair(b, k)] A[b] over the virtual
// legs (same P, Aux), and z[s] += sum_k A[k] (sum_b H[pair(k, b)] A[b]).
// One grid row per sample; sites dense (W,N,P,A,E,S), H dense [W][N][E][S].
template <typename T>
__global__ void hole_grad_kernel(const T* H, long long hstride, const double* const* A, int dW, int dN, int dP,
                                 int dA, int dE, int dS, int eW, int eN, int eE, int eS, int eP, int eAux,
                                 double* G, long long gstride, double* z) {
  const int s = blockIdx.y;
  const long long total = (long long)dW * dN * dP * dA * dE * dS;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const T* h = H + (long long)s * hstride;
  const double* a = A[s];
  __shared__ double part[256];
  double contrib = 0.0;
  if (t < total) {
    long long r = t;
    const int is = (int)(r % dS); r /= dS;
    const int ie = (int)(r % dE); r /= dE;
    const int ia = (int)(r % dA); r /= dA;
    const int ip = (int)(r % dP); r /= dP;
    const int in = (int)(r % dN); r /= dN;
    const int iw = (int)r;
    const int par = (iw >= eW) + (in >= eN) + (ie >= eE) + (is >= eS) + (ip >= eP) + (ia >= eAux);
    double g1 = 0.0, g2 = 0.0;
    if ((par & 1) == 0) {
      const int DN = dN * dN, DE = dE * dE, DS = dS * dS;
      const int vpar = (iw >= eW) + (in >= eN) + (ie >= eE) + (is >= eS);
      for (int w = 0; w < dW; ++w)
        for (int n = 0; n < dN; ++n)
          for (int e = 0; e < dE; ++e)
            for (int sv = 0; sv < dS; ++sv) {
              const int bpar = (w >= eW) + (n >= eN) + (e >= eE) + (sv >= eS);
              if (((bpar + vpar) & 1) != 0) continue;
              const double ab = a[((((long long)w * dN + n) * dP + ip) * dA + ia) * dE * dS + (long long)e * dS + sv];
              if (ab == 0.0) continue;
              const long long h1 = (((long long)pair_index(iw, dW, eW, w, dW, eW) * DN +
                                     pair_index(in, dN, eN, n, dN, eN)) * DE +
                                    pair_index(ie, dE, eE, e, dE, eE)) * DS +
                                   pair_index(is, dS, eS, sv, dS, eS);
              const long long h2 = (((long long)pair_index(w, dW, eW, iw, dW, eW) * DN +
                                     pair_index(n, dN, eN, in, dN, eN)) * DE +
                                    pair_index(e, dE, eE, ie, dE, eE)) * DS +
                                   pair_index(sv, dS, eS, is, dS, eS);
              g1 += (double)h[h1] * ab;
              g2 += (double)h[h2] * ab;
            }
    }
    G[(long long)s * gstride + t] = g1 + g2;
    contrib = a[t] * g1;
  }
  part[threadIdx.x] = contrib;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&z[s], part[0]);
}

// G[s][:] /= z[s]; z == 0 or nonfinite -> status 6 at the site
__global__ void grad_scale_kernel(double* G, long long gstride, long long size, const double* z, int* status,
                                  int* fail, int site, int nb) {
  const int s = blockIdx.y;
  if (s >= nb) return;
  const double zz = z[s];
  const bool ok = zz != 0.0 && isfinite(zz);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < size; t += (long long)gridDim.x * blockDim.x)
    G[(long long)s * gstride + t] = ok ? G[(long long)s * gstride + t] / zz : NAN;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0 && status[s] == 0) {
    status[s] = 6;
    fail[s] = site;
  }
}

long long count_legs(const std::vector<Leg>& l) {
  long long n = 1;
  for (auto& g : l) n *= g.dim;
  return n;
}

}  // namespace

template <typename T>
void lognorm_grad_batch(Context& ctx, const std::vector<const StateDev*>& states, int chi, GradOutputs& out) {
  hipStream_t st = ctx.stream;
  const int n = (int)states.size();
  if (n == 0) return;
  if (chi < 1) throw HipError("chi must be positive");
  const StateDev& s0 = *states[0];
  const int Lx = s0.Lx, Ly = s0.Ly, ns = Lx * Ly;
  for (const StateDev* sd : states) {
    if (!sd || sd->Lx != Lx || sd->Ly != Ly || (int)sd->sites.size() != ns)
      throw HipError("gradient batch: states must share the lattice");
    for (int i = 0; i < ns; ++i)
      for (int k = 0; k < 6; ++k)
        if (sd->sites[i].dim[k] != s0.sites[i].dim[k] || sd->sites[i].even[k] != s0.sites[i].even[k])
          throw HipError("gradient batch: states must share every leg space");
  }
  std::vector<LayerSpec> spec(ns);
  std::vector<long long> lsize(ns), ssize(ns), goff(ns + 1, 0);
  long long D = 1, maxout = 1;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = s0.sites[i];
    spec[i] = LayerSpec{pair_leg(s.dim[0], s.even[0]), pair_leg(s.dim[1], s.even[1]), pair_leg(s.dim[4], s.even[4]),
                        pair_leg(s.dim[5], s.even[5])};
    lsize[i] = (long long)spec[i].w.dim * spec[i].n.dim * spec[i].e.dim * spec[i].s.dim;
    ssize[i] = s.size();
    goff[i + 1] = goff[i] + ssize[i];
    maxout = std::max(maxout, lsize[i]);
    D = std::max<long long>(D, std::max(std::max(spec[i].w.dim, spec[i].n.dim), std::max(spec[i].e.dim, spec[i].s.dim)));
  }
  long long layer_elems = 0;
  for (long long v : lsize) layer_elems += v;
  long long pmax = 1;
  for (const LayerSpec& q : spec) pmax = std::max<long long>(pmax, std::max(q.n.dim, q.s.dim));
  const long long cap = (long long)chi * pmax * chi;
  const double ts = sizeof(T);
  const double small = (double)chi * D * chi, mid = small * D;
  const double hole = std::max(mid, (double)D * D * D * D);
  const double row_b = ts * ((Lx + 4) * small + 3.0 * hole);
  const size_t per_arena =
      (size_t)(1.15 * std::max(row_b, absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub))) + 8192;
  const size_t per_persist = (size_t)(ts * (layer_elems + (double)(Ly + 3) * Lx * cap)) +
                             sizeof(double) * goff[ns] + sizeof(LayerJob) * ns + 512;
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
    for (auto& b : south) b.release();
    NB[0].release();
    NB[1].release();
    for (void* p : scratch) (void)hipFree(p);
    scratch.clear();
  };
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
    std::vector<void*> hptr((size_t)ns * nb);
    for (int i = 0; i < ns; ++i)
      for (int s = 0; s < nb; ++s) hptr[(size_t)i * nb + s] = site_base[i] + (long long)s * lsize[i];
    upload(ptrs, hptr.data(), sizeof(void*) * hptr.size(), st);
    const double** aptr = (const double**)dev(sizeof(double*) * (size_t)ns * nb);
    LayerJob* djobs = (LayerJob*)dev(sizeof(LayerJob) * (size_t)ns * nb);
    double* G = (double*)dev(sizeof(double) * (size_t)goff[ns] * nb);
    for (int y = 0; y <= Ly; ++y) prepare_boundary(south[y], Lx, nb, cap);
    prepare_boundary(NB[0], Lx, nb, cap);
    prepare_boundary(NB[1], Lx, nb, cap);
    double* zsum = (double*)dev(sizeof(double) * nb);
    double* nrm = (double*)dev(sizeof(double) * nb);
    double* fac = (double*)dev(sizeof(double) * nb);
    int* status = (int*)dev(sizeof(int) * nb);
    int* fail = (int*)dev(sizeof(int) * nb);
    double* mant = (double*)dev(sizeof(double) * nb);
    long long* dsst = (long long*)dev(sizeof(long long) * Lx);
    T** dtens = (T**)dev(sizeof(T*) * Lx);
    std::vector<LayerJob> jobs;
    std::vector<const double*> haptr((size_t)ns * nb);
    std::vector<double> h_mant(nb), h_ls(nb), h_mtn(nb), h_mts(nb), h_g;
    std::vector<int> h_status(nb);
    for (long long c0 = 0; c0 < n; c0 += nb) {
      const int cur = (int)std::min<long long>(nb, n - c0);
      jobs.clear();
      for (int i = 0; i < ns; ++i)
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
          haptr[(size_t)i * nb + s] = sd.d;
        }
      upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
      SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
      upload(aptr, haptr.data(), sizeof(double*) * haptr.size(), st);
      SPP_CHECK(launch_fill<int>(status, nb, 0, st));
      SPP_CHECK(launch_fill<int>(fail, nb, -1, st));
      auto table = [&](int site) { return (void* const*)(ptrs + (size_t)site * nb); };
      // south boundaries, rotated frame
      for (auto& b : south) b.nb = cur;
      std::vector<Leg> physS(Lx);
      for (int xr = 0; xr < Lx; ++xr) physS[xr] = spec[(Lx - 1 - xr) + (Ly - 1) * Lx].s;
      set_trivial(ctx, south[Ly], physS);
      for (int y = Ly - 1; y >= 0; --y) {
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
      }
      NB[0].nb = cur;
      NB[1].nb = cur;
      std::vector<Leg> physN(Lx);
      for (int x = 0; x < Lx; ++x) physN[x] = spec[x].n;
      set_trivial(ctx, NB[0], physN);
      int which = 0;
      for (int y = 0; y < Ly; ++y) {
        const Boundary<T>& N = NB[which];
        const Boundary<T>& S = south[y + 1];
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
        // right messages
        std::vector<T*> R(Lx + 1);
        std::vector<std::vector<Leg>> Rl(Lx + 1);
        for (int x = 0; x <= Lx; ++x) {
          Leg wl = x < Lx ? spec[x + y * Lx].w : spec[(Lx - 1) + y * Lx].e;
          Rl[x] = {N.bond(x), wl, south_bond(S, Lx, x)};
          R[x] = take(count_legs(Rl[x]));
        }
        const long long big = std::max((long long)chi * D * D * chi, D * D * D * D);
        T* BX = take(big);
        T* BY = take(big);
        T* Hb = take(big);
        SPP_CHECK(launch_fill<T>(R[Lx], cur, T(1), st));
        auto Rref = [&](int x) { return dense_ref(R[x], count_legs(Rl[x]), "ced", Rl[x]); };
        for (int x = Lx - 1; x >= 0; --x) {
          const int site = x + y * Lx;
          const LayerSpec& sp = spec[site];
          std::vector<Leg> xl = {N.bond(x), Rl[x + 1][2], N.phys(x), Rl[x + 1][1]};
          TRef Xr = folded_ref(BX, count_legs(xl), "adne", xl);
          contract<T>(ctx, Xr, N.ref(x, "anc"), Rref(x + 1), cur);
          std::vector<Leg> yl = {N.bond(x), Rl[x + 1][2], sp.w, sp.s};
          TRef Yr = folded_ref(BY, count_legs(yl), "adws", yl);
          TRef lay = dense_ref(nullptr, 0, "wnes", {sp.w, sp.n, sp.e, sp.s});
          lay.ptrs = table(site);
          contract<T>(ctx, Yr, Xr, lay, cur);
          TRef Rr = dense_ref(R[x], count_legs(Rl[x]), "awb", Rl[x]);
          contract<T>(ctx, Rr, Yr, bottom_ref(S, Lx, x, "bsd", true), cur);
          normalize(Rr, site);
        }
        // left sweep with the site holes
        T* Lb = take((long long)chi * D * chi);
        T* O0 = take((long long)chi * D * chi);
        SPP_CHECK(launch_fill<T>(Lb, cur, T(1), st));
        std::vector<Leg> Ll = {N.bond(0), spec[y * Lx].w, south_bond(S, Lx, 0)};
        for (int x = 0; x < Lx; ++x) {
          const int site = x + y * Lx;
          const LayerSpec& sp = spec[site];
          TRef Lr = dense_ref(Lb, count_legs(Ll), "awb", Ll);
          // X[w,b,c,n] = L[a,w,b] T[a,n,c]
          std::vector<Leg> xl = {Ll[1], Ll[2], N.bond(x + 1), N.phys(x)};
          TRef Xr = dense_ref(BX, count_legs(xl), "wbcn", xl);
          contract<T>(ctx, Xr, Lr, N.ref(x, "anc"), cur);
          // Y[b,c,e,s] = B[b,s,d] R[c,e,d]
          std::vector<Leg> yl = {south_bond(S, Lx, x), N.bond(x + 1), Rl[x + 1][1], sp.s};
          TRef Yr = dense_ref(BY, count_legs(yl), "bces", yl);
          contract<T>(ctx, Yr, bottom_ref(S, Lx, x, "bsd", true), Rref(x + 1), cur);
          // H[w,n,e,s] = X Y over (b, c)
          std::vector<Leg> hl = {sp.w, sp.n, sp.e, sp.s};
          TRef Hr = dense_ref(Hb, count_legs(hl), "wnes", hl);
          contract<T>(ctx, Hr, Xr, Yr, cur);
          // gradient of this site for every sample of the chunk
          const SiteDev& sd = s0.sites[site];
          SPP_CHECK(launch_fill<double>(zsum, cur, 0.0, st));
          const long long sz = ssize[site];
          dim3 grid((unsigned)((sz + 255) / 256), (unsigned)cur);
          hipLaunchKernelGGL(hole_grad_kernel<T>, grid, dim3(256), 0, st, Hb, count_legs(hl),
                             aptr + (size_t)site * nb, sd.dim[0], sd.dim[1], sd.dim[2], sd.dim[3], sd.dim[4],
                             sd.dim[5], sd.even[0], sd.even[1], sd.even[4], sd.even[5], sd.even[2], sd.even[3],
                             G + goff[site], goff[ns], zsum);
          SPP_CHECK(hipGetLastError());
          dim3 sgrid((unsigned)std::min<long long>((sz + 255) / 256, 1024), (unsigned)cur);
          hipLaunchKernelGGL(grad_scale_kernel, sgrid, dim3(256), 0, st, G + goff[site], goff[ns], sz, zsum, status,
                             fail, site, cur);
          SPP_CHECK(hipGetLastError());
          if (x < Lx - 1) {
            // left <- safe_scale(left_step of the plain layer)
            std::vector<Leg> yl2 = {Ll[2], N.bond(x + 1), sp.e, sp.s};
            TRef Y2 = folded_ref(BY, count_legs(yl2), "bces", yl2);
            TRef lay = dense_ref(nullptr, 0, "wnes", {sp.w, sp.n, sp.e, sp.s});
            lay.ptrs = table(site);
            contract<T>(ctx, Y2, Xr, lay, cur);
            std::vector<Leg> ol = {N.bond(x + 1), sp.e, south_bond(S, Lx, x + 1)};
            TRef Or = dense_ref(O0, count_legs(ol), "ced", ol);
            contract<T>(ctx, Or, Y2, bottom_ref(S, Lx, x, "bsd", true), cur);
            normalize(Or, site);
            std::swap(Lb, O0);
            Ll = ol;
          }
        }
        ctx.arena.release(mark);
        RowLayers<T> row;
        row.L = Lx;
        row.spec.assign(spec.begin() + y * Lx, spec.begin() + (y + 1) * Lx);
        row.ptrs.resize(Lx);
        for (int x = 0; x < Lx; ++x) row.ptrs[x] = table(x + y * Lx);
        absorb_forward<T>(ctx, NB[which], row, chi, NB[1 - which]);
        which = 1 - which;
      }
      Boundary<T>& Nf = NB[which];
      for (int x = 0; x <= Lx; ++x)
        if (Nf.bdim[x] != 1) throw HipError("final boundary must have unit bonds");
      std::vector<long long> sst(Lx);
      for (int x = 0; x < Lx; ++x) sst[x] = Nf.tsize(x);
      upload(dsst, sst.data(), sizeof(long long) * Lx, st);
      upload(dtens, Nf.t.data(), sizeof(T*) * Lx, st);
      SPP_CHECK(launch_close_unit<T>(dtens, dsst, Lx, mant, cur, st));
      SPP_CHECK(hipMemcpyAsync(h_mant.data(), mant, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_ls.data(), Nf.log_scale, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mtn.data(), Nf.max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mts.data(), south[0].max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_status.data(), status, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      if (out.grads)
        SPP_CHECK(hipMemcpyAsync(out.grads + c0 * goff[ns], G, sizeof(double) * (size_t)goff[ns] * cur,
                                 hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s) {
        const long long g = c0 + s;
        const double m = h_mant[s];
        int stt = h_status[s];
        if (stt == 0 && !(m > 0.0 && std::isfinite(m))) stt = 6;
        const double tln = stt == 0 ? std::log(m) + h_ls[s] : NAN;
        if (out.tensor_log_norm) out.tensor_log_norm[g] = tln;
        if (out.log_norm) out.log_norm[g] = stt == 0 ? tln + 2.0 * states[g]->log_amplitude : NAN;
        if (out.max_trunc) out.max_trunc[g] = std::max(h_mtn[s], h_mts[s]);
        if (out.status) out.status[g] = stt;
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

template void lognorm_grad_batch<float>(Context&, const std::vector<const StateDev*>&, int, GradOutputs&);
template void lognorm_grad_batch<double>(Context&, const std::vector<const StateDev*>&, int, GradOutputs&);

}  // namespace spp
