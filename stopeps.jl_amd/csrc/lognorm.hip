// Batched lognorm (boundary_mps.jl:162-191): per sample, the double layers of
// its sites (ket (x) bra traced over the physical and auxiliary legs), Ly
// compressed north absorbs from the trivial boundary (contract_layers), and
// the closing of the final unit-bond boundary; log_norm adds the state's
// 2 log_amplitude. Samples are processed in chunks; each sample's layers are
// addressed through per-sample pointer tables, so the absorb is the one of the
// Born draw.
#include "lognorm.h"
#include "boundary_ops.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace spp {

template <typename T>
void lognorm_batch(Context& ctx, const std::vector<const StateDev*>& states, int chi, LognormOutputs& out,
                   char* pool, size_t pool_bytes) {
  hipStream_t st = ctx.stream;
  const int n = (int)states.size();
  if (n == 0) return;
  if (chi < 1) throw HipError("chi must be positive");
  const StateDev& s0 = *states[0];
  const int Lx = s0.Lx, Ly = s0.Ly, ns = Lx * Ly;
  for (const StateDev* sd : states) {
    if (!sd || sd->Lx != Lx || sd->Ly != Ly || (int)sd->sites.size() != ns)
      throw HipError("lognorm batch: states must share the lattice");
    for (int i = 0; i < ns; ++i)
      if (!batchable_sites(sd->sites[i], s0.sites[i]))
        throw HipError("lognorm batch: states must share every leg space (up to the auxiliary charge)");
  }
  std::vector<LayerSpec> spec(ns);
  std::vector<long long> lsize(ns);
  long long D = 1, maxout = 1;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = s0.sites[i];
    spec[i] = LayerSpec{pair_leg(s.dim[0], s.even[0]), pair_leg(s.dim[1], s.even[1]), pair_leg(s.dim[4], s.even[4]),
                        pair_leg(s.dim[5], s.even[5])};
    lsize[i] = (long long)spec[i].w.dim * spec[i].n.dim * spec[i].e.dim * spec[i].s.dim;
    maxout = std::max(maxout, lsize[i]);
    D = std::max<long long>(D, std::max(std::max(spec[i].w.dim, spec[i].n.dim), std::max(spec[i].e.dim, spec[i].s.dim)));
  }
  long long layer_elems = 0;
  for (long long v : lsize) layer_elems += v;
  long long pmax = 1;
  for (const LayerSpec& q : spec) pmax = std::max<long long>(pmax, q.s.dim);
  const long long cap = (long long)chi * pmax * chi;
  const double ts = sizeof(T);
  const size_t per_arena = (size_t)(1.15 * absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub)) + 4096;
  const size_t per_persist = (size_t)(ts * (layer_elems + 2.0 * Lx * cap)) + sizeof(LayerJob) * ns + 256;
  // device bytes of one chunk of nb samples carved from a pool (256-byte
  // aligned pieces: the buffers below, the two boundaries, the dot partials)
  auto pool_need = [&](long long nb) -> double {
    const double al = 256.0;
    const double bufs = ts * layer_elems * nb + 8.0 * ns * nb + sizeof(LayerJob) * (double)ns * nb + 8.0 * nb +
                        8.0 * Lx + 8.0 * Lx + 6 * al;
    const double bnd = 2.0 * (Lx * (ts * nb * cap + al) + (Lx - 1) * (4.0 * nb + al) + 2 * (8.0 * nb + al));
    const double dots = 8.0 * (double)nb * kDotChunks + al;
    return bufs + bnd + dots + (double)per_arena * nb;
  };
  long long chunk;
  if (pool) {
    if (pool_need(1) > (double)pool_bytes)
      throw HipError("lognorm without allocation: one state needs " + std::to_string((long long)pool_need(1)) +
                     " bytes of device scratch, the static scratch holds " + std::to_string((long long)pool_bytes));
    chunk = 1;
    while (chunk < n && pool_need(chunk + 1) <= (double)pool_bytes) ++chunk;
  } else {
    SPP_CHECK(hipStreamSynchronize(st));
    size_t freeb = 0, totalb = 0;
    SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
    const double avail = 0.6 * (double)freeb + (double)ctx.arena.capacity();
    chunk = std::max<long long>(1, (long long)(avail / (double)(per_arena + per_persist)));
  }
  if (ctx.chunk > 0) chunk = std::min<long long>(chunk, ctx.chunk);
  chunk = std::min<long long>(chunk, n);
  chunk = std::min<long long>(chunk, std::max(1, 65535 / ns));  // double-layer grid: one row per (sample, site)
  const int nb = (int)chunk;

  Boundary<T> NB[2];
  std::vector<void*> scratch;
  size_t pool_top = 0;
  auto dev = [&](size_t bytes) -> void* {
    bytes = std::max<size_t>(bytes, 16);
    if (pool) {
      const size_t at = (pool_top + 255) & ~size_t(255);
      if (at + bytes > pool_bytes) throw HipError("lognorm without allocation: scratch exhausted");
      pool_top = at + bytes;
      return pool + at;
    }
    void* p = nullptr;
    SPP_CHECK(hipMalloc(&p, bytes));
    scratch.push_back(p);
    return p;
  };
  auto make_boundary = [&](Boundary<T>& b) {
    if (!pool) {
      prepare_boundary(b, Lx, nb, cap);
      return;
    }
    b.L = Lx;
    b.nb = nb;
    b.bdim.assign(Lx + 1, 1);
    b.beven.assign(Lx + 1, 1);
    b.bsplit.assign(Lx + 1, nullptr);
    b.pdim.assign(Lx, 1);
    b.peven.assign(Lx, 1);
    b.t.assign(Lx, nullptr);
    for (int x = 0; x < Lx; ++x) b.t[x] = (T*)dev(sizeof(T) * nb * cap);
    for (int k = 1; k < Lx; ++k) b.bsplit[k] = (int*)dev(sizeof(int) * nb);
    b.log_scale = (double*)dev(sizeof(double) * nb);
    b.max_trunc = (double*)dev(sizeof(double) * nb);
    b.owns = false;
  };
  auto cleanup = [&]() {
    NB[0].release();
    NB[1].release();
    for (void* p : scratch) (void)hipFree(p);
    scratch.clear();
    if (pool) {
      ctx.arena.restore();
      ctx.noalloc_dot = nullptr;
      ctx.noalloc_dot_n = 0;
    }
  };
  try {
    if (pool) {
      ctx.noalloc_dot = (double*)dev(sizeof(double) * (size_t)nb * kDotChunks);
      ctx.noalloc_dot_n = (size_t)nb * kDotChunks;
    } else {
      ctx.arena.reserve(per_arena * (size_t)nb);
    }
    T* layers = (T*)dev(sizeof(T) * (size_t)layer_elems * nb);
    std::vector<T*> site_base(ns);
    {
      long long off = 0;
      for (int i = 0; i < ns; ++i) {
        site_base[i] = layers + off * nb;
        off += lsize[i];
      }
    }
    // per-sample pointer tables [site][sample]
    void** ptrs = (void**)dev(sizeof(void*) * (size_t)ns * nb);
    std::vector<void*> hptr((size_t)ns * nb);
    for (int i = 0; i < ns; ++i)
      for (int s = 0; s < nb; ++s) hptr[(size_t)i * nb + s] = site_base[i] + (long long)s * lsize[i];
    upload(ptrs, hptr.data(), sizeof(void*) * hptr.size(), st);
    LayerJob* djobs = (LayerJob*)dev(sizeof(LayerJob) * (size_t)ns * nb);
    make_boundary(NB[0]);
    make_boundary(NB[1]);
    double* mant = (double*)dev(sizeof(double) * nb);
    long long* dsst = (long long*)dev(sizeof(long long) * Lx);
    T** dtens = (T**)dev(sizeof(T*) * Lx);
    if (pool) {
      // the absorb workspace is the rest of the pool
      const size_t at = (pool_top + 255) & ~size_t(255);
      if (at >= pool_bytes || pool_bytes - at < per_arena * (size_t)nb)
        throw HipError("lognorm without allocation: scratch exhausted");
      ctx.arena.borrow(pool + at, pool_bytes - at);
    }
    std::vector<LayerJob> jobs;
    std::vector<double> h_mant(nb), h_ls(nb), h_mt(nb);
    for (long long s0i = 0; s0i < n; s0i += nb) {
      const int cur = (int)std::min<long long>(nb, n - s0i);
      jobs.clear();
      for (int i = 0; i < ns; ++i)
        for (int s = 0; s < cur; ++s) {
          const SiteDev& sd = states[s0i + s]->sites[i];
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
      upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
      SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
      NB[0].nb = cur;
      NB[1].nb = cur;
      std::vector<Leg> phys(Lx);
      for (int x = 0; x < Lx; ++x) phys[x] = spec[x].n;
      set_trivial(ctx, NB[0], phys);
      int which = 0;
      for (int y = 0; y < Ly; ++y) {
        RowLayers<T> row;
        row.L = Lx;
        row.spec.assign(spec.begin() + y * Lx, spec.begin() + (y + 1) * Lx);
        row.ptrs.resize(Lx);
        for (int x = 0; x < Lx; ++x) row.ptrs[x] = ptrs + (size_t)(x + y * Lx) * nb;
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
      SPP_CHECK(hipMemcpyAsync(h_mt.data(), Nf.max_trunc, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s) {
        const long long g = s0i + s;
        const double m = h_mant[s];
        const bool ok = m > 0.0 && std::isfinite(m) && std::isfinite(h_ls[s]);
        const double tln = ok ? std::log(m) + h_ls[s] : NAN;
        if (out.tensor_log_norm) out.tensor_log_norm[g] = tln;
        if (out.log_norm) out.log_norm[g] = ok ? tln + 2.0 * states[g]->log_amplitude : NAN;
        if (out.max_trunc) out.max_trunc[g] = h_mt[s];
        if (out.status) out.status[g] = ok ? 0 : 6;
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

template void lognorm_batch<float>(Context&, const std::vector<const StateDev*>&, int, LognormOutputs&, char*,
                                   size_t);
template void lognorm_batch<double>(Context&, const std::vector<const StateDev*>&, int, LognormOutputs&, char*,
                                    size_t);

}  // namespace spp
