// C ABI (include/stopeps_hip.h). Every entry point catches exceptions and
// reports them through spp_last_error.
#include "../../include/stopeps_hip.h"
#include "energy.h"
#include "lognorm.h"
#include "adjoint.h"
#include "fu.h"
#include "sampler.h"
#include "psd.h"
#include "tridiag.h"
#include <cstring>
#include <limits>
#include <memory>

using namespace spp;

// Device memory of spp_lognorm_batch_noalloc: one static array per GPU, held
// until the process ends. Host code never passes this name to a kernel or a
// copy; it uses the per-device address resolved by hipGetSymbolAddress in
// spp_create (Context::static_scratch).
static __device__ __align__(256) char scratch[SPP_NOALLOC_SCRATCH_BYTES];

struct spp_ctx {
  Context c;
};

struct spp_state {
  StateDev s;
};

struct spp_cache {
  int precision = 0;
  int split = 0;  // BF16 piece count of the FP32 products the cache was built with (0: FP32 matrix cores)
  std::unique_ptr<CacheDev<float>> f;
  std::unique_ptr<CacheDev<double>> d;
};

// A state on which a full-update step failed is refused by the compute entry
// points (it keeps its sites from the failing bond on).
static void usable(const StateDev& s) {
  if (s.failed)
    throw HipError(std::string("state invalidated by a failed full-update step: ") + fu_failure_message(s.failed));
}

// Calls through a cache compute their products the way the cache was built
// (the context's precision setting takes effect for caches built afterwards).
struct CacheProducts {
  Context& c;
  int saved;
  CacheProducts(Context& ctx, const spp_cache& cache) : c(ctx), saved(ctx.split) { ctx.split = cache.split; }
  ~CacheProducts() { c.split = saved; }
  CacheProducts(const CacheProducts&) = delete;
  CacheProducts& operator=(const CacheProducts&) = delete;
};

template <typename F>
static int guarded(spp_ctx* ctx, F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    if (ctx) ctx->c.last_error = e.what();
    return -1;
  } catch (...) {
    if (ctx) ctx->c.last_error = "unknown error";
    return -2;
  }
}

extern "C" {

int spp_create(int device, spp_ctx** out) {
  if (!out) return -3;
  *out = nullptr;
  spp_ctx* ctx = new spp_ctx();
  int rc = guarded(ctx, [&] {
    int n = 0;
    SPP_CHECK(hipGetDeviceCount(&n));
    if (n <= 0) throw HipError("no HIP device available");
    if (device < 0 || device >= n) throw HipError("invalid device index");
    ctx->c.device = device;
    SPP_CHECK(hipSetDevice(device));
    SPP_CHECK(hipStreamCreate(&ctx->c.stream));
    // this device's copy of the static scratch array (the no-allocation entry
    // point), resolved once per device with that device current
    static std::vector<char*> scratch_of_device;
    if ((int)scratch_of_device.size() < n) scratch_of_device.resize(n, nullptr);
    if (!scratch_of_device[device]) {
      void* ptr = nullptr;
      SPP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      if (!ptr) throw HipError("static scratch has no device address");
      scratch_of_device[device] = (char*)ptr;
    }
    ctx->c.static_scratch = scratch_of_device[device];
  });
  if (rc != 0) {
    static thread_local std::string err;
    err = ctx->c.last_error;
    delete ctx;
    return rc;
  }
  *out = ctx;
  return 0;
}

int spp_destroy(spp_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
  return 0;
}

const char* spp_last_error(const spp_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

int spp_synchronize(spp_ctx* ctx) {
  return guarded(ctx, [&] { SPP_CHECK(hipStreamSynchronize(ctx->c.stream)); });
}

int spp_set_precision(spp_ctx* ctx, int precision) {
  return guarded(ctx, [&] {
    if (precision < 0 || precision > 3)
      throw HipError("precision must be 0 (FP64), 1 (FP32), 2 (FP32 from six BF16 products) or 3 (FP32 from "
                     "three BF16 products)");
    ctx->c.precision = precision == 0 ? 0 : 1;
    ctx->c.split = precision == 2 ? 3 : precision == 3 ? 2 : 0;
  });
}

int spp_set_option(spp_ctx* ctx, const char* name, double value) {
  return guarded(ctx, [&] {
    std::string n = name ? name : "";
    if (n == "ksub") {
      if (value < 2 || value > 64) throw HipError("ksub must lie in [2, 64]");
      ctx->c.ksub = (int)value;
    } else if (n == "max_subspace_iter") {
      if (value < 1) throw HipError("max_subspace_iter must be positive");
      ctx->c.max_subspace_iter = (int)value;
    } else if (n == "subspace_tol") {
      if (!(value > 0)) throw HipError("subspace_tol must be positive");
      ctx->c.subspace_tol = value;
    } else if (n == "cheb_degree") {
      if (value < 1 || value > 32) throw HipError("cheb_degree must lie in [1, 32]");
      ctx->c.cheb_degree = (int)value;
    } else if (n == "profile") {
      ctx->c.flush_timer();
      ctx->c.profile = value != 0 ? 1 : 0;
    } else if (n == "boundary_cap_even" || n == "boundary_cap_odd") {
      if (!(value == value) || value > 1e9) throw HipError(n + " must be an integer (negative: no budget)");
      ctx->c.bcap[n == "boundary_cap_even" ? 0 : 1] = value < 0 ? -1 : (int)value;
    } else if (n == "fu_fail_sample") {
      ctx->c.fu_fail_sample = value < 0 ? -1 : (long long)value;
    } else if (n == "fu_fail_bond") {
      ctx->c.fu_fail_bond = (int)value;
    } else if (n == "fu_fail_code") {
      if (value < 0 || value > 10) throw HipError("fu_fail_code must lie in [0, 10]");
      ctx->c.fu_fail_code = (int)value;
    } else if (n == "chunk") {
      if (value < 0) throw HipError("chunk must be >= 0");
      ctx->c.chunk = (long long)value;
    } else if (n == "gram_chunk_bytes") {
      if (!(value >= 1)) throw HipError("gram_chunk_bytes must be >= 1");
      ctx->c.gram_chunk_bytes = value;
    } else if (n == "autotune") {
      if (value != 0 && value != 1 && value != -1)
        throw HipError("autotune must be 1 (on), 0 (off) or -1 (forget the tuned shapes)");
      if (value == -1) ctx->c.tile_table.clear();
      else ctx->c.autotune = (int)value;
    } else if (n == "graphs") {
      if (value != 0 && value != 1) throw HipError("graphs must be 0 or 1");
      ctx->c.graphs = (int)value;
    } else if (n == "kron") {
      if (value != 0 && value != 1 && value != 2) throw HipError("kron must be 0, 1 or 2");
      ctx->c.kron = (int)value;
    } else if (n == "kron_groups") {
      if (value < 1 || value > 64) throw HipError("kron_groups must lie in [1, 64]");
      ctx->c.kron_groups = (int)value;
    } else if (n == "kron_rows") {
      if (value < 0 || value > 32) throw HipError("kron_rows must lie in [0, 32]");
      ctx->c.kron_rows = (int)value;
    } else if (n == "gemm_tile") {
      if (value < -1 || value >= kMaxTileShapes) throw HipError("gemm_tile must lie in [-1, 5]");
      ctx->c.force_tile = (int)value;
    } else {
      throw HipError("unknown option " + n);
    }
  });
}

int spp_get_counters(spp_ctx* ctx, double* counters, int n) {
  if (!ctx || !counters) return -3;
  if (guarded(ctx, [&] { ctx->c.flush_timer(); }) != 0) return -1;
  const Counters& k = ctx->c.counters;
  double v[23] = {k.gemm_flops,        k.gemm_launches,    (double)ctx->c.last_iterations,
                  (double)ctx->c.arena.peak(), k.gemm_ms, k.gemm_flops_profiled,
                  k.prec_launches[0],  k.prec_ms[0],       k.prec_flops[0],
                  k.prec_launches[1],  k.prec_ms[1],       k.prec_flops[1],
                  k.eig_calls,         k.eig_iterations,   k.eig_filter_gemms,
                  k.fu_repairs,        k.tuned_classes,    k.tuned_changed,
                  k.tune_ms,           k.graph_captures,   k.graph_replays,
                  k.eig_unconverged,   k.kron_launches};
  int m = n < 23 ? n : 23;
  for (int i = 0; i < m; ++i) counters[i] = v[i];
  return m;
}

int spp_reset_counters(spp_ctx* ctx) {
  if (!ctx) return -3;
  if (guarded(ctx, [&] { ctx->c.flush_timer(); }) != 0) return -1;
  ctx->c.counters = Counters{};
  return 0;
}

int spp_state_upload(spp_ctx* ctx, int Lx, int Ly, int kind, const int* dims, const int* evens,
                     const double* const* data, double beta, double log_amplitude, spp_state** out) {
  if (out) *out = nullptr;
  std::unique_ptr<spp_state> st(new spp_state());
  int rc = guarded(ctx, [&] {
    if (!out || !dims || !evens || !data) throw HipError("null argument");
    if (Lx <= 0 || Ly <= 0) throw HipError("lattice dimensions must be positive");
    if (kind != 0 && kind != 1) throw HipError("kind must be 0 (PEPO) or 1 (PEPS)");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    StateDev& s = st->s;
    s.Lx = Lx;
    s.Ly = Ly;
    s.kind = kind;
    s.beta = beta;
    s.log_amplitude = log_amplitude;
    s.sites.resize((size_t)Lx * Ly);
    for (int i = 0; i < Lx * Ly; ++i) {
      SiteDev& d = s.sites[i];
      for (int k = 0; k < 6; ++k) {
        d.dim[k] = dims[6 * i + k];
        d.even[k] = evens[6 * i + k];
        if (d.dim[k] <= 0 || d.even[k] < 0 || d.even[k] > d.dim[k]) throw HipError("invalid leg description");
      }
      int x = i % Lx, y = i / Lx;
      if ((x == 0 && d.dim[0] != 1) || (y == 0 && d.dim[1] != 1) || (x == Lx - 1 && d.dim[4] != 1) ||
          (y == Ly - 1 && d.dim[5] != 1))
        throw HipError("nontrivial OBC leg at site " + std::to_string(i));
      if (x > 0 && (dims[6 * (i - 1) + 4] != d.dim[0] || evens[6 * (i - 1) + 4] != d.even[0]))
        throw HipError("horizontal space mismatch");
      if (y > 0 && (dims[6 * (i - Lx) + 5] != d.dim[1] || evens[6 * (i - Lx) + 5] != d.even[1]))
        throw HipError("vertical space mismatch");
      long long n = d.size();
      SPP_CHECK(hipMalloc(&d.d, sizeof(double) * n));
      upload(d.d, data[i], sizeof(double) * n, ctx->c.stream);
    }
    s.live.assign((size_t)Lx * Ly * 8, 0);
    for (int i = 0; i < Lx * Ly; ++i) {
      const int legs[4] = {0, 1, 4, 5};
      for (int v = 0; v < 4; ++v) {
        s.live[(size_t)i * 8 + 2 * v] = s.sites[i].even[legs[v]];
        s.live[(size_t)i * 8 + 2 * v + 1] = s.sites[i].dim[legs[v]] - s.sites[i].even[legs[v]];
      }
    }
    SPP_CHECK(hipStreamSynchronize(ctx->c.stream));
  });
  if (rc == 0) *out = st.release();
  return rc;
}

int spp_project_batch(spp_ctx* ctx, const spp_state* pepo, int n, const int* labels, const double* basis,
                      const int* capacity, spp_state** out, double* log_amplitude, double* normalizations) {
  return guarded(ctx, [&] {
    if (!pepo || (n > 0 && (!labels || !out))) throw HipError("null argument");
    if (n < 0) throw HipError("negative sample count");
    if (pepo->s.kind != 0) throw HipError("projection needs a PEPO");
    if (capacity && (capacity[0] < 0 || capacity[1] < 0 || capacity[0] + capacity[1] <= 0))
      throw HipError("invalid bond capacity");
    usable(pepo->s);
    SPP_CHECK(hipSetDevice(ctx->c.device));
    for (int i = 0; i < n; ++i) out[i] = nullptr;
    std::vector<StateDev*> made;
    project_batch(ctx->c, pepo->s, n, labels, basis, capacity, made, normalizations);
    for (int i = 0; i < n; ++i) {
      spp_state* h = new spp_state();
      // move the device state into the handle (the block owns every site)
      StateDev& d = h->s;
      d.Lx = made[i]->Lx;
      d.Ly = made[i]->Ly;
      d.kind = made[i]->kind;
      d.beta = made[i]->beta;
      d.log_amplitude = made[i]->log_amplitude;
      d.sites = made[i]->sites;
      d.live = made[i]->live;
      d.block = made[i]->block;
      made[i]->block = nullptr;
      made[i]->sites.clear();
      delete made[i];
      out[i] = h;
      if (log_amplitude) log_amplitude[i] = d.log_amplitude;
    }
  });
}

int spp_state_set_live(spp_ctx* ctx, spp_state* state, const int* live) {
  return guarded(ctx, [&] {
    if (!state || !live) throw HipError("null argument");
    StateDev& s = state->s;
    const int ns = s.Lx * s.Ly;
    const int legs[4] = {0, 1, 4, 5};
    for (int i = 0; i < ns; ++i)
      for (int v = 0; v < 4; ++v) {
        const int e = live[i * 8 + 2 * v], o = live[i * 8 + 2 * v + 1];
        const SiteDev& d = s.sites[i];
        if (e < 0 || o < 0 || e > d.even[legs[v]] || o > d.dim[legs[v]] - d.even[legs[v]])
          throw HipError("live counts exceed the leg capacity");
      }
    for (int i = 0; i < ns; ++i) {
      const int x = i % s.Lx, y = i / s.Lx;
      if (x + 1 < s.Lx && (live[i * 8 + 4] != live[(i + 1) * 8 + 0] || live[i * 8 + 5] != live[(i + 1) * 8 + 1]))
        throw HipError("live counts differ across a horizontal bond");
      if (y + 1 < s.Ly &&
          (live[i * 8 + 6] != live[(i + s.Lx) * 8 + 2] || live[i * 8 + 7] != live[(i + s.Lx) * 8 + 3]))
        throw HipError("live counts differ across a vertical bond");
    }
    s.live.assign(live, live + (size_t)ns * 8);
  });
}

int spp_state_download(spp_ctx* ctx, const spp_state* state, double* const* data, double* beta,
                       double* log_amplitude, int* live) {
  return guarded(ctx, [&] {
    if (!state || !data) throw HipError("null argument");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    const StateDev& s = state->s;
    for (size_t i = 0; i < s.sites.size(); ++i) {
      if (!data[i]) throw HipError("null site buffer");
      SPP_CHECK(hipMemcpyAsync(data[i], s.sites[i].d, sizeof(double) * s.sites[i].size(), hipMemcpyDeviceToHost,
                               ctx->c.stream));
    }
    SPP_CHECK(hipStreamSynchronize(ctx->c.stream));
    if (beta) *beta = s.beta;
    if (log_amplitude) *log_amplitude = s.log_amplitude;
    if (live) std::memcpy(live, s.live.data(), sizeof(int) * s.live.size());
  });
}

int spp_fu_evolution_step(spp_ctx* ctx, int n, spp_state* const* states, int chi, int order, double delta, int D,
                          const double* options, const double* gates_h, const double* gates_v, double* records,
                          double* normalizations, double* max_env_trunc, double* max_boundary_entropy,
                          int* status) {
  return guarded(ctx, [&] {
    if (n < 0) throw HipError("negative state count");
    if (n > 0 && !states) throw HipError("null state list");
    if (!options) throw HipError("null options");
    if (order != 1 && order != 2) throw HipError("order must be 1 or 2");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    std::vector<StateDev*> v(n);
    for (int i = 0; i < n; ++i) {
      if (!states[i]) throw HipError("null state");
      usable(states[i]->s);
      v[i] = &states[i]->s;
    }
    FuParams P;
    P.opts = fu::Opts{(int)options[0], options[1], options[2], (int)options[3], options[4], D > 0 ? D : 0};
    if (P.opts.max_iterations < 0 || P.opts.linear_max_iterations < 0) throw HipError("negative iteration limit");
    P.order = order;
    P.delta = delta;
    P.gates_h = gates_h;
    P.gates_v = gates_v;
    FuOutputs o;
    o.records = records;
    o.normalizations = normalizations;
    o.max_env_trunc = max_env_trunc;
    o.max_entropy = max_boundary_entropy;
    o.status = status;
    if (ctx->c.precision == 1) fu_evolution_step<float>(ctx->c, v, chi, P, o);
    else fu_evolution_step<double>(ctx->c, v, chi, P, o);
  });
}

int spp_fu_set_repair(spp_ctx* ctx, spp_fu_repair_fn fn, void* user) {
  if (!ctx) return -3;
  ctx->c.fu_repair_fn = (void*)fn;
  ctx->c.fu_repair_user = fn ? user : nullptr;
  return 0;
}

int spp_state_free(spp_state* state) {
  delete state;
  return 0;
}

int spp_sampling_cache_build_kind(spp_ctx* ctx, const spp_state* state, int chi, const double* basis, int kind,
                                  spp_cache** out) {
  if (out) *out = nullptr;
  std::unique_ptr<spp_cache> c(new spp_cache());
  int rc = guarded(ctx, [&] {
    if (!state || !out) throw HipError("null argument");
    if (chi < 1) throw HipError("chi must be positive");
    if (kind != 0 && kind != 1) throw HipError("unknown sampling kind");
    if (kind == 0 && state->s.kind != 0) throw HipError("the sampling cache needs a PEPO");
    if (kind == 1 && state->s.kind != 1) throw HipError("proposal requires PEPS");
    usable(state->s);
    SPP_CHECK(hipSetDevice(ctx->c.device));
    c->precision = ctx->c.precision;
    c->split = c->precision == 1 ? ctx->c.split : 0;
    if (c->precision == 1) {
      c->f.reset(new CacheDev<float>());
      build_cache<float>(ctx->c, state->s, chi, basis, *c->f, kind);
    } else {
      c->d.reset(new CacheDev<double>());
      build_cache<double>(ctx->c, state->s, chi, basis, *c->d, kind);
    }
  });
  if (rc == 0) *out = c.release();
  return rc;
}

int spp_sampling_cache_build(spp_ctx* ctx, const spp_state* pepo, int chi, const double* basis,
                             spp_cache** out) {
  return spp_sampling_cache_build_kind(ctx, pepo, chi, basis, 0, out);
}

int spp_sampling_cache_free(spp_cache* cache) {
  delete cache;
  return 0;
}

int spp_sampling_cache_info(const spp_cache* cache, double* south_max_trunc) {
  if (!cache || !south_max_trunc) return -3;
  const std::vector<double>& v = cache->precision == 1 ? cache->f->south_max_trunc : cache->d->south_max_trunc;
  for (size_t i = 0; i < v.size(); ++i) south_max_trunc[i] = v[i];
  return 0;
}

int spp_born_sample_batch(spp_ctx* ctx, spp_cache* cache, int n, const double* uniforms, const int* labels_in,
                          int* labels_out, double* log_prob, double* conditionals, double* max_trunc,
                          double* tensor_log_norm, double* log_amplitude, int* status, int* fail_site) {
  return spp_born_sample_batch_repairs(ctx, cache, n, uniforms, labels_in, labels_out, log_prob, conditionals,
                                       max_trunc, tensor_log_norm, log_amplitude, status, fail_site, nullptr);
}

int spp_born_sample_batch_repairs(spp_ctx* ctx, spp_cache* cache, int n, const double* uniforms,
                                  const int* labels_in, int* labels_out, double* log_prob, double* conditionals,
                                  double* max_trunc, double* tensor_log_norm, double* log_amplitude, int* status,
                                  int* fail_site, double* repairs) {
  return guarded(ctx, [&] {
    if (!cache) throw HipError("null cache");
    if (n < 0) throw HipError("negative sample count");
    if (!uniforms && !labels_in) throw HipError("an explicit RNG is required");
    if (labels_in) {
      // evaluation mode: every label indexes one of the two basis vectors
      const int ns = cache->precision == 1 ? cache->f->Lx * cache->f->Ly : cache->d->Lx * cache->d->Ly;
      for (long long i = 0; i < (long long)n * ns; ++i)
        if (labels_in[i] != 0 && labels_in[i] != 1) throw HipError("labels must be 0 or 1");
    }
    SPP_CHECK(hipSetDevice(ctx->c.device));
    DrawOutputs o;
    o.labels = labels_out;
    o.log_prob = log_prob;
    o.conditionals = conditionals;
    o.max_trunc = max_trunc;
    o.tensor_log_norm = tensor_log_norm;
    o.log_amplitude = log_amplitude;
    o.status = status;
    o.fail_site = fail_site;
    o.repairs = repairs;
    CacheProducts products(ctx->c, *cache);
    if (cache->precision == 1) draw_batch<float>(ctx->c, *cache->f, n, uniforms, labels_in, o);
    else draw_batch<double>(ctx->c, *cache->d, n, uniforms, labels_in, o);
  });
}

int spp_measure_energy_batch(spp_ctx* ctx, const spp_cache* cache, const spp_state* pepo, int chi, int n,
                             const int* labels, const double* X, const double* F, const double* G, int factor_odd,
                             double J, double h, double* energy_total, double* energy_per_site, double* log_norm,
                             double* tensor_log_norm, double* max_trunc, int* status, int* fail_site,
                             double* boundary_entropy, double* max_boundary_entropy) {
  return guarded(ctx, [&] {
    if (!cache || !pepo) throw HipError("null cache or state");
    if (!X || !F || !G) throw HipError("null operator");
    if (n < 0) throw HipError("negative sample count");
    if (n > 0 && !labels) throw HipError("labels are required");
    if (factor_odd != 0 && factor_odd != 1) throw HipError("factor_odd must be 0 or 1");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    EnergyOps ops;
    std::memcpy(ops.X, X, sizeof(ops.X));
    std::memcpy(ops.F, F, sizeof(ops.F));
    std::memcpy(ops.G, G, sizeof(ops.G));
    ops.odd = factor_odd;
    ops.J = J;
    ops.h = h;
    EnergyOutputs o;
    o.energy_total = energy_total;
    o.energy_per_site = energy_per_site;
    o.log_norm = log_norm;
    o.tensor_log_norm = tensor_log_norm;
    o.max_trunc = max_trunc;
    o.status = status;
    o.fail_site = fail_site;
    o.entropy_profile = boundary_entropy;
    o.entropy_max = max_boundary_entropy;
    CacheProducts products(ctx->c, *cache);
    if (cache->precision == 1) measure_energy_batch<float>(ctx->c, pepo->s, *cache->f, chi, n, labels, ops, o);
    else measure_energy_batch<double>(ctx->c, pepo->s, *cache->d, chi, n, labels, ops, o);
  });
}

int spp_measure_energy_states(spp_ctx* ctx, int n, const spp_state* const* states, int chi, const double* X,
                              const double* F, const double* G, int factor_odd, double J, double h,
                              double* energy_total, double* energy_per_site, double* log_norm,
                              double* tensor_log_norm, double* max_trunc, int* status, int* fail_site,
                              double* boundary_entropy, double* max_boundary_entropy) {
  return guarded(ctx, [&] {
    if (n < 0) throw HipError("negative state count");
    if (n > 0 && !states) throw HipError("null state list");
    if (!X || !F || !G) throw HipError("null operator");
    if (factor_odd != 0 && factor_odd != 1) throw HipError("factor_odd must be 0 or 1");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    std::vector<const StateDev*> v(n);
    for (int i = 0; i < n; ++i) {
      if (!states[i]) throw HipError("null state");
      usable(states[i]->s);
      v[i] = &states[i]->s;
    }
    EnergyOps ops;
    std::memcpy(ops.X, X, sizeof(ops.X));
    std::memcpy(ops.F, F, sizeof(ops.F));
    std::memcpy(ops.G, G, sizeof(ops.G));
    ops.odd = factor_odd;
    ops.J = J;
    ops.h = h;
    EnergyOutputs o;
    o.energy_total = energy_total;
    o.energy_per_site = energy_per_site;
    o.log_norm = log_norm;
    o.tensor_log_norm = tensor_log_norm;
    o.max_trunc = max_trunc;
    o.status = status;
    o.fail_site = fail_site;
    o.entropy_profile = boundary_entropy;
    o.entropy_max = max_boundary_entropy;
    if (ctx->c.precision == 1) measure_energy_states<float>(ctx->c, v, chi, ops, o);
    else measure_energy_states<double>(ctx->c, v, chi, ops, o);
  });
}

int spp_lognorm_batch(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* log_norm,
                      double* tensor_log_norm, double* max_trunc, int* status) {
  return guarded(ctx, [&] {
    if (n < 0) throw HipError("negative state count");
    if (n > 0 && !states) throw HipError("null state list");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    std::vector<const StateDev*> v(n);
    for (int i = 0; i < n; ++i) {
      if (!states[i]) throw HipError("null state");
      usable(states[i]->s);
      v[i] = &states[i]->s;
    }
    LognormOutputs o;
    o.log_norm = log_norm;
    o.tensor_log_norm = tensor_log_norm;
    o.max_trunc = max_trunc;
    o.status = status;
    if (ctx->c.precision == 1) lognorm_batch<float>(ctx->c, v, chi, o);
    else lognorm_batch<double>(ctx->c, v, chi, o);
  });
}

int spp_lognorm_batch_noalloc(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* log_norm,
                              double* tensor_log_norm, double* max_trunc, int* status) {
  return guarded(ctx, [&] {
    if (n < 0) throw HipError("negative state count");
    if (n > 0 && !states) throw HipError("null state list");
    if (!ctx->c.static_scratch) throw HipError("static scratch not resolved for this context");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    std::vector<const StateDev*> v(n);
    for (int i = 0; i < n; ++i) {
      if (!states[i]) throw HipError("null state");
      usable(states[i]->s);
      v[i] = &states[i]->s;
    }
    LognormOutputs o;
    o.log_norm = log_norm;
    o.tensor_log_norm = tensor_log_norm;
    o.max_trunc = max_trunc;
    o.status = status;
    // launch timing creates events: off for the duration of the call
    const int profile = ctx->c.profile;
    ctx->c.profile = 0;
    try {
      if (ctx->c.precision == 1)
        lognorm_batch<float>(ctx->c, v, chi, o, ctx->c.static_scratch, SPP_NOALLOC_SCRATCH_BYTES);
      else
        lognorm_batch<double>(ctx->c, v, chi, o, ctx->c.static_scratch, SPP_NOALLOC_SCRATCH_BYTES);
    } catch (...) {
      ctx->c.profile = profile;
      throw;
    }
    ctx->c.profile = profile;
  });
}

int spp_lognorm_grad_batch(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* grads,
                           double* log_norm, double* tensor_log_norm, double* max_trunc, int* status) {
  return guarded(ctx, [&] {
    if (n < 0) throw HipError("negative state count");
    if (n > 0 && !states) throw HipError("null state list");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    std::vector<const StateDev*> v(n);
    for (int i = 0; i < n; ++i) {
      if (!states[i]) throw HipError("null state");
      usable(states[i]->s);
      v[i] = &states[i]->s;
    }
    GradOutputs o;
    o.grads = grads;
    o.log_norm = log_norm;
    o.tensor_log_norm = tensor_log_norm;
    o.max_trunc = max_trunc;
    o.status = status;
    if (ctx->c.precision == 1) lognorm_grad_batch<float>(ctx->c, v, chi, o);
    else lognorm_grad_batch<double>(ctx->c, v, chi, o);
  });
}

int spp_selftest_fu_bond(const int* dims, const double* options, const double* metric, const double* target,
                         double* left, double* right, double* record) {
  if (!dims || !options || !metric || !target || !left || !right || !record) return -3;
  try {
    fu::Dims g{dims[0], dims[1], dims[2], dims[3], dims[4], dims[5], dims[6], dims[7]};
    fu::Opts o{(int)options[0], options[1], options[2], (int)options[3], options[4]};
    if (g.na <= 0 || g.nb <= 0 || g.d <= 0 || g.D <= 0 || g.ea < 0 || g.ea > g.na || g.eb < 0 || g.eb > g.nb ||
        g.ed < 0 || g.ed > g.d || g.eD < 0 || g.eD > g.D)
      return -3;
    fu::Result r = fu_bond_host(g, o, metric, target, left, right);
    const double v[12] = {r.initial_cost, r.final_cost, r.residual, r.target_norm, (double)r.iterations,
                          (double)r.limit_reached, (double)r.stop, (double)r.status, (double)r.kept_even,
                          (double)r.kept_odd, (double)r.floored, (double)r.nonconverged};
    for (int i = 0; i < 12; ++i) record[i] = v[i];
    return 0;
  } catch (...) {
    return -1;
  }
}

int spp_selftest_jacobi(int kind, int m, int n, const double* A, int max_sweeps, double* values, int* converged) {
  try {
    if ((kind != 0 && kind != 1) || m <= 0 || n <= 0 || (kind == 0 && m != n) || (kind == 1 && m < n) || !A ||
        !values || !converged || max_sweeps < 0)
      return -3;
    *converged = jacobi_host(kind, m, n, A, max_sweeps, values) ? 1 : 0;
    return 0;
  } catch (...) {
    return -1;
  }
}

int spp_selftest_fu_pair(int horizontal, const int* dims_a, const int* evens_a, const int* dims_b,
                         const int* evens_b, const double* site_a, const double* site_b, const int* live_a,
                         const int* live_b, const double* gate, double* qa, double* qb, double* ca, double* cb,
                         double* target, int* geometry) {
  if (!dims_a || !evens_a || !dims_b || !evens_b || !site_a || !site_b || !live_a || !live_b || !gate || !qa || !qb ||
      !ca || !cb || !target || !geometry)
    return -3;
  try {
    return fu_pair_host(horizontal, dims_a, evens_a, dims_b, evens_b, site_a, site_b, live_a, live_b, gate, qa, qb,
                        ca, cb, target, geometry);
  } catch (...) {
    return -1;
  }
}

int spp_selftest_gemm(spp_ctx* ctx, int precision, int M, int N, int K, int evenM, int evenN, int evenK,
                      int layout, const double* A, const double* B, double* C) {
  return guarded(ctx, [&] {
    if (M <= 0 || N <= 0 || K <= 0) throw HipError("dimensions must be positive");
    if (evenM < 0 || evenM > M || evenN < 0 || evenN > N || evenK < 0 || evenK > K)
      throw HipError("even counts must lie in [0, dim]");
    if (precision < 0 || precision > 3) throw HipError("precision must lie in [0, 3]");
    hipStream_t st = ctx->c.stream;
    const int split = ctx->c.split;
    ctx->c.split = precision == 2 ? 3 : precision == 3 ? 2 : 0;
    struct Restore {
      Context& c;
      int v;
      ~Restore() { c.split = v; }
    } restore{ctx->c, split};
    if (precision > 1) precision = 1;
    size_t es = precision == 1 ? sizeof(float) : sizeof(double);
    void *dA, *dB, *dC, *tA = nullptr, *tB = nullptr;
    SPP_CHECK(hipMalloc(&dA, sizeof(double) * M * K));
    SPP_CHECK(hipMalloc(&dB, sizeof(double) * K * N));
    SPP_CHECK(hipMalloc(&dC, sizeof(double) * M * N));
    upload(dA, A, sizeof(double) * M * K, st);
    upload(dB, B, sizeof(double) * K * N, st);
    Leg lm = fixed_leg(M, evenM), ln = fixed_leg(N, evenN), lk = fixed_leg(K, evenK);
    if (layout < 0 || layout > 3) throw HipError("layout must lie in [0, 3]");
    // storage order of the operands: A as [M][K] or [K][M], B as [K][N] or [N][K]
    const bool at = (layout & 1) != 0, bt = (layout & 2) != 0;
    auto aref = [&](void* p) {
      return at ? dense_ref(p, 0, "km", {lk, lm}) : dense_ref(p, 0, "mk", {lm, lk});
    };
    auto bref = [&](void* p) {
      return bt ? dense_ref(p, 0, "nk", {ln, lk}) : dense_ref(p, 0, "kn", {lk, ln});
    };
    const double qnan = std::numeric_limits<double>::quiet_NaN();
    if (precision == 1) {
      SPP_CHECK(hipMalloc(&tA, es * M * K));
      SPP_CHECK(hipMalloc(&tB, es * K * N));
      void* tC;
      SPP_CHECK(hipMalloc(&tC, es * M * N));
      SPP_CHECK(launch_fill<float>((float*)tC, (long long)M * N, (float)qnan, st));
      SPP_CHECK((launch_convert<float, double>((float*)tA, (const double*)dA, (long long)M * K, st)));
      SPP_CHECK((launch_convert<float, double>((float*)tB, (const double*)dB, (long long)K * N, st)));
      contract<float>(ctx->c, dense_ref(tC, 0, "mn", {lm, ln}), aref(tA), bref(tB), 1);
      SPP_CHECK((launch_convert<double, float>((double*)dC, (const float*)tC, (long long)M * N, st)));
      SPP_CHECK(hipStreamSynchronize(st));
      (void)hipFree(tC);
    } else {
      SPP_CHECK(launch_fill<double>((double*)dC, (long long)M * N, qnan, st));
      contract<double>(ctx->c, dense_ref(dC, 0, "mn", {lm, ln}), aref(dA), bref(dB), 1);
    }
    SPP_CHECK(hipMemcpyAsync(C, dC, sizeof(double) * M * N, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    if (tA) (void)hipFree(tA);
    if (tB) (void)hipFree(tB);
  });
}

int spp_selftest_spectrum(spp_ctx* ctx, int nprob, int ld, const int* n, const double* blocks, double* eigenvalues) {
  return guarded(ctx, [&] {
    if (!n || !blocks || !eigenvalues) throw HipError("null argument");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    spectrum_selftest(ctx->c, nprob, ld, n, blocks, eigenvalues);
  });
}

int spp_selftest_psd(spp_ctx* ctx, int nprob, int ld, const int* n, const double* blocks, const double* vectors,
                     double* eigenvalues, double* projections, int* failed) {
  return guarded(ctx, [&] {
    if (nprob < 0 || ld < 1) throw HipError("bad sizes");
    if (nprob > 0 && (!n || !blocks || !vectors || !eigenvalues || !projections || !failed))
      throw HipError("null argument");
    if ((size_t)(kTriWaves + 6 * (size_t)ld) * sizeof(double) > 160 * 1024) throw HipError("block too large");
    SPP_CHECK(hipSetDevice(ctx->c.device));
    psd_selftest(ctx->c, nprob, ld, n, blocks, vectors, eigenvalues, projections, failed);
  });
}

int spp_selftest_ql(int n, double* d, double* e, double* vectors, int nvec, int maxit) {
  if (n < 0 || nvec < 0 || (n > 0 && (!d || !e || (nvec > 0 && !vectors)))) return -3;
  return ql_project(n, d, e, vectors, n, nvec, maxit);
}

int spp_selftest_kron(spp_ctx* ctx, const int* d, const int* e, int dphi, int ephi, const double* S, int stage,
                      int nl, int el, int nr, int er, const float* x, float* y, int rows_per_block) {
  if (!ctx || !d || !e || !S || !x || !y || stage < 0 || stage >= kKronStages || nl < 1 || nr < 1) return -3;
  KronSite ks;
  if (!kron_site_layout(d, e, dphi, ephi, ks)) return -2;
  const int q = kron_layer_parity(d, e, dphi, ephi, S);
  if (q < 0) return -2;
  return guarded(ctx, [&] {
    hipStream_t st = ctx->c.stream;
    std::vector<char> tabs;
    kron_build_tables(ks, S, q, tabs);
    kron_prepare();
    // pair legs W=0, N=1, E=2, S=3 of the absorb's frame: the layer's own legs, or for
    // stage 3 (the first transfer of a rotated row) the rotated frame (E, S, W, N)
    static const int kFrame[2][4] = {{0, 1, 2, 3}, {2, 3, 0, 1}};
    const int* fr = kFrame[stage == 3 ? 1 : 0];
    long long pd[4], pf[4];
    int pe[4];
    for (int k = 0; k < 4; ++k) {
      const int dk = d[fr[k]], ek = e[fr[k]];
      pd[k] = (long long)dk * dk;
      pe[k] = ek * ek + (dk - ek) * (dk - ek);
      pf[k] = std::max<long long>(pe[k], pd[k] - pe[k]);
    }
    KronRows kr;
    kr.rc = nr;
    kr.leven = el;
    kr.reven = er;
    long long nx, ny;
    if (stage == 0 || stage == 3) {
      kr.nrows = (long long)nl * nr * pd[2];
      kr.nf = (int)pd[2];
      kr.ef = pe[2];
      kr.sx_slab = pd[2] * pd[1] * pf[2];
      kr.sxf = pd[1] * pf[2];
      kr.sx1 = pf[2];
      nx = (long long)nl * nr * pd[2] * pd[1] * pf[2];
      ny = (long long)nl * nr * pd[2] * pd[0] * pf[3];
    } else if (stage == 1) {
      kr.nrows = (long long)nl * nr * pd[0];
      kr.nf = (int)pd[0];
      kr.ef = pe[0];
      kr.sx_slab = pd[2] * pd[0] * pf[3];
      kr.sxf = pf[3];
      kr.sx1 = pd[0] * pf[3];
      nx = (long long)nl * nr * pd[2] * pd[0] * pf[3];
      ny = (long long)nl * nr * pd[0] * pd[0] * pf[1];
    } else {
      // x = X[W][nl][nr][fold N], rows (l, R); y = Y[nl][nr][E][fold S]
      kr.nrows = (long long)nl * nr;
      kr.nf = 1;
      kr.ef = 1;
      kr.sx_slab = pf[1];
      kr.sxf = 0;
      kr.sx1 = (long long)nl * nr * pf[1];
      nx = pd[0] * nl * nr * pf[1];
      ny = (long long)nl * nr * pd[2] * pf[3];
    }
    kr.sx2 = 1;
    kr.xs = nx;
    kr.ys = ny;
    char* dt = nullptr;
    float *dx = nullptr, *dy = nullptr;
    void** dp = nullptr;
    auto cleanup = [&] {
      for (void* p : {(void*)dt, (void*)dx, (void*)dy, (void*)dp})
        if (p) (void)hipFree(p);
    };
    try {
      SPP_CHECK(hipMalloc(&dt, tabs.size()));
      SPP_CHECK(hipMalloc(&dx, sizeof(float) * nx));
      SPP_CHECK(hipMalloc(&dy, sizeof(float) * ny));
      SPP_CHECK(hipMalloc(&dp, sizeof(void*)));
      upload(dt, tabs.data(), tabs.size(), st);
      upload(dx, x, sizeof(float) * nx, st);
      void* tp = dt;
      upload(dp, &tp, sizeof(void*), st);
      SPP_CHECK(launch_fill<float>(dy, ny, 0.f, st));
      SPP_CHECK(launch_kron_rows(ks, stage, dx, dy, kr, dp, 1, rows_per_block, st, ctx->c.kron_groups, nullptr,
                                 ctx->c.kron == 2));
      SPP_CHECK(hipMemcpyAsync(y, dy, sizeof(float) * ny, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int spp_selftest_copy(spp_ctx* ctx, long long n) {
  return guarded(ctx, [&] {
    if (n <= 0) throw HipError("n must be positive");
    hipStream_t st = ctx->c.stream;
    float *a = nullptr, *b = nullptr;
    SPP_CHECK(hipMalloc(&a, sizeof(float) * n));
    SPP_CHECK(hipMalloc(&b, sizeof(float) * n));
    SPP_CHECK(launch_fill<float>(a, n, 1.0f, st));
    SPP_CHECK(hipStreamSynchronize(st));
    SPP_CHECK((launch_convert<float, float>(b, a, n, st)));
    SPP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(a);
    (void)hipFree(b);
  });
}

}  // extern "C"
