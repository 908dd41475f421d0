// Batched perfect Born sampling (restates sample_configuration, auxiliary
// kind, src/sampling/born.jl:86-163, and SamplingCache, born.jl:45-73).
//
// Per row y, for all samples of a chunk at once: right row environments
// (_right_step, src/contraction/boundary_mps.jl:240-246), then for x = 1..Lx
// the two candidate left steps (_left_step, boundary_mps.jl:233-239) joined
// with the right environment (_row_scalar, boundary_mps.jl:274-284), the
// validity check and categorical draw with the injected uniform
// (born.jl:136-151, _categorical src/resample/selection.jl:68-80); the left
// step of the drawn candidate is reused instead of being recomputed. The row
// of drawn candidates is then absorbed into the per-sample north boundary
// (born.jl:157-160).
#include "sampler.h"
#include "boundary_ops.h"
#include "tridiag.h"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>

namespace spp {

StateDev::~StateDev() {
  if (block) {
    (void)hipFree(block);
    return;
  }
  for (auto& s : sites)
    if (s.d) (void)hipFree(s.d);
}

template <typename T>
CacheDev<T>::~CacheDev() {
  for (auto p : layers)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 2; ++k)
    for (auto p : cand[k])
      if (p) (void)hipFree(p);
  if (cand_lognorm) (void)hipFree(cand_lognorm);
  for (auto& b : south) b.release();
  for (auto& ps : phys) {
    if (ps.a) (void)hipFree(ps.a);
    if (ps.inv) (void)hipFree(ps.inv);
  }
  if (kron_buf) (void)hipFree(kron_buf);
}
template struct CacheDev<float>;
template struct CacheDev<double>;

// ---------------------------------------------------------------------------
// small kernels

// log || sum_a A[..a..] u[a] || of one FP64 site, for both basis vectors.
__global__ void proj_lognorm_kernel(const double* a, int n_before, int naux, int n_after, double u00, double u01,
                                    double u10, double u11, double* out) {
  __shared__ double part[2][256];
  double acc0 = 0.0, acc1 = 0.0;
  const long long total = (long long)n_before * n_after;
  for (long long i = threadIdx.x; i < total; i += blockDim.x) {
    long long b = i / n_after, c = i % n_after;
    double v0 = 0.0, v1 = 0.0;
    for (int k = 0; k < naux; ++k) {
      double x = a[(b * naux + k) * n_after + c];
      if (k == 0) { v0 += u00 * x; v1 += u01 * x; }
      else { v0 += u10 * x; v1 += u11 * x; }
    }
    acc0 += v0 * v0;
    acc1 += v1 * v1;
  }
  part[0][threadIdx.x] = acc0;
  part[1][threadIdx.x] = acc1;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      part[0][threadIdx.x] += part[0][threadIdx.x + w];
      part[1][threadIdx.x] += part[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = 0.5 * log(part[0][0]);
    out[1] = 0.5 * log(part[1][0]);
  }
}

// One (state, site) of project_batch: source PEPO site and destination PEPS
// site (dense (W,N,P,A,E,S)), the dimensions of both, the basis column of the
// site's label and, per virtual leg, the source index of every destination
// index (-1: a zero slot of the larger capacity).
struct ProjectJob {
  const double* src;
  double* dst;
  int sd[6], dd[6];
  int map_off[4];  // offset of each virtual leg's map in the job's map table
  const int* maps;
  double u[2];
};

// dst = the projected site; nrm2[job] = its squared norm (one workgroup per job).
__global__ void project_site_kernel(const ProjectJob* jobs, double* nrm2) {
  const ProjectJob& J = jobs[blockIdx.x];
  const int dW = J.dd[0], dN = J.dd[1], dP = J.dd[2], dE = J.dd[4], dS = J.dd[5];
  const long long total = (long long)dW * dN * dP * dE * dS;
  const int sA = J.sd[3];
  long long st[6];
  st[5] = 1;
  for (int k = 4; k >= 0; --k) st[k] = st[k + 1] * J.sd[k + 1];
  double acc = 0.0;
  for (long long t = threadIdx.x; t < total; t += blockDim.x) {
    long long r = t;
    const int is = (int)(r % dS); r /= dS;
    const int ie = (int)(r % dE); r /= dE;
    const int ip = (int)(r % dP); r /= dP;
    const int in = (int)(r % dN); r /= dN;
    const int iw = (int)r;
    const int w = J.maps[J.map_off[0] + iw], n = J.maps[J.map_off[1] + in];
    const int e = J.maps[J.map_off[2] + ie], sv = J.maps[J.map_off[3] + is];
    double v = 0.0;
    if (w >= 0 && n >= 0 && e >= 0 && sv >= 0) {
      const long long base = w * st[0] + n * st[1] + ip * st[2] + e * st[4] + sv * st[5];
      for (int a = 0; a < sA; ++a) v += J.src[base + a * st[3]] * J.u[a];
    }
    J.dst[t] = v;
    acc += v * v;
  }
  __shared__ double part[256];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) nrm2[blockIdx.x] = part[0];
}

__global__ void project_scale_kernel(const ProjectJob* jobs, const double* nrm2) {
  const ProjectJob& J = jobs[blockIdx.x];
  const long long total = (long long)J.dd[0] * J.dd[1] * J.dd[2] * J.dd[4] * J.dd[5];
  const double f = 1.0 / sqrt(nrm2[blockIdx.x]);
  for (long long t = threadIdx.x; t < total; t += blockDim.x) J.dst[t] *= f;
}

void project_batch(Context& ctx, const StateDev& pepo, int n, const int* labels, const double* basis, const int* cap,
                   std::vector<StateDev*>& out, double* normalizations) {
  hipStream_t st = ctx.stream;
  const int Lx = pepo.Lx, Ly = pepo.Ly, ns = Lx * Ly;
  const double U[4] = {basis ? basis[0] : 1.0, basis ? basis[1] : 0.0, basis ? basis[2] : 0.0,
                       basis ? basis[3] : 1.0};
  bool aux_graded = false;
  for (const SiteDev& s : pepo.sites) {
    if (s.dim[3] != 2) throw HipError("projection needs a PEPO with a two-index auxiliary leg");
    aux_graded = aux_graded || s.even[3] < s.dim[3];
  }
  const int legs[4] = {0, 1, 4, 5};
  // per site: destination leg dims / evens and the index maps of its virtual legs
  std::vector<std::array<int, 6>> ddim(ns), deven(ns);
  std::vector<int> maps;
  std::vector<std::array<int, 4>> moff(ns);
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = pepo.sites[i];
    const int x = i % Lx, y = i / Lx;
    for (int k = 0; k < 6; ++k) {
      ddim[i][k] = s.dim[k];
      deven[i][k] = s.even[k];
    }
    ddim[i][3] = 1;
    for (int v = 0; v < 4; ++v) {
      const int k = legs[v];
      const bool edge = (k == 0 && x == 0) || (k == 1 && y == 0) || (k == 4 && x == Lx - 1) || (k == 5 && y == Ly - 1);
      moff[i][v] = (int)maps.size();
      if (edge || !cap) {
        for (int j = 0; j < s.dim[k]; ++j) maps.push_back(j);
        continue;
      }
      const int le = pepo.live[(size_t)i * 8 + 2 * v], lo = pepo.live[(size_t)i * 8 + 2 * v + 1];
      if (le > cap[0] || lo > cap[1]) throw HipError("live bond sectors exceed the capacity");
      ddim[i][k] = cap[0] + cap[1];
      deven[i][k] = cap[0];
      for (int j = 0; j < cap[0]; ++j) maps.push_back(j < le ? j : -1);
      for (int j = 0; j < cap[1]; ++j) maps.push_back(j < lo ? s.even[k] + j : -1);
    }
  }
  std::vector<long long> dsize(ns), doff(ns + 1, 0);
  for (int i = 0; i < ns; ++i) {
    dsize[i] = 1;
    for (int k = 0; k < 6; ++k) dsize[i] *= ddim[i][k];
    doff[i + 1] = doff[i] + dsize[i];
  }
  int* dmaps = nullptr;
  ProjectJob* djobs = nullptr;
  double* nrm2 = nullptr;
  std::vector<StateDev*> made;
  auto cleanup = [&]() {
    for (void* p : {(void*)dmaps, (void*)djobs, (void*)nrm2})
      if (p) (void)hipFree(p);
  };
  try {
    SPP_CHECK(hipMalloc(&dmaps, sizeof(int) * std::max<size_t>(maps.size(), 1)));
    upload(dmaps, maps.data(), sizeof(int) * maps.size(), st);
    std::vector<ProjectJob> jobs((size_t)n * ns);
    for (int s0 = 0; s0 < n; ++s0) {
      StateDev* d = new StateDev();
      made.push_back(d);
      d->Lx = Lx;
      d->Ly = Ly;
      d->kind = 1;
      d->beta = pepo.beta;
      d->log_amplitude = pepo.log_amplitude;
      d->sites.resize(ns);
      SPP_CHECK(hipMalloc(&d->block, sizeof(double) * doff[ns]));
      d->live.assign((size_t)ns * 8, 0);
      for (int i = 0; i < ns; ++i) {
        const int k = labels[(size_t)s0 * ns + i];
        if (k != 0 && k != 1) throw HipError("labels must be 0 or 1");
        SiteDev& sd = d->sites[i];
        sd.d = d->block + doff[i];
        for (int q = 0; q < 6; ++q) {
          sd.dim[q] = ddim[i][q];
          sd.even[q] = deven[i][q];
        }
        // the auxiliary index carries the label's charge (Z2), even otherwise
        sd.even[3] = aux_graded && k == 1 ? 0 : 1;
        // the live sectors stay those of the PEPO (its live indices come first)
        for (int v = 0; v < 8; ++v) d->live[(size_t)i * 8 + v] = pepo.live[(size_t)i * 8 + v];
        ProjectJob& J = jobs[(size_t)s0 * ns + i];
        J.src = pepo.sites[i].d;
        J.dst = sd.d;
        for (int q = 0; q < 6; ++q) {
          J.sd[q] = pepo.sites[i].dim[q];
          J.dd[q] = ddim[i][q];
        }
        for (int v = 0; v < 4; ++v) J.map_off[v] = moff[i][v];
        J.maps = dmaps;
        J.u[0] = U[0 * 2 + k];
        J.u[1] = U[1 * 2 + k];
      }
    }
    SPP_CHECK(hipMalloc(&djobs, sizeof(ProjectJob) * std::max<size_t>(jobs.size(), 1)));
    SPP_CHECK(hipMalloc(&nrm2, sizeof(double) * std::max<size_t>(jobs.size(), 1)));
    upload(djobs, jobs.data(), sizeof(ProjectJob) * jobs.size(), st);
    const int nj = (int)jobs.size();
    for (int j0 = 0; j0 < nj; j0 += 65535) {
      const int cnt = std::min(65535, nj - j0);
      hipLaunchKernelGGL(project_site_kernel, dim3(cnt), dim3(256), 0, st, djobs + j0, nrm2 + j0);
      SPP_CHECK(hipGetLastError());
    }
    std::vector<double> h(nj);
    SPP_CHECK(hipMemcpyAsync(h.data(), nrm2, sizeof(double) * nj, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
    for (int j = 0; j < nj; ++j)
      if (!(std::isfinite(h[j]) && h[j] > 0.0)) {
        const int i = j % ns;
        throw HipError("zero or nonfinite tensor at (" + std::to_string(i % Lx + 1) + "," +
                       std::to_string(i / Lx + 1) + ")");
      }
    for (int j0 = 0; j0 < nj; j0 += 65535) {
      const int cnt = std::min(65535, nj - j0);
      hipLaunchKernelGGL(project_scale_kernel, dim3(cnt), dim3(256), 0, st, djobs + j0, nrm2 + j0);
      SPP_CHECK(hipGetLastError());
    }
    for (int s0 = 0; s0 < n; ++s0)
      for (int i = 0; i < ns; ++i) {
        const double ln = 0.5 * std::log(h[(size_t)s0 * ns + i]);
        made[s0]->log_amplitude += ln;
        if (normalizations) normalizations[(size_t)s0 * ns + i] = ln;
      }
    SPP_CHECK(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamSynchronize(st);
    for (StateDev* d : made) delete d;
    cleanup();
    throw;
  }
  cleanup();
  out = made;
}

// Born conditional and categorical draw for one site, all samples.
__global__ void born_select_kernel(const double* raw0, const double* raw1, const double* uniforms, const int* labels_in,
                                   long long sample0, int nsite, int site, int* choice, int* labels, double* logp,
                                   double* cond, int* status, int* fail_site, int* done, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  const long long gs = sample0 + s;
  double raw[2] = {raw0[s], raw1[s]};
  int k = 0;
  if (status[s] == 0 && !done[s]) {
    double scale = fmax(fabs(raw[0]), fabs(raw[1]));
    const double eps = 2.220446049250313e-16;
    bool ok = isfinite(raw[0]) && isfinite(raw[1]) && raw[0] >= -64.0 * eps * scale && raw[1] >= -64.0 * eps * scale;
    if (!ok) {
      status[s] = 1;
      fail_site[s] = site;
    } else {
      double v0 = fmax(raw[0], 0.0), v1 = fmax(raw[1], 0.0);
      double total = v0 + v1;
      if (!(total > 0.0)) {
        status[s] = 2;
        fail_site[s] = site;
      } else {
        double p[2] = {v0 / total, v1 / total};
        if (labels_in) {
          k = labels_in[gs * nsite + site];
        } else {
          // _categorical(log.(p), u): z = logsumexp, cumulative exp(lm - z)
          double lm[2] = {log(p[0]), log(p[1])};
          double pivot = fmax(lm[0], lm[1]);
          double z = pivot + log(exp(lm[0] - pivot) + exp(lm[1] - pivot));
          double u = uniforms[gs * nsite + site];
          double cum = 0.0;
          int last = -1;
          k = -1;
          for (int j = 0; j < 2; ++j) {
            if (lm[j] == -INFINITY) continue;
            last = j;
            cum += exp(lm[j] - z);
            if (u < cum) { k = j; break; }
          }
          if (k < 0) k = last;
        }
        const double pk = p[k];
        cond[(long long)s * nsite + site] = pk;
        if (pk == 0.0) {
          logp[s] = -INFINITY;
          done[s] = 1;
        } else {
          logp[s] += log(pk);
        }
      }
    }
  } else if (labels_in) {
    k = labels_in[gs * nsite + site];
  }
  choice[s] = k;
  labels[(long long)s * nsite + site] = k;
}

// max_environment_truncation of the samples still being drawn: a sample that
// ends early (p = 0, or the charge check) keeps what it had when it ended
// (born.jl:100-104, 146-148).
__global__ void live_trunc_kernel(const double* mt, const int* done, double* out, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  if (!done[s]) out[s] = mt[s];
}

// Per-sample pointers of the drawn candidate's layer and (k0, k1 non-null)
// of its single-layer tables.
__global__ void make_ptrs_kernel(const int* choice, void* c0, void* c1, void** ptrs, void* k0, void* k1,
                                 void** kptrs, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  ptrs[s] = choice[s] == 0 ? c0 : c1;
  if (k0) kptrs[s] = choice[s] == 0 ? k0 : k1;
}

// Single-layer tables of every (site, label) candidate of an FP32 cache: the
// candidate's layer is sum_phi S_k (x) S_k with S_k the aux projection
// sum_a U[a, k] A[w, n, phi, a, e, s] (auxiliary kind) or the physical one
// sum_p u_k[p] A[w, n, p, 0, e, s] (physical kind, one inner index). Left
// empty when a candidate carries both parities or a leg does not fit.
template <typename T>
static void build_kron(Context& ctx, const StateDev& st_, CacheDev<T>& c) {
  if (sizeof(T) != sizeof(float)) return;
  const int ns = (int)st_.sites.size();
  std::vector<KronSite> ks(ns);
  std::vector<std::vector<char>> tabs(2 * ns);
  for (int i = 0; i < ns; ++i) {
    const SiteDev& sd = st_.sites[i];
    std::vector<double> a(sd.size());
    SPP_CHECK(hipMemcpy(a.data(), sd.d, sizeof(double) * a.size(), hipMemcpyDeviceToHost));
    const int d[4] = {sd.dim[0], sd.dim[1], sd.dim[4], sd.dim[5]};
    const int e[4] = {sd.even[0], sd.even[1], sd.even[4], sd.even[5]};
    const int dp = sd.dim[2], da = sd.dim[3];
    const int dphi = c.kind == 0 ? dp : 1, ephi = c.kind == 0 ? sd.even[2] : 1;
    if (!kron_site_layout(d, e, dphi, ephi, ks[i])) return;
    const long long vol = (long long)d[0] * d[1] * d[2] * d[3];
    for (int k = 0; k < 2; ++k) {
      std::vector<double> S((size_t)vol * dphi, 0.0);
      for (int w = 0; w < d[0]; ++w)
        for (int n = 0; n < d[1]; ++n)
          for (int x = 0; x < d[2]; ++x)
            for (int y = 0; y < d[3]; ++y)
              for (int p = 0; p < dp; ++p)
                for (int q = 0; q < da; ++q) {
                  const double v = a[((((((long long)w * d[1] + n) * dp + p) * da + q) * d[2] + x) * d[3] + y)];
                  const int f = c.kind == 0 ? p : 0;
                  const double u = c.kind == 0 ? c.basis[q * 2 + k] : c.basis[p * 2 + k];
                  S[((((long long)w * d[1] + n) * dphi + f) * d[2] + x) * d[3] + y] += u * v;
                }
      const int q = kron_layer_parity(d, e, dphi, ephi, S.data());
      if (q < 0) return;
      kron_build_tables(ks[i], S.data(), q, tabs[2 * i + k]);
    }
  }
  size_t total = 0;
  for (auto& t : tabs) total += t.size();
  SPP_CHECK(hipMalloc(&c.kron_buf, total));
  c.kron_tab[0].assign(ns, nullptr);
  c.kron_tab[1].assign(ns, nullptr);
  size_t off = 0;
  for (int i = 0; i < ns; ++i)
    for (int k = 0; k < 2; ++k) {
      const auto& t = tabs[2 * i + k];
      SPP_CHECK(hipMemcpy(c.kron_buf + off, t.data(), t.size(), hipMemcpyHostToDevice));
      c.kron_tab[k][i] = c.kron_buf + off;
      off += t.size();
    }
  c.kron_site = ks;
  kron_prepare();
  (void)ctx;
}

// ---------------------------------------------------------------------------

template <typename T>
static size_t draw_bytes_per_sample(const CacheDev<T>& c, const Context& ctx) {
  long long D = 1, chi = c.chi;
  for (auto& sp : c.spec) D = std::max<long long>(D, std::max(std::max(sp.w.dim, sp.n.dim), std::max(sp.e.dim, sp.s.dim)));
  const long long Lx = c.Lx;
  const double ts = sizeof(T);
  double rowenv = ts * ((Lx + 4) * chi * D * chi + 3.0 * chi * D * D * chi) + 64 * 8;
  if (c.kind == 1)  // the site hole and its PSD workspace
    rowenv += ts * (double)D * D * D * D + 8.0 * (double)psd_scratch_doubles(c.psd_ld) + 8.0 * 6 * Lx * c.Ly;
  return (size_t)(1.15 * std::max(rowenv, absorb_bytes_per_sample(Lx, chi, D, sizeof(T), ctx.ksub))) + 4096;
}

static long long layer_count(const LayerSpec& sp) {
  return (long long)sp.w.dim * sp.n.dim * sp.e.dim * sp.s.dim;
}

static long long max_leg_dim(const std::vector<LayerSpec>& spec) {
  long long D = 1;
  for (auto& sp : spec) D = std::max<long long>(D, std::max(std::max(sp.w.dim, sp.n.dim), std::max(sp.e.dim, sp.s.dim)));
  return D;
}

// South boundaries of the unprojected PEPO (batch 1) in the rotated frame:
// south[y] holds rows y..Ly-1 (boundary_environments, boundary_mps.jl:193-209).
template <typename TS>
static void build_south(Context& ctx, const std::vector<LayerSpec>& spec, const std::vector<TS*>& layers, int Lx,
                        int Ly, int chi, std::vector<Boundary<TS>>& south) {
  south.resize(Ly + 1);
  std::vector<Leg> phys(Lx);
  for (int xr = 0; xr < Lx; ++xr) phys[xr] = spec[(Lx - 1 - xr) + (Ly - 1) * Lx].s;
  prepare_boundary(south[Ly], Lx, 1, 1);
  set_trivial(ctx, south[Ly], phys);
  freeze_splits(ctx, south[Ly]);
  for (int y = Ly - 1; y >= 0; --y) {
    RowLayers<TS> rr;
    rr.L = Lx;
    rr.rotated = true;
    rr.shared.resize(Lx);
    rr.spec.resize(Lx);
    long long capmax = 1;
    for (int xr = 0; xr < Lx; ++xr) {
      const int x = Lx - 1 - xr;
      const LayerSpec& sp = spec[x + y * Lx];
      rr.shared[xr] = layers[x + y * Lx];
      rr.spec[xr] = LayerSpec{sp.e, sp.s, sp.w, sp.n};
      capmax = std::max<long long>(capmax, (long long)chi * sp.n.dim * chi);
    }
    prepare_boundary(south[y], Lx, 1, capmax);
    absorb_forward<TS>(ctx, south[y + 1], rr, chi, south[y]);
    freeze_splits(ctx, south[y]);
  }
}

// Copy of a batch-1 boundary with frozen splits in another element type.
template <typename TO, typename TI>
static void convert_boundary(Context& ctx, const Boundary<TI>& in, Boundary<TO>& out) {
  hipStream_t st = ctx.stream;
  out.release();
  out.L = in.L;
  out.nb = in.nb;
  out.bdim = in.bdim;
  out.beven = in.beven;
  out.bsplit.assign(in.L + 1, nullptr);
  for (auto p : in.bsplit)
    if (p) throw HipError("convert_boundary: per-sample splits must be frozen");
  out.pdim = in.pdim;
  out.peven = in.peven;
  out.t.assign(in.L, nullptr);
  out.owns = true;
  for (int x = 0; x < in.L; ++x) {
    const long long n = in.tsize(x) * in.nb;
    SPP_CHECK(hipMalloc(&out.t[x], sizeof(TO) * n));
    SPP_CHECK((launch_convert<TO, TI>(out.t[x], in.t[x], n, st)));
  }
  SPP_CHECK(hipMalloc(&out.log_scale, sizeof(double) * in.nb));
  SPP_CHECK(hipMalloc(&out.max_trunc, sizeof(double) * in.nb));
  SPP_CHECK(hipMemcpyAsync(out.log_scale, in.log_scale, sizeof(double) * in.nb, hipMemcpyDeviceToDevice, st));
  SPP_CHECK(hipMemcpyAsync(out.max_trunc, in.max_trunc, sizeof(double) * in.nb, hipMemcpyDeviceToDevice, st));
}

// Physical-kind tables of a cache (born.jl:56-73): FP64 copies of the PEPS
// sites, the metric block tables, the state's total charge
// (`_state_charge`, basis.jl:89-92) and the reachable charges
// (`_charge_suffix`, basis.jl:93-100) for the symmetry basis.
template <typename T>
static void build_physical(Context& ctx, const StateDev& peps, CacheDev<T>& c) {
  hipStream_t st = ctx.stream;
  const int ns = peps.Lx * peps.Ly;
  bool symmetric = false;
  for (const SiteDev& s : peps.sites) symmetric = symmetric || s.even[2] < s.dim[2];
  if (symmetric) {
    const double* U = c.basis;
    if (!(U[0] == 1.0 && U[1] == 0.0 && U[2] == 0.0 && U[3] == 1.0))
      throw HipError("symmetric sampling requires SymmetryBasis()");
    for (const SiteDev& s : peps.sites)
      if (s.even[2] != 1) throw HipError("sampling requires one-dimensional irreps");
  }
  c.charges[0] = 0;
  c.charges[1] = symmetric ? 1 : 0;
  int target = 0;
  for (const SiteDev& s : peps.sites) target ^= s.even[3] == 0 ? 1 : 0;
  c.target_charge = symmetric ? target : -1;
  c.reach.assign(ns + 1, 0);
  c.reach[ns] = 1;
  for (int i = ns - 1; i >= 0; --i)
    for (int r = 0; r < 2; ++r)
      if ((c.reach[i + 1] >> r) & 1)
        for (int k = 0; k < 2; ++k) c.reach[i] |= 1 << (c.charges[k] ^ r);
  c.phys.resize(ns);
  int ld = 1;
  std::vector<std::vector<int>> blocks(ns);
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = peps.sites[i];
    PhysSite& ps = c.phys[i];
    for (int k = 0; k < 6; ++k) {
      ps.dim[k] = s.dim[k];
      ps.even[k] = s.even[k];
    }
    std::vector<int> byq[2];
    const int dW = s.dim[0], dN = s.dim[1], dE = s.dim[4], dS = s.dim[5];
    for (int w = 0; w < dW; ++w)
      for (int n = 0; n < dN; ++n)
        for (int e = 0; e < dE; ++e)
          for (int sv = 0; sv < dS; ++sv) {
            const int q = ((w >= s.even[0]) + (n >= s.even[1]) + (e >= s.even[4]) + (sv >= s.even[5])) & 1;
            byq[q].push_back(((w * dN + n) * dE + e) * dS + sv);
          }
    ps.nq[0] = (int)byq[0].size();
    ps.nq[1] = (int)byq[1].size();
    ld = std::max(ld, std::max(ps.nq[0], ps.nq[1]));
    blocks[i] = byq[0];
    blocks[i].insert(blocks[i].end(), byq[1].begin(), byq[1].end());
  }
  // a block must fit one workgroup's LDS vectors (6 ld doubles)
  if ((size_t)(kTriWaves + 6 * (size_t)ld) * sizeof(double) > 160 * 1024)
    throw HipError("physical-kind site metric block too large for the eigensolver (" + std::to_string(ld) +
                   " rows)");
  c.psd_ld = ld;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = peps.sites[i];
    PhysSite& ps = c.phys[i];
    std::vector<int> inv(2 * (size_t)ld, 0);
    for (int j = 0; j < ps.nq[0]; ++j) inv[j] = blocks[i][j];
    for (int j = 0; j < ps.nq[1]; ++j) inv[ld + j] = blocks[i][ps.nq[0] + j];
    SPP_CHECK(hipMalloc(&ps.inv, sizeof(int) * 2 * (size_t)ld));
    upload(ps.inv, inv.data(), sizeof(int) * 2 * (size_t)ld, st);
    SPP_CHECK(hipMalloc(&ps.a, sizeof(double) * s.size()));
    SPP_CHECK(hipMemcpyAsync(ps.a, s.d, sizeof(double) * s.size(), hipMemcpyDeviceToDevice, st));
  }
}

template <typename T>
void build_cache(Context& ctx, const StateDev& pepo, int chi, const double* basis, CacheDev<T>& c, int kind) {
  hipStream_t st = ctx.stream;
  const int Lx = pepo.Lx, Ly = pepo.Ly, ns = Lx * Ly;
  c.Lx = Lx;
  c.Ly = Ly;
  c.chi = chi;
  c.kind = kind;
  c.precision = sizeof(T) == 8 ? 0 : 1;
  if (basis) std::memcpy(c.basis, basis, sizeof(double) * 4);
  if (kind == 1) {
    if (pepo.kind != 1) throw HipError("proposal requires PEPS");
    for (const SiteDev& s : pepo.sites)
      if (s.dim[2] != 2 || s.dim[3] != 1)
        throw HipError("physical-kind sampling cache needs a PEPS with d = 2 and a one-index auxiliary leg");
    build_physical<T>(ctx, pepo, c);
  }
  c.pepo_log_amplitude = pepo.log_amplitude;
  c.pepo_beta = pepo.beta;
  c.layers.assign(ns, nullptr);
  c.cand[0].assign(ns, nullptr);
  c.cand[1].assign(ns, nullptr);
  c.spec.resize(ns);
  std::vector<LayerJob> jobs;
  long long maxout = 0;
  for (int i = 0; i < ns; ++i) {
    const SiteDev& s = pepo.sites[i];
    if (kind == 0 && (s.dim[3] != 2 || s.dim[2] != 2))
      throw HipError("sampling cache needs a PEPO with d = 2 physical and aux legs");
    LayerSpec sp{pair_leg(s.dim[0], s.even[0]), pair_leg(s.dim[1], s.even[1]), pair_leg(s.dim[4], s.even[4]),
                 pair_leg(s.dim[5], s.even[5])};
    c.spec[i] = sp;
    long long n = (long long)sp.w.dim * sp.n.dim * sp.e.dim * sp.s.dim;
    maxout = std::max(maxout, n);
    SPP_CHECK(hipMalloc(&c.layers[i], sizeof(T) * n));
    SPP_CHECK(hipMalloc(&c.cand[0][i], sizeof(T) * n));
    SPP_CHECK(hipMalloc(&c.cand[1][i], sizeof(T) * n));
    for (int v = 0; v < 3; ++v) {
      LayerJob J{};
      J.ket = s.d;
      J.bra = s.d;
      for (int k = 0; k < 6; ++k) {
        J.dk[k] = J.db[k] = s.dim[k];
        J.ek[k] = J.eb[k] = s.even[k];
      }
      J.use_op = 0;
      if (v == 0) {
        J.out = c.layers[i];
        J.use_aux = 0;
      } else if (kind == 0) {
        J.out = c.cand[v - 1][i];
        J.use_aux = 1;
        J.aux[0] = c.basis[0 * 2 + (v - 1)];
        J.aux[1] = c.basis[1 * 2 + (v - 1)];
      } else {
        // physical projector u_k u_k^T on the ket's physical leg (the bra is
        // contracted over it): the double layer of _project(A, B, k, :physical)
        J.out = c.cand[v - 1][i];
        J.use_aux = 0;
        J.use_op = 1;
        const double u0 = c.basis[0 * 2 + (v - 1)], u1 = c.basis[1 * 2 + (v - 1)];
        J.op[0] = u0 * u0;
        J.op[1] = u0 * u1;
        J.op[2] = u1 * u0;
        J.op[3] = u1 * u1;
      }
      jobs.push_back(J);
    }
  }
  LayerJob* djobs = nullptr;
  SPP_CHECK(hipMalloc(&djobs, sizeof(LayerJob) * jobs.size()));
  upload(djobs, jobs.data(), sizeof(LayerJob) * jobs.size(), st);
  SPP_CHECK(launch_double_layers<T>(djobs, (int)jobs.size(), maxout, st));
  build_kron<T>(ctx, pepo, c);
  SPP_CHECK(hipMalloc(&c.cand_lognorm, sizeof(double) * 2 * ns));
  if (kind == 1) SPP_CHECK(launch_fill<double>(c.cand_lognorm, 2 * ns, 0.0, st));
  for (int i = 0; i < ns && kind == 0; ++i) {
    const SiteDev& s = pepo.sites[i];
    int nbef = s.dim[0] * s.dim[1] * s.dim[2], naft = s.dim[4] * s.dim[5];
    hipLaunchKernelGGL(proj_lognorm_kernel, dim3(1), dim3(256), 0, st, s.d, nbef, s.dim[3], naft, c.basis[0],
                       c.basis[1], c.basis[2], c.basis[3], c.cand_lognorm + 2 * i);
    SPP_CHECK(hipGetLastError());
  }
  // South boundaries, rotated frame (batch 1). They are computed in FP64 for
  // every precision: an FP32 cache stores FP64-contracted boundaries.
  const double need = absorb_bytes_per_sample(Lx, chi, max_leg_dim(c.spec), sizeof(double), ctx.ksub);
  if (ctx.arena.mark() == 0)
    ctx.arena.reserve(std::max(ctx.arena.capacity(), std::max(draw_bytes_per_sample(c, ctx), (size_t)(1.15 * need) + 4096)));
  if constexpr (sizeof(T) == sizeof(double)) {
    build_south<double>(ctx, c.spec, c.layers, Lx, Ly, chi, c.south);
  } else {
    std::vector<double*> l64(ns, nullptr);
    std::vector<Boundary<double>> s64;
    try {
      std::vector<LayerJob> jobs64;
      for (int i = 0; i < ns; ++i) {
        const SiteDev& s = pepo.sites[i];
        SPP_CHECK(hipMalloc(&l64[i], sizeof(double) * layer_count(c.spec[i])));
        LayerJob J{};
        J.ket = s.d;
        J.bra = s.d;
        for (int k = 0; k < 6; ++k) {
          J.dk[k] = J.db[k] = s.dim[k];
          J.ek[k] = J.eb[k] = s.even[k];
        }
        J.out = l64[i];
        jobs64.push_back(J);
      }
      upload(djobs, jobs64.data(), sizeof(LayerJob) * ns, st);
      SPP_CHECK(launch_double_layers<double>(djobs, ns, maxout, st));
      build_south<double>(ctx, c.spec, l64, Lx, Ly, chi, s64);
      c.south.resize(Ly + 1);
      for (int y = 0; y <= Ly; ++y) convert_boundary<T, double>(ctx, s64[y], c.south[y]);
      SPP_CHECK(hipStreamSynchronize(st));
    } catch (...) {
      for (auto& b : s64) b.release();
      for (double* p : l64)
        if (p) (void)hipFree(p);
      throw;
    }
    for (auto& b : s64) b.release();
    for (double* p : l64)
      if (p) SPP_CHECK(hipFree(p));
  }
  c.south_max_trunc.assign(Ly + 1, 0.0);
  for (int y = 0; y <= Ly; ++y)
    SPP_CHECK(hipMemcpyAsync(&c.south_max_trunc[y], c.south[y].max_trunc, sizeof(double), hipMemcpyDeviceToHost, st));
  SPP_CHECK(hipStreamSynchronize(st));
  SPP_CHECK(hipFree(djobs));
}

template void build_cache<float>(Context&, const StateDev&, int, const double*, CacheDev<float>&, int);
template void build_cache<double>(Context&, const StateDev&, int, const double*, CacheDev<double>&, int);

template <typename T>
void DrawWork<T>::allocate(int nb_, int ns_, int lx_, long long capmax_, bool physical) {
  nb = nb_;
  ns = ns_;
  lx = lx_;
  capmax = capmax_;
  SPP_CHECK(hipMalloc(&choice, sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&labels, sizeof(int) * nb * ns));
  SPP_CHECK(hipMalloc(&status, sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&fail, sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&done, sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&prefix, sizeof(int) * nb));
  SPP_CHECK(hipMalloc(&dlab, sizeof(int) * nb * ns));
  SPP_CHECK(hipMalloc(&logp, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&cond, sizeof(double) * nb * ns));
  SPP_CHECK(hipMalloc(&raw0, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&raw1, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&nrm, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&fac, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&lnamp, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&mant, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&mtlive, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&duni, sizeof(double) * nb * ns));
  if (physical) SPP_CHECK(hipMalloc(&drep, sizeof(double) * (size_t)nb * ns * 6));
  SPP_CHECK(hipMalloc(&ptrs, sizeof(void*) * 2 * nb * lx));  // layers, then single-layer tables
  SPP_CHECK(hipMalloc(&dsst, sizeof(long long) * lx));
  SPP_CHECK(hipMalloc(&dtens, sizeof(T*) * lx));
  prepare_boundary(NB[0], lx, nb, capmax);
  prepare_boundary(NB[1], lx, nb, capmax);
}

template <typename T>
void DrawWork<T>::drop_graphs() {
  for (auto& g : graphs)
    if (g.second) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
  seen.clear();
}

template <typename T>
DrawWork<T>::~DrawWork() {
  drop_graphs();
  NB[0].release();
  NB[1].release();
  for (void* p : {(void*)choice, (void*)labels, (void*)status, (void*)fail, (void*)done, (void*)prefix, (void*)dlab,
                  (void*)logp, (void*)cond, (void*)raw0, (void*)raw1, (void*)nrm, (void*)fac, (void*)lnamp,
                  (void*)mant, (void*)mtlive, (void*)drep, (void*)duni, (void*)ptrs, (void*)dsst, (void*)dtens})
    if (p) (void)hipFree(p);
}
template struct DrawWork<float>;
template struct DrawWork<double>;

// A left step of the draw, Y[b,c,e,s] = sum_{w,n} X[w,b,c,n] E[w,n,e,s], with
// the layer E applied through its single-layer tables (kron.h, stage 2): per
// sample (tabs) or one candidate for all samples (shared). X is folded on n
// ([w][b][c][n']), Y on s ([b][c][e][s']). FP32 caches only.
template <typename T>
static void left_step_kron(Context& ctx, const KronSite& ks, const T* X, T* Y, long long xs, long long ys,
                           const Leg& sbl, const Leg& nb1, const LayerSpec& sp, void* const* tabs, const void* shared,
                           int cur) {
  if constexpr (sizeof(T) != sizeof(float)) {
    throw HipError("factorised left steps need an FP32 cache");
  } else {
    const long long fn = std::max(sp.n.even, sp.n.dim - sp.n.even);
    KronRows kr;
    kr.nrows = (long long)sbl.dim * nb1.dim;
    kr.nf = 1;
    kr.ef = 1;
    kr.rc = nb1.dim;
    kr.lsplit = sbl.split;
    kr.leven = sbl.even;
    kr.rsplit = nb1.split;
    kr.reven = nb1.even;
    kr.xs = xs;
    kr.ys = ys;
    kr.sx_slab = fn;
    kr.sxf = 0;
    kr.sx1 = (long long)sbl.dim * nb1.dim * fn;
    kr.sx2 = 1;
    const double fl =
        2.0 * cur * (ctx.kron == 2 ? kron_mfma_row_macs(ks, 2) : kron_row_macs(ks, 2)) * kr.nrows;
    timed_kron(ctx, fl, [&] {
      return launch_kron_rows(ks, 2, reinterpret_cast<const float*>(X), reinterpret_cast<float*>(Y), kr, tabs, cur,
                              ctx.kron_rows, ctx.stream, ctx.kron_groups, shared, ctx.kron == 2);
    });
    ctx.counters.kron_launches += 1;
    ctx.counters.gemm_flops += fl;
  }
}

template <typename T>
void draw_batch(Context& ctx, CacheDev<T>& c, int n, const double* uniforms, const int* labels_in,
                DrawOutputs& out) {
  hipStream_t st = ctx.stream;
  const int Lx = c.Lx, Ly = c.Ly, ns = Lx * Ly, chi = c.chi;
  if (n <= 0) return;
  // chunk size from the workspace budget
  size_t per = draw_bytes_per_sample(c, ctx);
  size_t freeb = 0, totalb = 0;
  SPP_CHECK(hipMemGetInfo(&freeb, &totalb));
  size_t budget = ctx.arena.capacity() > 0 ? std::max(ctx.arena.capacity(), (size_t)(0.6 * freeb))
                                            : (size_t)(0.6 * freeb);
  long long chunk = std::max<long long>(1, (long long)(budget / per));
  if (ctx.chunk > 0) chunk = std::min<long long>(chunk, ctx.chunk);
  chunk = std::min<long long>(chunk, n);
  chunk = std::min<long long>(chunk, 16384);
  ctx.arena.reserve(per * chunk);
  const int nb = (int)chunk;
  const bool phys = c.kind == 1;
  // buffers kept with the cache (reallocated when a larger chunk comes)
  long long capmax = 1;
  for (auto& sp : c.spec) capmax = std::max<long long>(capmax, (long long)chi * sp.s.dim * chi);
  if (!c.work || c.work->nb < nb || c.work->capmax < capmax) {
    c.work.reset();
    c.work.reset(new DrawWork<T>());
    c.work->allocate(nb, ns, Lx, capmax, phys);
  }
  DrawWork<T>& W = *c.work;
  (void)dot_partials(ctx, nb);  // grown now, outside any captured sweep
  if (W.arena_base != ctx.arena.base() || W.dot_base != ctx.dot_part.p) {
    W.drop_graphs();
    W.arena_base = ctx.arena.base();
    W.dot_base = ctx.dot_part.p;
  }
  const bool use_graphs = ctx.graphs && !ctx.profile && !ctx.noalloc_dot;
  Boundary<T>* NB = W.NB;
  int *choice = W.choice, *labels = W.labels, *status = W.status, *fail = W.fail, *done = W.done;
  double *logp = W.logp, *cond = W.cond, *raw0 = W.raw0, *raw1 = W.raw1, *nrm = W.nrm, *fac = W.fac;
  double *lnamp = W.lnamp, *mant = W.mant;
  void** ptrs = W.ptrs;
  void** kptrs = W.ptrs + (size_t)nb * Lx;
  const bool kron = !c.kron_site.empty();
  const bool kron_steps = kron && ctx.kron;  // candidate left steps through the single layers
  auto kron0 = [&](int site) -> void* { return kron ? c.kron_tab[0][site] : nullptr; };
  auto kron1 = [&](int site) -> void* { return kron ? c.kron_tab[1][site] : nullptr; };
  int* dlabels_in = labels_in ? W.dlab : nullptr;
  double* duni = uniforms ? W.duni : nullptr;
  long long* dsst = W.dsst;
  T** dtens = W.dtens;
  int* prefix = W.prefix;
  double* mtlive = W.mtlive;
  double* drep = W.drep;
  // evaluation mode of a charged physical cache: a configuration outside the
  // state's total charge has log p = -Inf before any contraction (born.jl:100-104)
  std::vector<int> h_rej(labels_in ? n : 0, 0);
  bool any_rej = false;
  if (labels_in && phys && c.target_charge >= 0)
    for (int i = 0; i < n; ++i) {
      int q = 0;
      for (int k = 0; k < ns; ++k) q ^= c.charges[labels_in[(long long)i * ns + k]];
      h_rej[i] = q != c.target_charge;
      any_rej = any_rej || h_rej[i];
    }
  const int tb = 256, gb = (nb + tb - 1) / tb;
  std::vector<int> h_status(nb), h_fail(nb), h_labels((size_t)nb * ns);
  std::vector<double> h_logp(nb), h_cond((size_t)nb * ns), h_mt(nb), h_mant(nb), h_ls(nb), h_amp(nb);
  std::vector<double> h_rep(phys ? (size_t)nb * ns * 6 : 0);
  std::vector<int> h_done0(nb);
  std::vector<double> h_logp0(nb);
  double south_mt = 0.0;
  for (double v : c.south_max_trunc) south_mt = std::max(south_mt, v);
  try {
    for (long long s0 = 0; s0 < n; s0 += nb) {
      const int cur = (int)std::min<long long>(nb, n - s0);
      NB[0].nb = cur;
      NB[1].nb = cur;
      // this chunk's uniforms or labels (sample indices relative to the chunk)
      if (labels_in) upload(dlabels_in, labels_in + s0 * ns, sizeof(int) * (size_t)cur * ns, st);
      if (uniforms) upload(duni, uniforms + s0 * ns, sizeof(double) * (size_t)cur * ns, st);
      // per-chunk initialisation
      std::vector<Leg> phys0(Lx);
      for (int x = 0; x < Lx; ++x) phys0[x] = c.spec[x].n;
      set_trivial(ctx, NB[0], phys0);
      int which = 0;
      SPP_CHECK(launch_fill<int>(status, nb, 0, st));
      SPP_CHECK(launch_fill<int>(fail, nb, -1, st));
      SPP_CHECK(launch_fill<int>(done, nb, 0, st));
      SPP_CHECK(launch_fill<double>(logp, nb, 0.0, st));
      SPP_CHECK(launch_fill<double>(cond, (long long)nb * ns, 0.0, st));
      SPP_CHECK(launch_fill<int>(choice, nb, 0, st));
      SPP_CHECK(launch_fill<int>(prefix, nb, 0, st));
      SPP_CHECK(launch_fill<double>(mtlive, nb, 0.0, st));
      if (phys) SPP_CHECK(launch_fill<double>(drep, (long long)nb * ns * 6, NAN, st));
      if (any_rej) {
        for (int s = 0; s < cur; ++s) {
          h_done0[s] = h_rej[s0 + s];
          h_logp0[s] = h_rej[s0 + s] ? -INFINITY : 0.0;
        }
        upload(done, h_done0.data(), sizeof(int) * cur, st);
        upload(logp, h_logp0.data(), sizeof(double) * cur, st);
      }
      for (int y = 0; y < Ly; ++y) {
        Boundary<T>& Nc = NB[which];
        const Boundary<T>& sb = c.south[y + 1];
        const size_t mark = ctx.arena.mark();
        // the row sweep below is replayed from its graph when one was captured
        std::vector<long long> gkey = {y, cur, labels_in ? 1 : 0, (long long)mark};
        for (int b = 0; b <= Lx; ++b) gkey.push_back(Nc.bdim[b]);
        bool replayed = false;
        if (use_graphs) {
          auto it = W.graphs.find(gkey);
          if (it != W.graphs.end() && it->second) {
            SPP_CHECK(hipGraphLaunch(it->second, st));
            replayed = true;
            ctx.counters.graph_replays += 1;
          }
        }
        const bool capture = use_graphs && !replayed && W.seen.count(gkey) && !W.graphs.count(gkey);
        if (capture) {
          SPP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
          ctx.capturing = true;
        }
        try {
        if (!replayed) {
        // ---------------- right environments ----------------
        std::vector<T*> R(Lx + 1);
        std::vector<std::vector<Leg>> Rl(Lx + 1);
        std::vector<long long> Rs(Lx + 1);
        for (int x = 0; x <= Lx; ++x) {
          Leg wl = x < Lx ? c.spec[x + y * Lx].w : c.spec[(Lx - 1) + y * Lx].e;
          Rl[x] = {Nc.bond(x), wl, south_bond(sb, Lx, x)};
          Rs[x] = (long long)Rl[x][0].dim * Rl[x][1].dim * Rl[x][2].dim;
          R[x] = (T*)ctx.arena.take(sizeof(T) * cur * Rs[x]);
        }
        SPP_CHECK(launch_fill<T>(R[Lx], cur, T(1), st));
        for (int x = Lx - 1; x >= 0; --x) {
          const LayerSpec& sp = c.spec[x + y * Lx];
          const size_t xm = ctx.arena.mark();
          long long xs = (long long)Nc.bdim[x] * sp.n.dim * sp.e.dim * Rl[x + 1][2].dim;
          T* X = (T*)ctx.arena.take(sizeof(T) * cur * xs);
          TRef Xr = folded_ref(X, xs, "adne", {Nc.bond(x), Rl[x + 1][2], sp.n, sp.e});
          contract<T>(ctx, Xr, Nc.ref(x, "anc"), dense_ref(R[x + 1], Rs[x + 1], "ced", Rl[x + 1]), cur);
          long long ys = (long long)Nc.bdim[x] * Rl[x + 1][2].dim * sp.w.dim * sp.s.dim;
          T* Y = (T*)ctx.arena.take(sizeof(T) * cur * ys);
          TRef Yr = folded_ref(Y, ys, "adws", {Nc.bond(x), Rl[x + 1][2], sp.w, sp.s});
          contract<T>(ctx, Yr, Xr, dense_ref(c.layers[x + y * Lx], 0, "wnes", {sp.w, sp.n, sp.e, sp.s}), cur);
          TRef Rr = dense_ref(R[x], Rs[x], "awb", Rl[x]);
          contract<T>(ctx, Rr, Yr, bottom_ref(sb, Lx, x, "bsd"), cur);
          FullView fv = full_view(Rr);
          SPP_CHECK(launch_dot<T>(fv, fv, nrm, cur, st, dot_partials(ctx, cur)));
          SPP_CHECK(launch_safe_factor(nrm, fac, status, fail, x + y * Lx, cur, st));
          SPP_CHECK(launch_scale<T>(fv, fac, cur, st));
          ctx.arena.release(xm);
        }
        // ---------------- left sweep with draws ----------------
        long long lmax = 1;
        for (int x = 0; x <= Lx; ++x) lmax = std::max(lmax, Rs[x]);
        T* Lb = (T*)ctx.arena.take(sizeof(T) * cur * lmax);
        std::vector<Leg> Ll = {Nc.bond(0), c.spec[y * Lx].w, south_bond(sb, Lx, 0)};
        long long Ls = 1;
        SPP_CHECK(launch_fill<T>(Lb, cur, T(1), st));
        for (int x = 0; x < Lx; ++x) {
          const int site = x + y * Lx;
          const LayerSpec& sp = c.spec[site];
          const size_t xm = ctx.arena.mark();
          Leg sbl = south_bond(sb, Lx, x), sbr = south_bond(sb, Lx, x + 1);
          long long xs = (long long)sp.w.dim * sbl.dim * sp.n.dim * Nc.bdim[x + 1];
          T* X = (T*)ctx.arena.take(sizeof(T) * cur * xs);
          TRef Xr = folded_ref(X, xs, "wbcn", {sp.w, sbl, Nc.bond(x + 1), sp.n});
          contract<T>(ctx, Xr, dense_ref(Lb, Ls, "awb", Ll), Nc.ref(x, "anc"), cur);
          long long ys = (long long)sbl.dim * Nc.bdim[x + 1] * sp.e.dim * sp.s.dim;
          long long os = (long long)Nc.bdim[x + 1] * sp.e.dim * sbr.dim;
          T* Yk = (T*)ctx.arena.take(sizeof(T) * cur * ys);
          T* O[2];
          O[0] = (T*)ctx.arena.take(sizeof(T) * cur * os);
          O[1] = (T*)ctx.arena.take(sizeof(T) * cur * os);
          std::vector<Leg> Ol = {Nc.bond(x + 1), sp.e, sbr};
          if (phys) {
            // the site hole H[w,n,e,s] = X[w,b,c,n] (B[b,s,d] R[c,e,d]) (born.jl:75-83)
            const long long y2s = (long long)sbl.dim * Nc.bdim[x + 1] * sp.e.dim * sp.s.dim;
            T* Y2 = (T*)ctx.arena.take(sizeof(T) * cur * y2s);
            TRef Y2r = dense_ref(Y2, y2s, "bces", {sbl, Nc.bond(x + 1), sp.e, sp.s});
            contract<T>(ctx, Y2r, bottom_ref(sb, Lx, x, "bsd"), dense_ref(R[x + 1], Rs[x + 1], "ced", Rl[x + 1]),
                        cur);
            const long long hs = layer_count(sp);
            T* Hb = (T*)ctx.arena.take(sizeof(T) * cur * hs);
            contract<T>(ctx, dense_ref(Hb, hs, "wnes", {sp.w, sp.n, sp.e, sp.s}), Xr, Y2r, cur);
            PsdScratch w;
            w.ld = c.psd_ld;
            const size_t l2 = (size_t)c.psd_ld;
            w.M = (double*)ctx.arena.take(sizeof(double) * cur * 2 * l2 * l2);
            w.taus = (double*)ctx.arena.take(sizeof(double) * cur * 2 * l2);
            w.d = (double*)ctx.arena.take(sizeof(double) * cur * 2 * l2);
            w.e = (double*)ctx.arena.take(sizeof(double) * cur * 2 * l2);
            w.y = (double*)ctx.arena.take(sizeof(double) * cur * 4 * l2);
            w.norms = (double*)ctx.arena.take(sizeof(double) * cur * 4);
            w.res = (double*)ctx.arena.take(sizeof(double) * cur * 16);
            physical_conditionals(ctx, Hb, sizeof(T) == sizeof(float), hs, c.phys[site], c.basis, cur, w);
            launch_physical_select(ctx, w, duni, dlabels_in, 0, ns, site, c.charges, c.target_charge,
                                   c.reach[site + 1], prefix, choice, labels, logp, cond, status, fail, done, drep,
                                   cur);
            hipLaunchKernelGGL(make_ptrs_kernel, dim3(gb), dim3(tb), 0, st, choice, (void*)c.cand[0][site],
                               (void*)c.cand[1][site], ptrs + (long long)x * cur, kron0(site), kron1(site),
                               kptrs + (long long)x * cur, cur);
            SPP_CHECK(hipGetLastError());
            if (x < Lx - 1) {
              // left <- safe_scale(left_step of the drawn candidate)
              TRef lay = dense_ref(nullptr, 0, "wnes", {sp.w, sp.n, sp.e, sp.s});
              lay.ptrs = (void* const*)(ptrs + (long long)x * cur);
              TRef Yr = folded_ref(Yk, ys, "bces", {sbl, Nc.bond(x + 1), sp.e, sp.s});
              if (kron_steps)
                left_step_kron(ctx, c.kron_site[site], X, Yk, xs, ys, sbl, Nc.bond(x + 1), sp,
                               (void* const*)(kptrs + (long long)x * cur), nullptr, cur);
              else
                contract<T>(ctx, Yr, Xr, lay, cur);
              TRef Lnew = dense_ref(Lb, os, "ced", Ol);
              contract<T>(ctx, Lnew, Yr, bottom_ref(sb, Lx, x, "bsd"), cur);
              FullView fv = full_view(Lnew);
              SPP_CHECK(launch_dot<T>(fv, fv, nrm, cur, st, dot_partials(ctx, cur)));
              SPP_CHECK(launch_safe_factor(nrm, fac, status, fail, site, cur, st));
              SPP_CHECK(launch_scale<T>(fv, fac, cur, st));
              Ll = Ol;
              Ls = os;
            }
            ctx.arena.release(xm);
            continue;
          }
          double* rawk[2] = {raw0, raw1};
          for (int k = 0; k < 2; ++k) {
            TRef Yr = folded_ref(Yk, ys, "bces", {sbl, Nc.bond(x + 1), sp.e, sp.s});
            if (kron_steps)
              left_step_kron(ctx, c.kron_site[site], X, Yk, xs, ys, sbl, Nc.bond(x + 1), sp, nullptr,
                             c.kron_tab[k][site], cur);
            else
              contract<T>(ctx, Yr, Xr, dense_ref(c.cand[k][site], 0, "wnes", {sp.w, sp.n, sp.e, sp.s}), cur);
            TRef Or = dense_ref(O[k], os, "ced", Ol);
            contract<T>(ctx, Or, Yr, bottom_ref(sb, Lx, x, "bsd"), cur);
            SPP_CHECK(launch_dot<T>(full_view(Or), full_view(dense_ref(R[x + 1], Rs[x + 1], "ced", Rl[x + 1])),
                                    rawk[k], cur, st, dot_partials(ctx, cur)));
          }
          hipLaunchKernelGGL(born_select_kernel, dim3(gb), dim3(tb), 0, st, raw0, raw1, duni, dlabels_in, 0, ns,
                             site, choice, labels, logp, cond, status, fail, done, cur);
          SPP_CHECK(hipGetLastError());
          hipLaunchKernelGGL(make_ptrs_kernel, dim3(gb), dim3(tb), 0, st, choice, (void*)c.cand[0][site],
                             (void*)c.cand[1][site], ptrs + (long long)x * cur, kron0(site), kron1(site),
                             kptrs + (long long)x * cur, cur);
          SPP_CHECK(hipGetLastError());
          if (x < Lx - 1) {
            // left <- safe_scale(left_step of the drawn candidate)
            TRef Lnew = dense_ref(Lb, os, "ced", Ol);
            SPP_CHECK(launch_select_copy<T>(full_view(Lnew), full_view(dense_ref(O[0], os, "ced", Ol)),
                                            full_view(dense_ref(O[1], os, "ced", Ol)), choice, nullptr, cur, st));
            FullView fv = full_view(Lnew);
            SPP_CHECK(launch_dot<T>(fv, fv, nrm, cur, st, dot_partials(ctx, cur)));
            SPP_CHECK(launch_safe_factor(nrm, fac, status, fail, site, cur, st));
            SPP_CHECK(launch_scale<T>(fv, fac, cur, st));
            Ll = Ol;
            Ls = os;
          }
          ctx.arena.release(xm);
        }
        }  // !replayed
        } catch (...) {
          ctx.capturing = false;
          if (capture) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            W.graphs[gkey] = nullptr;
          }
          throw;
        }
        ctx.capturing = false;
        if (capture) {
          hipGraph_t g = nullptr;
          hipGraphExec_t ex = nullptr;
          const hipError_t e1 = hipStreamEndCapture(st, &g);
          const hipError_t e2 = (e1 == hipSuccess && g) ? hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) : e1;
          if (g) (void)hipGraphDestroy(g);
          if (e2 != hipSuccess || !ex) {
            (void)hipGetLastError();
            throw HipError(std::string("row sweep capture failed: ") + hipGetErrorString(e2));
          }
          W.graphs[gkey] = ex;
          SPP_CHECK(hipGraphLaunch(ex, st));
          ctx.counters.graph_captures += 1;
        }
        if (!replayed) W.seen[gkey] = 1;
        ctx.arena.release(mark);
        // ---------------- absorb the drawn row ----------------
        // (the physical kind has no use for the last one, born.jl:157-160)
        if (phys && y == Ly - 1) break;
        RowLayers<T> row;
        row.L = Lx;
        row.spec.assign(c.spec.begin() + y * Lx, c.spec.begin() + (y + 1) * Lx);
        row.ptrs.resize(Lx);
        for (int x = 0; x < Lx; ++x) row.ptrs[x] = ptrs + (long long)x * cur;
        if (kron)
          for (int x = 0; x < Lx; ++x) {
            row.kron_site.push_back(&c.kron_site[y * Lx + x]);
            row.kron_tabs.push_back(kptrs + (long long)x * cur);
          }
        absorb_forward<T>(ctx, Nc, row, chi, NB[1 - which]);
        which = 1 - which;
        hipLaunchKernelGGL(live_trunc_kernel, dim3(gb), dim3(tb), 0, st, NB[which].max_trunc, done, mtlive, cur);
        SPP_CHECK(hipGetLastError());
      }
      // lognorm of the normalised projected sample from the final north boundary
      Boundary<T>& Nf = NB[which];
      if (!phys) {
        for (int x = 0; x <= Lx; ++x)
          if (Nf.bdim[x] != 1) throw HipError("final boundary must have unit bonds");
        std::vector<long long> sst(Lx);
        for (int x = 0; x < Lx; ++x) sst[x] = Nf.tsize(x);
        upload(dsst, sst.data(), sizeof(long long) * Lx, st);
        upload(dtens, Nf.t.data(), sizeof(T*) * Lx, st);
        SPP_CHECK(launch_close_unit<T>(dtens, dsst, Lx, mant, cur, st));
        SPP_CHECK(launch_gather_lognorm(c.cand_lognorm, labels, ns, lnamp, cur, st));
      } else {
        SPP_CHECK(launch_fill<double>(mant, cur, NAN, st));
        SPP_CHECK(launch_fill<double>(lnamp, cur, NAN, st));
        SPP_CHECK(hipMemcpyAsync(h_rep.data(), drep, sizeof(double) * (size_t)cur * ns * 6, hipMemcpyDeviceToHost,
                                 st));
      }
      SPP_CHECK(hipMemcpyAsync(h_status.data(), status, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_fail.data(), fail, sizeof(int) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_labels.data(), labels, sizeof(int) * cur * ns, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_logp.data(), logp, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_cond.data(), cond, sizeof(double) * cur * ns, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mt.data(), mtlive, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_mant.data(), mant, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_ls.data(), Nf.log_scale, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipMemcpyAsync(h_amp.data(), lnamp, sizeof(double) * cur, hipMemcpyDeviceToHost, st));
      SPP_CHECK(hipStreamSynchronize(st));
      for (int s = 0; s < cur; ++s) {
        long long g = s0 + s;
        if (out.labels) std::memcpy(out.labels + g * ns, h_labels.data() + (size_t)s * ns, sizeof(int) * ns);
        if (out.conditionals)
          std::memcpy(out.conditionals + g * ns, h_cond.data() + (size_t)s * ns, sizeof(double) * ns);
        if (out.log_prob) out.log_prob[g] = h_logp[s];
        if (out.max_trunc) out.max_trunc[g] = std::max(south_mt, h_mt[s]);
        if (out.status) out.status[g] = h_status[s];
        if (out.fail_site) out.fail_site[g] = h_fail[s];
        // normalised projected PEPS: log_amplitude = pepo + sum log||A_k||,
        // tensor_log_norm = log(mantissa) + log_scale - 2 sum log||A_k||
        if (out.log_amplitude) out.log_amplitude[g] = c.pepo_log_amplitude + h_amp[s];
        if (out.tensor_log_norm) {
          double m = h_mant[s];
          out.tensor_log_norm[g] = (m > 0.0 && std::isfinite(m)) ? std::log(m) + h_ls[s] - 2.0 * h_amp[s] : NAN;
        }
        if (phys && out.repairs)
          std::memcpy(out.repairs + g * ns * 6, h_rep.data() + (size_t)s * ns * 6, sizeof(double) * ns * 6);
      }
    }
  } catch (...) {
    (void)hipStreamSynchronize(st);
    ctx.arena.release(0);
    c.work.reset();  // buffers of a failed call are not reused
    throw;
  }
}

template void draw_batch<float>(Context&, CacheDev<float>&, int, const double*, const int*, DrawOutputs&);
template void draw_batch<double>(Context&, CacheDev<double>&, int, const double*, const int*, DrawOutputs&);

}  // namespace spp
