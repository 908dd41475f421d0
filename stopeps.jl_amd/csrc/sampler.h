// Perfect Born sampling of PEPO auxiliary indices, batched over samples.
#pragma once
#include "engine.h"
#include <map>
#include <memory>

namespace spp {

struct SiteDev {
  double* d = nullptr;   // dense (W,N,P,A,E,S), parity-sorted legs, FP64
  int dim[6] = {1, 1, 1, 1, 1, 1};
  int even[6] = {1, 1, 1, 1, 1, 1};
  long long size() const {
    long long n = 1;
    for (int k = 0; k < 6; ++k) n *= dim[k];
    return n;
  }
};

// Sites that one batched call can hold side by side: equal dimensions on every
// leg and equal parity splits on every leg but the auxiliary one. The double
// layers trace the auxiliary leg, so its parity split (the Z2 charge of a
// projected sample's label) may differ per sample.
inline bool batchable_sites(const SiteDev& a, const SiteDev& b) {
  for (int k = 0; k < 6; ++k)
    if (a.dim[k] != b.dim[k] || (k != 3 && a.even[k] != b.even[k])) return false;
  return true;
}

struct StateDev {
  int Lx = 0, Ly = 0, kind = 0;  // kind 0: PEPO, 1: PEPS
  std::vector<SiteDev> sites;    // x + y*Lx
  double beta = 0, log_amplitude = 0;
  // per site, per virtual leg (W, N, E, S): live even / odd counts of the
  // bond (the sectors in use; the rest of the leg's capacity is zero)
  std::vector<int> live;
  // nonzero: a full-update step failed on this state (its failure code); the
  // compute entry points refuse it
  int failed = 0;
  // one device block holding every site (sites[i].d point into it; freed as a
  // whole), or null when each site owns its own allocation
  double* block = nullptr;
  ~StateDev();
};

inline Leg pair_leg(int dim, int even) {  // (ket, bra) fused leg of a double layer
  int odd = dim - even;
  return fixed_leg(dim * dim, even * even + odd * odd);
}

// Physical-kind data of one site (resampling proposals, born.jl:116-134): the
// FP64 site tensor and the ket tuples (row-major over W, N, E, S) of the two
// parity blocks of its metric.
struct PhysSite {
  double* a = nullptr;   // site copy (W,N,P,A,E,S), A of dimension 1
  int dim[6] = {1, 1, 1, 1, 1, 1};
  int even[6] = {1, 1, 1, 1, 1, 1};
  int nq[2] = {0, 0};    // block sizes
  int* inv = nullptr;    // device [2][ld]: tuple index of (block, position)
};

// Device workspace of physical_conditionals for nb samples (blocks of at most
// ld rows): M [nb][2][ld][ld], taus / d / e [nb][2][ld], y [nb][2][2][ld],
// norms [nb][2][2], res [nb][2][8].
struct PsdScratch {
  int ld = 0;
  double *M = nullptr, *taus = nullptr, *d = nullptr, *e = nullptr, *y = nullptr, *norms = nullptr, *res = nullptr;
};
inline size_t psd_scratch_doubles(int ld) { return 2 * (size_t)ld * ld + 3 * 2 * (size_t)ld + 4 * (size_t)ld + 4 + 16; }

// Eigenvalue summaries and label weights of the PSD-projected metric of one
// site for nb samples: H [nb][hs] pair-leg holes (FP32 when h_fp32), U the 2x2
// basis (row-major U[p][k]). Results in w.res / w.norms for
// physical_select_kernel.
void physical_conditionals(Context& ctx, const void* H, bool h_fp32, long long hs, const PhysSite& ps,
                           const double* U, int nb, const PsdScratch& w);
// The site's report (repairs [nb][nsite][6]) and Born conditional of every
// sample from physical_conditionals, then the draw with the uniforms (or the
// given labels), as born_select_kernel does for the auxiliary kind. prefix
// [nb]: charge of the labels chosen so far; target < 0: no charge constraint.
void launch_physical_select(Context& ctx, const PsdScratch& w, const double* uniforms, const int* labels_in,
                            long long sample0, int nsite, int site, const int* charges, int target, int reach_next,
                            int* prefix, int* choice, int* labels, double* logp, double* cond, int* status,
                            int* fail_site, int* done, double* repairs, int nb);
// Host self-test of the PSD path: eigenvalues and label projections of
// symmetric blocks [nprob][ld][ld] with vectors [nprob][2][ld].
void psd_selftest(Context& ctx, int nprob, int ld, const int* n_h, const double* mats_h, const double* vecs_h,
                  double* eig_h, double* proj_h, int* fail_h);

// `_categorical(log.(p), u)` (src/resample/selection.jl:68-80) over npos
// probabilities with the uniform u: the first index whose cumulative mass
// exp(log p - logsumexp) exceeds u, skipping zero entries, else the last
// positive one.
__device__ inline int categorical_index(const double* p, int npos, double u) {
  double pivot = -INFINITY;
  for (int j = 0; j < npos; ++j) pivot = fmax(pivot, log(p[j]));
  double sum = 0.0;
  for (int j = 0; j < npos; ++j) sum += exp(log(p[j]) - pivot);
  const double z = pivot + log(sum);
  double cum = 0.0;
  int last = -1;
  for (int j = 0; j < npos; ++j) {
    const double lm = log(p[j]);
    if (lm == -INFINITY) continue;
    last = j;
    cum += exp(lm - z);
    if (u < cum) return j;
  }
  return last;
}

// Device buffers of the batched draw, kept with the sampling cache between
// calls (sized for nb samples per chunk), and the captured row sweeps: every
// row's right-environment and draw sweep is a fixed sequence of launches over
// these buffers, the context's workspace arena and its dot-product scratch, so
// it is captured once per key (row, samples in the chunk, mode, bond
// dimensions) and replayed while those addresses stay the same.
template <typename T>
struct DrawWork {
  int nb = 0, ns = 0, lx = 0;
  long long capmax = 0;
  int *choice = nullptr, *labels = nullptr, *status = nullptr, *fail = nullptr, *done = nullptr, *prefix = nullptr;
  int* dlab = nullptr;
  double *logp = nullptr, *cond = nullptr, *raw0 = nullptr, *raw1 = nullptr, *nrm = nullptr, *fac = nullptr;
  double *lnamp = nullptr, *mant = nullptr, *mtlive = nullptr, *drep = nullptr, *duni = nullptr;
  void** ptrs = nullptr;
  long long* dsst = nullptr;
  T** dtens = nullptr;
  Boundary<T> NB[2];
  std::map<std::vector<long long>, hipGraphExec_t> graphs;  // null exec: capture failed, run eagerly
  std::map<std::vector<long long>, int> seen;
  const char* arena_base = nullptr;  // addresses the graphs were captured with
  const double* dot_base = nullptr;
  void allocate(int nb_, int ns_, int lx, long long capmax_, bool physical);
  void drop_graphs();
  ~DrawWork();
};

template <typename T>
struct CacheDev {
  int Lx = 0, Ly = 0, chi = 0, precision = 0;
  double basis[4] = {1, 0, 0, 1};  // U[a][k] row-major
  int charges[2] = {0, 1};
  std::vector<T*> layers;            // unprojected double layers, x + y*Lx
  std::vector<T*> cand[2];           // aux-projected double layers
  std::vector<LayerSpec> spec;       // leg specs (W,N,E,S) of every layer
  double* cand_lognorm = nullptr;    // device [site][2]: log ||project(A, k)||
  std::vector<Boundary<T>> south;    // south[y], y = 0..Ly, rotated storage, shared
  std::vector<double> south_max_trunc;  // host copy per y
  double pepo_log_amplitude = 0, pepo_beta = 0;
  // 0: auxiliary kind (PEPO), 1: physical kind (PEPS, resampling proposals)
  int kind = 0;
  // physical kind: per-site tensors and metric tables, the largest block, the
  // state's total charge (-1: unconstrained) and the reachable-charge masks
  // reach[i] (bit q: charge q can still be fused by sites i..; reach[ns] = 1)
  std::vector<PhysSite> phys;
  int psd_ld = 0;
  int target_charge = -1;
  std::vector<int> reach;
  std::unique_ptr<DrawWork<T>> work;  // draw buffers and captured sweeps (first draw on)
  // Single-layer tables of the candidates for the factorised Gram transfers
  // of the draw's absorbs (FP32 caches whose every candidate fits, kron.h):
  // per site the layout, per (site, label) the tables inside kron_buf.
  std::vector<KronSite> kron_site;
  std::vector<void*> kron_tab[2];
  char* kron_buf = nullptr;
  ~CacheDev();
};

// `project_configuration` (born.jl:27-36) of n configurations of a resident
// PEPO on the device: site i of state s is sum_a A_i[w,n,p,a,e,s] U[a,k] with
// k = labels[s][i] (a one-index auxiliary leg of the label's charge), its
// interior bonds re-embedded in the capacity (even, odd) = cap (the PEPO's own
// when cap is null: the first live indices of each parity kept, zero slots
// appended), then `normalize_state!` (finite.jl:89-100): every site divided by
// its norm, whose logs (normalizations [n][ns]) add to the log amplitude.
void project_batch(Context& ctx, const StateDev& pepo, int n, const int* labels, const double* basis, const int* cap,
                   std::vector<StateDev*>& out, double* normalizations);

// kind 0: auxiliary kind of a PEPO; kind 1: physical kind of a PEPS.
template <typename T>
void build_cache(Context& ctx, const StateDev& pepo, int chi, const double* basis, CacheDev<T>& cache, int kind = 0);

struct DrawOutputs {
  int* labels = nullptr;           // [n][Lx*Ly] basis indices (0-based), row-major sites
  double* log_prob = nullptr;      // [n]
  double* conditionals = nullptr;  // [n][Lx*Ly]
  double* max_trunc = nullptr;     // [n]
  double* tensor_log_norm = nullptr;  // [n] log norm of the normalised projected PEPS (null: skip)
  double* log_amplitude = nullptr; // [n] log amplitude of the projected PEPS (null: skip)
  int* status = nullptr;           // [n] 0 ok, 1 invalid conditional, 2 zero mass, 3 nonfinite
                                   // environment; physical kind: 4 nonfinite norm metric,
                                   // 5 no label fits the total charge, 6 eigensolver cap
  int* fail_site = nullptr;        // [n] row-major site index of the failure (or -1)
  double* repairs = nullptr;       // physical kind [n][Lx*Ly][6] (null: skip): the site's
                                   // _positive_metric report and the zero-weight fallback
};

// uniforms: [n][Lx*Ly] (row-major, x fastest) or null in evaluation mode;
// labels_in: [n][Lx*Ly] basis indices for evaluation mode or null.
template <typename T>
void draw_batch(Context& ctx, CacheDev<T>& cache, int n, const double* uniforms, const int* labels_in,
                DrawOutputs& out);

}  // namespace spp
