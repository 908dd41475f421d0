// Perfect Born sampling of PEPO auxiliary indices, batched over samples.
#pragma once
#include "engine.h"

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
  ~StateDev();
};

inline Leg pair_leg(int dim, int even) {  // (ket, bra) fused leg of a double layer
  int odd = dim - even;
  return fixed_leg(dim * dim, even * even + odd * odd);
}

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
  ~CacheDev();
};

template <typename T>
void build_cache(Context& ctx, const StateDev& pepo, int chi, const double* basis, CacheDev<T>& cache);

struct DrawOutputs {
  int* labels = nullptr;           // [n][Lx*Ly] basis indices (0-based), row-major sites
  double* log_prob = nullptr;      // [n]
  double* conditionals = nullptr;  // [n][Lx*Ly]
  double* max_trunc = nullptr;     // [n]
  double* tensor_log_norm = nullptr;  // [n] log norm of the normalised projected PEPS (null: skip)
  double* log_amplitude = nullptr; // [n] log amplitude of the projected PEPS (null: skip)
  int* status = nullptr;           // [n] 0 ok, 1 invalid conditional, 2 zero mass, 3 nonfinite
  int* fail_site = nullptr;        // [n] row-major site index of the failure (or -1)
};

// uniforms: [n][Lx*Ly] (row-major, x fastest) or null in evaluation mode;
// labels_in: [n][Lx*Ly] basis indices for evaluation mode or null.
template <typename T>
void draw_batch(Context& ctx, CacheDev<T>& cache, int n, const double* uniforms, const int* labels_in,
                DrawOutputs& out);

}  // namespace spp
