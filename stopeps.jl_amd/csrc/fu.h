// Full-update evolution (src/evolution/full_update.jl) on the device.
#pragma once
#include "fu_core.h"
#include "sampler.h"

namespace spp {

// Host execution of the per-bond solve that the device kernel runs (selftest).
fu::Result fu_bond_host(const fu::Dims& g, const fu::Opts& o, const double* N4, const double* target, double* A,
                        double* B);

// Host run of the bond solve's Jacobi eigensolver (kind 0: symmetric n x n A,
// eigenvalues unordered) or one-sided Jacobi SVD (kind 1: m x n A with m >= n,
// singular values descending) with a sweep cap; returns whether it converged
// within max_sweeps (selftest).
bool jacobi_host(int kind, int m, int n, const double* A, int max_sweeps, double* values);

// Host run of the reduced-pair stage of the device FU for one sample
// (selftest): QR of both sites (live sectors only) into pseudo-site Q
// tensors (W,N,P,A,E,S slots with Aux in P and r in the bond leg's slot),
// cores C[r][p][k] and the gated target[r][p][q][s]. geometry: {rcap_a,
// even_a, rcap_b, even_b, live r counts (a even, a odd, b even, b odd)}.
int fu_pair_host(int horizontal, const int* dims_a, const int* evens_a, const int* dims_b, const int* evens_b,
                 const double* site_a, const double* site_b, const int* live_a, const int* live_b, const double* gate,
                 double* qa, double* qb, double* ca, double* cb, double* target, int* geometry);

struct FuParams {
  fu::Opts opts;
  int order = 2;             // Trotter order (2: sweeps H, V, V reversed, H reversed)
  double delta = 0.0;        // the step's delta: beta grows by 2 delta
  const double* gates_h = nullptr;  // host [Ly*(Lx-1)][d^4]: exp(-gate_delta h_b), bond (x,y)-(x+1,y) at y*(Lx-1)+x
  const double* gates_v = nullptr;  // host [(Ly-1)*Lx][d^4]: bond (x,y)-(x,y+1) at y*Lx+x
};

// Per-sample outcome of a full-update step: 0 ok, 1..8 the bond-solve status
// of the first bond that failed (fu::Status), or the codes below.
constexpr int kFuNonfiniteEnvironment = 9;  // a row environment of the sample became nonfinite
constexpr int kFuZeroTensor = 10;           // a site norm was zero or nonfinite at normalize_state!
constexpr int kFuRepairFailed = 11;         // the PSD repair did not remove the residual failure
inline const char* fu_failure_message(int code) {
  static const char* msg[] = {"ok",
                              "nonfinite FU norm metric",
                              "zero FU norm metric",
                              "gate produced a zero target",
                              "negative FU residual needs PSD repair",
                              "nonfinite FU residual/target",
                              "singular FU gauge",
                              "FU workspace exhausted",
                              "FU eigen/SVD solve did not converge",
                              "nonfinite row environment",
                              "zero or nonfinite tensor after the full update",
                              "PSD repair did not remove the FU residual failure"};
  return (code >= 0 && code <= 11) ? msg[code] : "unknown failure";
}

struct FuOutputs {
  double* records = nullptr;         // [n][bond updates][8]: initial_cost, final_cost, residual, iterations,
                                     // limit_reached, stop, kept_even, kept_odd (sweep order)
  double* normalizations = nullptr;  // [n][Lx*Ly]: log of each site norm divided out
  double* max_env_trunc = nullptr;   // [n]
  // [n] maximum pre-truncation Schmidt entropy over every boundary cut of the
  // step (record_entropy, full_update.jl:560-591); nullptr: not recorded
  double* max_entropy = nullptr;
  // [n] per-sample outcome (0 ok, else a failure code above); a failed sample
  // keeps its sites from the failing bond on and its state is marked failed.
  // nullptr: the call fails after finishing the step when any sample failed.
  int* status = nullptr;
};

// `evolution_step!` (full_update.jl:563-597) with FullUpdate on every state
// (same lattice and leg spaces; the states' sites are updated in place).
template <typename T>
void fu_evolution_step(Context& ctx, const std::vector<StateDev*>& states, int chi, const FuParams& P,
                       FuOutputs& out);

}  // namespace spp
