// Batched energy measurement of projected PEPO samples (restates
// measure_energy(::FiniteState, ::TFIM, ::AbstractRowBoundaryMPS),
// src/measurement/energy.jl:31-78).
#pragma once
#include "sampler.h"

namespace spp {

// The observables of a nearest-neighbour model with one onsite and one
// two-site term: E = -J sum_<ij> <F_i G_j> - h sum_i <X_i>. The two-site
// operator is given by its rank-one split F (x) G through a one-dimensional
// factor leg (_operator_factors, energy.jl:8-12); in a Z2-graded state that
// leg carries charge 1 when the factors flip the physical parity.
struct EnergyOps {
  double X[4] = {1, 0, 0, 1};  // onsite operator X[p'][p], row-major
  double F[4] = {1, 0, 0, 1};  // first factor F[p'][p]
  double G[4] = {1, 0, 0, 1};  // second factor G[q'][q]
  int odd = 0;                 // factor leg has odd parity
  double J = 1.0, h = 1.0;
};

struct EnergyOutputs {
  double* energy_total = nullptr;     // [n]
  double* energy_per_site = nullptr;  // [n]
  double* log_norm = nullptr;         // [n] log <psi|psi> incl. 2 log_amplitude of the normalised sample
  double* tensor_log_norm = nullptr;  // [n] log <psi|psi> of the normalised site tensors
  double* max_trunc = nullptr;        // [n] max discarded weight over the north and south boundaries
  // [n] 0 ok, 3 nonfinite row environment, 4 zero local norm environment,
  // 5 nonfinite observable, 6 norm not finite and positive
  int* status = nullptr;
  int* fail_site = nullptr;           // [n] row-major site of the failure, or -1
  // Boundary entropy (record_entropy, energy.jl:31-78; nullptr: not recorded):
  // [n][2][Ly-1][Lx-1] Se.north then Se.south (NaN where undefined) and [n]
  // the maximum over every cut of the south and north absorbs.
  double* entropy_profile = nullptr;
  double* entropy_max = nullptr;
};

// Energies of the PEPS samples obtained by projecting the auxiliary legs of
// `pepo` onto the basis vectors `labels` [n][Lx*Ly] (0/1, row-major sites) of
// `cache` (built from the same PEPO), contracted with boundary MPS of bond `chi`.
template <typename T>
void measure_energy_batch(Context& ctx, const StateDev& pepo, const CacheDev<T>& cache, int chi, int n,
                          const int* labels, const EnergyOps& ops, EnergyOutputs& out);

// Energies of uploaded states (PEPS or PEPO, all on one lattice with the same
// leg spaces; double layers trace the auxiliary leg), e.g. particles after
// full-update evolution. log_norm adds each state's 2 log_amplitude;
// tensor_log_norm is that of the stored tensors.
template <typename T>
void measure_energy_states(Context& ctx, const std::vector<const StateDev*>& states, int chi, const EnergyOps& ops,
                           EnergyOutputs& out);

}  // namespace spp
