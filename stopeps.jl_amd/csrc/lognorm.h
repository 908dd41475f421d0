// Batched lognorm of finite PEPS/PEPO states (restates
// lognorm(::FiniteState, ::AbstractRowBoundaryMPS) and contract_layers,
// src/contraction/boundary_mps.jl:162-191).
#pragma once
#include "sampler.h"

namespace spp {

struct LognormOutputs {
  double* log_norm = nullptr;         // [n] log <psi|psi> incl. 2 log_amplitude
  double* tensor_log_norm = nullptr;  // [n] log <psi|psi> of the stored tensors
  double* max_trunc = nullptr;        // [n] max discarded weight over the north absorbs
  int* status = nullptr;              // [n] 0 ok, 6 norm not finite and positive
};

// All states share the lattice and every site's leg dimensions and even counts.
template <typename T>
void lognorm_batch(Context& ctx, const std::vector<const StateDev*>& states, int chi, LognormOutputs& out);

}  // namespace spp
