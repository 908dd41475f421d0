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
// With pool != nullptr every device buffer of the call (layers, pointer
// tables, boundaries, the absorb workspace, dot partial sums) is carved from
// [pool, pool + pool_bytes) and the call makes no device allocation; a state
// batch whose single-state need exceeds pool_bytes is rejected before any
// launch.
template <typename T>
void lognorm_batch(Context& ctx, const std::vector<const StateDev*>& states, int chi, LognormOutputs& out,
                   char* pool = nullptr, size_t pool_bytes = 0);

}  // namespace spp
