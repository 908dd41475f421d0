// Gradient of log <psi|psi> with respect to every site tensor (environment
// adjoint of lognorm; the reference builds the same single-site norm matrix in
// _site_hole, src/sampling/born.jl:75-83).
#pragma once
#include "lognorm.h"

namespace spp {

struct GradOutputs {
  // [n][sum over sites of the site size], each site dense row-major (W,N,P,A,E,S)
  // in the parity-sorted leg order of the upload, FP64
  double* grads = nullptr;
  double* log_norm = nullptr;
  double* tensor_log_norm = nullptr;
  double* max_trunc = nullptr;   // over the north and south absorbs
  int* status = nullptr;         // 0 ok, 3 nonfinite row environment, 6 norm not finite and positive
};

template <typename T>
void lognorm_grad_batch(Context& ctx, const std::vector<const StateDev*>& states, int chi, GradOutputs& out);

}  // namespace spp
