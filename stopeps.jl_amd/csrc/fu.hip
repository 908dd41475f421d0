// Full-update evolution of a batch of states (src/evolution/full_update.jl):
// the per-bond solve of fu_core.h on the device, and its host selftest.
#include "fu.h"
#include "engine.h"
#include <vector>

namespace spp {

// Host execution of the per-bond solve (selftest of the code the device runs).
fu::Result fu_bond_host(const fu::Dims& g, const fu::Opts& o, const double* N4, const double* target, double* A,
                        double* B) {
  const long long cap = fu::workspace_doubles(g);
  std::vector<double> ws((size_t)cap);
  fu::Work w{ws.data(), 1, 0, cap};
  fu::View n{const_cast<double*>(N4), 1}, t{const_cast<double*>(target), 1}, a{A, 1}, b{B, 1};
  return fu::fu_bond(g, o, n, t, a, b, w);
}

}  // namespace spp
