// Full-update evolution (src/evolution/full_update.jl) on the device.
#pragma once
#include "fu_core.h"

namespace spp {

// Host execution of the per-bond solve that the device kernel runs (selftest).
fu::Result fu_bond_host(const fu::Dims& g, const fu::Opts& o, const double* N4, const double* target, double* A,
                        double* B);

}  // namespace spp
