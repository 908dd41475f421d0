"""stopeps.jl_amd - MI355X-native Born-sampling path of StoPEPS.jl.

The package directory name contains a dot, so load it with
``load_package()`` from ``__graft_entry__`` or the tests (it registers the
package as ``stopeps_jl_amd``). Compute runs in ``libstopeps_hip.so``
(hand-written HIP for gfx950, C ABI in include/stopeps_hip.h); this package
is the host-side mirror of the reference's Julia API for that path.
"""
from .states import (Lattice, Site, FiniteState, FinitePEPO, FinitePEPS, identity_pepo, parity_sorted,
                     synthetic_thermal_pepo, random_peps, pad_bonds, normalize_state, reembed_bonds,
                     save_state_npz, load_state_npz,
                     LocalBasis, SymmetryBasis, XBasis, resolve_basis)
from .sampling import (Context, DeviceState, SamplingCache, BornError, sample_configuration,
                       sample_configurations, sample_uniforms, project_configuration, project_configurations,
                       splitmix64)
from .measurement import (TFIM, spin_operators, operator_factors, EnergyError, measure_energies, measure_energy,
                          measure_energy_states)
from .contraction import NormError, lognorm, lognorms, lognorm_gradient, lognorm_gradients, lognorm_rrule
from .evolution import FullUpdate, FullUpdateError, trotter_gate, evolution_step, evolution_steps
from .particles import (Particle, ensemble_statistics, step_schedule, purify, stochastic_ensemble, new_particles,
                        advance_particles, statistics, group_configurations)
from .config import RunConfig, ResamplingOptions, validate_config, temperature_nodes, parameter_dict
from .driver import ThermalCheckpoint, run_purification, run_stochastic
from .results import save_result, load_result, save_checkpoint, load_checkpoint, RESULT_SCHEMA_VERSION
from . import _lib, particles, repair, resample

__all__ = ["synthetic_thermal_pepo", "random_peps", "Lattice", "Site", "FiniteState", "FinitePEPO", "FinitePEPS", "identity_pepo", "parity_sorted",
           "LocalBasis", "SymmetryBasis", "XBasis", "resolve_basis", "Context", "DeviceState",
           "SamplingCache", "BornError", "sample_configuration", "sample_configurations", "sample_uniforms",
           "project_configuration", "project_configurations", "splitmix64", "TFIM", "spin_operators", "operator_factors", "EnergyError",
           "measure_energies", "measure_energy", "measure_energy_states", "NormError", "lognorm", "lognorms",
           "lognorm_gradient", "lognorm_gradients", "lognorm_rrule", "pad_bonds", "FullUpdate", "trotter_gate",
           "evolution_step", "evolution_steps", "Particle", "ensemble_statistics", "temperature_nodes",
           "step_schedule", "purify", "stochastic_ensemble", "new_particles", "advance_particles", "statistics",
           "group_configurations", "normalize_state", "reembed_bonds", "RunConfig", "ResamplingOptions",
           "validate_config", "parameter_dict", "ThermalCheckpoint", "run_purification", "run_stochastic",
           "save_state_npz", "load_state_npz", "FullUpdateError", "save_result", "load_result",
           "save_checkpoint", "load_checkpoint", "RESULT_SCHEMA_VERSION"]
