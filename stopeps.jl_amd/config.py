"""Run configuration (src/driver/config.jl:1-170): ``RunConfig`` and
``ResamplingOptions`` with the reference's field names and defaults, their
validation, the output beta grid and the parameter dictionary of the result
schema.

Spaces are written as this library writes them: a Z2 bond space as its
(even, odd) sector dimensions, a dense one as an int. Fields that select
reference features this library does not provide (diagonal or variational
contraction) are kept so that a configuration reads the same; the drivers
reject them with a message naming the feature.
"""
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from .evolution import FullUpdate
from .measurement import TFIM
from .particles import same_beta, temperature_nodes as _nodes
from .states import Lattice, SymmetryBasis, XBasis, LocalBasis

RESULT_SCHEMA_VERSION = "2.7.0"
FU_ALGORITHM_REVISION = 2


@dataclass
class ResamplingOptions:
    """``ResamplingOptions`` (config.jl:1-18)."""
    enabled: bool = True
    eta_th: float = 0.7
    lambda_max: float = 0.999
    lambda_tolerance: float = 1e-8
    lambda_threshold: float = 0.2
    max_resample_sweep: int = 3
    score_batch_size: int = 32


@dataclass
class RunConfig:
    """``RunConfig`` (config.jl:20-61). ``chi_sampling`` and ``chi_measurement``
    default to ``chi_environment``, ``chi_proposal`` to ``chi_sampling``."""
    lattice: Lattice = field(default_factory=lambda: Lattice(2, 2))
    model: TFIM = field(default_factory=TFIM)
    beta0: float = 0.1
    beta_log: list = field(default_factory=lambda: [round(0.05 * k, 12) for k in range(7)])
    beta_measure: list = field(default_factory=lambda: [0.0, 0.1, 0.2, 0.3])
    delta_tau_pur: float = 0.025
    delta_tau_peps: float = 0.025
    D_pur: int = 4
    D_peps: int = 4
    bond_space_pur: object = None
    bond_space_peps: object = None
    chi_environment: int = 32
    chi_sampling: int = None
    chi_measurement: int = None
    chi_proposal: int = None
    environment_space: object = None
    sampling_space: object = None
    measurement_space: object = None
    proposal_space: object = None
    sampling_basis: object = "auto"
    trotter_order: int = 2
    update: FullUpdate = field(default_factory=FullUpdate)
    contraction: str = "BoundaryMPS"
    max_sweep: int = 4
    threshold_sweep: float = 1e-8
    target_cache_bytes: int = 256 * 1024 ** 2
    record_boundary_entropy: bool = True
    Ns: int = 16
    seed: int = 20260924
    resampling: ResamplingOptions = field(default_factory=ResamplingOptions)
    save_interval_seconds: float = 3600.0
    output_directory: str = None
    overwrite_output: bool = False

    def __post_init__(self):
        if self.chi_sampling is None:
            self.chi_sampling = self.chi_environment
        if self.chi_measurement is None:
            self.chi_measurement = self.chi_environment
        if self.chi_proposal is None:
            self.chi_proposal = self.chi_sampling
        if self.sampling_space is None:
            self.sampling_space = self.environment_space
        if self.measurement_space is None:
            self.measurement_space = self.environment_space
        if self.proposal_space is None:
            self.proposal_space = self.sampling_space


def measure_at(c, beta):
    """``_measure_at`` (config.jl:64)."""
    return any(same_beta(b, beta) for b in c.beta_measure)


def temperature_nodes(c):
    """``temperature_nodes(c::RunConfig)`` (config.jl:68-76)."""
    return _nodes(c.beta0, c.beta_log, c.beta_measure)


def sampling_basis(c):
    """``_basis(c)``: the local sampling basis (basis.jl:64-80): ``:auto`` is
    SymmetryBasis for a Z2 model and XBasis for a dense one."""
    b = c.sampling_basis
    if isinstance(b, LocalBasis):
        return b
    if b in ("auto", None):
        return SymmetryBasis() if c.model.symmetric else XBasis()
    if b in ("symmetry", "SymmetryBasis"):
        return SymmetryBasis()
    if b in ("x", "XBasis"):
        return XBasis()
    raise ValueError(f"unknown sampling basis {b!r}")


def bond_capacity(c, D, space):
    """The device bond capacity for a bond dimension ``D`` and sector budget
    ``space``: the budget itself, or (D, D) per parity for a Z2 model without
    one (truncrank(D) alone may put all D in one sector), D for a dense model."""
    if space is not None:
        return tuple(space) if c.model.symmetric else int(space)
    return (int(D), int(D)) if c.model.symmetric else (int(D), 0)


def validate_config(c):
    """``validate_config`` (config.jl:80-115): numerical controls, model and
    basis. Returns the configuration."""
    if not all(np.isfinite(b) and b >= 0 for b in [c.beta0, *c.beta_log, *c.beta_measure]):
        raise ValueError("invalid beta grid")
    if not all(np.isfinite(x) and x > 0 for x in (c.delta_tau_pur, c.delta_tau_peps)):
        raise ValueError("delta_tau must be positive")
    if min(c.D_pur, c.D_peps, c.chi_environment, c.chi_sampling, c.chi_measurement, c.chi_proposal, c.Ns) < 1:
        raise ValueError("dimensions and sample count must be positive")
    if not (c.seed >= 0 and c.trotter_order in (1, 2)):
        raise ValueError("invalid seed or Trotter order")
    if not (np.isfinite(c.save_interval_seconds) and c.save_interval_seconds >= 0):
        raise ValueError("invalid save interval")
    r = c.resampling
    if not (0 <= r.eta_th <= 1 and 0 <= r.lambda_max < 1 and r.lambda_tolerance > 0
            and np.isfinite(r.lambda_threshold) and 0 <= r.lambda_threshold < 1
            and r.max_resample_sweep >= 1 and r.score_batch_size > 0):
        raise ValueError("invalid resampling options")
    if r.enabled and c.Ns < 2:
        raise ValueError("resampling needs Ns >= 2")
    u = c.update
    if not (u.max_iterations >= 0 and u.linear_max_iterations > 0
            and all(np.isfinite(x) and x >= 0 for x in (u.error_tolerance, u.diagonal_shift, u.linear_tolerance))):
        raise ValueError("invalid full-update options")
    if not all(np.isfinite(x) for x in (c.model.J, c.model.h)):
        raise ValueError("nonfinite model parameter")
    for V in (c.bond_space_pur, c.bond_space_peps, c.environment_space, c.sampling_space, c.measurement_space,
              c.proposal_space):
        if V is None:
            continue
        if c.model.symmetric and not (len(V) == 2 and min(V) >= 0 and sum(V) > 0):
            raise ValueError("sector budgets must be nonempty (even, odd) spaces with the model symmetry")
        if not c.model.symmetric and not int(V) > 0:
            raise ValueError("sector budgets must be nonempty spaces with the model symmetry")
    sampling_basis(c)
    return c


def device_support(c, stochastic):
    """Raise for configuration choices this library does not run on the device."""
    if c.contraction not in ("BoundaryMPS", "BoundaryMPSRow"):
        raise NotImplementedError(f"contraction {c.contraction}: only the row boundary MPS is provided")


def parameter_dict(c):
    """``parameter_dict`` (config.jl:117-170): the effective configuration as
    plain values, plus the conventions this library follows."""
    def enc(x):
        if dataclasses.is_dataclass(x):
            return {"type": type(x).__name__, **{f.name: enc(getattr(x, f.name)) for f in dataclasses.fields(x)}}
        if isinstance(x, Lattice):
            return {"type": "Lattice", "Lx": x.Lx, "Ly": x.Ly}
        if isinstance(x, TFIM):
            return {"type": "TFIM", "J": x.J, "h": x.h,
                    "physical_space": "Z2Space(0=>1,1=>1)" if x.symmetric else "ComplexSpace(2)"}
        if isinstance(x, FullUpdate):
            return {"type": "FullUpdate", "max_iterations": x.max_iterations, "error_tolerance": x.error_tolerance,
                    "diagonal_shift": x.diagonal_shift, "linear_max_iterations": x.linear_max_iterations,
                    "linear_tolerance": x.linear_tolerance}
        if isinstance(x, LocalBasis):
            return {"labels": list(x.labels), "charges": list(x.charges), "U": np.asarray(x.U).tolist()}
        if isinstance(x, (list, tuple)):
            return [enc(v) for v in x]
        return x
    data = {f.name: enc(getattr(c, f.name)) for f in dataclasses.fields(c)}
    data["sampling_basis"] = enc(sampling_basis(c))
    data["contraction"] = "BoundaryMPS-Row"
    L = c.lattice
    data["sampling_site_order"] = [(x + 1, y + 1) for (x, y) in L.site_order()]
    data["contraction_order"] = "row_major"
    data["state_truncation_metric"] = "relative_squared_FU_residual"
    data["full_update_algorithm"] = "environment_gauge_SVD_ALS"
    data["full_update_algorithm_revision"] = FU_ALGORITHM_REVISION
    data["full_update_gauge_floor"] = ("max(sqrt(eps(real_scalar_type)),metric_dimension*eps(real_scalar_type))"
                                       "*spectral_scale; auxiliary_metric_only")
    data["full_update_cost_evaluation"] = "original_coordinates_and_full_metric; PSD_repair_not_provided"
    data["full_update_linear_system"] = "environment_and_internal_gauge_coordinates; relative_diagonal_shift_and_CG_residual"
    data["full_update_internal_gauge"] = "QR_after_left_solve; LQ_after_right_solve; best_center_SVD_without_truncation"
    data["environment_truncation_metric"] = "canonical_SVD_relative_discarded_weight"
    data["boundary_entropy_convention"] = "pretruncation_Schmidt_entropy; nats; sector_dimension_weighted"
    data["boundary_entropy_log_scope"] = "maximum_over_FU_boundary_cuts_in_output_interval"
    data["boundary_entropy_profile_scope"] = ("energy_measurement_boundaries; Se.north[y,b] and Se.south[y,b]; "
                                              "cut_after_row_y_and_column_b")
    data["sampling_rng"] = "counter_splitmix64(seed, sample_index, site)"
    return data
