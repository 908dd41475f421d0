"""Particles, weights and ensemble statistics of stochastic PEPS: the FP64 host
bookkeeping around the device path (src/driver/particles.jl:2-102,
src/statistics/ensemble.jl:8-40, src/driver/config.jl:63-76,
src/driver/run.jl:45-66 and 87-130, src/resample/selection.jl:122-140).

Every tensor operation runs on the device: Born draws (``sample_configurations``),
observation (``measure_energies`` on projected samples, ``measure_energy_states``
or ``lognorms`` on resident states) and evolution (``evolution_steps``). What
stays here is scalar per particle, kept in FP64 as the reference keeps it:
relative log weights, multiplicities, beta schedules and the ratio estimators.
Particles of one beta node are observed and evolved as one batch.
"""
import math
import time

import numpy as np

from .contraction import lognorms
from .evolution import FullUpdate, evolution_steps
from .measurement import measure_energies, measure_energy_states
from .sampling import (DeviceState, SamplingCache, project_configurations, sample_configurations,
                       sample_uniforms)
from .states import normalize_state, pad_bonds, reembed_bonds, resolve_basis

EPS = np.finfo(np.float64).eps


# ----------------------------------------------------------------------------
# beta grids (config.jl:63-76, particles.jl:55-64)
# ----------------------------------------------------------------------------

def same_beta(a, b):
    """``_same_beta`` (config.jl:63): equal up to 64 ulps of max(1, |a|, |b|)."""
    return abs(a - b) <= 64 * EPS * max(1.0, abs(a), abs(b))


def measure_at(beta_measure, beta):
    """``_measure_at`` (config.jl:64): energy is measured at this node."""
    return any(same_beta(b, beta) for b in beta_measure)


def temperature_nodes(beta0, beta_log, beta_measure):
    """``temperature_nodes(c)`` (config.jl:68-76): sorted grid of 0, beta0,
    beta_log and beta_measure, near-equal values merged (beta0 wins)."""
    out = []
    for b in sorted([0.0, float(beta0)] + [float(v) for v in beta_log] + [float(v) for v in beta_measure]):
        if same_beta(b, beta0):
            b = float(beta0)
        if out and same_beta(out[-1], b):
            continue
        out.append(b)
    return out


def step_schedule(start, target, delta_tau):
    """``_step_schedule`` (particles.jl:55-64): beta endpoints of the Trotter
    steps from ``start`` to ``target``, each advancing beta by 2 delta_tau, the
    last one shortened to land on the node."""
    steps = []
    b = float(start)
    while not same_beta(b, target):
        if not target > b:
            raise ValueError("evolution requires increasing beta endpoints")
        b = min(float(target), b + 2 * delta_tau)
        if same_beta(b, target):
            b = float(target)
        steps.append(b)
    return steps


# ----------------------------------------------------------------------------
# ensemble statistics (ensemble.jl:8-40)
# ----------------------------------------------------------------------------

def ensemble_statistics(relative, multiplicity, energy, Ns, log_scale=0.0, nsites=1):
    """``ensemble_statistics`` (ensemble.jl:8-25): self-normalised estimators
    with weights w = exp(relative - max(relative)) times multiplicities, the
    effective-sample fraction eta = clamp(mu1^2/mu2, 1/Nv, 1), delta-method
    standard errors, lnZ = log_scale + pivot + log mu1 and eta_p."""
    relative = np.asarray(relative, dtype=np.float64)
    m = np.asarray(multiplicity, dtype=np.float64)
    energy = np.asarray(energy, dtype=np.float64)
    if not (len(relative) == len(m) == len(energy) > 0):
        raise ValueError("empty or inconsistent ensemble")
    if not (np.all(np.isfinite(relative)) and np.all(m > 0)):
        raise ValueError("invalid particle weights/counts")
    Nv = float(m.sum())
    pivot = float(relative.max())
    w = np.exp(relative - pivot)
    mass = m * w
    total = float(mass.sum())
    mu1 = total / Nv
    mu2 = float(np.sum(m * w * w)) / Nv
    eta = min(max(mu1 * mu1 / mu2, 1.0 / Nv), 1.0)

    def sem(f):
        if Nv <= 1:
            return math.nan
        return math.sqrt(max(0.0, float(np.sum(m * np.abs(f) ** 2)) / (Nv * (Nv - 1))))

    measured = bool(np.all(np.isfinite(energy)))
    E = float(np.sum(mass * energy)) / total if measured else math.nan
    E_err = sem(w * (energy - E) / mu1) if measured else math.nan
    return {"energy_total": nsites * E, "energy_per_site": E, "energy_per_site_error": E_err,
            "lnZ": log_scale + pivot + math.log(mu1), "lnZ_error": sem((w - mu1) / mu1), "eta": eta,
            "eta_error": sem((2 * mu1 / mu2) * (w - mu1) - (mu1 * mu1 / (mu2 * mu2)) * (w * w - mu2)),
            "ESS": Nv * eta,
            "eta_p": min(max(total * total / (Ns * float(np.sum(mass * mass))), 1.0 / Ns), len(relative) / Ns),
            "target_samples": Ns, "virtual_samples": Nv, "representatives": len(relative)}


def group_configurations(configurations):
    """``_group_configurations`` (selection.jl:122-140): identical draws share
    one particle. Returns first_draw (index of each group's first draw),
    multiplicity and draw_to_group, groups in order of first appearance."""
    first, mult, mapping, lookup = [], [], [], {}
    for d, cfg in enumerate(configurations):
        key = np.asarray(cfg).tobytes()
        j = lookup.get(key)
        if j is None:
            j = len(first)
            lookup[key] = j
            first.append(d)
            mult.append(0)
        mult[j] += 1
        mapping.append(j)
    return {"first_draw": first, "multiplicity": mult, "draw_to_group": mapping}


# ----------------------------------------------------------------------------
# particles (particles.jl:2-24, 42-53, 86-102)
# ----------------------------------------------------------------------------

class Particle:
    """``Particle`` (particles.jl:2-24): a Born sample with its resident state
    and estimator weight. ``state`` is a ``DeviceState``."""

    __slots__ = ("id", "generation", "sample_index", "seed", "count", "configuration", "state", "log_p0",
                 "log_q_proposal", "relative_log_weight", "weight_log_scale", "physical_log_norm",
                 "tensor_log_norm", "energy_total", "energy_per_site", "max_state_truncation",
                 "max_environment_truncation", "max_boundary_entropy", "Se", "time_seconds", "measurement_seconds")

    def __init__(self, **kw):
        for k in self.__slots__:
            setattr(self, k, kw.get(k))

    @property
    def beta(self):
        return self.state.state.beta

    def record(self, owner=0):
        """``_particle_scalar`` (particles.jl:103-118): the scalar record."""
        return {"id": self.id, "generation": self.generation, "sample_index": self.sample_index,
                "seed": self.seed, "count": self.count, "configuration": np.array(self.configuration),
                "owner": owner, "beta": self.beta, "log_p0": self.log_p0, "log_q_proposal": self.log_q_proposal,
                "relative_log_weight": self.relative_log_weight, "weight_log_scale": self.weight_log_scale,
                "log_weight": self.weight_log_scale + self.relative_log_weight,
                "physical_log_norm": self.physical_log_norm, "energy_total": self.energy_total,
                "energy_per_site": self.energy_per_site, "max_state_truncation": self.max_state_truncation,
                "max_environment_truncation": self.max_environment_truncation,
                "max_boundary_entropy": self.max_boundary_entropy,
                "Se": None if self.Se is None else {k: np.array(v) for k, v in self.Se.items()},
                "time_seconds": self.time_seconds, "measurement_seconds": self.measurement_seconds}


def nan_max(values, init=None):
    """Julia's ``max``/``maximum`` on floats: NaN if any operand is NaN."""
    vals = list(values) if init is None else [init, *values]
    if not vals:
        raise ValueError("empty maximum")
    return math.nan if any(math.isnan(v) for v in vals) else max(vals)


def observe(ctx, states, model, chi, energy, record_entropy=False, sector_space=None):
    """``_observe`` (particles.jl:35-41) for a batch of resident states:
    ``measure_energy`` (with the ``Se`` profiles when ``record_entropy``) when
    ``energy``, else ``lognorm`` with NaN energies and no ``Se``."""
    t0 = time.perf_counter()
    if energy:
        r = measure_energy_states(ctx, states, model, chi, record_entropy=record_entropy, sector_space=sector_space)
    else:
        r = lognorms(ctx, states, chi, sector_space=sector_space)
        r["energy_total"] = np.full(len(states), np.nan)
        r["energy_per_site"] = np.full(len(states), np.nan)
        r["Se"] = [None] * len(states)
    r["measurement_seconds"] = (time.perf_counter() - t0) / max(len(states), 1)
    return r


def project_particles(ctx, cache, configurations, ids, counts, log_p0, log_q, sampling_error, model,
                      chi_measurement, energy, bond_space=None, pepo_live=None, seed=0, generation=0,
                      record_entropy=False, measurement_space=None):
    """``_new_particle`` (particles.jl:42-53) for each given configuration:
    project the resident PEPO of ``cache`` on the device
    (``project_configurations``: aux projection and ``normalize_state!``,
    bonds re-embedded in ``bond_space`` = (even, odd) capacities when given,
    with the PEPO's live sectors), observe the batch and set
    relative_log_weight = log_norm - log_q. When ``energy`` is set the
    projected samples are measured through the sampling cache's shared layers
    (the same values as measuring the projected states). ``pepo_live`` is
    accepted for compatibility; the resident PEPO's live sectors are used."""
    if len(configurations) == 0:
        return []
    states, _ = project_configurations(ctx, cache._dev, np.asarray(configurations), cache.basis, bond_space)
    t0 = time.perf_counter()
    if energy:
        m = measure_energies(cache, np.asarray(configurations), model, chi_measurement,
                             record_entropy=record_entropy, sector_space=measurement_space)
    else:
        m = lognorms(ctx, states, chi_measurement, sector_space=measurement_space)
        m["energy_total"] = np.full(len(states), np.nan)
        m["energy_per_site"] = np.full(len(states), np.nan)
        m["Se"] = [None] * len(states)
    mt_sec = (time.perf_counter() - t0) / max(len(states), 1)
    out = []
    for u, cfg in enumerate(configurations):
        lq = float(log_q[u])
        out.append(Particle(
            id=int(ids[u]), generation=generation, sample_index=int(ids[u]), seed=seed, count=int(counts[u]),
            configuration=np.array(cfg), state=states[u], log_p0=float(log_p0[u]), log_q_proposal=lq,
            relative_log_weight=float(m["log_norm"][u]) - lq, weight_log_scale=0.0,
            physical_log_norm=float(m["log_norm"][u]), tensor_log_norm=float(m["tensor_log_norm"][u]),
            energy_total=float(m["energy_total"][u]), energy_per_site=float(m["energy_per_site"][u]),
            max_state_truncation=0.0,
            max_environment_truncation=max(float(sampling_error[u]), float(m["max_environment_truncation"][u])),
            max_boundary_entropy=math.nan, Se=m["Se"][u], time_seconds=0.0, measurement_seconds=mt_sec))
    return out


def new_particles(ctx, cache, draws, model, chi_measurement, energy, bond_space=None, pepo_live=None, seed=0,
                  first_id=1, record_entropy=False, measurement_space=None):
    """``_initial_particles`` after the draws (run.jl:45-66): one particle per
    distinct configuration (``project_particles``), id = sample index of its
    first draw, multiplicity = its number of draws, log q = log p0 = the draw's
    log probability. ``draws`` is the dict of ``sample_configurations``."""
    grouping = group_configurations(draws["configuration"])
    reps = grouping["first_draw"]
    logp = [float(draws["log_probability"][i]) for i in reps]
    out = project_particles(ctx, cache, [draws["configuration"][i] for i in reps], [first_id + i for i in reps],
                            grouping["multiplicity"], logp, logp,
                            [float(draws["max_environment_truncation"][i]) for i in reps], model, chi_measurement,
                            energy, bond_space=bond_space, pepo_live=pepo_live, seed=seed,
                            record_entropy=record_entropy, measurement_space=measurement_space)
    return out, grouping


def advance_particles(ctx, particles, model, steps, chi_environment, chi_measurement, method, energy, order=2,
                      on_step=None, D=None, record_entropy=False, environment_space=None, measurement_space=None):
    """``_advance_particle!`` (particles.jl:86-102) with ``_evolve!``
    (particles.jl:66-85) for a batch of particles at the same beta: one
    full-update step per endpoint in ``steps`` (delta = (beta - state.beta)/2),
    then ``_observe``; relative_log_weight += 2 sum(log_normalization) +
    (tensor_log_norm_new - tensor_log_norm_old). ``D`` is the bonds' truncrank
    (None: the resident bond capacity's dimension). With ``record_entropy``
    every particle's max_boundary_entropy is the maximum over the FU boundary
    cuts of these steps and ``Se`` holds the measurement's profiles; without
    recording, or without steps, max_boundary_entropy is NaN (``_evolve!``)."""
    if not particles:
        return particles
    devs = [p.state for p in particles]
    normalization = np.zeros(len(particles))
    maxstate = np.zeros(len(particles))
    maxenv = np.zeros(len(particles))
    maxentropy = [0.0 if record_entropy and steps else math.nan] * len(particles)
    total = 0.0
    for beta in steps:
        b0 = devs[0].state.beta
        if any(not same_beta(d.state.beta, b0) for d in devs):
            raise ValueError("particles of one batch must share beta")
        t0 = time.perf_counter()
        r = evolution_steps(ctx, devs, model, (beta - b0) / 2, chi_environment, method, order, D,
                            record_entropy=record_entropy, boundary_space=environment_space)
        elapsed = time.perf_counter() - t0
        for d in devs:
            d.state.beta = beta
        normalization += r["normalizations"].sum(axis=1)
        if r["records"].shape[1]:
            maxstate = np.maximum(maxstate, r["records"][:, :, 1].max(axis=1))
        maxenv = np.maximum(maxenv, r["max_environment_truncation"])
        maxentropy = [nan_max([a, float(b)]) for a, b in zip(maxentropy, r["max_boundary_entropy"])]
        total += elapsed
        if on_step is not None:
            on_step({"beta": beta, "time_seconds": elapsed, **r})
    m = observe(ctx, devs, model, chi_measurement, energy, record_entropy, measurement_space)
    for k, p in enumerate(particles):
        old = p.tensor_log_norm
        p.relative_log_weight += 2.0 * float(normalization[k]) + (float(m["tensor_log_norm"][k]) - old)
        p.physical_log_norm = float(m["log_norm"][k])
        p.tensor_log_norm = float(m["tensor_log_norm"][k])
        p.energy_total = float(m["energy_total"][k])
        p.energy_per_site = float(m["energy_per_site"][k])
        p.max_state_truncation = float(maxstate[k])
        p.max_environment_truncation = max(float(maxenv[k]), float(m["max_environment_truncation"][k]))
        p.max_boundary_entropy = maxentropy[k]
        p.Se = m["Se"][k]
        p.time_seconds = total / len(particles)
        p.measurement_seconds = m["measurement_seconds"]
    return particles


def aligned_weights(records):
    """``_aligned_weights`` (ensemble.jl:26-30): relative weights on a common scale."""
    scale = max(r["weight_log_scale"] for r in records)
    return [r["relative_log_weight"] + (r["weight_log_scale"] - scale) for r in records], scale


def statistics(records, Ns, nsites):
    """``_statistics`` (ensemble.jl:31-40) over particle records."""
    relative, scale = aligned_weights(records)
    s = ensemble_statistics(relative, [r["count"] for r in records], [r["energy_per_site"] for r in records], Ns,
                            log_scale=scale, nsites=nsites)
    s.update({"beta": records[0]["beta"],
              "max_state_truncation": max(r["max_state_truncation"] for r in records),
              "max_environment_truncation": max(r["max_environment_truncation"] for r in records),
              "max_boundary_entropy": nan_max(r["max_boundary_entropy"] for r in records),
              "distinct_configurations": len({np.asarray(r["configuration"]).tobytes() for r in records})})
    return s


# ----------------------------------------------------------------------------
# ensemble drivers (run.jl:2-43, 45-66, 87-130) on explicit arguments; the
# RunConfig entry points are in driver.py
# ----------------------------------------------------------------------------

def purify(ctx, pepo, model, beta0, beta_log, beta_measure, delta_tau, chi_environment, chi_measurement, bond_space,
           method=None, order=2, D=None, record_entropy=False, environment_space=None, measurement_space=None,
           on_step=None):
    """The evolution of ``run_purification`` (run.jl:2-43) from the identity PEPO
    (``identity_pepo``; normalised here as the reference does), on the device:
    FU steps to every node <= beta0 and an observation there. Returns the
    resident PEPO (``DeviceState``, bonds at the ``bond_space`` capacity with
    their live sectors) and the rows (beta, energy_total, energy_per_site,
    lnZ, truncations, max_boundary_entropy, Se, seconds). ``on_step`` is
    called after every full-update step with its beta, time and outputs (the
    progress the reference writes to its run log)."""
    method = method or FullUpdate()
    st, live = pad_bonds(pepo, bond_space)
    normalize_state(st)
    dev = DeviceState(ctx, st, live=live)
    rows = []
    for beta in [b for b in temperature_nodes(beta0, beta_log, beta_measure) if b <= beta0]:
        steps = step_schedule(dev.state.beta, beta, delta_tau)
        p = Particle(state=dev, tensor_log_norm=0.0, relative_log_weight=0.0)
        advance_particles(ctx, [p], model, steps, chi_environment, chi_measurement, method,
                          measure_at(beta_measure, beta), order, on_step=on_step, D=D,
                          record_entropy=record_entropy, environment_space=environment_space,
                          measurement_space=measurement_space)
        rows.append({"beta": beta, "source": "purification", "energy_total": p.energy_total,
                     "energy_per_site": p.energy_per_site,
                     "energy_per_site_error": 0.0 if measure_at(beta_measure, beta) else math.nan,
                     "lnZ": p.physical_log_norm, "lnZ_error": 0.0, "max_state_truncation": p.max_state_truncation,
                     "max_environment_truncation": p.max_environment_truncation,
                     "max_boundary_entropy": p.max_boundary_entropy, "Se": p.Se, "time_seconds": p.time_seconds,
                     "measurement_seconds": p.measurement_seconds})
    return dev, rows


def stochastic_ensemble(ctx, pepo, model, beta0, beta_log, beta_measure, delta_tau, Ns, seed, chi_sampling,
                        chi_environment, chi_measurement, bond_space=None, method=None, order=2, basis=None,
                        uniforms=None, on_node=None, D=None, draws_out=None, record_entropy=False,
                        environment_space=None, sampling_space=None, measurement_space=None, resampling=None,
                        chi_proposal=None, proposal_space=None, out=None):
    """The ensemble of ``run_stochastic`` (run.jl:87-130): Ns Born draws from
    the beta0 PEPO (``pepo``: a host ``FiniteState`` or a resident
    ``DeviceState``), one particle per distinct configuration, then FU
    evolution through every node >= beta0 with an observation and the
    ensemble statistics at each. Uniforms default to ``sample_uniforms(seed,
    i, Lx*Ly)`` for sample indices i = 1..Ns. With ``resampling``
    (``ResamplingOptions``, enabled) an annealed Born resampling event runs at
    every node whose eta falls below eta_th (run.jl:117-122, ``resample.py``;
    proposals at ``chi_proposal`` / ``proposal_space``, audit and replay rows
    into the result dictionary ``out``). Returns (particles, rows)."""
    method = method or FullUpdate()
    live = None
    if isinstance(pepo, DeviceState):
        # the cache reads the resident PEPO (its live sectors, advanced beta and log amplitude)
        live = pepo.live()
        host = pepo.state
    else:
        host = pepo
    if not same_beta(host.beta, beta0):
        raise ValueError("checkpoint beta0 mismatch")
    L = host.lattice
    ns = L.Lx * L.Ly
    cache = SamplingCache(ctx, pepo, chi_sampling, basis=resolve_basis(host, basis), sector_space=sampling_space)
    if uniforms is None:
        uniforms = np.stack([sample_uniforms(seed, i, ns) for i in range(1, Ns + 1)])
    draws = sample_configurations(cache, uniforms=uniforms)
    if draws_out is not None:
        draws_out.update(draws)
    nodes = [b for b in temperature_nodes(beta0, beta_log, beta_measure) if b >= beta0]
    particles, grouping = new_particles(ctx, cache, draws, model, chi_measurement,
                                        measure_at(beta_measure, nodes[0]), bond_space=bond_space,
                                        pepo_live=live, seed=seed, record_entropy=record_entropy,
                                        measurement_space=measurement_space)
    events = None
    if resampling is not None and resampling.enabled:
        from .resample import EventSettings
        events = EventSettings(ctx, cache, Ns, L, seed, resampling,
                               chi_sampling if chi_proposal is None else chi_proposal, proposal_space, model,
                               chi_environment, chi_measurement, method, order, D, delta_tau, beta_measure,
                               bond_space=bond_space, pepo_live=live, record_entropy=record_entropy,
                               environment_space=environment_space, measurement_space=measurement_space)
    rows = []
    replay = []
    event = 0
    next_sample_id = Ns + 1
    for node, beta in enumerate(nodes):
        steps = [] if node == 0 else step_schedule(nodes[node - 1], beta, delta_tau)
        replay.append({"beta": beta, "steps": list(steps)})
        if node > 0:
            advance_particles(ctx, particles, model, steps, chi_environment, chi_measurement, method,
                              measure_at(beta_measure, beta), order, D=D, record_entropy=record_entropy,
                              environment_space=environment_space, measurement_space=measurement_space)
        records = [p.record() for p in particles]
        stats = statistics(records, Ns, ns)
        stats["source"] = "sampling"
        stats["time_seconds"] = sum(p.time_seconds for p in particles)
        if events is not None and stats["eta"] < resampling.eta_th:
            from .resample import resample
            event += 1
            particles, records, stats, next_sample_id = resample(particles, records, events, replay, event,
                                                                 next_sample_id, out)
        rows.append(stats)
        if on_node is not None:
            on_node(stats, records)
    return particles, rows
