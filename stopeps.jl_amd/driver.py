"""``run_purification`` / ``run_stochastic`` on a ``RunConfig``
(src/driver/run.jl:1-135) with the result record of the reference
(src/io/run_output.jl:4-60, docs/result-schema.md, schema 2.7.0): the
checkpoint is a ``ThermalCheckpoint`` object whose PEPO stays resident on the
GPU, and ``run_stochastic`` returns the result dictionary (parameters, model,
metadata, observables {purification, curve, samples, replay}, diagnostics
{history, events, born_draws, born_grouping}). With ``output_directory`` set,
the checkpoint is written to ``results/checkpoint.npz`` and the result to
``results/result.npz`` (``results.py``: the reference's groups and atomic
publication in an npz container instead of JLD2). Annealed Born resampling
events run as the reference runs them (``resample.py``). MPI and the text
logs are not provided; rows and records are plain dicts where the reference
has NamedTuples.
"""
import os

import datetime
import time

from .config import (RESULT_SCHEMA_VERSION, bond_capacity, device_support, measure_at, parameter_dict,
                     sampling_basis, temperature_nodes, validate_config)
from .particles import purify, same_beta, stochastic_ensemble
from .sampling import Context, DeviceState
from .states import identity_pepo

from .results import CHECKPOINT_FORMAT, load_checkpoint, save_checkpoint, save_result


class ThermalCheckpoint:
    """``ThermalCheckpoint`` (run_output.jl:4-11): the beta0 PEPO (resident,
    ``DeviceState``), the preparation parameters, the purification rows, the
    history and the run id."""

    def __init__(self, state, preparation, purification, history, run_id):
        self.state = state
        self.preparation = preparation
        self.purification = purification
        self.history = history
        self.run_id = run_id


def _new_record(c, run_id, stage):
    """``_new_record`` (run_output.jl:51-60)."""
    params = parameter_dict(c)
    return {"schema_version": RESULT_SCHEMA_VERSION,
            "created_at": datetime.datetime.now(datetime.timezone.utc).isoformat(),
            "method": "stoPEPS", "parameters": params, "model": params["model"],
            "observables": {"curve": [], "purification": [], "samples": [], "replay": []},
            "diagnostics": {"events": [], "history": []},
            "metadata": {"run_id": run_id, "stage": stage, "complete": False, "mpi_size": 1,
                         "restart_supported": False, "progress": {},
                         "uncertainty": "conditional_particle_delta_method; shared_history_scale_fixed",
                         "device": "MI355X (HIP, gfx950)"}}


def _context(ctx):
    return ctx if ctx is not None else Context(0, "fp64")


def run_purification(c, ctx=None, on_step=None):
    """``run_purification(c)`` (run.jl:2-43): the identity PEPO, normalised,
    evolved by full update (D_pur, bond_space_pur, delta_tau_pur) to every node
    up to beta0 with an observation at each. Returns the ``ThermalCheckpoint``
    at beta0; its ``result`` attribute holds the preparation's result record.
    ``on_step(info)`` is called after every full-update step (beta, seconds
    and the step's outputs)."""
    validate_config(c)
    device_support(c, stochastic=False)
    ctx = _context(ctx)
    run_id = f"{datetime.datetime.now(datetime.timezone.utc).isoformat()}-{time.time_ns()}"
    out = _new_record(c, run_id, "purification")
    pepo = identity_pepo(c.lattice, symmetric=c.model.symmetric)
    cap = bond_capacity(c, c.D_pur, c.bond_space_pur)
    history = [{"phase": "initialization", "beta": 0.0}]
    dev, rows = purify(ctx, pepo, c.model, c.beta0, c.beta_log, c.beta_measure, c.delta_tau_pur, c.chi_environment,
                       c.chi_measurement, cap, c.update, c.trotter_order, D=c.D_pur,
                       record_entropy=c.record_boundary_entropy, environment_space=c.environment_space,
                       measurement_space=c.measurement_space, on_step=on_step)
    for r in rows:
        r.update({"eta": float("nan"), "eta_error": float("nan")})
        out["observables"]["purification"].append(r)
        out["observables"]["curve"].append(r)
    out["diagnostics"]["history"] = history
    out["metadata"]["stage_complete"] = True
    cp = ThermalCheckpoint(dev, out["parameters"], list(rows), history, run_id)
    cp.result = out
    if c.output_directory is not None:
        path = os.path.join(c.output_directory, "results", "checkpoint.npz")
        if os.path.lexists(path) and not c.overwrite_output:
            raise FileExistsError("checkpoint already exists; choose another directory or overwrite_output=true")
        save_checkpoint(path, cp, overwrite=c.overwrite_output)
    return cp


def run_stochastic(c, checkpoint=None, ctx=None, uniforms=None, checkpoint_path=None):
    """``run_stochastic(c; checkpoint, checkpoint_path)`` (run.jl:76-135):
    Ns Born draws from the checkpoint's beta0 PEPO (sample indices 1..Ns,
    uniforms from ``sample_uniforms(seed, i, Lx*Ly)`` unless given), one
    particle per distinct configuration, full update (D_peps,
    bond_space_peps, delta_tau_peps) through every node from beta0, an
    observation and the ensemble statistics at each. Without ``checkpoint``
    the checkpoint is loaded from ``checkpoint_path`` or the output
    directory's ``results/checkpoint.npz`` (run.jl:80-88). With
    ``c.resampling.enabled`` (the reference's default) a node whose eta falls
    below eta_th runs an annealed Born resampling event (``resample.py``),
    recorded under diagnostics.events and observables.replay. Returns the result
    dictionary, also saved to ``results/result.npz`` of the output directory."""
    validate_config(c)
    device_support(c, stochastic=True)
    ctx = _context(ctx)
    if checkpoint is None:
        if checkpoint_path is None:
            if c.output_directory is None:
                raise ValueError("supply a checkpoint or checkpoint_path")
            checkpoint_path = os.path.join(c.output_directory, "results", "checkpoint.npz")
        checkpoint = load_checkpoint(checkpoint_path, c, ctx)
    p = parameter_dict(c)
    pre = checkpoint.preparation
    if not (pre["model"] == p["model"] and pre["lattice"] == p["lattice"]):
        raise ValueError("checkpoint model, lattice or beta0 mismatch")
    state = checkpoint.state
    beta_cp = state.state.beta if isinstance(state, DeviceState) else state.beta
    if not same_beta(beta_cp, c.beta0):
        raise ValueError("checkpoint model, lattice or beta0 mismatch")
    out = _new_record(c, checkpoint.run_id, "stochastic")
    out["observables"]["purification"].extend(checkpoint.purification)
    out["observables"]["curve"].extend(r for r in checkpoint.purification
                                       if r["beta"] < c.beta0 and not same_beta(r["beta"], c.beta0))
    out["diagnostics"]["history"].extend(checkpoint.history)
    out["metadata"]["checkpoint_preparation"] = pre
    nodes = [b for b in temperature_nodes(c) if b >= c.beta0]
    cap = bond_capacity(c, c.D_peps, c.bond_space_peps)
    draws = {}

    def on_node(stats, records):
        out["observables"]["curve"].append(stats)
        out["observables"]["samples"].extend(records)
        out["metadata"]["progress"] = {"phase": "stochastic", "beta": stats["beta"],
                                       "event": len(out["diagnostics"]["events"])}

    particles, _ = stochastic_ensemble(ctx, state, c.model, c.beta0, c.beta_log, c.beta_measure, c.delta_tau_peps,
                                       c.Ns, c.seed, c.chi_sampling, c.chi_environment, c.chi_measurement,
                                       bond_space=cap, method=c.update, order=c.trotter_order,
                                       basis=sampling_basis(c), uniforms=uniforms, on_node=on_node, D=c.D_peps,
                                       draws_out=draws, record_entropy=c.record_boundary_entropy,
                                       environment_space=c.environment_space, sampling_space=c.sampling_space,
                                       measurement_space=c.measurement_space, resampling=c.resampling,
                                       chi_proposal=c.chi_proposal, proposal_space=c.proposal_space, out=out)
    out["diagnostics"]["born_draws"] = [
        {"index": i + 1, "configuration": draws["configuration"][i], "log_probability": float(draws["log_probability"][i]),
         "max_environment_truncation": float(draws["max_environment_truncation"][i])}
        for i in range(len(draws["log_probability"]))]
    out["metadata"]["complete"] = True
    out["metadata"]["stage_complete"] = True
    if c.output_directory is not None:
        save_result(os.path.join(c.output_directory, "results", "result.npz"), method="stoPEPS", model=out["model"],
                    parameters=out["parameters"], observables=out["observables"], diagnostics=out["diagnostics"],
                    metadata=out["metadata"], overwrite=True)
    out["particles"] = particles
    return out
