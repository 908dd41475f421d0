"""Full-update imaginary-time evolution on the device: the host-side mirror
of ``FullUpdate`` and ``evolution_step!`` (src/evolution/full_update.jl:11-17,
563-597) with the TFIM Trotter gates (src/models/tfim.jl:35-48).

The sweeps, environments, bond metrics and bond solves run in
``spp_fu_evolution_step``; this module builds the gates and marshals arrays.
"""
import ctypes

import numpy as np

from . import _lib
from .repair import RepairCallback
from .measurement import spin_operators
from .sampling import DeviceState
from .states import FiniteState, Site


class FullUpdate:
    """``FullUpdate(; max_iterations=4, error_tolerance=1e-8, diagonal_shift=1e-10,
    linear_max_iterations=24, linear_tolerance=1e-9)`` (full_update.jl:11-17)."""

    def __init__(self, max_iterations=4, error_tolerance=1e-8, diagonal_shift=1e-10, linear_max_iterations=24,
                 linear_tolerance=1e-9):
        self.max_iterations = int(max_iterations)
        self.error_tolerance = float(error_tolerance)
        self.diagonal_shift = float(diagonal_shift)
        self.linear_max_iterations = int(linear_max_iterations)
        self.linear_tolerance = float(linear_tolerance)

    def options(self):
        return np.array([self.max_iterations, self.error_tolerance, self.diagonal_shift,
                         self.linear_max_iterations, self.linear_tolerance], dtype=np.float64)


STOP_REASONS = ("disabled", "iteration_limit", "error_tolerance", "residual_fallback")

# per-sample outcomes of spp_fu_evolution_step (the reference's messages, full_update.jl:355-369, 491-501)
FAILURES = ("ok", "nonfinite FU norm metric", "zero FU norm metric", "gate produced a zero target",
            "negative FU residual needs PSD repair", "nonfinite FU residual/target", "singular FU gauge",
            "FU workspace exhausted", "FU eigen/SVD solve did not converge", "nonfinite row environment",
            "zero or nonfinite tensor after the full update", "PSD repair did not remove the FU residual failure")


class FullUpdateError(RuntimeError):
    """A full-update step failed for some samples; ``status`` holds every
    sample's outcome (0 ok). Failed states keep their sites from the failing
    bond on and are refused by later calls."""

    def __init__(self, status):
        self.status = np.asarray(status)
        bad = np.nonzero(self.status)[0]
        code = int(self.status[bad[0]])
        super().__init__(f"full update failed for sample(s) {list(map(int, bad))}: "
                         f"{FAILURES[code] if 0 <= code < len(FAILURES) else 'unknown failure'}")


def _degree(lattice, i):
    """``_degree`` (tfim.jl:43), 0-based coordinates."""
    x, y = i
    return (x > 0) + (x < lattice.Lx - 1) + (y > 0) + (y < lattice.Ly - 1)


def bond_hamiltonian(model, lattice, a, b):
    """``bond_hamiltonian`` (tfim.jl:44-47) as a 4x4 matrix over (p, q)."""
    X, ZZ = spin_operators(model)
    I = np.eye(2)
    return (-model.J * ZZ.reshape(4, 4) - model.h / _degree(lattice, a) * np.kron(X, I)
            - model.h / _degree(lattice, b) * np.kron(I, X))


def trotter_gate(model, lattice, a, b, delta):
    """``trotter_gate = exp(-delta bond_hamiltonian)`` (tfim.jl:48) as G[p,q,u,v]
    (symmetric matrix exponential through its eigendecomposition)."""
    w, V = np.linalg.eigh(bond_hamiltonian(model, lattice, a, b))
    return ((V * np.exp(-delta * w)) @ V.T).reshape(2, 2, 2, 2)


def onsite_gate(model, delta):
    """``exp(-delta onsite_hamiltonian)`` (tfim.jl:43, full_update.jl:568-570)
    with ``onsite_hamiltonian = -h X``, as G[p, u]."""
    X, _ = spin_operators(model)
    w, V = np.linalg.eigh(-model.h * X)
    return (V * np.exp(-delta * w)) @ V.T


def _gates(model, lattice, delta, ppar):
    perm = np.argsort(ppar, kind="stable")
    def sorted_gate(a, b):
        G = trotter_gate(model, lattice, a, b, delta)
        return np.ascontiguousarray(G[np.ix_(perm, perm, perm, perm)])
    Lx, Ly = lattice.Lx, lattice.Ly
    gh = [sorted_gate((x, y), (x + 1, y)) for y in range(Ly) for x in range(Lx - 1)]
    gv = [sorted_gate((x, y), (x, y + 1)) for y in range(Ly - 1) for x in range(Lx)]
    gh = np.ascontiguousarray(np.array(gh).reshape(-1)) if gh else np.zeros(1)
    gv = np.ascontiguousarray(np.array(gv).reshape(-1)) if gv else np.zeros(1)
    return gh, gv


def evolution_steps(ctx, states, model, delta, chi, method=None, order=2, D=None, record_entropy=False,
                    boundary_space=None, check=True, repair=True):
    """``evolution_step!(state, model, delta, D, FullUpdate(...), BoundaryMPS(chi);
    order, sector_space, record_entropy)`` for every uploaded state (``DeviceState``; same lattice
    and leg spaces up to the auxiliary charge), in place on the device.
    ``sector_space`` is the states' uploaded bond space (the per-parity
    capacity); ``D`` the truncrank of every bond (None: the capacity's
    dimension). Returns a dict: records [n, updates, 8]
    (initial_cost, final_cost, residual, iterations, iteration_limit_reached,
    stop, kept_even, kept_odd), normalizations [n, Lx*Ly],
    max_environment_truncation [n], max_boundary_entropy [n] (the largest
    pre-truncation Schmidt entropy over every boundary cut of the step; NaN
    unless ``record_entropy``). ``boundary_space``: the sector budget of the
    environment boundaries (``BoundaryMPS(chi, sector_space)``). ``status`` [n]
    is every sample's outcome (0 ok, else the index into ``FAILURES``); with
    ``check`` a failure raises :class:`FullUpdateError` after the step (the
    other samples have finished it). ``repair``: a bond whose residual check
    requests the PSD repair of its metric (full_update.jl:491-501) is repaired
    on the host (:mod:`repair`; a ``RepairCallback`` may be passed to observe or
    alter it) and its records are returned as ``repairs``; False: such a bond
    fails its sample."""
    method = method or FullUpdate()
    devs = list(states)
    n = len(devs)
    if n == 0:
        return {"records": np.zeros((0, 0, 8)), "normalizations": np.zeros((0, 0)),
                "max_environment_truncation": np.zeros(0), "max_boundary_entropy": np.zeros(0),
                "status": np.zeros(0, dtype=np.int32)}
    L = devs[0].state.lattice
    if any(not isinstance(d, DeviceState) for d in devs):
        raise TypeError("evolution_steps takes DeviceState objects (the states evolve in place on the device)")
    ppar = np.asarray(devs[0].state[(0, 0)].parity[2])
    if model.symmetric != bool(np.any(ppar)):
        raise ValueError("model physical space does not match the state")
    gate_delta = delta / 2 if order == 2 else delta
    gh, gv = _gates(model, L, gate_delta, ppar)
    if L.Lx * L.Ly == 1:  # one site: the onsite gate at the full delta
        perm = np.argsort(ppar, kind="stable")
        gh = np.ascontiguousarray(onsite_gate(model, delta)[np.ix_(perm, perm)])
    nh, nv = L.Ly * (L.Lx - 1), (L.Ly - 1) * L.Lx
    nup = (2 if order == 2 else 1) * (nh + nv)
    rec = np.zeros((n, nup, 8))
    norms = np.zeros((n, L.Lx * L.Ly))
    mt = np.zeros(n)
    ent = np.zeros(n) if record_entropy else None
    status = np.zeros(n, dtype=np.int32)
    handles = (ctypes.c_void_p * n)(*[d.handle.value for d in devs])
    o = method.options()
    cb = None
    if repair:
        cb = repair if isinstance(repair, RepairCallback) else RepairCallback()
        cb.install(ctx)
    try:
        with ctx.boundary_space(boundary_space):
            rc = ctx._lib.spp_fu_evolution_step(
                ctx.handle, n, handles, int(chi), int(order), float(delta), int(D or 0), _lib.dptr(o),
                _lib.dptr(gh), _lib.dptr(gv), _lib.dptr(rec), _lib.dptr(norms), _lib.dptr(mt), _lib.dptr(ent),
                _lib.iptr(status))
    finally:
        if cb is not None:
            RepairCallback.uninstall(ctx)
    if rc != 0 and cb is not None and cb.error is not None:
        raise RuntimeError(f"PSD repair callback failed: {cb.error!r}") from cb.error
    _lib.check(ctx.handle, rc)
    for d in devs:
        d.state.beta += 2 * delta
    out = {"records": rec, "normalizations": norms, "max_environment_truncation": mt,
           "max_boundary_entropy": ent if ent is not None else np.full(n, np.nan), "status": status,
           "repairs": list(cb.records) if cb is not None else []}
    if check and np.any(status):
        raise FullUpdateError(status)
    return out


def evolution_step(ctx, state, model, delta, chi, method=None, order=2, D=None, record_entropy=False,
                   boundary_space=None):
    """One state: the report of :func:`evolution_steps` for it."""
    r = evolution_steps(ctx, [state], model, delta, chi, method, order, D, record_entropy, boundary_space)
    recs = r["records"][0]
    return {"records": [{"initial_cost": float(v[0]), "final_cost": float(v[1]), "residual": float(v[2]),
                         "iterations": int(v[3]), "iteration_limit_reached": bool(v[4]),
                         "stop_reason": STOP_REASONS[int(v[5])], "kept": (int(v[6]), int(v[7]))} for v in recs],
            "normalizations": r["normalizations"][0], "max_environment_truncation": float(r["max_environment_truncation"][0]),
            "max_boundary_entropy": float(r["max_boundary_entropy"][0]),
            "max_state_truncation": float(np.max(recs[:, 1])) if len(recs) else 0.0}
