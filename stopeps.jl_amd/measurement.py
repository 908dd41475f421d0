"""Energy measurement of Born samples on the device: the host-side mirror of
``measure_energy`` (src/measurement/energy.jl:31-78) as ``_observe`` runs it on
each projected particle (src/driver/particles.jl:35-52), and of the TFIM
operators it reads (src/models/tfim.jl:1-34).

The contraction runs in ``spp_measure_energy_batch``; this module only turns
the model into its 2x2 operator factors and maps status codes to the
reference's errors.
"""
import numpy as np

from . import _lib

EPS = np.finfo(np.float64).eps


class TFIM:
    """``TFIM(; J, h, physical_space)`` (src/models/tfim.jl:9-19):
    H = -J sum ZZ - h sum X. ``symmetric=True`` is ``Z2Space(0=>1,1=>1)`` in the X
    eigenbasis, ``False`` is ``ComplexSpace(2)``."""

    def __init__(self, J=1.0, h=0.7, symmetric=False):
        self.J, self.h, self.symmetric = float(J), float(h), bool(symmetric)


def spin_operators(model):
    """``spin_operators(model)`` (tfim.jl:22-34): X [2,2] and ZZ [p,q,u,v]."""
    if model.symmetric:
        X = np.diag([1.0, -1.0])
        z = np.array([[0.0, 1.0], [1.0, 0.0]])
    else:
        X = np.array([[0.0, 1.0], [1.0, 0.0]])
        z = np.diag([1.0, -1.0])
    return X, np.einsum("pu,qv->pquv", z, z)


def operator_factors(ZZ, even=2):
    """``_operator_factors`` (energy.jl:8-12): split ZZ[p,q,u,v] through
    (p,u) | (q,v) by a truncated SVD with atol 64 eps |M|. ``even`` is the
    number of even physical indices (2: dense). Returns (F[p][u], G[q][v], odd)
    where ``odd`` is the parity of the one-dimensional factor leg."""
    M = np.asarray(ZZ, dtype=np.float64).transpose(0, 2, 1, 3).reshape(4, 4)
    atol = 64 * EPS * np.linalg.norm(M)
    par = np.array([(p >= even) ^ (u >= even) for p in range(2) for u in range(2)], dtype=int)
    terms = []
    for q in ((0, 1) if even < 2 else (0,)):
        rows = np.nonzero(par == q)[0] if even < 2 else np.arange(4)
        if np.any(M[np.ix_(rows, np.setdiff1d(np.arange(4), rows))]):
            raise ValueError("two-site operator is not parity-neutral")
        U, S, Vh = np.linalg.svd(M[np.ix_(rows, rows)])
        for k in range(len(S)):
            if S[k] > atol:
                f = np.zeros(4)
                g = np.zeros(4)
                f[rows] = U[:, k] * S[k]
                g[rows] = Vh[k]
                terms.append((f.reshape(2, 2), g.reshape(2, 2), int(q)))
    if len(terms) != 1:
        raise NotImplementedError("the device energy supports two-site operators of rank one (TFIM ZZ)")
    return terms[0]


class EnergyError(ValueError):
    """A numerical condition of ``measure_energy`` with the reference's message."""


def _raise_status(lattice, status, fail):
    bad = np.nonzero(status)[0]
    if len(bad) == 0:
        return
    i = int(bad[0])
    code = int(status[i])
    if code == 3:
        raise EnergyError("nonfinite row environment")
    if code == 4:
        raise EnergyError("zero local norm environment")
    if code == 5:
        raise EnergyError("nonreal observable")
    raise EnergyError("norm requires a finite positive real value")


def _device_operators(model, physical_parity):
    """X and the rank-one split F (x) G of the model's ZZ in the device's
    parity-sorted physical basis, and the factor leg's parity."""
    ppar = np.asarray(physical_parity)
    pe = int(np.sum(ppar == 0))
    X, ZZ = spin_operators(model)
    if model.symmetric != (pe < 2):
        raise ValueError("model physical space does not match the state")
    # the device holds the physical leg parity-sorted (even indices first)
    perm = np.argsort(ppar, kind="stable")
    X = X[np.ix_(perm, perm)]
    ZZ = ZZ[np.ix_(perm, perm, perm, perm)]
    F, G, odd = operator_factors(ZZ, pe)
    return np.ascontiguousarray(X), np.ascontiguousarray(F), np.ascontiguousarray(G), odd


def _entropy_buffers(lattice, n, record_entropy):
    if not record_entropy:
        return None, None
    return (np.full((n, 2, lattice.Ly - 1, lattice.Lx - 1), np.nan), np.full(n, np.nan))


def _entropy_outputs(out, prof, emax):
    """``Se`` (north, south [Ly-1, Lx-1] arrays indexed [row_cut-1, bond-1],
    NaN where undefined) and ``max_boundary_entropy`` per sample."""
    if prof is None:
        out["Se"] = [None] * len(out["energy_total"])
        out["max_boundary_entropy"] = np.full(len(out["energy_total"]), np.nan)
    else:
        out["Se"] = [{"north": prof[i, 0], "south": prof[i, 1]} for i in range(prof.shape[0])]
        out["max_boundary_entropy"] = emax


def measure_energies(cache, configurations, model, chi=None, check=True, record_entropy=False, sector_space=None):
    """Batched ``measure_energy(project_configuration(pepo, c), model,
    BoundaryMPS(chi); record_entropy)`` for configurations [n, Lx, Ly] of basis
    labels. ``chi`` defaults to the cache's (sampling) chi. Returns a dict of
    arrays: energy_total, energy_per_site, log_norm, tensor_log_norm,
    max_environment_truncation, status, fail_site, and per sample ``Se``
    ({"north", "south"} pre-truncation Schmidt entropies [Ly-1, Lx-1], or None)
    and max_boundary_entropy (the maximum over every boundary cut; NaN when
    not recorded). ``sector_space``: the boundary sector budget of the
    measurement method (``BoundaryMPS(chi, sector_space)``)."""
    L = cache.lattice
    ns = L.Lx * L.Ly
    cfg = np.asarray(configurations)
    if cfg.ndim != 3 or cfg.shape[1:] != (L.Lx, L.Ly):
        raise ValueError("configurations must have shape [n, Lx, Ly]")
    n = cfg.shape[0]
    idx = np.vectorize(cache.basis.index)(cfg) if n else np.zeros((0, L.Lx, L.Ly), dtype=int)
    labels = np.ascontiguousarray(np.transpose(idx, (0, 2, 1)).reshape(n, ns), dtype=np.int32)
    X, F, G, odd = _device_operators(model, cache.state[(0, 0)].parity[2])
    out = {k: np.full(n, np.nan) for k in ("energy_total", "energy_per_site", "log_norm", "tensor_log_norm",
                                            "max_environment_truncation")}
    status = np.zeros(n, dtype=np.int32)
    fail = np.full(n, -1, dtype=np.int32)
    prof, emax = _entropy_buffers(L, n, record_entropy)
    ctx = cache.ctx
    with ctx.boundary_space(sector_space):
        rc = ctx._lib.spp_measure_energy_batch(
            ctx.handle, cache.handle, cache._dev.handle, int(chi if chi is not None else cache.chi), n,
            _lib.iptr(labels), _lib.dptr(X), _lib.dptr(F), _lib.dptr(G), odd, model.J, model.h,
            _lib.dptr(out["energy_total"]), _lib.dptr(out["energy_per_site"]), _lib.dptr(out["log_norm"]),
            _lib.dptr(out["tensor_log_norm"]), _lib.dptr(out["max_environment_truncation"]), _lib.iptr(status),
            _lib.iptr(fail), _lib.dptr(prof), _lib.dptr(emax))
    _lib.check(ctx.handle, rc)
    if check:
        _raise_status(L, status, fail)
    out["status"] = status
    out["fail_site"] = fail
    _entropy_outputs(out, prof, emax)
    return out


def measure_energy(cache, configuration, model, chi=None, record_entropy=False, sector_space=None):
    """One sample: the NamedTuple fields of ``measure_energy`` (energy.jl:73-77)
    as a dict (``Se`` is None unless ``record_entropy``)."""
    r = measure_energies(cache, np.asarray(configuration)[None], model, chi, record_entropy=record_entropy,
                         sector_space=sector_space)
    return {"energy_total": float(r["energy_total"][0]), "energy_per_site": float(r["energy_per_site"][0]),
            "log_norm": float(r["log_norm"][0]), "tensor_log_norm": float(r["tensor_log_norm"][0]),
            "max_environment_truncation": float(r["max_environment_truncation"][0]), "Se": r["Se"][0],
            "max_boundary_entropy": float(r["max_boundary_entropy"][0])}


def measure_energy_states(ctx, states, model, chi, check=True, record_entropy=False, sector_space=None):
    """Batched ``measure_energy(state, model, BoundaryMPS(chi); record_entropy)``
    (energy.jl:31-78) of resident states (``DeviceState``, or host states that
    are uploaded for the call): PEPS or PEPO on one lattice with the same leg
    spaces, e.g. the particles after ``_evolve!`` (particles.jl:86-95). Returns
    the dict of ``measure_energies``; log_norm includes each state's
    2 log_amplitude."""
    from .sampling import DeviceState
    import ctypes
    devs = [s if isinstance(s, DeviceState) else DeviceState(ctx, s) for s in states]
    n = len(devs)
    out = {k: np.full(n, np.nan) for k in ("energy_total", "energy_per_site", "log_norm", "tensor_log_norm",
                                            "max_environment_truncation")}
    status = np.zeros(n, dtype=np.int32)
    fail = np.full(n, -1, dtype=np.int32)
    if n == 0:
        out["status"], out["fail_site"] = status, fail
        _entropy_outputs(out, None, None)
        return out
    L = devs[0].state.lattice
    prof, emax = _entropy_buffers(L, n, record_entropy)
    X, F, G, odd = _device_operators(model, devs[0].state[(0, 0)].parity[2])
    handles = (ctypes.c_void_p * n)(*[d.handle.value for d in devs])
    with ctx.boundary_space(sector_space):
        rc = ctx._lib.spp_measure_energy_states(
            ctx.handle, n, handles, int(chi), _lib.dptr(X), _lib.dptr(F), _lib.dptr(G), odd, model.J, model.h,
            _lib.dptr(out["energy_total"]), _lib.dptr(out["energy_per_site"]), _lib.dptr(out["log_norm"]),
            _lib.dptr(out["tensor_log_norm"]), _lib.dptr(out["max_environment_truncation"]), _lib.iptr(status),
            _lib.iptr(fail), _lib.dptr(prof), _lib.dptr(emax))
    _lib.check(ctx.handle, rc)
    if check:
        _raise_status(L, status, fail)
    out["status"] = status
    out["fail_site"] = fail
    _entropy_outputs(out, prof, emax)
    return out
