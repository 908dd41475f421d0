"""Norms of finite states on the device: the host-side mirror of
``lognorm(::FiniteState, ::AbstractRowBoundaryMPS)`` (src/contraction/boundary_mps.jl:162-191).
The contraction runs in ``spp_lognorm_batch``."""
import ctypes

import numpy as np

from . import _lib
from .sampling import DeviceState


class NormError(ValueError):
    """``_positive_real`` failure (boundary_mps.jl:187-190)."""


def lognorms(ctx, states, chi, check=True, sector_space=None):
    """Batched ``lognorm(state, BoundaryMPS(chi, sector_space))`` of states on
    one lattice with the same leg spaces (PEPS or PEPO). Returns a dict of
    arrays: log_norm, tensor_log_norm, max_environment_truncation, status."""
    devs = [s if isinstance(s, DeviceState) else DeviceState(ctx, s) for s in states]
    n = len(devs)
    handles = (ctypes.c_void_p * max(n, 1))(*[d.handle.value for d in devs])
    out = {k: np.full(n, np.nan) for k in ("log_norm", "tensor_log_norm", "max_environment_truncation")}
    status = np.zeros(n, dtype=np.int32)
    with ctx.boundary_space(sector_space):
        rc = ctx._lib.spp_lognorm_batch(
            ctx.handle, n, handles, int(chi), _lib.dptr(out["log_norm"]), _lib.dptr(out["tensor_log_norm"]),
            _lib.dptr(out["max_environment_truncation"]), _lib.iptr(status))
    _lib.check(ctx.handle, rc)
    if check and np.any(status):
        raise NormError("norm requires a finite positive real value")
    out["status"] = status
    return out


def lognorm(ctx, state, chi):
    """``lognorm(state, BoundaryMPS(chi))``: the NamedTuple fields as a dict."""
    r = lognorms(ctx, [state], chi)
    return {"log_norm": float(r["log_norm"][0]), "tensor_log_norm": float(r["tensor_log_norm"][0]),
            "max_environment_truncation": float(r["max_environment_truncation"][0])}


def _unsort(a, parity):
    """Undo ``parity_sorted`` on every leg of a dense site array."""
    for k, p in enumerate(parity):
        order = np.argsort(p, kind="stable")
        a = np.take(a, np.argsort(order), axis=k)
    return a


def lognorm_gradients(ctx, states, chi, check=True):
    """Gradient of ``lognorm(state, BoundaryMPS(chi))`` with respect to every
    site tensor, for a batch of states on one lattice with the same leg
    spaces. There is no reference implementation; the gradient is assembled
    from the single-site norm matrices of ``_site_hole`` (src/sampling/born.jl:75-83)
    as (N A + N^T A) / (A N A), i.e. the environment gradient with the
    boundary truncations held fixed (exact when chi reaches the exact bond).
    Returns (grads, info): grads[i][(x, y)] is an array shaped like the site's
    data in its own leg order; info holds log_norm, tensor_log_norm,
    max_environment_truncation and status arrays."""
    devs = [s if isinstance(s, DeviceState) else DeviceState(ctx, s) for s in states]
    n = len(devs)
    info = {k: np.full(n, np.nan) for k in ("log_norm", "tensor_log_norm", "max_environment_truncation")}
    status = np.zeros(n, dtype=np.int32)
    if n == 0:
        info["status"] = status
        return [], info
    L = devs[0].state.lattice
    sites = L.site_order()
    sizes = [devs[0].state[xy].data.size for xy in sites]
    total = int(np.sum(sizes))
    flat = np.zeros((n, total))
    handles = (ctypes.c_void_p * n)(*[d.handle.value for d in devs])
    _lib.check(ctx.handle, ctx._lib.spp_lognorm_grad_batch(
        ctx.handle, n, handles, int(chi), _lib.dptr(flat), _lib.dptr(info["log_norm"]),
        _lib.dptr(info["tensor_log_norm"]), _lib.dptr(info["max_environment_truncation"]), _lib.iptr(status)))
    if check and np.any(status):
        raise NormError("norm requires a finite positive real value")
    info["status"] = status
    offs = np.concatenate([[0], np.cumsum(sizes)])
    grads = []
    for i, d in enumerate(devs):
        g = {}
        for j, xy in enumerate(sites):
            site = d.state[xy]
            a = flat[i, offs[j]:offs[j + 1]].reshape(site.data.shape)
            g[xy] = _unsort(a, site.parity)
        grads.append(g)
    return grads, info


def lognorm_gradient(ctx, state, chi):
    """One state: (grads dict (x, y) -> array, log_norm)."""
    grads, info = lognorm_gradients(ctx, [state], chi)
    return grads[0], float(info["log_norm"][0])


def lognorm_rrule(ctx, state, chi):
    """Reverse rule of ``lognorm(state, BoundaryMPS(chi))`` in the ChainRules
    form: returns (log_norm, pullback) with pullback(dy) -> {(x, y): dy * grad}
    the cotangent of every site tensor. One forward-and-backward sweep
    (``spp_lognorm_grad_batch``) serves both; the gradient holds the boundary
    truncations fixed, so it is exact once chi reaches the exact boundary bond."""
    grads, log_norm = lognorm_gradient(ctx, state, chi)

    def pullback(dy):
        return {xy: float(dy) * g for xy, g in grads.items()}
    return log_norm, pullback
