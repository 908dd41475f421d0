"""Norms of finite states on the device: the host-side mirror of
``lognorm(::FiniteState, ::AbstractRowBoundaryMPS)`` (src/contraction/boundary_mps.jl:162-191).
The contraction runs in ``spp_lognorm_batch``."""
import ctypes

import numpy as np

from . import _lib
from .sampling import DeviceState


class NormError(ValueError):
    """``_positive_real`` failure (boundary_mps.jl:187-190)."""


def lognorms(ctx, states, chi, check=True):
    """Batched ``lognorm(state, BoundaryMPS(chi))`` of states on one lattice
    with the same leg spaces (PEPS or PEPO). Returns a dict of arrays:
    log_norm, tensor_log_norm, max_environment_truncation, status."""
    devs = [s if isinstance(s, DeviceState) else DeviceState(ctx, s) for s in states]
    n = len(devs)
    handles = (ctypes.c_void_p * max(n, 1))(*[d.handle.value for d in devs])
    out = {k: np.full(n, np.nan) for k in ("log_norm", "tensor_log_norm", "max_environment_truncation")}
    status = np.zeros(n, dtype=np.int32)
    _lib.check(ctx.handle, ctx._lib.spp_lognorm_batch(
        ctx.handle, n, handles, int(chi), _lib.dptr(out["log_norm"]), _lib.dptr(out["tensor_log_norm"]),
        _lib.dptr(out["max_environment_truncation"]), _lib.iptr(status)))
    if check and np.any(status):
        raise NormError("norm requires a finite positive real value")
    out["status"] = status
    return out


def lognorm(ctx, state, chi):
    """``lognorm(state, BoundaryMPS(chi))``: the NamedTuple fields as a dict."""
    r = lognorms(ctx, [state], chi)
    return {"log_norm": float(r["log_norm"][0]), "tensor_log_norm": float(r["tensor_log_norm"][0]),
            "max_environment_truncation": float(r["max_environment_truncation"][0])}
