"""Energy measurement with row boundary MPS (oracle restatement of
src/measurement/energy.jl)."""
import numpy as np

from .z2 import Z2, ein, fuse, tsvd, EPS
from . import boundary as bm


def real_ratio(a, b):
    """``_real_ratio`` (energy.jl:1-6)."""
    if b == 0:
        raise ValueError("zero local norm environment")
    z = a / b
    if not (np.isfinite(z) and abs(np.imag(z)) <= 1e-9 * max(1.0, abs(z))):
        raise ValueError(f"nonreal observable: {z}")
    return float(np.real(z))


def operator_factors(ZZ, pp):
    """``_operator_factors`` (energy.jl:8-12): ZZ[p,q,u,v] split through (p,u)|(q,v)
    by a truncated SVD with atol 64 eps |M|. Returns F[p,u,r], G[r,q,v]."""
    op = Z2(ZZ, [pp, pp, pp, pp])
    M = ein("pquv->puqv", op)
    Mn = M.norm()
    U, S, V, bp, err, _ = tsvd(M, 2, atol=64 * EPS * Mn)
    F = Z2(U.a * S[None, None, :], U.p)
    return F, V


def operator_pair(A, B, factors, direction):
    """``_operator_pair`` (energy.jl:13-23)."""
    F, G = factors
    if direction == "horizontal":
        C = ein("wnuaes,pur->wnpaers", A, F)
        D = ein("wnvaes,rqv->wrnqaes", B, G)
        return fuse(C, 4, 2), fuse(D, 0, 2)
    C = ein("wnuaes,pur->wnpaesr", A, F)
    D = ein("wnvaes,rqv->wnrqaes", B, G)
    return fuse(C, 5, 2), fuse(D, 1, 2)


def measure_energy(state, model, chi, record_entropy=False):
    """``measure_energy(state, model, ::AbstractRowBoundaryMPS)`` (energy.jl:31-78)."""
    Lx, Ly = state.lattice.Lx, state.lattice.Ly
    layers = bm.double_layers(state)
    entropy = None
    if record_entropy:
        entropy = {"maximum": 0.0, "north": np.full((Ly - 1, Lx - 1), np.nan),
                   "south": np.full((Ly - 1, Lx - 1), np.nan)}
    north, south, mt = bm.boundary_environments(layers, chi, entropy=entropy)
    X, ZZ = model.operators()
    pp = state[(0, 0)].p[2]
    Xt = Z2(X, [pp, pp])
    Xl = [[bm.double_layer(state[(x, y)], physical=Xt) for y in range(Ly)] for x in range(Lx)]
    factors = operator_factors(ZZ, pp)
    xsum = 0.0
    zzsum = 0.0
    for y in range(Ly):
        rowl = bm.row(layers, y)
        top = north[y].tensors
        bottom = south[y + 1].tensors
        left, right = bm.row_environments(top, rowl, bottom)
        for x in range(Lx):
            den = bm.row_scalar(left[x], top[x], rowl[x], bottom[x], right[x + 1])
            num = bm.row_scalar(left[x], top[x], Xl[x][y], bottom[x], right[x + 1])
            xsum += real_ratio(num, den)
        for x in range(Lx - 1):
            A, B = state[(x, y)], state[(x + 1, y)]
            C, D = operator_pair(A, B, factors, "horizontal")
            lft = bm.left_step(left[x], top[x], bm.double_layer(C, A), bottom[x])
            num = bm.row_scalar(lft, top[x + 1], bm.double_layer(D, B), bottom[x + 1], right[x + 2])
            lft = bm.left_step(left[x], top[x], rowl[x], bottom[x])
            den = bm.row_scalar(lft, top[x + 1], rowl[x + 1], bottom[x + 1], right[x + 2])
            zzsum += real_ratio(num, den)
    for y in range(Ly - 1):
        rowl = [bm.Slab(layers[x][y], layers[x][y + 1]) for x in range(Lx)]
        top = north[y].tensors
        bottom = south[y + 2].tensors
        left, right = bm.row_environments(top, rowl, bottom)
        for x in range(Lx):
            A, B = state[(x, y)], state[(x, y + 1)]
            C, D = operator_pair(A, B, factors, "vertical")
            imp = bm.Slab(bm.double_layer(C, A), bm.double_layer(D, B))
            den = bm.row_scalar(left[x], top[x], rowl[x], bottom[x], right[x + 1])
            num = bm.row_scalar(left[x], top[x], imp, bottom[x], right[x + 1])
            zzsum += real_ratio(num, den)
    final = north[Ly]
    log_norm = np.log(bm.positive_real(bm.close_boundary(final.tensors), "norm")) + final.log_scale
    energy_total = -model.J * zzsum - model.h * xsum
    return {"energy_total": energy_total, "energy_per_site": energy_total / (Lx * Ly),
            "log_norm": float(log_norm + 2 * state.log_amplitude), "tensor_log_norm": float(log_norm),
            "max_environment_truncation": mt,
            "Se": None if entropy is None else {"north": entropy["north"], "south": entropy["south"]},
            "max_boundary_entropy": None if entropy is None else entropy["maximum"]}
