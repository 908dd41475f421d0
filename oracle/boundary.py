"""Row boundary-MPS contraction (oracle restatement of
src/contraction/boundary_mps.jl, literal algorithm: absorb, LQ sweep, SVD sweep).

Boundary tensors have legs (l, p, r); row layers have legs (W, N, E, S).
"""
import numpy as np

from .z2 import Z2, ein, fuse, right_orth, left_orth, tsvd, permute, split


def act(A, O, axis):
    """``_act`` (boundary_mps.jl:59-62): A'[.., i, ..] = sum_j A[.., j, ..] O[i, j]."""
    letters = "abcdefgh"[:A.ndim]
    out = letters[:axis] + "z" + letters[axis + 1:]
    return ein(f"{letters},z{letters[axis]}->{out}", A, O)


def double_layer(ket, bra=None, physical=None, auxiliary=None):
    """``double_layer`` (boundary_mps.jl:65-73): legs (W,N,E,S), pairs fused (ket,bra)."""
    if bra is None:
        bra = ket
    A = ket if physical is None else act(ket, physical, 2)
    A = A if auxiliary is None else act(A, auxiliary, 3)
    raw = ein("wnpaes,WNpaES->wWnNeEsS", A, bra.conj())
    for axis in (6, 4, 2, 0):
        raw = fuse(raw, axis, 2)
    return raw


def double_layers(state):
    return [[double_layer(state[(x, y)]) for y in range(state.lattice.Ly)]
            for x in range(state.lattice.Lx)]


def row(layers, y):
    return [layers[x][y] for x in range(len(layers))]


class Boundary:
    """``BoundaryEnvironment`` (boundary_mps.jl:79-83)."""

    def __init__(self, tensors, log_scale=0.0, max_truncation=0.0):
        self.tensors = tensors
        self.log_scale = log_scale
        self.max_truncation = max_truncation


def trivial_boundary(rowlayers, side):
    """``_trivial_boundary`` (boundary_mps.jl:86-91)."""
    axis = 1 if side == "north" else 3
    out = []
    for E in rowlayers:
        if E.shape[axis] != 1:
            raise ValueError("nontrivial OBC leg")
        out.append(Z2(np.ones((1, 1, 1)), [np.array([0]), E.p[axis], np.array([0])]))
    return Boundary(out, 0.0, 0.0)


def schmidt_entropy(spectra):
    """``_schmidt_entropy`` (boundary_mps.jl:21-30): nats, pre-truncation spectrum."""
    s = np.concatenate([np.asarray(x) for x in spectra]) if spectra else np.zeros(0)
    scale = np.linalg.norm(s)
    if scale == 0:
        return float("nan")
    p = (s / scale) ** 2
    p = p[p > 0]
    return max(0.0, float(-np.sum(p * np.log(p))))


def record_entropy(entropy, side, row_cut, bond, value):
    """``_record_entropy!`` (boundary_mps.jl:50-56): ``entropy`` is a dict with
    "maximum" and, when a profile is kept, "north"/"south" arrays
    [Ly-1, Lx-1]. Julia's ``max`` propagates NaN."""
    m = entropy["maximum"]
    entropy["maximum"] = float("nan") if (np.isnan(m) or np.isnan(value)) else max(m, value)
    prof = entropy.get(side)
    if prof is not None and 1 <= row_cut <= prof.shape[0]:
        prof[row_cut - 1, bond - 1] = value


class BoundaryMPS:
    """``BoundaryMPS(chi, sector_space)`` (boundary_mps.jl:11-18): every
    compression keeps ``truncrank(chi) & truncspace(sector_space)``; the
    sector space is an (even, odd) budget, or None. Every function of this
    oracle that takes ``chi`` also takes a ``BoundaryMPS``."""

    def __init__(self, chi, sector_space=None):
        self.chi = int(chi)
        self.caps = None if sector_space is None else tuple(int(c) for c in sector_space)


def _method(chi):
    return chi if isinstance(chi, BoundaryMPS) else BoundaryMPS(chi)


def compress_boundary(B, chi, reverse=False, observer=None):
    """``compress_boundary!`` (boundary_mps.jl:94-134). Modifies the list ``B``.
    ``chi``: the bond, or a ``BoundaryMPS`` with a sector budget."""
    method = _method(chi)
    chi, caps = method.chi, method.caps
    if chi < 1:
        raise ValueError("chi must be positive")
    N = len(B)
    max_error = 0.0
    if not reverse:
        for x in range(N - 1, 0, -1):
            L, Q = right_orth(B[x], 1)
            B[x] = Q
            B[x - 1] = ein("lpk,kr->lpr", B[x - 1], L)
        for x in range(N - 1):
            U, S, V, bp, err, spectra = tsvd(B[x], 2, chi=chi, caps=caps)
            nM = B[x].norm()
            if observer is not None:
                observer(x + 1, schmidt_entropy(spectra))
            max_error = max(max_error, 0.0 if nM == 0 else (err / nM) ** 2)
            B[x] = U
            R = Z2(S[:, None] * V.a.reshape(V.shape[0], -1), [bp, V.p[1]])
            B[x + 1] = ein("lk,kpr->lpr", R, B[x + 1])
    else:
        for x in range(N - 1):
            Q, R = left_orth(B[x], 2)
            B[x] = Q
            B[x + 1] = ein("lk,kpr->lpr", R, B[x + 1])
        for x in range(N - 1, 0, -1):
            U, S, V, bp, err, spectra = tsvd(B[x], 1, chi=chi, caps=caps)
            nM = B[x].norm()
            if observer is not None:
                observer(x, schmidt_entropy(spectra))
            max_error = max(max_error, 0.0 if nM == 0 else (err / nM) ** 2)
            B[x] = V
            L = Z2(U.a * S[None, :], [U.p[0], bp])
            B[x - 1] = ein("lpk,kr->lpr", B[x - 1], L)
    center = 0 if reverse else N - 1
    scale = B[center].norm()
    if not np.isfinite(scale):
        raise ValueError("nonfinite boundary norm")
    if scale != 0:
        B[center] = B[center].scale(1.0 / scale)
    return max_error, (-np.inf if scale == 0 else float(np.log(scale)))


def absorb(env, rowlayers, chi, side="north", observer=None):
    """``_absorb`` (boundary_mps.jl:136-151)."""
    B = []
    for A, E in zip(env.tensors, rowlayers):
        if side == "north":
            C = ein("lnr,wnep->lwpre", A, E)
        else:
            C = ein("lsr,wpes->lwpre", A, E)
        B.append(fuse(fuse(C, 3, 2), 0, 2))
    max_trunc, log_scale = compress_boundary(B, chi, reverse=(side == "south"), observer=observer)
    return Boundary(B, env.log_scale + log_scale, max(env.max_truncation, max_trunc))


def close_boundary(B):
    """``_close_boundary`` (boundary_mps.jl:153-163)."""
    V = np.ones(1)
    for A in B:
        if A.shape[0] != V.shape[0] or A.shape[1] != 1:
            raise ValueError("closing requires unit legs")
        V = V @ A.a[:, 0, :]
    if V.shape[0] != 1:
        raise ValueError("closing requires a unit right leg")
    return float(V[0])


def contract_layers(layers, chi):
    """``contract_layers`` (boundary_mps.jl:165-172)."""
    Ly = len(layers[0])
    env = trivial_boundary(row(layers, 0), "north")
    for y in range(Ly):
        env = absorb(env, row(layers, y), chi)
    return close_boundary(env.tensors), env.log_scale, env.max_truncation


def positive_real(z, name):
    """``_positive_real`` (boundary_mps.jl:187-191)."""
    if not (np.isfinite(z) and np.real(z) > 0 and abs(np.imag(z)) <= 1e-9 * abs(z)):
        raise ValueError(f"{name} requires a finite positive real value; received {z}")
    return float(np.real(z))


def lognorm(state, chi):
    """``lognorm(::FiniteState, ::BoundaryMPS)`` (boundary_mps.jl:180-186)."""
    m, ls, mt = contract_layers(double_layers(state), chi)
    v = positive_real(m, "norm")
    return {"log_norm": float(np.log(v) + ls + 2 * state.log_amplitude),
            "tensor_log_norm": float(np.log(v) + ls), "max_environment_truncation": mt}


def boundary_environments(layers, chi, side="both", entropy=None):
    """``boundary_environments`` (boundary_mps.jl:193-209). ``entropy`` is an
    optional dict with "north"/"south" arrays [Ly-1, Lx-1] and "maximum"."""
    Lx, Ly = len(layers), len(layers[0])
    north = [None] * (Ly + 1)
    south = [None] * (Ly + 1)
    north[0] = trivial_boundary(row(layers, 0), "north")
    south[Ly] = trivial_boundary(row(layers, Ly - 1), "south")

    def observer_for(s, row_cut):
        if entropy is None:
            return None
        return lambda bond, value: record_entropy(entropy, s, row_cut, bond, value)

    if side != "south":
        for y in range(Ly):
            north[y + 1] = absorb(north[y], row(layers, y), chi, "north", observer_for("north", y + 1))
    if side != "north":
        for y in range(Ly - 1, -1, -1):
            south[y] = absorb(south[y + 1], row(layers, y), chi, "south", observer_for("south", y))
    mt = max(0.0 if side == "south" else north[Ly].max_truncation,
             0.0 if side == "north" else south[0].max_truncation)
    return north, south, mt


class Slab:
    """``SlabColumn`` (boundary_mps.jl:212-215)."""

    def __init__(self, upper, lower):
        self.upper, self.lower = upper, lower


def left_step(L, T, E, B):
    """``_left_step`` (boundary_mps.jl:216-223, 233-239)."""
    if isinstance(E, Slab):
        X = ein("awvb,anc->wvbnc", L, T)
        Y = ein("wvbnc,wnek->vbcek", X, E.upper)
        Z = ein("vbcek,vkfs->bcefs", Y, E.lower)
        return ein("bcefs,bsd->cefd", Z, B)
    X = ein("awb,anc->wbnc", L, T)
    Y = ein("wbnc,wnes->bces", X, E)
    return ein("bces,bsd->ced", Y, B)


def right_step(T, E, B, R):
    """``_right_step`` (boundary_mps.jl:224-231, 240-246)."""
    if isinstance(E, Slab):
        X = ein("anc,cefd->anefd", T, R)
        Y = ein("anefd,wnek->awkfd", X, E.upper)
        Z = ein("awkfd,vkfs->awvsd", Y, E.lower)
        return ein("awvsd,bsd->awvb", Z, B)
    X = ein("anc,ced->aned", T, R)
    Y = ein("aned,wnes->awsd", X, E)
    return ein("awsd,bsd->awb", Y, B)


def row_edge(top, rowl, bottom, side):
    """``_row_edge`` (boundary_mps.jl:247-256): all-ones tensor over unit legs."""
    i = 0 if side == "left" else len(rowl) - 1
    a = 0 if side == "left" else 2
    if isinstance(rowl[i], Slab):
        legs = [top[i].p[a], rowl[i].upper.p[a], rowl[i].lower.p[a], bottom[i].p[a]]
    else:
        legs = [top[i].p[a], rowl[i].p[a], bottom[i].p[a]]
    for p in legs:
        if len(p) != 1:
            raise ValueError("nontrivial OBC leg")
    return Z2(np.ones([1] * len(legs)), legs)


def safe_scale(t):
    """``_safe_scale`` (boundary_mps.jl:257-260)."""
    s = t.norm()
    if not np.isfinite(s):
        raise ValueError("nonfinite row environment")
    return t if s == 0 else t.scale(1.0 / s)


def row_environments(top, rowl, bottom, side="both"):
    """``row_environments`` (boundary_mps.jl:261-273)."""
    N = len(rowl)
    left = [None] * (N + 1)
    right = [None] * (N + 1)
    left[0] = row_edge(top, rowl, bottom, "left")
    right[N] = row_edge(top, rowl, bottom, "right")
    if side != "right":
        for x in range(N):
            left[x + 1] = safe_scale(left_step(left[x], top[x], rowl[x], bottom[x]))
    if side != "left":
        for x in range(N - 1, -1, -1):
            right[x] = safe_scale(right_step(top[x], rowl[x], bottom[x], right[x + 1]))
    return left, right


def join_scalars(L, R):
    """``_join_scalars`` (boundary_mps.jl:277-284)."""
    _ = ein("abc,abc->", L, R) if L.ndim == 3 else ein("abcd,abcd->", L, R)
    return complex(_.a) if np.iscomplexobj(_.a) else float(_.a)


def row_scalar(L, T, E, B, R):
    """``_row_scalar`` (boundary_mps.jl:274-276)."""
    return join_scalars(left_step(L, T, E, B), R)
