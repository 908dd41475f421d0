"""Perfect Born sampling of PEPO auxiliary indices (oracle restatement of
src/sampling/born.jl, src/sampling/basis.jl, src/resample/selection.jl:1-7,68-80).

The reference draws one ``rand(rng)`` per site from a Julia MersenneTwister
(born.jl:13-22); that stream cannot be produced without Julia, so the uniforms
are an input here (one per site, row-major with x fastest).
"""
import numpy as np

from .z2 import Z2, ein
from .model import State, normalize_state
from . import boundary as bm

MASK64 = (1 << 64) - 1


def splitmix64(v):
    """``_splitmix64`` (born.jl:1-6)."""
    v = (v + 0x9E3779B97F4A7C15) & MASK64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK64
    return v ^ (v >> 31)


def sample_rng_words(seed, index):
    """UInt32 seed words of ``sample_rng`` (born.jl:13-22), scheme version 2."""
    if seed < 0 or index < 1:
        raise ValueError("invalid random stream identity")
    words = [0x424F524E, 2]
    for v in (seed, index):
        m = splitmix64(v & MASK64)
        words += [m & 0xFFFFFFFF, m >> 32]
    return words


def counter_uniforms(seed, index, nsites):
    """Documented counter-based uniforms used in place of the MersenneTwister
    stream: u_k = (splitmix64(splitmix64(splitmix64(seed) ^ index) ^ k) >> 11) * 2^-53,
    k = 0..nsites-1 in row-major site order. Values lie in [0, 1)."""
    base = splitmix64(splitmix64(seed & MASK64) ^ (index & MASK64))
    return np.array([(splitmix64(base ^ k) >> 11) * 2.0 ** -53 for k in range(nsites)])


def logsumexp(x):
    """``_rs_logsumexp`` (selection.jl:1-7)."""
    x = np.asarray(x, dtype=np.float64)
    if x.size == 0:
        return -np.inf
    pivot = np.max(x)
    if pivot == -np.inf:
        return -np.inf
    if not np.isfinite(pivot):
        raise ValueError("nonfinite positive log mass")
    return float(pivot + np.log(np.sum(np.exp(x - pivot))))


def categorical(log_mass, u):
    """``_categorical`` (selection.jl:68-80) with the uniform ``u`` given.
    Returns a 0-based index."""
    z = logsumexp(log_mass)
    if not np.isfinite(z):
        raise ValueError("proposal requires positive total probability")
    cumulative = 0.0
    last_positive = -1
    for j, lm in enumerate(log_mass):
        if lm == -np.inf:
            continue
        last_positive = j
        cumulative += np.exp(lm - z)
        if u < cumulative:
            return j
    return last_positive


class LocalBasis:
    """``LocalSamplingBasis`` built by ``_local_basis`` (basis.jl:36-63)."""

    def __init__(self, U, labels, charges):
        self.U = np.asarray(U)
        self.labels = list(labels)
        self.charges = list(charges)

    def index(self, label):
        """``_basis_index`` (basis.jl:88-92), 0-based."""
        try:
            return self.labels.index(label)
        except ValueError:
            raise ValueError("configuration contains an unknown basis label")


def symmetry_basis():
    """``SymmetryBasis()`` on Z2Space(0=>1,1=>1): U = I, labels 1,2 <-> charges 0,1."""
    return LocalBasis(np.eye(2), [1, 2], [0, 1])


def x_basis():
    """``XBasis()`` (basis.jl:19): columns (1,-1)/sqrt2 -> -1, (1,1)/sqrt2 -> +1."""
    return LocalBasis(np.array([[1.0, 1.0], [-1.0, 1.0]]) / np.sqrt(2.0), [-1, 1], [0, 0])


def resolve_basis(symmetric):
    """``_resolve_basis(:auto)`` (basis.jl:64-80)."""
    return symmetry_basis() if symmetric else x_basis()


def project(A, basis, k, kind="auxiliary"):
    """``_project`` (basis.jl:87) for the auxiliary kind: A'[w,n,p,0,e,s] = sum_a A[..a..] U[a,k]."""
    if kind != "auxiliary":
        raise NotImplementedError("physical-kind projection is out of scope")
    aux_par = A.p[3]
    O = Z2(basis.U[:, k][None, :].astype(A.a.dtype), [np.array([basis.charges[k] if aux_par.any() else 0]), aux_par])
    return bm.act(A, O, 3)


class SamplingCache:
    """``SamplingCache(state, BoundaryMPS(chi); kind=:auxiliary)`` (born.jl:45-73)."""

    def __init__(self, pepo, chi, basis=None):
        symmetric = bool(np.any(pepo[(0, 0)].p[2]))
        self.state = pepo
        self.chi = chi
        self.basis = basis if basis is not None else resolve_basis(symmetric)
        self.layers = bm.double_layers(pepo)
        Lx, Ly = pepo.lattice.Lx, pepo.lattice.Ly
        self.candidates = [[[bm.double_layer(project(pepo[(x, y)], self.basis, k))
                             for k in range(len(self.basis.labels))]
                            for y in range(Ly)] for x in range(Lx)]
        south = [None] * (Ly + 1)
        south[Ly] = bm.trivial_boundary(bm.row(self.layers, Ly - 1), "south")
        for y in range(Ly - 1, -1, -1):
            south[y] = bm.absorb(south[y + 1], bm.row(self.layers, y), chi, "south")
        self.south = south


class BornError(ValueError):
    pass


def sample_configuration(cache, uniforms=None, configuration=None):
    """``sample_configuration`` auxiliary kind (born.jl:86-163).

    ``uniforms``: one per site, row-major (x fastest). ``configuration``:
    labels [x][y] for evaluation mode. Returns a dict with the reference's fields.
    """
    Lx, Ly = len(cache.layers), len(cache.layers[0])
    if configuration is None and uniforms is None:
        raise ValueError("an explicit RNG is required")
    B = cache.basis
    selected = [[None] * Ly for _ in range(Lx)]
    if configuration is not None:
        for x in range(Lx):
            for y in range(Ly):
                B.index(configuration[x][y])
                selected[x][y] = configuration[x][y]
    north = bm.trivial_boundary(bm.row(cache.layers, 0), "north")
    log_probability = 0.0
    conditionals = []
    max_env = max(e.max_truncation for e in cache.south)
    for y in range(Ly):
        rowl = list(bm.row(cache.layers, y))
        top = north.tensors
        bottom = cache.south[y + 1].tensors
        _, right = bm.row_environments(top, rowl, bottom, side="right")
        left = bm.row_edge(top, rowl, bottom, "left")
        for x in range(Lx):
            cands = cache.candidates[x][y]
            steps = [bm.left_step(left, top[x], E, bottom[x]) for E in cands]
            raw = np.array([bm.join_scalars(s, right[x + 1]) for s in steps])
            scale = float(np.max(np.abs(raw), initial=0.0))
            ok = all(np.isfinite(z) and abs(np.imag(z)) <= 1e-9 * max(scale, np.finfo(float).tiny)
                     and np.real(z) >= -64 * np.finfo(float).eps * scale for z in raw)
            if not ok:
                raise BornError(f"invalid Born conditional at ({x + 1},{y + 1}); increase chi_sampling")
            values = np.maximum(np.real(raw), 0.0)
            total = float(np.sum(values))
            if not total > 0:
                raise BornError(f"Born conditional has zero total mass at ({x + 1},{y + 1})")
            probs = values / total
            if configuration is None:
                with np.errstate(divide="ignore"):
                    k = categorical(np.log(probs), uniforms[x + y * Lx])
            else:
                k = B.index(selected[x][y])
            selected[x][y] = B.labels[k]
            p = float(probs[k])
            conditionals.append(p)
            if p == 0:
                return {"configuration": selected, "log_probability": -np.inf,
                        "conditionals": conditionals, "max_environment_truncation": max_env}
            log_probability += np.log(p)
            rowl[x] = cands[k]
            if x < Lx - 1:
                left = bm.safe_scale(bm.left_step(left, top[x], rowl[x], bottom[x]))
        north = bm.absorb(north, rowl, cache.chi, "north")
        max_env = max(max_env, north.max_truncation)
    return {"configuration": selected, "log_probability": float(log_probability),
            "conditionals": conditionals, "max_environment_truncation": max_env}


def project_configuration(pepo, configuration, basis=None):
    """``project_configuration`` (born.jl:27-36): aux projection then normalize_state!."""
    symmetric = bool(np.any(pepo[(0, 0)].p[2]))
    basis = basis if basis is not None else resolve_basis(symmetric)
    Lx, Ly = pepo.lattice.Lx, pepo.lattice.Ly
    sites = [[project(pepo[(x, y)], basis, basis.index(configuration[x][y])) for y in range(Ly)]
             for x in range(Lx)]
    st = State(pepo.lattice, sites, "peps", beta=pepo.beta, log_amplitude=pepo.log_amplitude)
    records = normalize_state(st)
    return st, records
