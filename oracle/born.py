"""Perfect Born sampling (oracle restatement of src/sampling/born.jl,
src/sampling/basis.jl, src/states/tensors.jl:65-94 and
src/resample/selection.jl:1-7,68-80): PEPO auxiliary indices
(``kind="auxiliary"``) and, for resampling proposals, PEPS physical indices
(``kind="physical"``) with the PSD-projected site metric.

The reference draws one ``rand(rng)`` per site from a Julia MersenneTwister
(born.jl:13-22); that stream cannot be produced without Julia, so the uniforms
are an input here (one per site, row-major with x fastest).
"""
import numpy as np

from .z2 import Z2, ein, permute, split, fused_parity
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
    """``_project`` (basis.jl:87) with the maps of ``_local_basis`` (basis.jl:53-61).
    Auxiliary kind: A'[w,n,p,0,e,s] = sum_a A[..a..] U[a,k] (the conjugate of
    the physical map). Physical kind: A'[w,n,0,a,e,s] = sum_p conj(U[p,k]) A[..p..];
    the unit leg carries the label's charge."""
    if kind == "auxiliary":
        aux_par = A.p[3]
        O = Z2(basis.U[:, k][None, :].astype(A.a.dtype),
               [np.array([basis.charges[k] if aux_par.any() else 0]), aux_par])
        return bm.act(A, O, 3)
    if kind != "physical":
        raise ValueError("unknown sampling kind")
    phys_par = A.p[2]
    O = Z2(np.conj(basis.U[:, k])[None, :].astype(A.a.dtype),
           [np.array([basis.charges[k] if phys_par.any() else 0]), phys_par])
    return bm.act(A, O, 2)


def state_charge(state):
    """``_state_charge`` (basis.jl:89-92): fused (Z2: summed) charge of every
    site's auxiliary leg."""
    q = 0
    for i in state.lattice.site_order():
        q ^= int(state[i].p[3][0])
    return q


def charge_suffix(charges_per_site):
    """``_charge_suffix`` (basis.jl:93-100): reachable[i] = charges that sites
    i.. can still fuse to (reachable[N] = {0})."""
    N = len(charges_per_site)
    reachable = [None] * (N + 1)
    reachable[N] = {0}
    for i in range(N - 1, -1, -1):
        reachable[i] = {c ^ r for c in charges_per_site[i] for r in reachable[i + 1]}
    return reachable


def site_hole(L, T, B, R, A):
    """``_site_hole`` (born.jl:75-83): the single-site metric of site A from the
    left and right row messages and the boundary tensors above and below, as
    a Z2 matrix (rows: the bra copies of W, N, E, S; columns: the ket copies)."""
    X = ein("awb,anc->wbnc", L, T)
    Y = ein("bsd,ced->bsce", B, R)
    H = ein("wbnc,bsce->wnes", X, Y)
    legs = (0, 1, 4, 5)
    for axis in (3, 2, 1, 0):
        V = A.p[legs[axis]]
        H = split(H, axis, (len(V), len(V)), (V, V))
    return permute(H, (1, 3, 5, 7, 0, 2, 4, 6))


def positive_metric(N, nrows, square_root=False):
    """``_positive_metric`` (tensors.jl:65-94) of a Z2 tensor seen as a matrix
    with its first ``nrows`` legs as rows: scale by the norm, average with the
    transpose, eigendecompose every parity block, clip negative eigenvalues.
    Returns (repaired matrix, or the root sqrt(D+) V' with ``square_root``,
    as a dense matrix over the fused row / column indices; report)."""
    M = N.a.reshape(int(np.prod(N.shape[:nrows])), -1)
    if not np.all(np.isfinite(M)):
        raise ValueError("nonfinite norm metric")
    rp = fused_parity(N.p[:nrows])
    cp = fused_parity(N.p[nrows:])
    scale = float(np.linalg.norm(M))
    scaled = M if scale == 0 else M / scale
    H = (scaled + scaled.conj().T) / 2
    vals = []
    out = np.zeros_like(H)
    for q in (0, 1):
        ri = np.nonzero(rp == q)[0]
        ci = np.nonzero(cp == q)[0]
        if len(ri) == 0:
            continue
        w, V = np.linalg.eigh(H[np.ix_(ri, ci)])
        vals.extend(float(v) for v in w)
        clipped = np.maximum(w, 0.0)
        if square_root:
            # rows of the root are eigen-indices; keep them in the row slots of the block
            out[np.ix_(ri, ci)] = (np.sqrt(clipped)[:, None] * V.conj().T)
        else:
            R = (V * clipped[None, :]) @ V.conj().T
            out[np.ix_(ri, ci)] = (R + R.conj().T) / 2
    vals = np.asarray(vals)
    tiny = np.finfo(np.float64).tiny
    minimum = float(min(np.min(vals, initial=0.0), 0.0))
    spectral = max(float(np.max(np.abs(vals), initial=0.0)), tiny)
    neg = vals < 0
    report = {"minimum_relative_eigenvalue": minimum / spectral,
              "negative_eigenvalues": float(np.sum(neg)),
              "negative_multiplets": int(np.sum(neg)),
              "negative_spectral_fraction": float(np.sqrt(np.sum(np.minimum(vals, 0) ** 2))
                                                  / max(np.sqrt(np.sum(vals ** 2)), tiny)),
              "hermiticity_defect": float(np.linalg.norm(scaled - scaled.conj().T)
                                          / max(np.linalg.norm(scaled), tiny))}
    return out, report


class SamplingCache:
    """``SamplingCache(state, BoundaryMPS(chi); kind, basis)`` (born.jl:45-73).
    ``kind="auxiliary"`` samples a PEPO's auxiliary indices, ``kind="physical"``
    a PEPS's physical indices (resampling proposals); a symmetric physical
    cache carries the state's charge and the reachable charge suffixes."""

    def __init__(self, pepo, chi, basis=None, kind="auxiliary"):
        if kind not in ("auxiliary", "physical"):
            raise ValueError("unknown sampling kind")
        if kind == "physical" and pepo.kind != "peps":
            raise ValueError("proposal requires PEPS")
        symmetric = bool(np.any(pepo[(0, 0)].p[2]))
        self.state = pepo
        self.chi = chi
        self.kind = kind
        self.basis = basis if basis is not None else resolve_basis(symmetric)
        self.layers = bm.double_layers(pepo)
        Lx, Ly = pepo.lattice.Lx, pepo.lattice.Ly
        self.candidates = [[[bm.double_layer(project(pepo[(x, y)], self.basis, k, kind))
                             for k in range(len(self.basis.labels))]
                            for y in range(Ly)] for x in range(Lx)]
        charged = kind == "physical" and symmetric
        self.target_charge = state_charge(pepo) if charged else None
        self.reachable = charge_suffix([self.basis.charges] * (Lx * Ly)) if charged else []
        south = [None] * (Ly + 1)
        south[Ly] = bm.trivial_boundary(bm.row(self.layers, Ly - 1), "south")
        for y in range(Ly - 1, -1, -1):
            south[y] = bm.absorb(south[y + 1], bm.row(self.layers, y), chi, "south")
        self.south = south


class BornError(ValueError):
    pass


def sample_configuration(cache, uniforms=None, configuration=None):
    """``sample_configuration`` (born.jl:86-163), both kinds.

    ``uniforms``: one per site, row-major (x fastest). ``configuration``:
    labels [x][y] for evaluation mode. Returns a dict with the reference's
    fields (``repairs``: the physical kind's PSD diagnostics per site and its
    zero-weight fallbacks)."""
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
    repairs = []
    max_env = max(e.max_truncation for e in cache.south)

    def result(lp):
        return {"configuration": selected, "log_probability": lp, "conditionals": conditionals,
                "max_environment_truncation": max_env, "repairs": repairs}

    prefix = 0
    if configuration is not None and cache.target_charge is not None:
        charge = 0
        for y in range(Ly):
            for x in range(Lx):
                charge ^= B.charges[B.index(configuration[x][y])]
        if charge != cache.target_charge:
            return result(-np.inf)
    for y in range(Ly):
        rowl = list(bm.row(cache.layers, y))
        top = north.tensors
        bottom = cache.south[y + 1].tensors
        _, right = bm.row_environments(top, rowl, bottom, side="right")
        left = bm.row_edge(top, rowl, bottom, "left")
        for x in range(Lx):
            cands = cache.candidates[x][y]
            position = x + y * Lx
            allowed = [cache.target_charge is None
                       or any((prefix ^ c ^ r) == cache.target_charge for r in cache.reachable[position + 1])
                       for c in B.charges]
            if not any(allowed):
                raise BornError("sampling prefix violates the total symmetry sector")
            if cache.kind == "physical":
                A = cache.state[(x, y)]
                metric = site_hole(left, top[x], bottom[x], right[x + 1], A)
                root, diagnostic = positive_metric(metric, 4, square_root=True)
                repairs.append({**diagnostic, "site": (x + 1, y + 1)})
                values = np.zeros(len(B.labels))
                for k in range(len(B.labels)):
                    if not allowed[k]:
                        continue
                    P = project(A, B, k, "physical")
                    vec = np.transpose(P.a, (0, 1, 4, 5, 2, 3)).reshape(root.shape[1], -1)
                    values[k] = float(np.sum(np.abs(root @ vec) ** 2))
                total = float(np.sum(values))
                if total == 0:
                    probs = np.asarray(allowed, dtype=np.float64) / sum(allowed)
                    repairs.append({"site": (x + 1, y + 1), "reason": "zero_proposal_weights",
                                    "allowed_outcomes": int(sum(allowed))})
                else:
                    probs = values / total
            else:
                steps = [bm.left_step(left, top[x], E, bottom[x]) for E in cands]
                raw = np.array([bm.join_scalars(st, right[x + 1]) for st in steps])
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
                return result(-np.inf)
            log_probability += np.log(p)
            rowl[x] = cands[k]
            prefix ^= B.charges[k]
            if x < Lx - 1:
                left = bm.safe_scale(bm.left_step(left, top[x], rowl[x], bottom[x]))
        if cache.kind == "auxiliary" or y < Ly - 1:
            north = bm.absorb(north, rowl, cache.chi, "north")
            max_env = max(max_env, north.max_truncation)
    return result(float(log_probability))


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
