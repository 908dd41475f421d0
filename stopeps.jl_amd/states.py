"""Host-side state model mirroring StoPEPS.jl's L1 layer (src/states/finite.jl,
src/models/tfim.jl, src/sampling/basis.jl).

A site tensor is a dense ``numpy`` array over the legs (W, N, P, Aux, E, S)
with one Z2 parity vector per leg (all zeros for ``ComplexSpace``). The device
layout wants every leg sorted by parity; :func:`parity_sorted` produces it.
"""
import numpy as np


class Lattice:
    """``Lattice(Lx, Ly)`` (src/states/finite.jl:7-16)."""

    def __init__(self, Lx, Ly):
        if int(Lx) <= 0 or int(Ly) <= 0:
            raise ValueError("lattice dimensions must be positive")
        self.Lx, self.Ly = int(Lx), int(Ly)

    def __len__(self):
        return self.Lx * self.Ly

    def site_order(self):
        """Row-major order, x fastest (src/states/finite.jl:16)."""
        return [(x, y) for y in range(self.Ly) for x in range(self.Lx)]


class Site:
    """One site tensor: ``data`` over (W,N,P,Aux,E,S) and per-leg parity vectors."""

    __slots__ = ("data", "parity")

    def __init__(self, data, parity):
        self.data = np.asarray(data, dtype=np.float64)
        self.parity = [np.asarray(p, dtype=np.int64) % 2 for p in parity]
        if self.data.ndim != 6 or len(self.parity) != 6:
            raise ValueError("expected six local legs")
        for p, n in zip(self.parity, self.data.shape):
            if len(p) != n:
                raise ValueError("parity vector length mismatch")


class FiniteState:
    """``FiniteState{K}`` (src/states/finite.jl:27-44); kind "pepo" or "peps"."""

    def __init__(self, lattice, sites, kind, beta=0.0, log_amplitude=0.0):
        if kind not in ("pepo", "peps"):
            raise ValueError("kind must be 'pepo' or 'peps'")
        self.lattice = lattice
        self.sites = sites  # sites[x][y] -> Site
        self.kind = kind
        self.beta = float(beta)
        self.log_amplitude = float(log_amplitude)
        validate_state(self)

    def __getitem__(self, i):
        return self.sites[i[0]][i[1]]


FinitePEPO = lambda lattice, sites, **kw: FiniteState(lattice, sites, "pepo", **kw)  # noqa: E731
FinitePEPS = lambda lattice, sites, **kw: FiniteState(lattice, sites, "peps", **kw)  # noqa: E731


def validate_state(s):
    """``validate_state`` (src/states/finite.jl:53-73)."""
    l = s.lattice
    for (x, y) in l.site_order():
        A = s[(x, y)]
        if not np.all(np.isfinite(A.data)):
            raise ValueError(f"nonfinite tensor at {(x + 1, y + 1)}")
        if s.kind == "peps" and A.data.shape[3] != 1:
            raise ValueError("PEPS auxiliary leg must contain one irreducible carrier")
        for edge, axis in ((x == 0, 0), (y == 0, 1), (x == l.Lx - 1, 4), (y == l.Ly - 1, 5)):
            if edge and A.data.shape[axis] != 1:
                raise ValueError(f"nontrivial OBC leg at {(x + 1, y + 1)}")
        if x < l.Lx - 1 and not np.array_equal(A.parity[4], s[(x + 1, y)].parity[0]):
            raise ValueError("horizontal space mismatch")
        if y < l.Ly - 1 and not np.array_equal(A.parity[5], s[(x, y + 1)].parity[1]):
            raise ValueError("vertical space mismatch")
    return s


def identity_pepo(lattice, symmetric=True):
    """``identity_pepo`` (src/states/finite.jl:80-86)."""
    pp = np.array([0, 1]) if symmetric else np.array([0, 0])
    u = np.array([0])
    sites = []
    for x in range(lattice.Lx):
        col = []
        for y in range(lattice.Ly):
            a = np.zeros((1, 1, 2, 2, 1, 1))
            a[0, 0, :, :, 0, 0] = np.eye(2)
            col.append(Site(a, [u, u, pp, pp, u, u]))
        sites.append(col)
    return FiniteState(lattice, sites, "pepo")


def normalize_state(state):
    """``normalize_state!`` (src/states/finite.jl:89-100): divide every site by
    its Frobenius norm in row-major order and add the logs to log_amplitude,
    in place. Returns the records {"site": (x+1, y+1), "log_normalization"}."""
    records = []
    for (x, y) in state.lattice.site_order():
        v = float(np.linalg.norm(state[(x, y)].data))
        if not (np.isfinite(v) and v > 0):
            raise ValueError(f"zero or nonfinite tensor at {(x + 1, y + 1)}")
        state.sites[x][y] = Site(state[(x, y)].data / v, state[(x, y)].parity)
        state.log_amplitude += np.log(v)
        records.append({"site": (x + 1, y + 1), "log_normalization": float(np.log(v))})
    return records


def synthetic_thermal_pepo(lattice, D=(3, 3), eps=0.3, seed=1):
    """Synthetic Z2 PEPO at production shapes: the beta = 0 identity PEPO
    embedded in bond spaces Z2Space(0=>D[0], 1=>D[1]) plus a seeded Gaussian
    perturbation of relative size ``eps`` in every allowed block, each site
    scaled to unit Frobenius norm. Used as the timing input of the cfg
    "8x8, D=6, chi=32" workload (no beta0 checkpoint ships with the reference)."""
    rng = np.random.default_rng(seed)
    bp = np.array([0] * D[0] + [1] * D[1])
    pp = np.array([0, 1])
    u = np.array([0])
    sites = []
    for x in range(lattice.Lx):
        col = []
        for y in range(lattice.Ly):
            legs = [u if x == 0 else bp, u if y == 0 else bp, pp, pp,
                    u if x == lattice.Lx - 1 else bp, u if y == lattice.Ly - 1 else bp]
            shape = [len(p) for p in legs]
            a = rng.standard_normal(shape)
            par = np.zeros(shape, dtype=np.int64)
            for k, p in enumerate(legs):
                sh = [1] * 6
                sh[k] = len(p)
                par = par + p.reshape(sh)
            a[par % 2 != 0] = 0.0
            a /= np.linalg.norm(a)
            a *= eps
            a[0, 0, :, :, 0, 0] += np.eye(2) / np.sqrt(2.0)
            a /= np.linalg.norm(a)
            col.append(Site(a, legs))
        sites.append(col)
    return FiniteState(lattice, sites, "pepo")


def random_peps(lattice, D=(3, 3), seed=7, kind="peps"):
    """Seeded random Z2 state: i.i.d. N(0,1) entries in every allowed block of
    bond spaces Z2Space(0=>D[0], 1=>D[1]), physical Z2Space(0=>1, 1=>1), each
    site scaled to unit Frobenius norm; ``kind`` "peps" has an auxiliary leg of
    dimension 1, "pepo" one equal to the physical leg. The input of the
    log-norm gradient workload (8x8 PEPS, D=6, chi=36, seed 7). Sites are drawn
    column by column (x outer, y inner), one full-shape Gaussian array each."""
    rng = np.random.default_rng(seed)
    bp = np.array([0] * D[0] + [1] * D[1])
    pp = np.array([0, 1])
    ap = pp if kind == "pepo" else np.array([0])
    u = np.array([0])
    sites = []
    for x in range(lattice.Lx):
        col = []
        for y in range(lattice.Ly):
            legs = [u if x == 0 else bp, u if y == 0 else bp, pp, ap,
                    u if x == lattice.Lx - 1 else bp, u if y == lattice.Ly - 1 else bp]
            shape = [len(p) for p in legs]
            a = rng.standard_normal(shape)
            par = np.zeros(shape, dtype=np.int64)
            for k, p in enumerate(legs):
                sh = [1] * 6
                sh[k] = len(p)
                par = par + p.reshape(sh)
            a[par % 2 != 0] = 0.0
            a /= np.linalg.norm(a)
            col.append(Site(a, legs))
        sites.append(col)
    return FiniteState(lattice, sites, kind)


def pad_bonds(state, capacity):
    """Embed every interior virtual leg in a bond space of (even, odd) =
    ``capacity`` by appending zero slots (the original entries come first in
    each parity, so they are the live ones on the device). Returns the padded
    state and the live counts [Lx*Ly, 4, 2] for ``DeviceState(..., live=...)``."""
    L = state.lattice
    ce, co = capacity
    legs = (0, 1, 4, 5)
    sites = [[None] * L.Ly for _ in range(L.Lx)]
    live = np.zeros((L.Lx * L.Ly, 4, 2), dtype=np.int32)
    for (x, y) in L.site_order():
        s = state[(x, y)]
        a = s.data
        par = list(s.parity)
        for v, k in enumerate(legs):
            p = par[k]
            e, o = int(np.sum(p == 0)), int(np.sum(p == 1))
            live[x + y * L.Lx, v] = (e, o)
            if len(p) == 1 and ((k == 0 and x == 0) or (k == 1 and y == 0) or (k == 4 and x == L.Lx - 1)
                                or (k == 5 and y == L.Ly - 1)):
                continue
            if e > ce or o > co:
                raise ValueError("bond space exceeds the capacity")
            extra = np.array([0] * (ce - e) + [1] * (co - o), dtype=np.int64)
            if len(extra):
                shape = list(a.shape)
                shape[k] = len(extra)
                a = np.concatenate([a, np.zeros(shape)], axis=k)
                par[k] = np.concatenate([p, extra])
        sites[x][y] = Site(a, par)
    return FiniteState(L, sites, state.kind, beta=state.beta, log_amplitude=state.log_amplitude), live


def reembed_bonds(state, live, capacity):
    """Re-embed every interior virtual leg in a bond space of (even, odd) =
    ``capacity``: keep its live slots (the first live[site, leg, q] indices of
    parity q, in leg order) and append zero slots. ``live`` [Lx*Ly, 4, 2] as in
    :func:`pad_bonds`. Returns the state and its live counts; a live sector
    larger than the capacity raises."""
    L = state.lattice
    ce, co = capacity
    legs = (0, 1, 4, 5)
    live = np.asarray(live).reshape(L.Lx * L.Ly, 4, 2)
    sites = [[None] * L.Ly for _ in range(L.Lx)]
    for (x, y) in L.site_order():
        s = state[(x, y)]
        a = s.data
        par = list(s.parity)
        for v, k in enumerate(legs):
            p = par[k]
            edge = len(p) == 1 and ((k == 0 and x == 0) or (k == 1 and y == 0) or (k == 4 and x == L.Lx - 1)
                                    or (k == 5 and y == L.Ly - 1))
            if edge:
                continue
            e, o = int(live[x + y * L.Lx, v, 0]), int(live[x + y * L.Lx, v, 1])
            if e > ce or o > co:
                raise ValueError("live bond sectors exceed the capacity")
            keep = np.concatenate([np.nonzero(p == 0)[0][:e], np.nonzero(p == 1)[0][:o]])
            keep.sort(kind="stable")
            a = np.take(a, keep, axis=k)
            p = p[keep]
            extra = np.array([0] * (ce - e) + [1] * (co - o), dtype=np.int64)
            if len(extra):
                shape = list(a.shape)
                shape[k] = len(extra)
                a = np.concatenate([a, np.zeros(shape)], axis=k)
                p = np.concatenate([p, extra])
            par[k] = p
        sites[x][y] = Site(a, par)
    return FiniteState(L, sites, state.kind, beta=state.beta, log_amplitude=state.log_amplitude), live.copy()


def parity_sorted(site):
    """Dense array with every leg sorted by parity (stable), and even counts."""
    a = site.data
    evens = []
    for k, p in enumerate(site.parity):
        order = np.argsort(p, kind="stable")
        a = np.take(a, order, axis=k)
        evens.append(int(np.sum(p == 0)))
    return np.ascontiguousarray(a), evens


class LocalBasis:
    """Local sampling basis (``LocalSamplingBasis``, src/sampling/basis.jl:29-63):
    columns of ``U`` are basis kets; ``labels`` their configuration labels."""

    def __init__(self, U, labels, charges):
        self.U = np.asarray(U, dtype=np.float64)
        self.labels = list(labels)
        self.charges = list(charges)

    def index(self, label):
        """``_basis_index`` (basis.jl:88-92), 0-based."""
        try:
            return self.labels.index(label)
        except ValueError:
            raise ValueError("configuration contains an unknown basis label")


def SymmetryBasis():
    """Z2 parity sectors: U = I, labels 1, 2 for charges 0, 1 (basis.jl:15, 51-57)."""
    return LocalBasis(np.eye(2), [1, 2], [0, 1])


def XBasis():
    """``XBasis()`` (basis.jl:19)."""
    return LocalBasis(np.array([[1.0, 1.0], [-1.0, 1.0]]) / np.sqrt(2.0), [-1, 1], [0, 0])


def resolve_basis(state, basis=None):
    """``_resolve_basis(:auto)`` (basis.jl:64-80): Z2 -> SymmetryBasis, dense -> XBasis."""
    if basis is not None:
        return basis
    symmetric = bool(np.any(state[(0, 0)].parity[2]))
    return SymmetryBasis() if symmetric else XBasis()


def save_state_npz(path, state, **extra):
    """Write a ``FiniteState`` to an ``.npz`` file: one dense array and six
    parity vectors per site (row-major order), kind, beta and log_amplitude,
    plus ``extra`` arrays. Loaded by :func:`load_state_npz`."""
    arrs = {}
    for i, (x, y) in enumerate(state.lattice.site_order()):
        s = state[(x, y)]
        arrs[f"site{i}"] = s.data
        for k in range(6):
            arrs[f"par{i}_{k}"] = s.parity[k].astype(np.int8)
    np.savez_compressed(path, Lx=state.lattice.Lx, Ly=state.lattice.Ly, kind=state.kind, beta=state.beta,
                        log_amplitude=state.log_amplitude, **arrs, **extra)


def load_state_npz(path):
    """(``FiniteState``, dict of the remaining arrays) of a file written by
    :func:`save_state_npz` (read without unpickling)."""
    f = np.load(path, allow_pickle=False)
    Lx, Ly = int(f["Lx"]), int(f["Ly"])
    lat = Lattice(Lx, Ly)
    sites = [[None] * Ly for _ in range(Lx)]
    used = {"Lx", "Ly", "kind", "beta", "log_amplitude"}
    for i in range(Lx * Ly):
        x, y = i % Lx, i // Lx
        sites[x][y] = Site(f[f"site{i}"], [f[f"par{i}_{k}"].astype(np.int64) for k in range(6)])
        used.add(f"site{i}")
        used.update(f"par{i}_{k}" for k in range(6))
    kind = str(f["kind"]) if "kind" in f.files else "pepo"
    st = FiniteState(lat, sites, kind, beta=float(f["beta"]), log_amplitude=float(f["log_amplitude"]))
    return st, {k: f[k] for k in f.files if k not in used}
