"""Lattice, finite states and the TFIM (oracle restatement).

Follows src/states/finite.jl, src/models/tfim.jl and src/states/tensors.jl.
Legs of every site tensor are (W, N, P, Aux, E, S), all outgoing.
"""
import numpy as np
import scipy.linalg

from .z2 import Z2, ein


class Lattice:
    """``Lattice(Lx, Ly)`` (src/states/finite.jl:7-16); sites [x][y], row-major order."""

    def __init__(self, Lx, Ly):
        if Lx <= 0 or Ly <= 0:
            raise ValueError("lattice dimensions must be positive")
        self.Lx, self.Ly = int(Lx), int(Ly)

    def __len__(self):
        return self.Lx * self.Ly

    def site_order(self):
        return [(x, y) for y in range(self.Ly) for x in range(self.Lx)]

    def bonds(self, direction):
        """``bonds`` (src/models/tfim.jl:35-41), 0-based coordinates."""
        if direction == "horizontal":
            return [((x, y), (x + 1, y)) for y in range(self.Ly) for x in range(self.Lx - 1)]
        if direction == "vertical":
            return [((x, y), (x, y + 1)) for y in range(self.Ly - 1) for x in range(self.Lx)]
        raise ValueError("unknown bond direction")

    def degree(self, i):
        x, y = i
        return (x > 0) + (x < self.Lx - 1) + (y > 0) + (y < self.Ly - 1)


class State:
    """``FiniteState`` (src/states/finite.jl:27-44). ``kind`` is "pepo" or "peps"."""

    def __init__(self, lattice, sites, kind, beta=0.0, log_amplitude=0.0):
        self.lattice = lattice
        self.sites = sites  # list over x of list over y of Z2 tensors
        self.kind = kind
        self.beta = float(beta)
        self.log_amplitude = float(log_amplitude)

    def __getitem__(self, i):
        return self.sites[i[0]][i[1]]

    def __setitem__(self, i, v):
        self.sites[i[0]][i[1]] = v

    def copy(self):
        return State(self.lattice, [[t.copy() for t in col] for col in self.sites], self.kind,
                     self.beta, self.log_amplitude)


def validate_state(s):
    """``validate_state`` (src/states/finite.jl:53-73): ranks, OBC legs, matching bonds."""
    l = s.lattice
    for (x, y) in l.site_order():
        A = s[(x, y)]
        if A.ndim != 6:
            raise ValueError("expected six local legs")
        if not np.all(np.isfinite(A.a)):
            raise ValueError(f"nonfinite tensor at {(x, y)}")
        if s.kind == "peps" and A.shape[3] != 1:
            raise ValueError("PEPS auxiliary leg must contain one irreducible carrier")
        for edge, axis in ((x == 0, 0), (y == 0, 1), (x == l.Lx - 1, 4), (y == l.Ly - 1, 5)):
            if edge and A.shape[axis] != 1:
                raise ValueError(f"nontrivial OBC leg at {(x, y)}")
        if x < l.Lx - 1 and not np.array_equal(A.p[4], s[(x + 1, y)].p[0]):
            raise ValueError("horizontal space mismatch")
        if y < l.Ly - 1 and not np.array_equal(A.p[5], s[(x, y + 1)].p[1]):
            raise ValueError("vertical space mismatch")
    return s


def physical_parity(symmetric):
    """Z2Space(0=>1,1=>1) in the X eigenbasis, or ComplexSpace(2) (all even)."""
    return np.array([0, 1]) if symmetric else np.array([0, 0])


def identity_pepo(lattice, symmetric=True):
    """``identity_pepo`` (src/states/finite.jl:80-86): unit bonds, A[p,a] = delta."""
    pp = physical_parity(symmetric)
    u = np.array([0])
    sites = []
    for x in range(lattice.Lx):
        col = []
        for y in range(lattice.Ly):
            a = np.zeros((1, 1, 2, 2, 1, 1))
            a[0, 0, :, :, 0, 0] = np.eye(2)
            col.append(Z2(a, [u, u, pp, pp, u, u]))
        sites.append(col)
    return State(lattice, sites, "pepo")


def normalize_state(s):
    """``normalize_state!`` (src/states/finite.jl:89-100), row-major site order."""
    records = []
    for i in s.lattice.site_order():
        value = s[i].norm()
        if not (np.isfinite(value) and value > 0):
            raise ValueError(f"zero or nonfinite tensor at {i}")
        s[i] = s[i].scale(1.0 / value)
        inc = float(np.log(value))
        s.log_amplitude += inc
        records.append({"site": i, "log_normalization": inc})
    return records


class TFIM:
    """``TFIM`` (src/models/tfim.jl:9-19); ``symmetric`` selects Z2Space(0=>1,1=>1)."""

    def __init__(self, J=1.0, h=0.7, symmetric=False):
        self.J, self.h, self.symmetric = float(J), float(h), bool(symmetric)

    def operators(self):
        """``spin_operators`` (src/models/tfim.jl:22-34) as dense 2x2 / 4-index arrays.
        Z2: X = diag(1,-1), ZZ = z (x) z with z the parity flip. Dense: X, Z Pauli."""
        if self.symmetric:
            X = np.diag([1.0, -1.0])
            z = np.array([[0.0, 1.0], [1.0, 0.0]])
            ZZ = np.einsum("pu,qv->pquv", z, z)
        else:
            X = np.array([[0.0, 1.0], [1.0, 0.0]])
            Z = np.diag([1.0, -1.0])
            ZZ = np.einsum("pu,qv->pquv", Z, Z)
        return X, ZZ

    def bond_hamiltonian(self, lattice, a, b):
        """``bond_hamiltonian`` (src/models/tfim.jl:44-47): -J ZZ - h/deg(a) X1 - h/deg(b) X2."""
        X, ZZ = self.operators()
        I = np.eye(2)
        H = -self.J * ZZ.reshape(4, 4)
        H = H - self.h / lattice.degree(a) * np.kron(X, I) - self.h / lattice.degree(b) * np.kron(I, X)
        return H

    def trotter_gate(self, lattice, a, b, delta):
        """``trotter_gate = exp(-delta h_bond)`` (src/models/tfim.jl:48) as G[p,q,u,v]."""
        return scipy.linalg.expm(-delta * self.bond_hamiltonian(lattice, a, b)).reshape(2, 2, 2, 2)

    def onsite_gate(self, delta):
        X, _ = self.operators()
        return scipy.linalg.expm(delta * self.h * X)


def dense_hamiltonian(lattice, model):
    """``dense_hamiltonian`` (src/models/tfim.jl:51-65) in the Z basis, N <= 16."""
    N = len(lattice)
    if N > 16:
        raise ValueError("ED is limited to 16 sites")
    dim = 2 ** N
    H = np.zeros((dim, dim))
    idx = lambda i: i[0] + i[1] * lattice.Lx
    pairs = lattice.bonds("horizontal") + lattice.bonds("vertical")
    b = np.arange(dim)
    zs = [1 - 2 * ((b >> k) & 1) for k in range(N)]
    diagv = np.zeros(dim)
    for (i, j) in pairs:
        diagv -= model.J * zs[idx(i)] * zs[idx(j)]
    H[b, b] = diagv
    for k in range(N):
        H[b ^ (1 << k), b] -= model.h
    return H


def exact_thermal_observables(lattice, model, beta):
    """``exact_thermal_observables`` (src/models/tfim.jl:72-76)."""
    e = np.linalg.eigvalsh(dense_hamiltonian(lattice, model))
    shift = e.min()
    w = np.exp(-beta * (e - shift))
    Z = w.sum()
    total = float(np.dot(w, e) / Z)
    return {"energy_total": total, "energy_per_site": total / len(lattice),
            "lnZ": float(np.log(Z) - beta * shift)}


def random_state(lattice, D, kind, seed, symmetric=True, d=2, dtype=np.float64):
    """Seeded i.i.d. N(0,1) entries in every allowed block, each site scaled to unit
    Frobenius norm (the random input of the timing and differentiation workloads). ``D`` is an
    (even, odd) pair for symmetric states or an int for dense ones. ``kind``
    "pepo": aux leg = physical space; "peps": aux leg of dim 1 and parity 0."""
    rng = np.random.default_rng(seed)
    if symmetric:
        De, Do = D
        bp = np.array([0] * De + [1] * Do)
    else:
        bp = np.zeros(int(D), dtype=np.int64)
    pp = physical_parity(symmetric)
    ap = pp if kind == "pepo" else np.array([0])
    u = np.array([0])
    sites = []
    for x in range(lattice.Lx):
        col = []
        for y in range(lattice.Ly):
            legs = [u if x == 0 else bp, u if y == 0 else bp, pp, ap,
                    u if x == lattice.Lx - 1 else bp, u if y == lattice.Ly - 1 else bp]
            t = Z2(rng.standard_normal([len(p) for p in legs]).astype(dtype), legs)
            from .z2 import symmetrize
            t = symmetrize(t)
            t = t.scale(1.0 / t.norm())
            col.append(t)
        sites.append(col)
    return State(lattice, sites, kind)


def embed_identity_pepo(lattice, D, eps, seed, symmetric=True):
    """Identity PEPO (beta=0) embedded in bond spaces of dimension ``D`` plus a
    seeded Gaussian perturbation of relative size ``eps`` in every allowed block.
    A well-conditioned synthetic thermal-like input at production shapes."""
    base = random_state(lattice, D, "pepo", seed, symmetric)
    for (x, y) in lattice.site_order():
        t = base[(x, y)]
        a = eps * t.a
        a[0, 0, :, :, 0, 0] += np.eye(2) / np.sqrt(2.0)
        base[(x, y)] = Z2(a / np.linalg.norm(a), t.p)
    return base
