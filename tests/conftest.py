import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def spp():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def gpu_ctx(spp):
    ctx = spp.Context(0, "fp64")
    yield ctx
    ctx.close()


def to_product_state(spp, ostate):
    """Oracle state -> product FiniteState (same dense arrays and parities)."""
    L = ostate.lattice
    sites = [[spp.Site(ostate[(x, y)].a, ostate[(x, y)].p) for y in range(L.Ly)] for x in range(L.Lx)]
    return spp.FiniteState(spp.Lattice(L.Lx, L.Ly), sites, ostate.kind, beta=ostate.beta,
                           log_amplitude=ostate.log_amplitude)


def oracle_cfg(res, L):
    return np.array([[res["configuration"][x][y] for y in range(L.Ly)] for x in range(L.Lx)])


def charged_peps(L, D, seed, aux_parities, symmetric=True):
    """Random oracle PEPS whose site i carries auxiliary charge aux_parities[i]
    (row-major): a projected particle's leg structure."""
    from oracle.model import State
    from oracle.z2 import Z2, symmetrize
    rng = np.random.default_rng(seed)
    if symmetric:
        bp = np.array([0] * D[0] + [1] * D[1])
        pp = np.array([0, 1])
    else:
        bp = np.zeros(D, dtype=np.int64)
        pp = np.zeros(2, dtype=np.int64)
    u = np.array([0])
    sites = []
    for x in range(L.Lx):
        col = []
        for y in range(L.Ly):
            ap = np.array([aux_parities[x + y * L.Lx] if symmetric else 0])
            legs = [u if x == 0 else bp, u if y == 0 else bp, pp, ap,
                    u if x == L.Lx - 1 else bp, u if y == L.Ly - 1 else bp]
            t = symmetrize(Z2(rng.standard_normal([len(p) for p in legs]), legs))
            col.append(t.scale(1.0 / t.norm()))
        sites.append(col)
    return State(L, sites, "peps")


def to_oracle_state(pstate):
    """Product FiniteState -> oracle State (same dense arrays and parities)."""
    from oracle.model import Lattice, State
    from oracle.z2 import Z2
    L = pstate.lattice
    sites = [[Z2(pstate[(x, y)].data, pstate[(x, y)].parity) for y in range(L.Ly)] for x in range(L.Lx)]
    return State(Lattice(L.Lx, L.Ly), sites, pstate.kind, beta=pstate.beta, log_amplitude=pstate.log_amplitude)


def boundary_distance(basis, ref, cfg_ref, cfg_got, u, L):
    """Distance of the uniform from the oracle's cumulative-probability
    boundary at the first site (row-major) where two draws differ: with two
    outcomes, ``_categorical`` (src/resample/selection.jl:68-80) picks the
    first when u < p_first, so a draw can flip only if u lies within the
    conditional's error of p_first."""
    for j in range(L.Lx * L.Ly):
        x, y = j % L.Lx, j // L.Lx
        if cfg_ref[x][y] != cfg_got[x][y]:
            c = ref["conditionals"][j]
            p_first = c if basis.index(cfg_ref[x][y]) == 0 else 1.0 - c
            return abs(u[j] - p_first)
    raise AssertionError("draws do not differ")
