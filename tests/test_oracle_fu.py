"""Pins of the full-update oracle (oracle/fu.py) by dense state vectors:
exact Trotter evolution when no truncation binds, the FU cost against the
brute-force distance ||Psi_new - G Psi||^2 / ||G Psi||^2 when it does, two
contraction orders of the bond metric, and the Jacobi-PCG solve."""
import itertools

import numpy as np
import pytest

from oracle.model import Lattice, TFIM, random_state
from oracle.z2 import Z2
from oracle import boundary as bm
from oracle import fu


def dense_vector(state):
    """Amplitudes over the physical indices (aux dims 1), row-major sites, times exp(log_amplitude)."""
    L = state.lattice
    counter = itertools.count()
    hb, vb = {}, {}

    def g(d, k):
        if k not in d:
            d[k] = next(counter)
        return d[k]
    args, out = [], []
    for y in range(L.Ly):
        for x in range(L.Lx):
            A = state[(x, y)]
            if A.shape[3] != 1:
                raise ValueError("PEPS expected")
            w, e = g(hb, (x, y)), g(hb, (x + 1, y))
            n, s = g(vb, (x, y)), g(vb, (x, y + 1))
            p, a = next(counter), next(counter)
            args += [A.a, [w, n, p, a, e, s]]
            out += [p]
            _ = a
    return np.einsum(*args, out, optimize=True) * np.exp(state.log_amplitude)


def apply_gate(psi, L, i, j, G):
    """psi'[.. p_i .. p_j ..] = sum G[p_i, p_j, u, v] psi[.. u .. v ..]."""
    a = i[0] + i[1] * L.Lx
    b = j[0] + j[1] * L.Lx
    psi = np.moveaxis(psi, (a, b), (0, 1))
    psi = np.einsum("pquv,uv...->pq...", G, psi)
    return np.moveaxis(psi, (0, 1), (a, b))


def sweep_order(L):
    """Bond order of the second-order step (full_update.jl:566-567, 510-561)."""
    Lx, Ly = L.Lx, L.Ly
    h_fwd = [((x, y), (x + 1, y)) for y in range(Ly) for x in range(Lx - 1)]
    v_fwd = [((x, y), (x, y + 1)) for y in range(Ly - 1) for x in range(Lx)]
    v_rev = [((x, y), (x, y + 1)) for y in range(Ly - 2, -1, -1) for x in range(Lx - 1, -1, -1)]
    h_rev = [((x, y), (x + 1, y)) for y in range(Ly - 1, -1, -1) for x in range(Lx - 2, -1, -1)]
    return h_fwd + v_fwd + v_rev + h_rev


@pytest.mark.parametrize("shape,D", [((2, 2), (1, 1)), ((3, 2), (1, 0)), ((2, 3), (1, 1))])
def test_evolution_step_is_exact_trotter_without_truncation(shape, D):
    """With D and chi above every exact rank, one second-order FU step applies
    exp(-delta/2 h_b) on every bond in the sweep order, exactly."""
    L = Lattice(*shape)
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    st = random_state(L, D, "peps", seed=3)
    psi = dense_vector(st)
    delta = 0.1
    for (i, j) in sweep_order(L):
        psi = apply_gate(psi, L, i, j, m.trotter_gate(L, i, j, delta / 2))
    rep = fu.evolution_step(st, m, delta, D=256, caps=None, opts=fu.FullUpdate(), chi=4096)
    got = dense_vector(st)
    assert abs(st.beta - 2 * delta) < 1e-15
    assert np.linalg.norm(got - psi) <= 1e-10 * np.linalg.norm(psi)
    assert all(r["final_cost"] <= 1e-12 for r in rep["records"])


def _single_bond(shape, direction, x, y, D, caps, opts, seed=5):
    L = Lattice(*shape)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    st = random_state(L, (2, 2), "peps", seed=seed)
    i = (x, y)
    j = (x + 1, y) if direction == "horizontal" else (x, y + 1)
    G = m.trotter_gate(L, i, j, 0.3)
    target_vec = apply_gate(dense_vector(st), L, i, j, G)
    layers = bm.double_layers(st)
    chi = 4096
    north, south, _ = bm.boundary_environments(layers, chi)
    top = north[y].tensors
    if direction == "horizontal":
        bottom = south[y + 1].tensors
        rowl = bm.row(layers, y)
    else:
        bottom = south[y + 2].tensors
        rowl = [bm.Slab(layers[k][y], layers[k][y + 1]) for k in range(L.Lx)]
    left, right = bm.row_environments(top, rowl, bottom)
    finalx = x + 1 if direction == "horizontal" else x
    env = fu.bond_environment(top, bottom, x, direction, left[x], right[finalx + 1])
    Gz = Z2(G, [st[i].p[2], st[j].p[2], st[i].p[2], st[j].p[2]])
    rec = fu.update_bond(st, i, j, Gz, D, caps, opts, env, direction)
    return st, target_vec, rec, env


@pytest.mark.parametrize("direction,x,y", [("horizontal", 0, 0), ("horizontal", 1, 1), ("vertical", 1, 0),
                                           ("vertical", 0, 1)])
def test_fu_cost_is_the_dense_distance(direction, x, y):
    """Exact environment, truncated bond: the relative FU cost equals
    ||Psi_new - G Psi||^2 / ||G Psi||^2 of the dense vectors, and ALS does not
    end above the gauged SVD start."""
    opts = fu.FullUpdate(max_iterations=40, linear_max_iterations=100)
    st, tv, rec, _ = _single_bond((3, 3), direction, x, y, 2, (1, 1), opts)
    got = dense_vector(st)
    dist = np.sum((got - tv) ** 2) / np.sum(tv ** 2)
    assert dist > 1e-6  # the truncation binds
    assert abs(rec["final_cost"] - dist) <= 1e-9 * max(dist, 1e-12) + 1e-13
    assert rec["final_cost"] <= rec["initial_cost"] + 1e-15
    assert rec["iterations"] >= 1


@pytest.mark.parametrize("direction", ["horizontal", "vertical"])
def test_reduced_metric_two_contraction_orders(direction):
    opts = fu.FullUpdate()
    L = Lattice(3, 3)
    st = random_state(L, (2, 1), "peps", seed=9)
    layers = bm.double_layers(st)
    north, south, _ = bm.boundary_environments(layers, 64)
    y, x = 0, 1
    i = (x, y)
    j = (x + 1, y) if direction == "horizontal" else (x, y + 1)
    top = north[y].tensors
    bottom = south[y + (1 if direction == "horizontal" else 2)].tensors
    rowl = bm.row(layers, y) if direction == "horizontal" else \
        [bm.Slab(layers[k][y], layers[k][y + 1]) for k in range(L.Lx)]
    left, right = bm.row_environments(top, rowl, bottom)
    finalx = x + 1 if direction == "horizontal" else x
    env = fu.bond_environment(top, bottom, x, direction, left[x], right[finalx + 1])
    m = TFIM(J=1.0, h=1.0, symmetric=True)
    G = Z2(m.trotter_gate(L, i, j, 0.1), [st[i].p[2], st[j].p[2], st[i].p[2], st[j].p[2]])
    pair = fu.reduced_pair(st[i], st[j], G, direction)
    N1 = fu.reduced_metric(env, pair["Qa"], pair["Qb"])
    N2 = fu.bond_hole_metric(env, pair["Qa"], pair["Qb"])
    assert np.max(np.abs(N1.a - N2.a)) <= 1e-12 * np.max(np.abs(N1.a))
    _ = opts


def test_solve_core_reaches_the_linear_solution():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((30, 30))
    M = A @ A.T + 0.5 * np.eye(30)
    b = rng.standard_normal((30, 2))
    opts = fu.FullUpdate(linear_max_iterations=200, linear_tolerance=1e-12, diagonal_shift=0.0)
    x, res = fu.solve_core(M, None, b, np.zeros_like(b), opts)
    assert res <= 1e-12
    assert np.allclose(x, np.linalg.solve(M, b), atol=1e-9)


def test_gauge_metric_is_the_congruence():
    """working metric = (Li x Ri)^T N (Li x Ri) and left @ Li = I."""
    rng = np.random.default_rng(2)
    p = np.array([0, 0, 1, 1, 1])
    X = rng.standard_normal((25, 25))
    M = X @ X.T
    par = (p[:, None] + p[None, :]).reshape(-1) % 2
    M[par[:, None] != par[None, :]] = 0.0
    N = Z2(M.reshape(5, 5, 5, 5), [p, p, p, p])
    g = fu.environment_gauge(N)
    Li, Ri = g["left_inverse"].a, g["right_inverse"].a
    ref = np.einsum("ijkl,ix,jy,kz,lw->xyzw", N.a, Li, Ri, Li, Ri)
    assert np.allclose(g["metric"].a, ref, atol=1e-12)
    assert np.allclose(g["left"].a @ Li, np.eye(5), atol=1e-10)
