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


# ---------------------------------------------------------------------------
# PSD repair of the bond metric (full_update.jl:177-340, 491-501)
# ---------------------------------------------------------------------------

def _env_problem(direction, x, y, seed, shape=(3, 3), D=(1, 1), noise=0.0):
    L = Lattice(*shape)
    st = random_state(L, D, "peps", seed=seed)
    layers = bm.double_layers(st)
    north, south, _ = bm.boundary_environments(layers, 4096)
    top = north[y].tensors
    if direction == "horizontal":
        bottom = south[y + 1].tensors
        rowl = bm.row(layers, y)
    else:
        bottom = south[y + 2].tensors
        rowl = [bm.Slab(layers[k][y], layers[k][y + 1]) for k in range(L.Lx)]
    left, right = bm.row_environments(top, rowl, bottom)
    finalx = x + 1 if direction == "horizontal" else x
    lm = left[x]
    if noise:
        # a perturbed environment: the hole is no longer a Gram matrix
        rng = np.random.default_rng(seed + 1)
        lm = Z2(lm.a + noise * np.linalg.norm(lm.a) * rng.standard_normal(lm.shape), lm.p)
        from oracle.z2 import symmetrize
        lm = symmetrize(lm)
    env = fu.bond_environment(top, bottom, x, direction, lm, right[finalx + 1])
    i = (x, y)
    j = (x + 1, y) if direction == "horizontal" else (x, y + 1)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    G = Z2(m.trotter_gate(L, i, j, 0.3), [st[i].p[2], st[j].p[2], st[i].p[2], st[j].p[2]])
    return env, fu.reduced_pair(st[i], st[j], G, direction)


@pytest.mark.parametrize("direction,x,y", [("horizontal", 0, 1), ("vertical", 1, 0)])
def test_repair_leaves_a_psd_hole_unchanged(direction, x, y):
    """The hole of an exact environment is a Gram matrix (PSD): its projection
    is itself, so the repaired metric is the reduced metric over ||H||."""
    env, pair = _env_problem(direction, x, y, seed=5)
    N = fu.reduced_metric(env, pair["Qa"], pair["Qb"])
    Nr, rep = fu.positive_reduced_metric(env, pair["Qa"], pair["Qb"])
    h = fu.bond_hole(env).norm()
    np.testing.assert_allclose(Nr.a * h, N.a, rtol=0, atol=1e-13 * np.max(np.abs(N.a)))
    assert rep["negative_spectral_fraction"] < 1e-12


@pytest.mark.parametrize("direction,x,y", [("horizontal", 1, 1), ("vertical", 1, 0)])
def test_repaired_metric_is_psd_and_costs_are_nonnegative(direction, x, y):
    """With a perturbed (indefinite) environment the unrepaired metric gives
    negative residual costs for some cores; the repaired one is PSD, so every
    residual cost is >= 0."""
    env, pair = _env_problem(direction, x, y, seed=6, noise=0.5)
    N = fu.reduced_metric(env, pair["Qa"], pair["Qb"])
    Nr, rep = fu.positive_reduced_metric(env, pair["Qa"], pair["Qb"])
    assert rep["negative_multiplets"] > 0

    def as_matrix(T):  # [rb, sb, rk, sk] -> [(rk, sk), (rb, sb)]
        ra, sa = T.shape[0], T.shape[1]
        return np.transpose(T.a, (2, 3, 0, 1)).reshape(ra * sa, ra * sa)
    Mn, Mr = as_matrix(N), as_matrix(Nr)
    assert np.min(np.linalg.eigvalsh((Mn + Mn.T) / 2)) < -1e-6 * np.max(np.abs(Mn))
    assert np.min(np.linalg.eigvalsh((Mr + Mr.T) / 2)) > -1e-12 * np.max(np.abs(Mr))
    rng = np.random.default_rng(3)
    for _ in range(20):
        th = rng.standard_normal(Mr.shape[0])
        assert th @ Mr @ th >= -1e-12 * np.abs(th) @ np.abs(Mr) @ np.abs(th)


def test_psd_projection_is_the_frobenius_nearest_psd_matrix():
    """Pins of the host projection (stopeps.jl_amd/repair.py): closed form on
    a diagonal matrix, idempotence on a PSD matrix, block structure, and no
    random PSD matrix closer to the input in the Frobenius norm."""
    import __graft_entry__ as ge
    rep = ge.load_package().repair
    P, r = rep.psd_projection(np.diag([2.0, -1.0, 0.5]), np.array([0, 0, 1]))
    s = np.linalg.norm([2.0, -1.0, 0.5])
    np.testing.assert_allclose(P, np.diag([2.0, 0.0, 0.5]) / s, atol=1e-15)
    assert r["negative_multiplets"] == 1
    rng = np.random.default_rng(8)
    A = rng.standard_normal((6, 6))
    par = np.array([0, 1, 0, 1, 0, 1])
    A[par[:, None] != par[None, :]] = 0.0
    G = A @ A.T
    P, _ = rep.psd_projection(G, par)
    np.testing.assert_allclose(P, G / np.linalg.norm(G), atol=1e-14)
    M = (A + A.T) / 2
    P, _ = rep.psd_projection(M, par)
    Ms = M / np.linalg.norm(M)
    assert np.all(P[par[:, None] != par[None, :]] == 0.0)
    assert np.min(np.linalg.eigvalsh(P)) > -1e-14
    d0 = np.linalg.norm(Ms - P)
    for _ in range(200):
        B = rng.standard_normal((6, 6)) * 0.3
        Q = P + B @ B.T * rng.uniform(0, 0.2)
        assert np.linalg.norm(Ms - Q) >= d0 - 1e-14


@pytest.mark.parametrize("direction,x,y,noise", [("horizontal", 1, 1, 0.5), ("vertical", 1, 0, 0.5),
                                                 ("horizontal", 0, 2, 0.0)])
def test_host_repair_matches_the_oracle(direction, x, y, noise):
    """The product's host repair (stopeps.jl_amd/repair.py), fed the hole and Q
    factors in the library's layouts (parity-sorted legs, pair legs in the
    library's pair order, Q as pseudo sites), against the oracle's restatement."""
    import __graft_entry__ as ge
    rep = ge.load_package().repair
    env, pair = _env_problem(direction, x, y, seed=9, noise=noise)
    Nr, _ = fu.positive_reduced_metric(env, pair["Qa"], pair["Qb"])
    H = fu.bond_hole(env)
    Qa, Qb = pair["Qa"], pair["Qb"]
    ext_a = {"horizontal": (0, 1, 5), "vertical": (0, 1, 4)}[direction]
    ext_b = {"horizontal": (1, 4, 5), "vertical": (0, 4, 5)}[direction]
    core = {"horizontal": (4, 0), "vertical": (5, 1)}[direction]

    def sort_perm(p):
        return np.argsort(p, kind="stable"), int(np.sum(p == 0))
    # hole legs: ket parities from the oracle's fused pair parities
    hole_dims, hole_evens, perms = [], [], []
    Hd = H.a
    for k in range(6):
        m = int(np.sqrt(H.shape[k]))
        pk = H.p[k].reshape(m, m)[:, 0] ^ H.p[k].reshape(m, m)[0, 0]
        perm, e = sort_perm(pk)
        P = rep.pair_order(m, e)
        src = np.empty(m * m, dtype=np.int64)
        for i in range(m):
            for j in range(m):
                src[P[i, j]] = perm[i] * m + perm[j]
        Hd = np.take(Hd, src, axis=k)
        hole_dims.append(m)
        hole_evens.append(e)

    def pseudo(Q, ext, core_slot):
        # Q legs (ext1, ext2, ext3, aux, r) -> pseudo site (W,N,P=aux,A=1,E,S), legs sorted by parity
        perms_ = [sort_perm(p) for p in Q.p]
        arr = Q.a
        for k, (pm, _) in enumerate(perms_):
            arr = np.take(arr, pm, axis=k)
        arr = arr[..., None]  # unit A leg as axis 5
        src = {slot: k for k, slot in enumerate(ext)}
        src[2], src[3], src[core_slot] = 3, 5, 4
        arr = np.transpose(arr, [src[sl] for sl in range(6)])
        dims = [arr.shape[sl] for sl in range(6)]
        evens = [1 if sl == 3 else perms_[src[sl]][1] for sl in range(6)]
        return np.ascontiguousarray(arr), dims, evens, perms_[4][0]
    qa, qad, qae, pra = pseudo(Qa, ext_a, core[0])
    qb, qbd, qbe, prb = pseudo(Qb, ext_b, core[1])
    got, _ = rep.repaired_metric(Hd.reshape(-1), hole_dims, hole_evens, [0, 0, 0, 1, 1, 1], list(ext_a) + list(ext_b),
                                 qa.reshape(-1), qad, qae, core[0], qb.reshape(-1), qbd, qbe, core[1])
    ra, rb = Qa.shape[4], Qb.shape[4]
    Pa = rep.pair_order(ra, qae[core[0]])
    Pb = rep.pair_order(rb, qbe[core[1]])
    want = np.zeros_like(got)
    No = Nr.a  # [rb, sb, rk, sk] in the oracle's leg order
    for rk in range(ra):
        for rbb in range(ra):
            for sk in range(rb):
                for sb in range(rb):
                    want[Pa[rk, rbb], Pb[sk, sb]] = No[pra[rbb], prb[sb], pra[rk], prb[sk]]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.max(np.abs(want)))
