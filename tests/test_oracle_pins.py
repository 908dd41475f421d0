"""Pins of the FP64 oracle that do not come from the oracle itself: closed
forms, brute-force contraction / enumeration, exact diagonalisation,
invariants and published test vectors. The reference's own tests pin nothing
numerically (test/runtests.jl:4-7), so these are the parity anchors."""
import itertools

import numpy as np
import pytest

from oracle.model import (Lattice, TFIM, identity_pepo, random_state, embed_identity_pepo, normalize_state,
                          dense_hamiltonian, exact_thermal_observables, State)
from oracle.z2 import Z2, ein, select_kept, svd_trunc, fused_parity, symmetrize
from oracle import boundary as bm
from oracle import born as ob
from oracle import energy as oe
from oracle import ensemble as oens
from conftest import charged_peps


def dense_network(state):
    """Contract every virtual bond: amplitude over (p, a) of each site, row-major."""
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
            w, e = g(hb, (x, y)), g(hb, (x + 1, y))
            n, s = g(vb, (x, y)), g(vb, (x, y + 1))
            p, a = next(counter), next(counter)
            args += [A.a, [w, n, p, a, e, s]]
            out += [p, a]
    return np.einsum(*args, out, optimize=True)


# ----------------------------- closed forms ---------------------------------

@pytest.mark.parametrize("shape", [(2, 2), (3, 2), (3, 3)])
def test_identity_pepo_lognorm_closed_form(shape):
    """beta = 0: <<I|I>> = Tr(I) = 2^N (src/states/finite.jl:80-86)."""
    L = Lattice(*shape)
    r = bm.lognorm(identity_pepo(L), 4)
    assert abs(r["log_norm"] - len(L) * np.log(2)) < 1e-12


def test_identity_pepo_conditionals_are_half():
    """beta = 0: every Born conditional of the identity PEPO is exactly 1/2."""
    L = Lattice(3, 3)
    cache = ob.SamplingCache(identity_pepo(L), 4)
    res = ob.sample_configuration(cache, uniforms=ob.counter_uniforms(1, 1, 9))
    np.testing.assert_allclose(res["conditionals"], 0.5, atol=1e-14)
    assert abs(res["log_probability"] - 9 * np.log(0.5)) < 1e-12


def test_splitmix64_published_vectors():
    """SplitMix64 (Steele, Lea, Flood 2014) from state 0: the first outputs are
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4 (born.jl:1-6 mixes value + gamma)."""
    gamma = 0x9E3779B97F4A7C15
    assert ob.splitmix64(0) == 0xE220A8397B1DCDAF
    assert ob.splitmix64(gamma) == 0x6E789E6AA1B965F4
    words = ob.sample_rng_words(1, 1)
    assert words[:2] == [0x424F524E, 2]


# ------------------------------ brute force ----------------------------------

@pytest.mark.parametrize("kind,D,sym", [("pepo", (1, 1), True), ("peps", (1, 1), True), ("pepo", 2, False)])
def test_lognorm_equals_brute_force(kind, D, sym):
    L = Lattice(3, 2) if sym else Lattice(2, 3)
    for seed in range(3):
        st = random_state(L, D, kind, seed, symmetric=sym)
        amp = dense_network(st)
        exact = np.log(np.sum(amp ** 2))
        r = bm.lognorm(st, 64)
        assert abs(r["log_norm"] - exact) < 1e-11
        assert r["max_environment_truncation"] < 1e-20


def test_lognorm_gauge_invariance():
    """An invertible matrix and its inverse on one internal bond leave the norm unchanged."""
    L = Lattice(3, 3)
    st = random_state(L, (2, 1), "peps", 4)
    r0 = bm.lognorm(st, 64)["log_norm"]
    g = st.copy()
    A, B = g[(1, 1)], g[(2, 1)]
    p = A.p[4]
    rng = np.random.default_rng(0)
    M = np.zeros((len(p), len(p)))
    for q in (0, 1):
        idx = np.nonzero(p == q)[0]
        M[np.ix_(idx, idx)] = rng.standard_normal((len(idx), len(idx))) + 3 * np.eye(len(idx))
    Mt = Z2(M, [p, p])
    Mi = Z2(np.linalg.inv(M), [p, p])
    g[(1, 1)] = ein("wnpaes,eE->wnpaEs", A, Mt)
    g[(2, 1)] = ein("Ww,wnpaes->Wnpaes", Mi, B)
    assert abs(bm.lognorm(g, 64)["log_norm"] - r0) < 1e-10


def test_born_probabilities_brute_force():
    """At chi >= exact bond: p(alpha) = ||<alpha|Psi>>||^2 / Z for every alpha,
    the conditionals multiply to p, and sum_alpha p = 1 (born.jl:86-163)."""
    L = Lattice(2, 3)
    st = random_state(L, (1, 1), "pepo", 5)
    cache = ob.SamplingCache(st, 64)
    amp = dense_network(st)
    Z = np.sum(amp ** 2)
    total = 0.0
    for bits in itertools.product([1, 2], repeat=6):
        cfg = [[bits[x + 2 * y] for y in range(3)] for x in range(2)]
        r = ob.sample_configuration(cache, configuration=cfg)
        idx = []
        for y in range(3):
            for x in range(2):
                idx += [slice(None), cfg[x][y] - 1]
        exact = np.sum(amp[tuple(idx)] ** 2) / Z
        assert abs(np.exp(r["log_probability"]) - exact) < 1e-12
        assert abs(np.sum(np.log(r["conditionals"])) - r["log_probability"]) < 1e-12
        total += np.exp(r["log_probability"])
    assert abs(total - 1.0) < 1e-12


def test_draw_frequencies_follow_conditionals():
    """Drawing with the uniform u picks the first label whose cumulative mass exceeds u."""
    L = Lattice(2, 2)
    st = random_state(L, (1, 1), "pepo", 6)
    cache = ob.SamplingCache(st, 16)
    for i in range(1, 6):
        u = ob.counter_uniforms(3, i, 4)
        r = ob.sample_configuration(cache, uniforms=u)
        ev = ob.sample_configuration(cache, configuration=r["configuration"])
        assert abs(ev["log_probability"] - r["log_probability"]) < 1e-13


def test_projection_norm_sum_rule():
    """sum_alpha ||P_alpha Psi||^2 = ||Psi||^2 with the projected, normalised PEPS
    carrying log_amplitude (born.jl:27-36, finite.jl:89-100)."""
    L = Lattice(2, 2)
    st = random_state(L, (1, 1), "pepo", 7)
    tot = 0.0
    for bits in itertools.product([1, 2], repeat=4):
        cfg = [[bits[x + 2 * y] for y in range(2)] for x in range(2)]
        ps, rec = ob.project_configuration(st, cfg)
        tot += np.exp(bm.lognorm(ps, 16)["log_norm"])
    assert abs(np.log(tot) - bm.lognorm(st, 16)["log_norm"]) < 1e-11


def test_truncated_north_and_south_contractions_agree():
    """Row-north and row-south contractions of the same network agree to the
    order of their truncation error (no closed form for chi < exact)."""
    L = Lattice(4, 4)
    st = embed_identity_pepo(L, (1, 1), 0.6, 9)
    layers = bm.double_layers(st)
    north, south, mt = bm.boundary_environments(layers, 6)
    vn = np.log(bm.close_boundary(north[4].tensors)) + north[4].log_scale
    vs = np.log(bm.close_boundary(south[0].tensors)) + south[0].log_scale
    assert mt > 0
    assert abs(vn - vs) < 1e2 * np.sqrt(mt) + 1e-12


def test_compression_matches_exact_mps_svd():
    """The first cut of the compression equals the SVD truncation of the exact
    absorbed MPS (canonical-SVD optimality, boundary_mps.jl:94-134)."""
    L = Lattice(3, 3)
    st = random_state(L, (1, 1), "peps", 10)
    layers = bm.double_layers(st)
    env = bm.trivial_boundary(bm.row(layers, 0), "north")
    env = bm.absorb(env, bm.row(layers, 0), 64)
    exact = bm.absorb(env, bm.row(layers, 1), 64)
    trunc = bm.absorb(env, bm.row(layers, 1), 3)

    def vec(b):
        v = np.ones((1,))
        for A in b.tensors:
            v = np.tensordot(v, A.a, axes=([v.ndim - 1], [0]))
        return v.reshape(-1) * np.exp(b.log_scale)
    ve, vt = vec(exact), vec(trunc)
    # the truncated MPS is no better than the best rank-3 approximation at cut 1
    M = ve.reshape(exact.tensors[0].shape[1], -1)
    s = np.linalg.svd(M, compute_uv=False)
    best = np.sqrt(np.sum(s[3:] ** 2))
    err = np.linalg.norm(ve - vt)
    assert err >= best * (1 - 1e-9)
    assert err < 10 * np.linalg.norm(ve) * np.sqrt(trunc.max_truncation) + 1e-12


# ------------------------------ boundary entropy -----------------------------

def test_schmidt_entropy_closed_forms():
    """``_schmidt_entropy`` (boundary_mps.jl:20-30): k equal values give log k,
    one value 0, two values the binary entropy, a zero spectrum NaN."""
    for k in (1, 2, 5, 16):
        assert bm.schmidt_entropy([np.full(k, 0.3)]) == pytest.approx(np.log(k), abs=1e-14)
    assert bm.schmidt_entropy([np.array([2.0]), np.zeros(3)]) == 0.0
    s = np.array([np.sqrt(0.8), np.sqrt(0.2)])
    h = -(0.8 * np.log(0.8) + 0.2 * np.log(0.2))
    assert bm.schmidt_entropy([s[:1], s[1:]]) == pytest.approx(h, abs=1e-15)
    assert np.isnan(bm.schmidt_entropy([np.zeros(2)]))
    e = {"maximum": 0.0}
    bm.record_entropy(e, "north", 1, 1, 0.5)
    bm.record_entropy(e, "north", 1, 1, float("nan"))
    bm.record_entropy(e, "north", 1, 1, 0.7)
    assert np.isnan(e["maximum"])  # Julia's max propagates NaN


def _rows_vector(layers, ys, open_side):
    """Dense contraction of the double-layer rows ``ys`` (top to bottom) with
    np.einsum; open legs: the south legs of the last row (open_side "south")
    or the north legs of the first row ("north"), in site order."""
    Lx = len(layers)
    counter = itertools.count()
    lab = {}

    def g(k):
        if k not in lab:
            lab[k] = next(counter)
        return lab[k]
    args = []
    for y in ys:
        for x in range(Lx):
            E = layers[x][y].a
            args += [E, [g(("h", x, y)), g(("v", x, y)), g(("h", x + 1, y)), g(("v", x, y + 1))]]
    yo = ys[-1] + 1 if open_side == "south" else ys[0]
    out = [g(("v", x, yo)) for x in range(Lx)]
    return np.einsum(*args, out, optimize=True)


def _exact_entropy(v, b):
    s = np.linalg.svd(v.reshape(int(np.prod(v.shape[:b])), -1), compute_uv=False)
    p = s ** 2 / np.sum(s ** 2)
    p = p[p > 0]
    return float(-np.sum(p * np.log(p)))


@pytest.mark.parametrize("kind,seed", [("peps", 21), ("pepo", 22)])
def test_boundary_entropy_profile_is_exact_schmidt_entropy(kind, seed):
    """Without truncation the pre-truncation spectrum of every cut is the exact
    Schmidt spectrum of the boundary vector, so Se.north[r-1, b-1] is the
    entropy of rows 1..r contracted densely and cut after column b, and
    Se.south[r-1, b-1] that of rows r+1..Ly (boundary_mps.jl:193-209,
    energy.jl:31-78)."""
    L = Lattice(3, 3)
    st = random_state(L, (1, 1), kind, seed)
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    r = oe.measure_energy(st, m, 256, record_entropy=True)
    layers = bm.double_layers(st)
    Se = r["Se"]
    assert Se["north"].shape == (2, 2) and Se["south"].shape == (2, 2)
    for rc in (1, 2):
        vn = _rows_vector(layers, list(range(rc)), "south")
        vs = _rows_vector(layers, list(range(rc, 3)), "north")
        for b in (1, 2):
            assert Se["north"][rc - 1, b - 1] == pytest.approx(_exact_entropy(vn, b), abs=1e-10)
            assert Se["south"][rc - 1, b - 1] == pytest.approx(_exact_entropy(vs, b), abs=1e-10)
    # the first and last absorbs close on unit legs: their cuts carry entropy 0
    assert r["max_boundary_entropy"] == pytest.approx(max(Se["north"].max(), Se["south"].max()), abs=1e-14)


def test_boundary_sector_budget():
    """``BoundaryMPS(chi, sector_space)`` (boundary_mps.jl:11-18): a budget of
    chi per parity changes nothing, an active one bounds every bond's sector
    counts and leaves the total within chi."""
    L = Lattice(3, 3)
    st = random_state(L, (2, 1), "pepo", 31)
    layers = bm.double_layers(st)
    plain = bm.boundary_environments(layers, 8)
    wide = bm.boundary_environments(layers, bm.BoundaryMPS(8, (8, 8)))
    for a, b in zip(plain[0] + plain[1], wide[0] + wide[1]):
        for x, y in zip(a.tensors, b.tensors):
            assert np.array_equal(x.a, y.a)
    north, south, _ = bm.boundary_environments(layers, bm.BoundaryMPS(8, (4, 2)))
    narrowed = False
    for env in north[1:] + south[:-1]:
        for A in env.tensors[1:]:
            p = A.p[0]
            assert np.sum(p == 0) <= 4 and np.sum(p == 1) <= 2 and len(p) <= 8
            narrowed |= len(p) == 6
    assert narrowed  # the unbudgeted bonds here are (4, 4) and (5, 3)
    ln_plain = bm.lognorm(st, 8)["log_norm"]
    ln_caps = bm.lognorm(st, bm.BoundaryMPS(8, (4, 2)))["log_norm"]
    assert abs(ln_caps - ln_plain) > 1e-4


# ------------------------------ truncation rules -----------------------------

def test_select_kept_truncrank_and_intersection():
    v0, v1 = np.array([10, 9, 8, 7, 6, 5.0]), np.array([4, 3, 2.0])
    assert select_kept([v0, v1], chi=6) == [6, 0]
    assert select_kept([v0, v1], chi=7) == [6, 1]
    assert select_kept([v0, v1], caps=[3, 3]) == [3, 3]
    # MatrixAlgebraKit `&` keeps the intersection of the kept index sets
    assert select_kept([v0, v1], chi=6, caps=[3, 3]) == [3, 0]
    assert select_kept([v0, v1], atol=4.5) == [6, 0]


def test_svd_trunc_error_is_discarded_two_norm():
    rng = np.random.default_rng(1)
    pr = np.array([0, 1, 0, 1, 0])
    pc = np.array([1, 0, 0, 1])
    M = rng.standard_normal((5, 4))
    M[pr[:, None] != pc[None, :]] = 0
    U, S, Vh, bp, err, spectra = svd_trunc(M, pr, pc, chi=2)
    allv = np.sort(np.concatenate(spectra))[::-1]
    assert abs(err - np.sqrt(np.sum(allv[2:] ** 2))) < 1e-13
    assert np.allclose(U.T @ U, np.eye(2))


# ------------------------------ categorical ---------------------------------

def test_categorical_inverse_cdf():
    """``_categorical`` (selection.jl:68-80): first j with u < cumulative, -Inf skipped."""
    lp = np.log(np.array([0.25, 0.75]))
    assert ob.categorical(lp, 0.0) == 0
    assert ob.categorical(lp, 0.2499999) == 0
    assert ob.categorical(lp, 0.25) == 1
    assert ob.categorical(lp, 0.9999999) == 1
    assert ob.categorical(np.array([-np.inf, 0.0]), 0.3) == 1
    assert ob.categorical(np.array([0.0, -np.inf]), 0.999) == 0
    # u beyond the rounded total returns the last positive entry
    assert ob.categorical(np.array([np.log(0.5), -np.inf]), 0.9999999999) == 0


# ------------------------------ TFIM / ED ------------------------------------

def test_exact_thermal_observables_closed_forms():
    L = Lattice(2, 2)
    beta, h = 0.7, 1.3
    r = exact_thermal_observables(L, TFIM(J=0.0, h=h), beta)
    assert abs(r["energy_per_site"] + h * np.tanh(beta * h)) < 1e-12
    assert abs(r["lnZ"] - 4 * np.log(2 * np.cosh(beta * h))) < 1e-12
    # h = 0: classical Ising on the 2x2 open ring (4 bonds): Z = sum over configs
    r = exact_thermal_observables(L, TFIM(J=1.0, h=0.0), beta)
    zs = [np.exp(beta * sum(s[a] * s[b] for a, b in [(0, 1), (2, 3), (0, 2), (1, 3)]))
          for s in itertools.product([1, -1], repeat=4)]
    assert abs(r["lnZ"] - np.log(sum(zs))) < 1e-12


def test_trotter_gate_is_exponential_of_bond_hamiltonian():
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    G = m.trotter_gate(L, (0, 0), (1, 0), 0.05).reshape(4, 4)
    H = m.bond_hamiltonian(L, (0, 0), (1, 0))
    w, V = np.linalg.eigh(H)
    assert np.allclose(G, V @ np.diag(np.exp(-0.05 * w)) @ V.T, atol=1e-14)


def _dense_charge_hamiltonian(L, m):
    """TFIM in the Z2 charge basis (charge 0 = |+x>): X -> diag(1,-1), Z -> flip."""
    N = len(L)
    dim = 2 ** N
    idx = lambda i: i[0] + i[1] * L.Lx
    b = np.arange(dim)
    H = np.zeros((dim, dim))
    xs = [1 - 2 * ((b >> (N - 1 - k)) & 1) for k in range(N)]
    H[b, b] -= m.h * sum(xs)
    for (i, j) in L.bonds("horizontal") + L.bonds("vertical"):
        H[b ^ (1 << (N - 1 - idx(i))) ^ (1 << (N - 1 - idx(j))), b] -= m.J
    return H


def test_measure_energy_equals_dense_expectation():
    L = Lattice(3, 2)
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    st = random_state(L, (1, 1), "peps", 11)
    r = oe.measure_energy(st, m, 64)
    amp = dense_network(st)  # axes p1 a1 p2 a2 ... (aux dims 1)
    psi = amp.reshape(-1)
    H = _dense_charge_hamiltonian(L, m)
    E = psi @ H @ psi / (psi @ psi)
    assert abs(r["energy_total"] - E) < 1e-10
    assert abs(r["log_norm"] - np.log(psi @ psi)) < 1e-10


def test_energy_of_beta0_sample_is_product_state():
    """At beta = 0 a projected sample is an X-eigenbasis product state:
    E = -h sum_i s_i with s = +1 (charge 0) / -1 (charge 1), <ZZ> = 0."""
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=0.7, symmetric=True)
    pepo = identity_pepo(L)
    cfg = [[1 + ((x + 2 * y) % 2) for y in range(3)] for x in range(3)]
    ps, _ = ob.project_configuration(pepo, cfg)
    r = oe.measure_energy(ps, m, 8)
    s = sum(1 if cfg[x][y] == 1 else -1 for x in range(3) for y in range(3))
    assert abs(r["energy_total"] + m.h * s) < 1e-12


# ------------------------------ statistics -----------------------------------

def test_ensemble_statistics_closed_forms():
    e = np.array([-1.0, -2.0, -3.0])
    r = oens.ensemble_statistics(np.zeros(3), [1, 1, 1], e, 3)
    assert abs(r["eta"] - 1) < 1e-15 and abs(r["energy_per_site"] + 2) < 1e-15
    assert abs(r["energy_per_site_error"] - np.std(e, ddof=1) / np.sqrt(3)) < 1e-14
    r = oens.ensemble_statistics(np.array([0.0, -800.0, -800.0]), [1, 1, 1], e, 3)
    assert abs(r["eta"] - 1 / 3) < 1e-12 and abs(r["energy_per_site"] + 1) < 1e-12
    r = oens.ensemble_statistics(np.zeros(2), [2, 1], np.array([-1.0, -4.0]), 3)
    assert abs(r["energy_per_site"] + 2.0) < 1e-15 and r["virtual_samples"] == 3


def test_temperature_nodes_and_step_schedule():
    assert oens.temperature_nodes(0.5, [0.0, 0.25, 1.0], [0.5, 2.0]) == [0.0, 0.25, 0.5, 1.0, 2.0]
    steps = oens.step_schedule(0.5, 2.0, 0.05)
    assert len(steps) == 15 and abs(steps[-1] - 2.0) == 0
    assert np.allclose(np.diff([0.5] + steps), 0.1)
    assert oens.step_schedule(0.0, 0.25, 0.1) == [0.2, 0.25]


@pytest.mark.parametrize("kind,shape,D,sym", [("peps", (2, 2), (1, 1), True), ("pepo", (3, 2), (1, 1), True),
                                              ("peps", (2, 3), 2, False)])
def test_lognorm_gradient_matches_finite_differences(kind, shape, D, sym):
    """Environment gradient of log<psi|psi> (exact chi) against central finite
    differences h = 1e-6 of the exact lognorm, on neutral entries of every site."""
    from oracle import adjoint as oa
    from oracle.z2 import Z2, total_parity_grid
    L = Lattice(*shape)
    state = random_state(L, D, kind, seed=71, symmetric=sym)
    chi = 64
    grads, _ = oa.lognorm_gradient(state, chi)
    rng = np.random.default_rng(3)
    h = 1e-6
    for (x, y) in L.site_order():
        A = state[(x, y)]
        allowed = np.argwhere(total_parity_grid(A) == 0)
        for idx in allowed[rng.choice(len(allowed), size=min(3, len(allowed)), replace=False)]:
            idx = tuple(idx)
            vals = []
            for sgn in (1, -1):
                a = A.a.copy()
                a[idx] += sgn * h
                state.sites[x][y] = Z2(a, A.p)
                vals.append(bm.lognorm(state, chi)["log_norm"])
            state.sites[x][y] = A
            fd = (vals[0] - vals[1]) / (2 * h)
            g = grads[(x, y)].a[idx]
            assert abs(fd - g) <= 1e-6 * max(1.0, abs(g)), ((x, y), idx, fd, g)


# ----------------------- physical-kind sampling (born.jl) -------------------

@pytest.mark.parametrize("symmetric", [True, False])
def test_physical_born_probabilities_brute_force(symmetric):
    """Physical kind at chi >= exact bond: p(sigma) = |<sigma|Psi>|^2 / Z over
    every physical configuration in the sampling basis (the exact site metric
    is PSD, so the projection only clips round-off), the conditionals multiply
    to p, sum p = 1, and configurations outside the state's total Z2 charge get
    log p = -Inf from the charge check (born.jl:100-104, 114-115)."""
    L = Lattice(2, 3)
    aux = [1, 0, 1, 1, 0, 0]
    st = charged_peps(L, (1, 1) if symmetric else 2, 11, aux, symmetric)
    basis = ob.resolve_basis(symmetric)
    cache = ob.SamplingCache(st, 64, basis=basis, kind="physical")
    amp = dense_network(st)
    idx_p = []
    for y in range(3):
        for x in range(2):
            idx_p += [slice(None), 0]
    amp = amp[tuple(idx_p)]  # physical legs only (unit aux legs)
    # rotate every physical leg to the sampling basis: <u_k| psi>
    for k in range(6):
        amp = np.moveaxis(np.tensordot(basis.U.conj().T, amp, axes=([1], [k])), 0, k)
    Z = np.sum(np.abs(amp) ** 2)
    target = 0
    for q in aux:
        target ^= q
    assert cache.target_charge == (target if symmetric else None)
    total = 0.0
    for ks in itertools.product([0, 1], repeat=6):
        cfg = [[basis.labels[ks[x + 2 * y]] for y in range(3)] for x in range(2)]
        r = ob.sample_configuration(cache, configuration=cfg)
        exact = np.abs(amp[tuple(ks[x + 2 * y] for y in range(3) for x in range(2))]) ** 2 / Z
        charge = 0
        for k in ks:
            charge ^= basis.charges[k]
        if symmetric and charge != target:
            assert r["log_probability"] == -np.inf and r["conditionals"] == []
            assert exact == 0.0
            continue
        assert abs(np.exp(r["log_probability"]) - exact) < 1e-12
        assert abs(np.sum(np.log(r["conditionals"])) - r["log_probability"]) < 1e-12
        assert len([d for d in r["repairs"] if "hermiticity_defect" in d]) == 6
        total += np.exp(r["log_probability"])
    assert abs(total - 1.0) < 1e-12
    # draws land on charge-allowed configurations and evaluate to their own log p
    for i in range(1, 6):
        d = ob.sample_configuration(cache, uniforms=ob.counter_uniforms(4, i, 6))
        ev = ob.sample_configuration(cache, configuration=d["configuration"])
        assert abs(ev["log_probability"] - d["log_probability"]) < 1e-13


def test_positive_metric_closed_form():
    """``_positive_metric`` (tensors.jl:65-94) on a Z2 matrix with a known
    eigendecomposition per parity block: the repaired matrix is V max(D, 0) V^T
    of the scaled input, the root satisfies root^T root = repaired, and the
    report has its closed-form values (eigenvalues of M / ||M||)."""
    rng = np.random.default_rng(3)
    par = np.array([0, 0, 1, 1, 1])  # one leg, rows and columns
    lam = {0: np.array([2.0, -0.5]), 1: np.array([1.0, 0.25, -0.1])}
    M = np.zeros((5, 5))
    Vs = {}
    for q in (0, 1):
        idx = np.nonzero(par == q)[0]
        Q, _ = np.linalg.qr(rng.standard_normal((len(idx), len(idx))))
        Vs[q] = (idx, Q)
        M[np.ix_(idx, idx)] = (Q * lam[q][None, :]) @ Q.T
    N = Z2(M, [par, par])
    s = np.linalg.norm(M)
    rep, report = ob.positive_metric(N, 1)
    root, report2 = ob.positive_metric(N, 1, square_root=True)
    exp = np.zeros((5, 5))
    for q in (0, 1):
        idx, Q = Vs[q]
        exp[np.ix_(idx, idx)] = (Q * np.maximum(lam[q], 0)[None, :]) @ Q.T / s
    assert np.allclose(rep, exp, atol=1e-14)
    assert np.allclose(root.T @ root, exp, atol=1e-14)
    allv = np.concatenate([lam[0], lam[1]]) / s
    assert report == report2
    assert abs(report["minimum_relative_eigenvalue"] - (-0.5 / 2.0)) < 1e-14
    assert report["negative_eigenvalues"] == 2.0 and report["negative_multiplets"] == 2
    assert abs(report["negative_spectral_fraction"]
               - np.sqrt(np.sum(np.minimum(allv, 0) ** 2)) / np.sqrt(np.sum(allv ** 2))) < 1e-14
    assert report["hermiticity_defect"] < 1e-15
    # an antisymmetric part K of a symmetric S: defect = ||2K|| / ||S + K||, spectrum of S
    K = np.zeros((5, 5))
    K[0, 1], K[1, 0] = 0.3, -0.3
    _, r3 = ob.positive_metric(Z2(M + K, [par, par]), 1)
    assert abs(r3["hermiticity_defect"] - np.linalg.norm(2 * K) / np.linalg.norm(M + K)) < 1e-14
    assert abs(r3["minimum_relative_eigenvalue"] - report["minimum_relative_eigenvalue"]) < 1e-14
