"""CPU tests of the full-update bond solve the device runs (the same source
built for the host: spp_selftest_fu_bond) against the oracle's
_initialize_pair + _optimize_pair on bond problems from exact environments."""
import numpy as np
import pytest

from oracle.model import Lattice, TFIM, random_state
from oracle.z2 import Z2
from oracle import boundary as bm
from oracle import fu


def bond_problem(shape, D, direction, x, y, seed, delta=0.3, h=2.0):
    L = Lattice(*shape)
    m = TFIM(J=1.0, h=h, symmetric=True)
    st = random_state(L, D, "peps", seed=seed)
    i = (x, y)
    j = (x + 1, y) if direction == "horizontal" else (x, y + 1)
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
    env = fu.bond_environment(top, bottom, x, direction, left[x], right[finalx + 1])
    G = Z2(m.trotter_gate(L, i, j, delta), [st[i].p[2], st[j].p[2], st[i].p[2], st[j].p[2]])
    pair = fu.reduced_pair(st[i], st[j], G, direction)
    N = fu.reduced_metric(env, pair["Qa"], pair["Qb"])
    return pair, N


def run_device_solve(spp, pair, N, Dcap, opts):
    ra = pair["Qa"].shape[-1]
    rb = pair["Qb"].shape[-1]
    ea = int(np.sum(pair["Qa"].p[-1] == 0))
    eb = int(np.sum(pair["Qb"].p[-1] == 0))
    D = sum(Dcap)
    dims = np.array([ra, ea, rb, eb, 2, 1, D, Dcap[0]], dtype=np.int32)
    o = np.array([opts.max_iterations, opts.error_tolerance, opts.diagonal_shift, opts.linear_max_iterations,
                  opts.linear_tolerance])
    A = np.zeros((ra, 2, D))
    B = np.zeros((D, 2, rb))
    rec = np.zeros(12)
    lib = spp._lib.load()
    rc = lib.spp_selftest_fu_bond(spp._lib.iptr(dims), spp._lib.dptr(o), spp._lib.dptr(np.ascontiguousarray(N.a)),
                                  spp._lib.dptr(np.ascontiguousarray(pair["target"].a)), spp._lib.dptr(A),
                                  spp._lib.dptr(B), spp._lib.dptr(rec))
    assert rc == 0
    return A, B, rec


STOP = {"disabled": 0, "iteration_limit": 1, "error_tolerance": 2, "residual_fallback": 3}


@pytest.mark.parametrize("direction,x,y,Dst,Dcap,its", [
    ("horizontal", 1, 1, (2, 2), (1, 1), 40),
    ("horizontal", 0, 0, (2, 2), (2, 1), 4),
    ("vertical", 1, 0, (2, 2), (1, 1), 40),
    ("vertical", 2, 1, (2, 1), (2, 2), 40),
    ("horizontal", 1, 2, (1, 1), (2, 2), 40),
])
def test_device_bond_solve_matches_oracle(spp, direction, x, y, Dst, Dcap, its):
    pair, N = bond_problem((3, 3), Dst, direction, x, y, seed=17)
    opts = fu.FullUpdate(max_iterations=its, linear_max_iterations=100)
    init = fu.initialize_pair(pair, N, sum(Dcap), Dcap)
    ref = fu.optimize_pair(init, N, opts)
    A, B, rec = run_device_solve(spp, pair, N, Dcap, opts)
    assert int(rec[7]) == 0, rec
    P_ref = np.einsum("rpk,kqs->rpqs", ref["left"].a, ref["right"].a)
    P = np.einsum("rpk,kqs->rpqs", A, B)
    assert np.max(np.abs(P - P_ref)) <= 1e-8 * np.max(np.abs(P_ref))
    assert abs(rec[0] - ref["initial_cost"]) <= 1e-9 * max(ref["initial_cost"], 1e-12) + 1e-14
    assert abs(rec[1] - ref["final_cost"]) <= 1e-7 * max(ref["final_cost"], 1e-12) + 1e-14
    assert int(rec[4]) == ref["iterations"]
    assert int(rec[6]) == STOP[ref["stop_reason"]]


@pytest.mark.parametrize("kind", [0, 1])
def test_jacobi_solves_report_the_sweep_cap(spp, kind):
    """The bond solve's eigensolver (kind 0) and SVD (kind 1), compiled for
    the host: converged within the cap they match LAPACK; a solve stopped at
    a cap too small to converge is reported as not converged (the device
    turns that into bond status 8, where the reference raises after its
    gesdd -> gesvd retry, src/states/svd.jl:1-11)."""
    lib = spp._lib.load()
    rng = np.random.default_rng(11 + kind)
    m, n = (40, 40) if kind == 0 else (50, 30)
    A = rng.standard_normal((m, n))
    if kind == 0:
        A = A + A.T
    vals = np.zeros(n)
    conv = np.zeros(1, dtype=np.int32)
    assert lib.spp_selftest_jacobi(kind, m, n, spp._lib.dptr(np.ascontiguousarray(A)), 80, spp._lib.dptr(vals),
                                   spp._lib.iptr(conv)) == 0
    assert conv[0] == 1
    ref = np.linalg.eigvalsh(A) if kind == 0 else np.linalg.svd(A, compute_uv=False)
    got = np.sort(vals) if kind == 0 else vals
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.max(np.abs(ref)))
    assert lib.spp_selftest_jacobi(kind, m, n, spp._lib.dptr(np.ascontiguousarray(A)), 1, spp._lib.dptr(vals),
                                   spp._lib.iptr(conv)) == 0
    assert conv[0] == 0
    # bad shapes are rejected
    assert lib.spp_selftest_jacobi(kind, 3, 5, spp._lib.dptr(np.zeros(15)), 80, spp._lib.dptr(np.zeros(5)),
                                   spp._lib.iptr(conv)) == -3
