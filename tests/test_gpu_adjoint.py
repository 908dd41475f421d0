"""GPU parity of the log-norm gradient (spp_lognorm_grad_batch, through the C
ABI) against the FP64 oracle's environment gradient, against central finite
differences of the device's own lognorm at exact chi, and the product-state
closed form."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, embed_identity_pepo, identity_pepo, random_state
from oracle import adjoint as oa
from oracle.z2 import total_parity_grid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(spp, ctx, states, chi, tol, precision="fp64"):
    """Every site gradient within tol * max|reference gradient| of the oracle's."""
    ctx.set_precision(precision)
    try:
        grads, info = spp.lognorm_gradients(ctx, [to_product_state(spp, s) for s in states], chi)
    finally:
        ctx.set_precision("fp64")
    assert np.all(info["status"] == 0)
    for i, s in enumerate(states):
        ref, ln = oa.lognorm_gradient(s, chi)
        scale = max(np.max(np.abs(r.a)) for r in ref.values())
        for xy, r in ref.items():
            err = np.max(np.abs(grads[i][xy] - r.a))
            assert err <= tol * scale, (i, xy, err, scale)
        assert abs(info["log_norm"][i] - ln["log_norm"]) <= 10 * tol * max(1.0, abs(ln["log_norm"]))
    return grads, info


@pytest.mark.parametrize("kind", ["peps", "pepo"])
def test_gradient_exact_chi(spp, gpu_ctx, kind):
    say("lognorm gradient exact chi, 3x3", kind)
    L = Lattice(3, 3)
    _compare(spp, gpu_ctx, [random_state(L, (1, 1), kind, seed=80 + k) for k in range(2)], 16, 1e-10)


def test_gradient_dense_and_uneven(spp, gpu_ctx):
    say("lognorm gradient dense 2x3 D=2 and uneven Z2 D=(2,1)")
    _compare(spp, gpu_ctx, [random_state(Lattice(2, 3), 2, "peps", seed=82, symmetric=False)], 16, 1e-10)
    _compare(spp, gpu_ctx, [random_state(Lattice(3, 3), (2, 1), "peps", seed=83)], 81, 1e-10)


def test_gradient_truncated_environment(spp, gpu_ctx):
    """Truncated chi: the same environment gradient (truncations fixed) as the oracle."""
    say("lognorm gradient truncated chi, 4x4 PEPO batch")
    L = Lattice(4, 4)
    _compare(spp, gpu_ctx, [embed_identity_pepo(L, (1, 1), 0.6, seed=84 + k) for k in range(3)], 6, 1e-8)


def test_gradient_fp32(spp, gpu_ctx):
    say("lognorm gradient fp32, 3x3 PEPS")
    _compare(spp, gpu_ctx, [random_state(Lattice(3, 3), (1, 1), "peps", seed=86)], 16, 1e-4, precision="fp32")


def test_gradient_4x4_d3_truncated_environment(spp, gpu_ctx):
    """4x4 Z2 PEPS, D = 3 as (2, 1), seed 7, chi = 36 < 81 (the exact centre bond):
    the truncated environment gradient against the oracle's."""
    say("lognorm gradient 4x4 D=3 chi=36 vs oracle environments")
    _compare(spp, gpu_ctx, [random_state(Lattice(4, 4), (2, 1), "peps", seed=7)], 36, 1e-8)


def test_gradient_matches_device_finite_differences(spp, gpu_ctx):
    """4x4 Z2 PEPS, D = 3 as (2, 1), seed 7, at the exact chi = 81: central
    differences (h = 1e-6) of the device's own lognorm, and the oracle."""
    say("lognorm gradient vs device finite differences, 4x4 Z2 D=(2,1) chi=81")
    L = Lattice(4, 4)
    ostate = random_state(L, (2, 1), "peps", seed=7)
    _compare(spp, gpu_ctx, [ostate], 81, 1e-10)
    state = to_product_state(spp, ostate)
    chi = 81
    g, _ = spp.lognorm_gradient(gpu_ctx, state, chi)
    rng = np.random.default_rng(5)
    h = 1e-6
    for (x, y) in [(0, 0), (1, 2), (2, 1), (3, 3)]:
        A = ostate[(x, y)]
        allowed = np.argwhere(total_parity_grid(A) == 0)
        for idx in allowed[rng.choice(len(allowed), size=3, replace=False)]:
            idx = tuple(idx)
            pert = []
            for sgn in (1, -1):
                a = A.a.copy()
                a[idx] += sgn * h
                site = spp.Site(a, A.p)
                sites = [[site if (xx, yy) == (x, y) else state[(xx, yy)] for yy in range(L.Ly)]
                         for xx in range(L.Lx)]
                pert.append(spp.FiniteState(state.lattice, sites, state.kind, beta=state.beta,
                                            log_amplitude=state.log_amplitude))
            ln = spp.lognorms(gpu_ctx, pert, chi)["log_norm"]
            fd = (ln[0] - ln[1]) / (2 * h)
            assert abs(fd - g[(x, y)][idx]) <= 1e-6 * max(1.0, abs(g[(x, y)][idx])), ((x, y), idx, fd)


def test_gradient_norm_identity(spp, gpu_ctx):
    """Euler: log<psi|psi> is homogeneous of degree 2 in each site, so
    <A_i, grad_i> = 2 for every site (any chi, truncations fixed)."""
    say("lognorm gradient Euler identity, 4x4 PEPS truncated")
    L = Lattice(4, 4)
    ostate = random_state(L, (1, 1), "peps", seed=88)
    g, _ = spp.lognorm_gradient(gpu_ctx, to_product_state(spp, ostate), 8)
    for xy in L.site_order():
        assert abs(np.sum(ostate[xy].a * g[xy]) - 2.0) < 1e-10, xy


def test_gradient_product_state_closed_form(spp, gpu_ctx):
    """Identity PEPO (unit bonds): <<I|I>> = prod_i |A_i|^2, so grad_i = 2 A_i / |A_i|^2."""
    say("lognorm gradient identity PEPO closed form")
    L = Lattice(3, 2)
    ostate = identity_pepo(L, symmetric=True)
    g, ln = spp.lognorm_gradient(gpu_ctx, to_product_state(spp, ostate), 4)
    for xy in L.site_order():
        a = ostate[xy].a
        np.testing.assert_allclose(g[xy], 2.0 * a / np.sum(a * a), atol=1e-12)
    assert abs(ln - 6 * np.log(2.0)) < 1e-12


def test_lognorm_rrule_pullback(spp, gpu_ctx):
    """The ChainRules form: the value is lognorm and pullback(dy) scales the
    gradient (identity PEPO closed form: grad_i = 2 A_i / |A_i|^2)."""
    say("lognorm rrule pullback")
    L = Lattice(3, 2)
    ostate = identity_pepo(L, symmetric=True)
    y, pullback = spp.lognorm_rrule(gpu_ctx, to_product_state(spp, ostate), 4)
    assert abs(y - 6 * np.log(2.0)) < 1e-12
    ct = pullback(-0.5)
    for xy in L.site_order():
        a = ostate[xy].a
        np.testing.assert_allclose(ct[xy], -0.5 * 2.0 * a / np.sum(a * a), atol=1e-12)


def test_gradient_empty_batch(spp, gpu_ctx):
    say("lognorm gradient empty batch")
    grads, info = spp.lognorm_gradients(gpu_ctx, [], 4)
    assert grads == [] and info["log_norm"].shape == (0,)


def test_gradient_8x8_d6_chi36_against_oracle_fixture(spp, gpu_ctx):
    """The gradient workload itself (8x8 Z2 PEPS, D=6 as (3,3), seed 7, chi=36,
    ~68% discarded weight per cut): six site gradients against the oracle's
    explicit environments, stored by tests/golden/make_adjoint_8x8.py."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "adjoint_8x8_d6_chi36.npz")
    if not os.path.exists(path):
        pytest.skip("fixture not generated (python tests/golden/make_adjoint_8x8.py)")
    ref = np.load(path)
    say("lognorm gradient 8x8 D=6 chi=36 vs oracle fixture")
    gpu_ctx.set_precision("fp64")
    state = spp.random_peps(spp.Lattice(8, 8), (3, 3), seed=7)
    g, ln = spp.lognorm_gradient(gpu_ctx, state, 36)
    assert abs(ln - float(ref["log_norm"])) <= 1e-10 * abs(float(ref["log_norm"]))
    for x, y in ref["sites"]:
        r = ref[f"g_{x}_{y}"]
        err = np.max(np.abs(g[(int(x), int(y))] - r))
        assert err <= 1e-8 * np.max(np.abs(r)), ((x, y), err)
