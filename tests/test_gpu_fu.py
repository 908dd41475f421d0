"""GPU parity of the full-update evolution step (spp_fu_evolution_step,
through the C ABI) against the FP64 oracle's evolution_step: the evolved
state as a dense vector (gauge invariant), the per-bond FU costs in sweep
order, and the accumulated log amplitude."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, TFIM, State, embed_identity_pepo, random_state
from oracle import born as ob
from oracle.z2 import Z2
from oracle import fu as ofu
from test_oracle_fu import dense_vector

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_oracle(spp_state):
    L = spp_state.lattice
    sites = [[Z2(spp_state[(x, y)].data, spp_state[(x, y)].parity) for y in range(L.Ly)] for x in range(L.Lx)]
    return State(Lattice(L.Lx, L.Ly), sites, spp_state.kind, beta=spp_state.beta,
                 log_amplitude=spp_state.log_amplitude)


def _compare(spp, ctx, ostates, m, delta, chi, caps, opts, pad=None, steps=1, tol=1e-8):
    ctx.set_precision("fp64")
    devs = []
    for o in ostates:
        s = to_product_state(spp, o)
        if pad is not None:
            s, live = spp.pad_bonds(s, pad)
            devs.append(spp.DeviceState(ctx, s, live=live))
        else:
            devs.append(spp.DeviceState(ctx, s))
    model = spp.TFIM(m.J, m.h, symmetric=m.symmetric)
    method = spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                            opts.linear_max_iterations, opts.linear_tolerance)
    D = sum(caps)
    for step in range(steps):
        got = spp.evolution_steps(ctx, devs, model, delta, chi, method)
        for k, o in enumerate(ostates):
            ref = ofu.evolution_step(o, m, delta, D, caps, opts, chi)
            costs = np.array([r["final_cost"] for r in ref["records"]])
            assert got["records"].shape[1] == len(costs)
            # relative squared residuals: below ~1e-9 they sit at the linear-solve / roundoff floor
            np.testing.assert_allclose(got["records"][k, :, 1], costs, rtol=1e-6, atol=1e-9)
            its = np.array([r["iterations"] for r in ref["records"]])
            assert np.mean(got["records"][k, :, 3] == its) > 0.9
            ev = to_oracle(devs[k].download())
            v_dev, v_ref = dense_vector(ev), dense_vector(o)
            err = np.linalg.norm(v_dev - v_ref) / np.linalg.norm(v_ref)
            say(f"  step {step} state {k}: |dpsi|/|psi| = {err:.2e}, max FU cost {costs.max():.3e}")
            assert err <= tol
            assert abs(ev.log_amplitude - o.log_amplitude) <= 1e-8 * max(1.0, abs(o.log_amplitude))
            assert abs(ev.beta - o.beta) < 1e-14
    return devs


def test_fu_step_2x2_peps(spp, gpu_ctx):
    say("FU step 2x2 Z2 PEPS D=(1,1)")
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    _compare(spp, gpu_ctx, [random_state(Lattice(2, 2), (1, 1), "peps", seed=21)], m, 0.1, 64, (1, 1),
             ofu.FullUpdate())


def test_fu_step_3x3_batch(spp, gpu_ctx):
    say("FU step 3x3 Z2 PEPS D=(1,1), batch of 3, 40 ALS iterations")
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    st = [random_state(Lattice(3, 3), (1, 1), "peps", seed=22 + k) for k in range(3)]
    _compare(spp, gpu_ctx, st, m, 0.1, 64, (1, 1), ofu.FullUpdate(max_iterations=40, linear_max_iterations=100))


def test_fu_step_padded_capacity_two_steps(spp, gpu_ctx):
    """Bonds uploaded at capacity (1,2) while the state lives in (1,1): the
    device works in the live sectors exactly as the reference's growing bond
    space. Truncation binds (costs 1e-5..1e-2); the configuration is one whose
    oracle costs move by < 1e-11 under 1e-13 input noise (many small FU
    problems are not: a flat ALS landscape or metric eigenvalues at the gauge
    floor make the result depend on roundoff, in the reference as well)."""
    say("FU two steps 3x2 Z2 PEPS live (1,1) in capacity (1,2)")
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    _compare(spp, gpu_ctx, [random_state(Lattice(3, 2), (1, 1), "peps", seed=35)], m, 0.3, 64, (1, 2),
             ofu.FullUpdate(max_iterations=40, linear_max_iterations=200), pad=(1, 2), steps=2)


def test_fu_step_pepo(spp, gpu_ctx):
    say("FU step 2x2 Z2 PEPO D=(1,1) (purification)")
    m = TFIM(J=1.0, h=0.7, symmetric=True)
    st = random_state(Lattice(2, 2), (1, 1), "pepo", seed=26)
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    dev = spp.DeviceState(ctx, to_product_state(spp, st))
    got = spp.evolution_step(ctx, dev, spp.TFIM(m.J, m.h, symmetric=True), 0.05, 64)
    ref = ofu.evolution_step(st, m, 0.05, 2, (1, 1), ofu.FullUpdate(), 64)
    np.testing.assert_allclose([r["final_cost"] for r in got["records"]],
                               [r["final_cost"] for r in ref["records"]], rtol=1e-6, atol=1e-9)
    ev = to_oracle(dev.download())
    # PEPO amplitude over (p, a) pairs
    from test_oracle_pins import dense_network
    a1 = dense_network(ev) * np.exp(ev.log_amplitude)
    a2 = dense_network(st) * np.exp(st.log_amplitude)
    assert np.linalg.norm(a1 - a2) <= 1e-8 * np.linalg.norm(a2)


def test_fu_step_dense_states(spp, gpu_ctx):
    say("FU step 2x3 dense PEPS D=2")
    m = TFIM(J=1.0, h=0.8, symmetric=False)
    _compare(spp, gpu_ctx, [random_state(Lattice(2, 3), 2, "peps", seed=27, symmetric=False)], m, 0.1, 64, (2, 0),
             ofu.FullUpdate(max_iterations=20, linear_max_iterations=60))


def test_fu_step_projected_particles_mixed_charges(spp, gpu_ctx):
    """Projected samples of one PEPO carry their label's Z2 charge on the
    auxiliary leg, so a batch of the reference's particles mixes auxiliary
    charges; each evolves as the oracle's single state."""
    say("FU step on projected particles with mixed auxiliary charges, 3x3")
    L = Lattice(3, 3)
    pepo = embed_identity_pepo(L, (1, 1), 0.5, seed=31)
    rng = np.random.default_rng(8)
    cfgs = np.asarray([1, 2])[rng.integers(0, 2, size=(3, 3, 3))]
    states = [ob.project_configuration(pepo, c)[0] for c in cfgs]
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    _compare(spp, gpu_ctx, states, m, 0.1, 64, (1, 1), ofu.FullUpdate(max_iterations=40, linear_max_iterations=100))


def test_fu_step_rank_below_sector_budget(spp, gpu_ctx):
    """truncrank(D) without a sector budget (the reference's bond_space =
    nothing): the device holds each bond at (D, D) per parity and truncates to
    rank D, as the oracle with caps None does. Two steps, 3x2, D=2."""
    say("FU two steps, truncrank D=2 with capacity (2,2) vs no sector budget")
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=40, linear_max_iterations=200)
    o = random_state(Lattice(3, 2), (1, 1), "peps", seed=35)
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    s, live = spp.pad_bonds(to_product_state(spp, o), (2, 2))
    dev = spp.DeviceState(ctx, s, live=live)
    model = spp.TFIM(m.J, m.h, symmetric=True)
    method = spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                            opts.linear_max_iterations, opts.linear_tolerance)
    for step in range(2):
        got = spp.evolution_steps(ctx, [dev], model, 0.3, 64, method, D=2)
        ref = ofu.evolution_step(o, m, 0.3, 2, None, opts, 64)
        np.testing.assert_allclose(got["records"][0, :, 1], [r["final_cost"] for r in ref["records"]],
                                   rtol=1e-6, atol=1e-9)
        assert np.all(got["records"][0, :, 6] + got["records"][0, :, 7] <= 2)
        ev = to_oracle(dev.download())
        err = np.linalg.norm(dense_vector(ev) - dense_vector(o)) / np.linalg.norm(dense_vector(o))
        say(f"  step {step}: |dpsi|/|psi| = {err:.2e}")
        assert err <= 1e-8


def test_fu_step_fp32_contractions(spp, gpu_ctx):
    """Mixed precision: environment and metric contractions in FP32 on the
    matrix cores, the bond solve (gauge, SVD, ALS, PCG) in FP64. The evolved
    state stays within 1e-5 of the FP64 oracle's."""
    say("FU step with FP32 contractions, 3x3 batch of 2")
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=40, linear_max_iterations=100)
    st = [random_state(Lattice(3, 3), (1, 1), "peps", seed=22 + k) for k in range(2)]
    ctx = gpu_ctx
    ctx.set_precision("fp32")
    try:
        devs = [spp.DeviceState(ctx, to_product_state(spp, o)) for o in st]
        got = spp.evolution_steps(ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), 0.1, 64,
                                  spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                                                 opts.linear_max_iterations, opts.linear_tolerance))
    finally:
        ctx.set_precision("fp64")
    for k, o in enumerate(st):
        ref = ofu.evolution_step(o, m, 0.1, 2, (1, 1), opts, 64)
        costs = np.array([r["final_cost"] for r in ref["records"]])
        np.testing.assert_allclose(got["records"][k, :, 1], costs, rtol=1e-3, atol=1e-7)
        ev = to_oracle(devs[k].download())
        err = np.linalg.norm(dense_vector(ev) - dense_vector(o)) / np.linalg.norm(dense_vector(o))
        say(f"  state {k}: |dpsi|/|psi| = {err:.2e}")
        assert err <= 1e-5
        assert abs(ev.log_amplitude - o.log_amplitude) <= 1e-5 * max(1.0, abs(o.log_amplitude))


@pytest.mark.parametrize("kind", ["peps", "pepo"])
def test_fu_step_single_site(spp, gpu_ctx, kind):
    """One-site lattice (full_update.jl:568-570): the onsite gate
    exp(-delta h_onsite) on the physical leg at the full delta, then
    normalize_state!, against the oracle's step."""
    say(f"full-update step on a 1x1 {kind}")
    m = TFIM(J=1.0, h=0.8, symmetric=True)
    o = random_state(Lattice(1, 1), (1, 1), kind, seed=5)
    dev = spp.DeviceState(gpu_ctx, to_product_state(spp, o))
    got = spp.evolution_steps(gpu_ctx, [dev], spp.TFIM(m.J, m.h, symmetric=True), 0.1, 8, record_entropy=True)
    ref = ofu.evolution_step(o, m, 0.1, 2, (1, 1), ofu.FullUpdate(), 8)
    ev = to_oracle(dev.download())
    np.testing.assert_allclose(ev[(0, 0)].a, o[(0, 0)].a, atol=1e-13)
    assert abs(ev.log_amplitude - o.log_amplitude) < 1e-13 and abs(ev.beta - o.beta) < 1e-15
    assert got["records"].shape == (1, 0, 8)
    assert abs(got["normalizations"][0, 0] - ref["normalizations"][0]["log_normalization"]) < 1e-13
    assert got["max_environment_truncation"][0] == 0.0 and got["max_boundary_entropy"][0] == 0.0


def test_results_do_not_depend_on_batch_mates(spp, gpu_ctx):
    """With the subspace truncation eigensolver in use (ksub below the block
    size), a state's lognorm and full-update step come out bit for bit the same
    whether it is stepped alone or in a batch with other states."""
    say("batch independence of the iterative truncation: lognorm and FU step")
    m = TFIM(J=1.0, h=1.5, symmetric=True)
    model = spp.TFIM(m.J, m.h, symmetric=True)
    method = spp.FullUpdate(10, 1e-8, 1e-10, 40, 1e-9)
    ostates = [random_state(Lattice(3, 3), (2, 1), "peps", seed=38 + k) for k in range(3)]
    gpu_ctx.set_option("ksub", 12)
    try:
        batch = spp.lognorms(gpu_ctx, [to_product_state(spp, o) for o in ostates], 8)
        alone = spp.lognorms(gpu_ctx, [to_product_state(spp, ostates[2])], 8)
        assert batch["log_norm"][2] == alone["log_norm"][0]
        devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, o)) for o in ostates]
        rb = spp.evolution_steps(gpu_ctx, devs, model, 0.05, 8, method)
        solo = spp.DeviceState(gpu_ctx, to_product_state(spp, ostates[2]))
        rs = spp.evolution_steps(gpu_ctx, [solo], model, 0.05, 8, method)
    finally:
        gpu_ctx.set_option("ksub", 64)
    np.testing.assert_array_equal(rb["records"][2], rs["records"][0])
    a, b = devs[2].download(), solo.download()
    assert a.log_amplitude == b.log_amplitude
    for xy in a.lattice.site_order():
        np.testing.assert_array_equal(a[xy].data, b[xy].data)


def test_failed_sample_is_isolated_and_refused(spp, gpu_ctx):
    """Failure handling of a batched step, forced from the host (the test hook
    marks one bond update of one sample as failed, every kernel runs
    normally): the other samples finish the step exactly as without the
    failure, the failed sample reports its code and keeps its sites, and later
    calls refuse it."""
    say("full update: a failed sample in a batch of 3")
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    model = spp.TFIM(m.J, m.h, symmetric=True)
    method = spp.FullUpdate(10, 1e-8, 1e-10, 50, 1e-9)
    ostates = [random_state(L, (1, 1), "peps", seed=70 + k) for k in range(3)]
    ref = [spp.DeviceState(ctx, to_product_state(spp, o)) for o in ostates]
    got = [spp.DeviceState(ctx, to_product_state(spp, o)) for o in ostates]
    r_ref = spp.evolution_steps(ctx, ref, model, 0.05, 8, method)
    try:
        ctx.set_option("fu_fail_sample", 1)
        ctx.set_option("fu_fail_bond", 3)
        ctx.set_option("fu_fail_code", 4)
        with pytest.raises(spp.FullUpdateError) as err:
            spp.evolution_steps(ctx, got, model, 0.05, 8, method, repair=False)
    finally:
        ctx.set_option("fu_fail_sample", -1)
    assert list(err.value.status) == [0, 4, 0] and "needs PSD repair" in str(err.value)
    for k in (0, 2):
        a, b = got[k].download(), ref[k].download()
        for xy in L.site_order():
            np.testing.assert_allclose(a[xy].data, b[xy].data, rtol=1e-12, atol=1e-14)
        assert abs(a.log_amplitude - b.log_amplitude) <= 1e-12 * max(1.0, abs(b.log_amplitude))
    # the failed sample is refused by the compute entry points
    with pytest.raises(spp._lib.HIPLibraryError, match="invalidated"):
        spp.lognorms(ctx, [got[1]], 8)
    with pytest.raises(spp._lib.HIPLibraryError, match="invalidated"):
        spp.evolution_steps(ctx, [got[1]], model, 0.05, 8, method)
    # it kept finite sites (no write after the failing bond)
    assert all(np.all(np.isfinite(got[1].download()[xy].data)) for xy in L.site_order())


def test_psd_repair_branch_matches_the_oracle(spp, gpu_ctx):
    """The PSD-repair branch of update_bond! (full_update.jl:491-501), taken
    at one bond update of one sample (forced from the host through the test
    hook; every kernel runs normally): the device contracts that sample's
    complete hole, the host projects it and the bond is solved again on the
    repaired metric. Checked against the oracle taking the same branch at the
    same bond, the other sample against the plain oracle step."""
    say("full update: PSD repair of bond update 5 of sample 1")
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    model = spp.TFIM(m.J, m.h, symmetric=True)
    opts = ofu.FullUpdate(10, 1e-8, 1e-10, 50, 1e-9)
    method = spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                            opts.linear_max_iterations, opts.linear_tolerance)
    ostates = [random_state(L, (1, 1), "peps", seed=90 + k) for k in range(2)]
    devs = [spp.DeviceState(ctx, to_product_state(spp, o)) for o in ostates]
    ctx.reset_counters()
    try:
        ctx.set_option("fu_fail_sample", 1)
        ctx.set_option("fu_fail_bond", 5)
        ctx.set_option("fu_fail_code", 4)
        r = spp.evolution_steps(ctx, devs, model, 0.05, 16, method)
    finally:
        ctx.set_option("fu_fail_sample", -1)
    assert list(r["status"]) == [0, 0]
    assert len(r["repairs"]) == 1 and r["repairs"][0]["sample"] == 1 and r["repairs"][0]["bond"] == 5
    assert ctx.counters()["fu_repairs"] == 1
    for k, o in enumerate(ostates):
        ref = ofu.evolution_step(o, m, 0.05, 2, (1, 1), opts, 16, force_repair=(5,) if k == 1 else ())
        if k == 1:
            assert len(ref["records"][5]["repairs"]) == 1
        costs = np.array([rr["final_cost"] for rr in ref["records"]])
        np.testing.assert_allclose(r["records"][k, :, 1], costs, rtol=1e-6, atol=1e-9)
        v_dev, v_ref = dense_vector(to_oracle(devs[k].download())), dense_vector(o)
        err = np.linalg.norm(v_dev - v_ref) / np.linalg.norm(v_ref)
        say(f"  state {k}: |dpsi|/|psi| = {err:.2e}")
        assert err <= 1e-8


def test_psd_repair_that_does_not_remove_the_failure(spp, gpu_ctx):
    """A repaired metric on which the residual check still fails (here the
    host returns the negated projection) fails the sample with the
    reference's message; the other sample finishes."""
    say("full update: a repair that does not remove the residual failure")
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    L = Lattice(3, 3)
    model = spp.TFIM(1.0, 2.0, symmetric=True)
    method = spp.FullUpdate(10, 1e-8, 1e-10, 50, 1e-9)
    devs = [spp.DeviceState(ctx, to_product_state(spp, random_state(L, (1, 1), "peps", seed=95 + k)))
            for k in range(2)]

    class Negated(spp.repair.RepairCallback):
        def transform(self, metric):
            metric *= -1.0
    try:
        ctx.set_option("fu_fail_sample", 0)
        ctx.set_option("fu_fail_bond", 2)
        ctx.set_option("fu_fail_code", 4)
        r = spp.evolution_steps(ctx, devs, model, 0.05, 16, method, repair=Negated(), check=False)
    finally:
        ctx.set_option("fu_fail_sample", -1)
    assert list(r["status"]) == [11, 0]
    err = spp.FullUpdateError(r["status"])
    assert "PSD repair did not remove the FU residual failure" in str(err)
