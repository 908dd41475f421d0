"""GPU parity of the boundary sector budget (``BoundaryMPS(chi,
sector_space)``, boundary_mps.jl:11-18: truncrank(chi) & truncspace(V) in
every compression) against the FP64 oracle: lognorm, Born draws from a cache
built with a sampling budget, the energy of resident states and a
full-update step with an environment budget. The budget (4, 2) at chi = 8 is
active on these states (their unbudgeted bonds are (4, 4) and (5, 3))."""
import sys

import numpy as np
import pytest

from conftest import to_product_state, oracle_cfg
from oracle.model import Lattice, TFIM, embed_identity_pepo, random_state
from oracle import born as ob
from oracle import boundary as obm
from oracle import energy as oe
from oracle import fu as ofu
from test_gpu_fu import to_oracle
from test_oracle_fu import dense_vector

pytestmark = pytest.mark.gpu

CHI, CAPS = 8, (4, 2)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_lognorm_with_budget(spp, gpu_ctx):
    say("lognorm with a boundary sector budget")
    states = [random_state(Lattice(3, 3), (2, 1), "pepo", seed=31 + k) for k in range(2)]
    got = spp.lognorms(gpu_ctx, [to_product_state(spp, s) for s in states], CHI, sector_space=CAPS)
    plain = spp.lognorms(gpu_ctx, [to_product_state(spp, s) for s in states], CHI)
    for i, s in enumerate(states):
        ref = obm.lognorm(s, obm.BoundaryMPS(CHI, CAPS))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 1e-9 * max(1.0, abs(ref["log_norm"]))
        assert abs(got["max_environment_truncation"][i] - ref["max_environment_truncation"]) <= 1e-9
        assert abs(plain["log_norm"][i] - got["log_norm"][i]) > 1e-5  # the budget is active


def test_draws_with_budget(spp, gpu_ctx):
    say("Born draws with a sampling budget")
    L = Lattice(3, 3)
    pepo = embed_identity_pepo(L, (2, 1), 0.6, seed=33)
    cache = spp.SamplingCache(gpu_ctx, to_product_state(spp, pepo), CHI, sector_space=CAPS)
    uni = np.stack([spp.sample_uniforms(7, i, 9) for i in range(1, 7)])
    got = spp.sample_configurations(cache, uniforms=uni)
    ocache = ob.SamplingCache(pepo, obm.BoundaryMPS(CHI, CAPS))
    np.testing.assert_allclose(cache.south_max_truncation, [e.max_truncation for e in ocache.south], atol=1e-10)
    for i in range(len(uni)):
        ref = ob.sample_configuration(ocache, uniforms=uni[i])
        assert np.array_equal(oracle_cfg(ref, L), got["configuration"][i])
        assert abs(ref["log_probability"] - got["log_probability"][i]) <= 1e-9 * max(1, abs(ref["log_probability"]))
        assert abs(ref["max_environment_truncation"] - got["max_environment_truncation"][i]) <= 1e-9


def test_energy_with_budget(spp, gpu_ctx):
    say("energy of resident states with a measurement budget")
    m = TFIM(J=1.0, h=1.1, symmetric=True)
    states = [random_state(Lattice(3, 3), (2, 1), "pepo", seed=35)]
    devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, s)) for s in states]
    got = spp.measure_energy_states(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), CHI, sector_space=CAPS,
                                    record_entropy=True)
    for i, s in enumerate(states):
        ref = oe.measure_energy(s, m, obm.BoundaryMPS(CHI, CAPS), record_entropy=True)
        assert abs(got["energy_total"][i] - ref["energy_total"]) <= 1e-9 * max(1.0, abs(ref["energy_total"]))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 1e-9 * max(1.0, abs(ref["log_norm"]))
        assert abs(got["max_boundary_entropy"][i] - ref["max_boundary_entropy"]) <= 1e-9


def test_fu_step_with_budget(spp, gpu_ctx):
    say("full-update step with an environment budget")
    m = TFIM(J=1.0, h=1.5, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=20, linear_max_iterations=60)
    o = random_state(Lattice(3, 3), (2, 1), "peps", seed=38)
    dev = spp.DeviceState(gpu_ctx, to_product_state(spp, o))
    method = spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                            opts.linear_max_iterations, opts.linear_tolerance)
    got = spp.evolution_steps(gpu_ctx, [dev], spp.TFIM(m.J, m.h, symmetric=True), 0.05, CHI, method,
                              boundary_space=CAPS)
    ref = ofu.evolution_step(o, m, 0.05, 3, None, opts, obm.BoundaryMPS(CHI, CAPS))
    costs = np.array([r["final_cost"] for r in ref["records"]])
    np.testing.assert_allclose(got["records"][0, :, 1], costs, rtol=1e-6, atol=1e-9)
    ev = to_oracle(dev.download())
    v_dev, v_ref = dense_vector(ev), dense_vector(o)
    err = np.linalg.norm(v_dev - v_ref) / np.linalg.norm(v_ref)
    say(f"  |dpsi|/|psi| = {err:.2e}, env truncation {got['max_environment_truncation'][0]:.3e} "
        f"vs {ref['max_environment_truncation']:.3e}")
    assert err <= 1e-8
    assert abs(got["max_environment_truncation"][0] - ref["max_environment_truncation"]) <= 1e-9
