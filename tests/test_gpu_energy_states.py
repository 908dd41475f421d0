"""GPU parity of the energy measurement of resident states
(spp_measure_energy_states, through the C ABI): against the FP64 oracle's
measure_energy on the same states, against the dense expectation value
<psi|H|psi>/<psi|psi> at exact chi, against the projected-sample path on the
same samples, and on states evolved by a full-update step."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, TFIM, embed_identity_pepo, random_state
from oracle import born as ob
from oracle import energy as oe
from oracle import fu as ofu
from test_oracle_pins import dense_network, _dense_charge_hamiltonian
from test_gpu_fu import to_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(spp, ctx, states, m, chi, tol, precision="fp64"):
    ctx.set_precision(precision)
    try:
        devs = [spp.DeviceState(ctx, to_product_state(spp, s)) for s in states]
        got = spp.measure_energy_states(ctx, devs, spp.TFIM(m.J, m.h, symmetric=m.symmetric), chi)
    finally:
        ctx.set_precision("fp64")
    for i, st in enumerate(states):
        ref = oe.measure_energy(st, m, chi)
        e_ref = ref["energy_total"]
        assert abs(got["energy_total"][i] - e_ref) <= tol * max(1.0, abs(e_ref)), (i, got["energy_total"][i], e_ref)
        assert abs(got["energy_per_site"][i] - ref["energy_per_site"]) <= tol * max(1.0, abs(e_ref))
        assert abs(got["tensor_log_norm"][i] - ref["tensor_log_norm"]) <= 10 * tol * max(1.0, abs(ref["tensor_log_norm"]))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 10 * tol * max(1.0, abs(ref["log_norm"]))
        mt = ref["max_environment_truncation"]
        assert abs(got["max_environment_truncation"][i] - mt) <= max(1e-9 if precision == "fp64" else 1e-6, 1e-3 * mt)
    assert np.all(got["status"] == 0)
    return got


def test_energy_states_dense_expectation_z2(spp, gpu_ctx):
    """Exact chi on 3x2 Z2 PEPS: the device energy is <psi|H|psi>/<psi|psi> of
    the dense state, and log_norm is log <psi|psi>."""
    say("energy of resident states vs dense expectation, Z2 3x2")
    L = Lattice(3, 2)
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    states = [random_state(L, (1, 1), "peps", seed=41 + k) for k in range(3)]
    got = _compare(spp, gpu_ctx, states, m, 64, 1e-10)
    H = _dense_charge_hamiltonian(L, m)
    for i, st in enumerate(states):
        psi = dense_network(st).reshape(-1)
        E = psi @ H @ psi / (psi @ psi)
        assert abs(got["energy_total"][i] - E) < 1e-10
        assert abs(got["log_norm"][i] - (np.log(psi @ psi) + 2 * st.log_amplitude)) < 1e-10


def test_energy_states_truncated_z2(spp, gpu_ctx):
    say("energy of resident states, truncated chi, Z2 4x4 D=(2,1)")
    L = Lattice(4, 4)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    states = [random_state(L, (2, 1), "peps", seed=51 + k) for k in range(2)]
    _compare(spp, gpu_ctx, states, m, 6, 1e-8)


def test_energy_states_dense_peps(spp, gpu_ctx):
    say("energy of resident states, dense D=2 3x3")
    L = Lattice(3, 3)
    m = TFIM(J=0.8, h=1.1, symmetric=False)
    states = [random_state(L, 2, "peps", seed=61 + k, symmetric=False) for k in range(2)]
    _compare(spp, gpu_ctx, states, m, 8, 1e-9)


def test_energy_states_pepo_traces_aux(spp, gpu_ctx):
    say("energy of a resident PEPO (aux traced), Z2 2x3")
    L = Lattice(2, 3)
    m = TFIM(J=1.0, h=0.7, symmetric=True)
    _compare(spp, gpu_ctx, [random_state(L, (1, 1), "pepo", seed=71)], m, 32, 1e-10)


def test_energy_states_match_projected_samples(spp, gpu_ctx):
    """Projected samples uploaded as states give the cache path's energies."""
    say("energy of projected samples: resident-state path vs sampling-cache path, 3x3")
    L = Lattice(3, 3)
    pepo = embed_identity_pepo(L, (1, 1), 0.5, seed=31)
    rng = np.random.default_rng(5)
    cfgs = np.asarray([1, 2])[rng.integers(0, 2, size=(4, 3, 3))]
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    cache = spp.SamplingCache(ctx, to_product_state(spp, pepo), 16)
    model = spp.TFIM(1.0, 2.0, symmetric=True)
    a = spp.measure_energies(cache, cfgs, model, chi=16)
    proj = [ob.project_configuration(pepo, c)[0] for c in cfgs]
    b = spp.measure_energy_states(ctx, [to_product_state(spp, p) for p in proj], model, 16)
    np.testing.assert_allclose(b["energy_total"], a["energy_total"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(b["log_norm"], a["log_norm"], rtol=1e-10, atol=1e-10)


def test_energy_states_after_full_update(spp, gpu_ctx):
    """A full-update step on the device, then the energy of the evolved states,
    against the oracle's step and energy."""
    say("energy after a full-update step, Z2 3x3 D=(1,1)")
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=40, linear_max_iterations=100)
    st = [random_state(L, (1, 1), "peps", seed=22 + k) for k in range(2)]
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    devs = [spp.DeviceState(ctx, to_product_state(spp, s)) for s in st]
    spp.evolution_steps(ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), 0.1, 64,
                        spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                                       opts.linear_max_iterations, opts.linear_tolerance))
    got = spp.measure_energy_states(ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), 64)
    for k, s in enumerate(st):
        ofu.evolution_step(s, m, 0.1, 2, (1, 1), opts, 64)
        ref = oe.measure_energy(s, m, 64)
        assert abs(got["energy_total"][k] - ref["energy_total"]) <= 1e-8 * max(1.0, abs(ref["energy_total"]))
        assert abs(got["log_norm"][k] - ref["log_norm"]) <= 1e-8 * max(1.0, abs(ref["log_norm"]))
        # the evolved device state itself, measured on the host
        ev = to_oracle(devs[k].download())
        assert abs(oe.measure_energy(ev, m, 64)["energy_total"] - got["energy_total"][k]) < 1e-9


def test_energy_states_fp32(spp, gpu_ctx):
    say("energy of resident states, FP32 contractions, Z2 3x3")
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=2.0, symmetric=True)
    _compare(spp, gpu_ctx, [random_state(L, (1, 1), "peps", seed=81)], m, 16, 2e-5, precision="fp32")


def test_energy_states_empty(spp, gpu_ctx):
    r = spp.measure_energy_states(gpu_ctx, [], spp.TFIM(1.0, 1.0, symmetric=True), 8)
    assert r["energy_total"].shape == (0,)
