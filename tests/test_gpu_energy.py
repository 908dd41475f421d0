"""GPU parity of the batched energy measurement (spp_measure_energy_batch,
through the C ABI) against the FP64 oracle's measure_energy on the same
projected samples, plus the beta = 0 closed form."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, TFIM, embed_identity_pepo, identity_pepo, random_state
from oracle import born as ob
from oracle import energy as oe

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _random_configs(pepo, n, seed):
    L = pepo.lattice
    labels = ob.resolve_basis(bool(np.any(pepo[(0, 0)].p[2]))).labels
    rng = np.random.default_rng(seed)
    return np.asarray(labels)[rng.integers(0, 2, size=(n, L.Lx, L.Ly))]


def _compare(spp, ctx, pepo, chi_sample, chi, cfgs, J, h, tol, precision="fp64"):
    sym = bool(np.any(pepo[(0, 0)].p[2]))
    ctx.set_precision(precision)
    try:
        cache = spp.SamplingCache(ctx, to_product_state(spp, pepo), chi_sample)
        got = spp.measure_energies(cache, cfgs, spp.TFIM(J, h, symmetric=sym), chi=chi)
    finally:
        ctx.set_precision("fp64")
    model = TFIM(J, h, symmetric=sym)
    for i, cfg in enumerate(cfgs):
        ps, _ = ob.project_configuration(pepo, cfg)
        ref = oe.measure_energy(ps, model, chi)
        e_ref = ref["energy_total"]
        assert abs(got["energy_total"][i] - e_ref) <= tol * max(1.0, abs(e_ref)), (i, got["energy_total"][i], e_ref)
        assert abs(got["energy_per_site"][i] - ref["energy_per_site"]) <= tol * max(1.0, abs(e_ref))
        assert abs(got["tensor_log_norm"][i] - ref["tensor_log_norm"]) <= 10 * tol * max(1.0, abs(ref["tensor_log_norm"]))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 10 * tol * max(1.0, abs(ref["log_norm"]))
        # discarded weights are (trace - kept) / trace of FP64 eigenvalues of a
        # rho accumulated in the contraction precision: FP32 leaves ~1e-7 of noise
        mt = ref["max_environment_truncation"]
        mt_floor = 1e-9 if precision == "fp64" else 1e-6
        assert abs(got["max_environment_truncation"][i] - mt) <= max(mt_floor, 1e-3 * mt)
    assert np.all(got["status"] == 0)
    return got


def test_energy_exact_chi_z2(spp, gpu_ctx):
    say("energy exact chi, Z2, 3x3")
    L = Lattice(3, 3)
    pepo = embed_identity_pepo(L, (1, 1), 0.5, seed=31)
    _compare(spp, gpu_ctx, pepo, 16, 16, _random_configs(pepo, 5, 1), 1.0, 2.0, 1e-10)


def test_energy_truncated_z2(spp, gpu_ctx):
    say("energy truncated chi, Z2, 4x4")
    L = Lattice(4, 4)
    pepo = embed_identity_pepo(L, (1, 1), 0.6, seed=32)
    _compare(spp, gpu_ctx, pepo, 6, 6, _random_configs(pepo, 4, 2), 1.0, 2.0, 1e-8)


def test_energy_uneven_bonds_measurement_chi(spp, gpu_ctx):
    say("energy uneven D=(2,1), measurement chi != sampling chi, 3x4")
    L = Lattice(3, 4)
    pepo = embed_identity_pepo(L, (2, 1), 0.4, seed=33)
    _compare(spp, gpu_ctx, pepo, 10, 7, _random_configs(pepo, 4, 3), 0.7, 1.3, 1e-8)


def test_energy_dense_xbasis(spp, gpu_ctx):
    say("energy dense X basis, 3x3")
    L = Lattice(3, 3)
    pepo = embed_identity_pepo(L, 2, 0.5, seed=34, symmetric=False)
    _compare(spp, gpu_ctx, pepo, 8, 8, _random_configs(pepo, 4, 4), 1.0, 2.0, 1e-9)


def test_energy_random_pepo_2x3(spp, gpu_ctx):
    say("energy random PEPO, Z2, 2x3")
    L = Lattice(2, 3)
    pepo = random_state(L, (1, 1), "pepo", seed=35)
    _compare(spp, gpu_ctx, pepo, 16, 16, _random_configs(pepo, 6, 5), 1.0, 0.5, 1e-10)


def test_energy_fp32(spp, gpu_ctx):
    say("energy truncated chi, Z2, 4x4, fp32")
    L = Lattice(4, 4)
    pepo = embed_identity_pepo(L, (1, 1), 0.6, seed=32)
    _compare(spp, gpu_ctx, pepo, 6, 6, _random_configs(pepo, 4, 2), 1.0, 2.0, 2e-5, precision="fp32")


@pytest.mark.parametrize("symmetric", [True, False])
def test_energy_beta0_closed_form(spp, gpu_ctx, symmetric):
    """beta = 0: every projected sample is an X-basis product state, so
    E = -h sum_i s_i with s_i = +-1 the X eigenvalue of its label, and <ZZ> = 0."""
    say("energy beta=0 closed form, symmetric =", symmetric)
    gpu_ctx.set_precision("fp64")
    L = Lattice(3, 4)
    pepo = identity_pepo(L, symmetric=symmetric)
    cache = spp.SamplingCache(gpu_ctx, to_product_state(spp, pepo), 4)
    cfgs = _random_configs(pepo, 6, 6)
    h = 1.7
    got = spp.measure_energies(cache, cfgs, spp.TFIM(1.0, h, symmetric=symmetric))
    s = np.where(cfgs == 1, 1.0, -1.0) if symmetric else cfgs.astype(float)
    np.testing.assert_allclose(got["energy_total"], -h * s.reshape(len(cfgs), -1).sum(axis=1), atol=1e-12)
    # unit-norm product state: the normalised projected tensors contract to 1
    np.testing.assert_allclose(got["tensor_log_norm"], 0.0, atol=1e-12)


def test_energy_empty_batch(spp, gpu_ctx):
    say("energy empty batch")
    L = Lattice(2, 2)
    pepo = identity_pepo(L, symmetric=True)
    cache = spp.SamplingCache(gpu_ctx, to_product_state(spp, pepo), 4)
    got = spp.measure_energies(cache, np.zeros((0, 2, 2), dtype=int), spp.TFIM(1.0, 1.0, symmetric=True))
    assert got["energy_total"].shape == (0,)
