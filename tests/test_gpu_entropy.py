"""GPU parity of the boundary-entropy observer (record_entropy of
measure_energy and evolution_step!, boundary_mps.jl:20-57, 193-209,
energy.jl:31-78, full_update.jl:560-591) through the C ABI: the eigenvalue
kernel against numpy, the Se profiles and their maximum against the FP64
oracle for resident states and projected samples (direct-Jacobi cuts and
cuts sent to the batched tridiagonal solve), and the full-update step's
maximum against the oracle's step."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, TFIM, random_state
from oracle import born as ob
from oracle import energy as oe
from oracle import fu as ofu

pytestmark = pytest.mark.gpu

# entropies are smooth functions of FP64 spectra that agree to ~1e-13
TOL = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_spectrum_kernel_matches_numpy(spp, gpu_ctx):
    """The batched Householder + bisection eigensolver on random PSD and
    indefinite blocks of mixed sizes inside a padded leading dimension, up to
    blocks larger than one workgroup's thread count."""
    lib = gpu_ctx._lib
    rng = np.random.default_rng(5)
    for ld, sizes in [(8, [0, 1, 2, 3, 8]), (70, [70, 64, 33, 5]), (600, [600, 513, 257])]:
        say(f"  spectrum kernel: ld {ld}, blocks {sizes}")
        mats = np.zeros((len(sizes), ld, ld))
        for i, n in enumerate(sizes):
            if n == 0:
                continue
            Y = rng.normal(size=(n, max(1, n // 2)))
            M = Y @ Y.T if i % 2 == 0 else Y @ Y.T - 0.3 * np.eye(n)
            M = M + 1e-13 * rng.normal(size=(n, n))  # slightly asymmetric, as rho is
            mats[i, :n, :n] = M
            mats[i, n:, :] = 7.0  # padding outside the block is never read
            mats[i, :, n:] = 7.0
        n = np.array(sizes, dtype=np.int32)
        eig = np.zeros((len(sizes), ld))
        from stopeps_jl_amd import _lib
        _lib.check(gpu_ctx.handle, lib.spp_selftest_spectrum(gpu_ctx.handle, len(sizes), ld, _lib.iptr(n),
                                                             _lib.dptr(np.ascontiguousarray(mats)),
                                                             _lib.dptr(eig)))
        for i, k in enumerate(sizes):
            if k == 0:
                continue
            B = mats[i, :k, :k]
            ref = np.linalg.eigvalsh(0.5 * (B + B.T))
            scale = max(1.0, np.max(np.abs(ref)))
            assert np.max(np.abs(eig[i, :k] - ref)) <= 1e-12 * scale * max(1, k) ** 0.5, (ld, k)


def _check_profiles(got, refs):
    for i, ref in enumerate(refs):
        for side in ("north", "south"):
            a, b = got["Se"][i][side], ref["Se"][side]
            assert a.shape == b.shape
            assert np.array_equal(np.isnan(a), np.isnan(b))
            ok = ~np.isnan(b)
            assert np.max(np.abs(a[ok] - b[ok]), initial=0.0) <= TOL, (i, side, a, b)
        assert abs(got["max_boundary_entropy"][i] - ref["max_boundary_entropy"]) <= TOL


@pytest.mark.parametrize("ksub", [64, 12])
def test_entropy_profiles_resident_states(spp, gpu_ctx, ksub):
    """Se of measure_energy_states against the oracle: a truncating chi on 3x3
    PEPOs and a 4x3 PEPS. ksub = 12 (still >= chi, so the same bonds are kept)
    sends every cut whose parity blocks exceed 12 rows to the subspace
    truncation and the batched tridiagonal solve."""
    say(f"boundary entropy of resident states, ksub {ksub}")
    m = TFIM(J=1.0, h=1.3, symmetric=True)
    cases = [([random_state(Lattice(3, 3), (1, 1), "pepo", seed=71 + k) for k in range(2)], 8),
             ([random_state(Lattice(4, 3), (2, 1), "peps", seed=75)], 6)]
    gpu_ctx.set_option("ksub", ksub)
    try:
        for states, chi in cases:
            devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, s)) for s in states]
            got = spp.measure_energy_states(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), chi,
                                            record_entropy=True)
            refs = [oe.measure_energy(s, m, chi, record_entropy=True) for s in states]
            for i, r in enumerate(refs):
                assert abs(got["energy_total"][i] - r["energy_total"]) <= 1e-8 * max(1.0, abs(r["energy_total"]))
                say(f"  state {i}: max entropy {got['max_boundary_entropy'][i]:.12f} "
                    f"vs {r['max_boundary_entropy']:.12f}")
            _check_profiles(got, refs)
            plain = spp.measure_energy_states(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), chi)
            assert plain["Se"] == [None] * len(states) and np.all(np.isnan(plain["max_boundary_entropy"]))
            np.testing.assert_array_equal(plain["energy_total"], got["energy_total"])
    finally:
        gpu_ctx.set_option("ksub", 64)


@pytest.mark.parametrize("ksub", [64, 8])
def test_entropy_profiles_projected_samples(spp, gpu_ctx, ksub):
    """Se of the projected-sample path (measure_energies through the sampling
    cache) against the oracle on the projected states."""
    say(f"boundary entropy of projected samples, ksub {ksub}")
    pepo = random_state(Lattice(3, 3), (1, 1), "pepo", seed=81)
    m = TFIM(J=1.0, h=0.8, symmetric=True)
    labels = ob.resolve_basis(True).labels
    cfgs = np.asarray(labels)[np.random.default_rng(3).integers(0, 2, size=(3, 3, 3))]
    gpu_ctx.set_option("ksub", ksub)
    try:
        cache = spp.SamplingCache(gpu_ctx, to_product_state(spp, pepo), 8)
        got = spp.measure_energies(cache, cfgs, spp.TFIM(m.J, m.h, symmetric=True), chi=6, record_entropy=True)
    finally:
        gpu_ctx.set_option("ksub", 64)
    refs = [oe.measure_energy(ob.project_configuration(pepo, c)[0], m, 6, record_entropy=True) for c in cfgs]
    _check_profiles(got, refs)


def test_fu_step_max_entropy(spp, gpu_ctx):
    """max_boundary_entropy of one second-order full-update step on 3x3 PEPOs
    (every absorb of the four sweeps) against the oracle's step."""
    say("boundary entropy maximum of a full-update step, 3x3 PEPO")
    m = TFIM(J=1.0, h=1.5, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=20, linear_max_iterations=60)
    ostates = [random_state(Lattice(3, 3), (1, 1), "pepo", seed=91 + k) for k in range(2)]
    devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, o)) for o in ostates]
    method = spp.FullUpdate(opts.max_iterations, opts.error_tolerance, opts.diagonal_shift,
                            opts.linear_max_iterations, opts.linear_tolerance)
    gpu_ctx.set_option("ksub", 12)
    try:
        got = spp.evolution_steps(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), 0.05, 8, method,
                                  record_entropy=True)
    finally:
        gpu_ctx.set_option("ksub", 64)
    for k, o in enumerate(ostates):
        ref = ofu.evolution_step(o, m, 0.05, 2, (1, 1), opts, 8, record_entropy=True)
        say(f"  state {k}: {got['max_boundary_entropy'][k]:.12f} vs {ref['max_boundary_entropy']:.12f}")
        assert abs(got["max_boundary_entropy"][k] - ref["max_boundary_entropy"]) <= 1e-7


def test_entropy_profiles_dense_model(spp, gpu_ctx):
    """Se on a dense (non-symmetric) TFIM state: one trivial sector, so the
    spectrum sits in a single block per cut."""
    say("boundary entropy of a dense-model PEPO")
    m = TFIM(J=0.8, h=1.1, symmetric=False)
    states = [random_state(Lattice(3, 3), 2, "pepo", seed=95, symmetric=False)]
    devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, s)) for s in states]
    got = spp.measure_energy_states(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=False), 6, record_entropy=True)
    _check_profiles(got, [oe.measure_energy(s, m, 6, record_entropy=True) for s in states])
