"""GPU parity of the device projection of configurations
(spp_project_batch: project_configuration, src/sampling/born.jl:27-36, with
normalize_state!, src/states/finite.jl:89-100) against the host projection:
site tensors, normalisation logs and log amplitudes, the bond capacity
re-embedding with the PEPO's live sectors, and the log norms of the projected
states; Z2 (SymmetryBasis) and dense (XBasis)."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, embed_identity_pepo

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("symmetric", [True, False])
def test_device_projection_matches_host(spp, gpu_ctx, symmetric):
    from stopeps_jl_amd.states import parity_sorted, pad_bonds
    gpu_ctx.set_precision("fp64")
    D = (2, 1) if symmetric else 3
    pepo = to_product_state(spp, embed_identity_pepo(Lattice(4, 3), D, 0.4, seed=12, symmetric=symmetric))
    L = pepo.lattice
    dev = spp.DeviceState(gpu_ctx, pepo)
    basis = spp.sampling.resolve_basis(pepo, None)
    rng = np.random.default_rng(3)
    cfgs = np.array([[[basis.labels[rng.integers(2)] for _ in range(L.Ly)] for _ in range(L.Lx)]
                     for _ in range(5)])
    cap = (2, 2) if symmetric else (4, 0)
    say("device projection", symmetric, "capacity", cap)
    states, norms = spp.project_configurations(gpu_ctx, dev, cfgs, basis, bond_space=cap if symmetric else None)
    for i, cfg in enumerate(cfgs):
        host, rec = spp.project_configuration(pepo, cfg, basis)
        if symmetric:
            host, live = pad_bonds(host, cap)
        got = states[i].download()
        assert abs(got.log_amplitude - host.log_amplitude) < 1e-12
        assert np.allclose(norms[i], [r["log_normalization"] for r in rec], atol=1e-13, rtol=0)
        for (x, y) in L.site_order():
            a, ev = parity_sorted(host[(x, y)])
            assert got[(x, y)].data.shape == a.shape
            assert np.max(np.abs(got[(x, y)].data - a)) < 1e-14
            assert [int(np.sum(p == 0)) for p in got[(x, y)].parity] == ev
    # the projected states evaluate like uploaded ones
    uploaded = []
    for cfg in cfgs:
        host, _ = spp.project_configuration(pepo, cfg, basis)
        if symmetric:
            host, live = pad_bonds(host, cap)
            uploaded.append(spp.DeviceState(gpu_ctx, host, live=live))
        else:
            uploaded.append(spp.DeviceState(gpu_ctx, host))
    a = spp.lognorms(gpu_ctx, states, 8)
    b = spp.lognorms(gpu_ctx, uploaded, 8)
    assert np.max(np.abs(a["log_norm"] - b["log_norm"])) < 1e-10
    if symmetric:
        assert np.array_equal(states[0].live(), uploaded[0].live())
