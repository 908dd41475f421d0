"""GPU parity on one-row, one-column and one-site lattices (Lattice(Lx, Ly)
with Lx or Ly = 1), which the reference allows: lognorm, Born draws, energy
and the full-update step against the FP64 oracle. A one-column lattice has no
horizontal bonds and its boundaries are one-site MPS; a one-row lattice has
no vertical bonds and a single absorbed row."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, TFIM, embed_identity_pepo, random_state
from oracle import boundary as obm
from oracle import energy as oe
from oracle import fu as ofu
from test_gpu_sampling import _compare_draws
from test_gpu_fu import _compare as _compare_fu

pytestmark = pytest.mark.gpu

SHAPES = [(1, 3), (3, 1), (1, 1)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("shape", SHAPES)
def test_thin_lognorm(spp, gpu_ctx, shape):
    say(f"lognorm on {shape}")
    states = [random_state(Lattice(*shape), (1, 1), kind, seed=61) for kind in ("pepo",)]
    got = spp.lognorms(gpu_ctx, [to_product_state(spp, s) for s in states], 8)
    for i, s in enumerate(states):
        ref = obm.lognorm(s, 8)
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 1e-11 * max(1.0, abs(ref["log_norm"]))


@pytest.mark.parametrize("shape", SHAPES)
def test_thin_draws(spp, gpu_ctx, shape):
    say(f"Born draws on {shape}")
    pepo = embed_identity_pepo(Lattice(*shape), (1, 1), 0.5, seed=62)
    assert _compare_draws(spp, gpu_ctx, pepo, 8, 6, 5, 1e-11, 1e-11) == []


@pytest.mark.parametrize("shape", SHAPES)
def test_thin_energy(spp, gpu_ctx, shape):
    say(f"energy of resident states on {shape}")
    m = TFIM(J=1.0, h=0.9, symmetric=True)
    states = [random_state(Lattice(*shape), (1, 1), "pepo", seed=63)]
    devs = [spp.DeviceState(gpu_ctx, to_product_state(spp, s)) for s in states]
    got = spp.measure_energy_states(gpu_ctx, devs, spp.TFIM(m.J, m.h, symmetric=True), 8, record_entropy=True)
    for i, s in enumerate(states):
        ref = oe.measure_energy(s, m, 8, record_entropy=True)
        assert abs(got["energy_total"][i] - ref["energy_total"]) <= 1e-10 * max(1.0, abs(ref["energy_total"]))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= 1e-10 * max(1.0, abs(ref["log_norm"]))
        assert abs(got["max_boundary_entropy"][i] - ref["max_boundary_entropy"]) <= 1e-9


@pytest.mark.parametrize("shape", [(1, 3), (3, 1)])
def test_thin_fu_step(spp, gpu_ctx, shape):
    say(f"full-update step on {shape}")
    m = TFIM(J=1.0, h=1.2, symmetric=True)
    st = [random_state(Lattice(*shape), (1, 1), "peps", seed=64)]
    _compare_fu(spp, gpu_ctx, st, m, 0.1, 16, (1, 1), ofu.FullUpdate(max_iterations=40, linear_max_iterations=100))
