"""GPU parity of the batched lognorm of finite states (spp_lognorm_batch,
through the C ABI) against the FP64 oracle, plus the beta = 0 closed form."""
import sys

import numpy as np
import pytest

from conftest import to_product_state
from oracle.model import Lattice, embed_identity_pepo, identity_pepo, random_state
from oracle import boundary as obm

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(spp, ctx, states, chi, tol, precision="fp64"):
    ctx.set_precision(precision)
    try:
        got = spp.lognorms(ctx, [to_product_state(spp, s) for s in states], chi)
    finally:
        ctx.set_precision("fp64")
    mt_floor = 1e-9 if precision == "fp64" else 1e-6
    for i, s in enumerate(states):
        ref = obm.lognorm(s, chi)
        assert abs(got["tensor_log_norm"][i] - ref["tensor_log_norm"]) <= tol * max(1.0, abs(ref["tensor_log_norm"]))
        assert abs(got["log_norm"][i] - ref["log_norm"]) <= tol * max(1.0, abs(ref["log_norm"]))
        mt = ref["max_environment_truncation"]
        assert abs(got["max_environment_truncation"][i] - mt) <= max(mt_floor, 1e-3 * mt)
    return got


@pytest.mark.parametrize("kind", ["peps", "pepo"])
def test_lognorm_exact_chi(spp, gpu_ctx, kind):
    say("lognorm exact chi", kind)
    L = Lattice(3, 3)
    states = [random_state(L, (1, 1), kind, seed=40 + k) for k in range(3)]
    _compare(spp, gpu_ctx, states, 16, 1e-11)


def test_lognorm_truncated_batch(spp, gpu_ctx):
    say("lognorm truncated chi, 4x4 PEPO batch")
    L = Lattice(4, 4)
    states = [embed_identity_pepo(L, (1, 1), 0.6, seed=50 + k) for k in range(4)]
    _compare(spp, gpu_ctx, states, 6, 1e-9)


def test_lognorm_uneven_dense(spp, gpu_ctx):
    say("lognorm uneven D=(2,1) and dense")
    _compare(spp, gpu_ctx, [embed_identity_pepo(Lattice(3, 4), (2, 1), 0.4, seed=60)], 7, 1e-9)
    _compare(spp, gpu_ctx, [random_state(Lattice(3, 3), 2, "peps", seed=61, symmetric=False)], 8, 1e-10)


def test_lognorm_fp32(spp, gpu_ctx):
    say("lognorm fp32")
    L = Lattice(4, 4)
    _compare(spp, gpu_ctx, [embed_identity_pepo(L, (1, 1), 0.6, seed=52)], 6, 2e-5, precision="fp32")


@pytest.mark.parametrize("symmetric", [True, False])
def test_lognorm_identity_pepo_closed_form(spp, gpu_ctx, symmetric):
    """beta = 0: <<I|I>> = 2^N for the identity PEPO (finite.jl:80-86)."""
    say("lognorm identity PEPO closed form, symmetric =", symmetric)
    L = Lattice(3, 4)
    got = spp.lognorm(gpu_ctx, to_product_state(spp, identity_pepo(L, symmetric=symmetric)), 4)
    assert abs(got["log_norm"] - 12 * np.log(2.0)) < 1e-12
    assert got["max_environment_truncation"] == 0.0
