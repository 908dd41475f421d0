"""GPU parity at the headline configuration: 8x8 open TFIM (J=1, h=2.0), Z2
bond space Z2Space(0=>3,1=>3) (D=6), PEPO purified to beta0=0.5, boundary
chi=32. Input: the frozen beta0 PEPO tests/golden/cfg2_pepo_beta05.npz;
expected values: the FP64 oracle's draws of sample indices 1..8 of seed 1
(configurations, log probabilities, conditionals, max_environment_truncation,
src/sampling/born.jl:86-163), the projected samples' log norms
(src/contraction/boundary_mps.jl:180-191) and ``measure_energy`` of samples
1..4 (src/measurement/energy.jl:31-78), in tests/golden/cfg2_oracle.npz
(written by tests/golden/make_cfg2_fixture.py).

Bars: FP64 configurations identical, log p and conditionals within 1e-9,
tensor log norms within 1e-9 relative, energies within 1e-8; FP32 (mixed
precision, with the products on the FP32 matrix cores or as six BF16 piece
products) energy per site within 1e-6 of the oracle on the weighted ensemble
exp(log_norm - log q) (the estimate the metric quotes) and within 2e-6 per
sample (measured: 4.3e-7 .. 1.14e-6), and any FP32 draw that differs from the
oracle's differs first at a site whose uniform lies within 1e-5 of the
cumulative boundary."""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT
from test_gpu_sampling import boundary_distance

pytestmark = pytest.mark.gpu

PEPO = os.path.join(ROOT, "tests", "golden", "cfg2_pepo_beta05.npz")
ORACLE = os.path.join(ROOT, "tests", "golden", "cfg2_oracle.npz")
CHI = 32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fixture(spp):
    pepo, _ = spp.load_state_npz(PEPO)
    ref = np.load(ORACLE, allow_pickle=False)
    return pepo, ref


def _draws(spp, ctx, pepo, ref, precision):
    ctx.set_precision(precision)
    cache = spp.SamplingCache(ctx, pepo, CHI)
    ns = 64
    uni = np.stack([spp.sample_uniforms(int(ref["seed"]), int(i), ns) for i in ref["sample_index"]])
    return cache, uni, spp.sample_configurations(cache, uniforms=uni)


def _energy_ensemble(e, log_norm, log_q):
    lw = log_norm - log_q
    w = np.exp(lw - np.max(lw))
    return float(np.sum(w * e) / np.sum(w))


class _Ref:
    """The oracle's draw record in the form boundary_distance reads."""

    def __init__(self, conditionals):
        self.c = conditionals

    def __getitem__(self, key):
        return self.c


@pytest.mark.parametrize("gram_bytes", [None, 3.7e7])
def test_cfg2_fp64_draws_and_energies_match_oracle(spp, gpu_ctx, fixture, gram_bytes):
    """gram_bytes 3.7e7: the Gram transfers of every boundary absorb take the
    chi=32 bra bond in index ranges of 3 (32 x 36^3 FP64 entries per index,
    11.9 MB; the last range holds 2), which must not change the results."""
    pepo, ref = fixture
    say(f"8x8 TFIM h=2 beta0=0.5 D=6 chi=32: FP64 draws of samples 1..8 against the oracle (gram {gram_bytes})")
    if gram_bytes:
        gpu_ctx.set_option("gram_chunk_bytes", gram_bytes)
    try:
        _fp64_checks(spp, gpu_ctx, pepo, ref)
    finally:
        gpu_ctx.set_option("gram_chunk_bytes", 2.0 ** 30)


def _fp64_checks(spp, gpu_ctx, pepo, ref):
    cache, uni, got = _draws(spp, gpu_ctx, pepo, ref, "fp64")
    np.testing.assert_allclose(cache.south_max_truncation, ref["south_max_truncation"], rtol=1e-6, atol=1e-12)
    for k in range(len(ref["sample_index"])):
        assert np.array_equal(got["configuration"][k], ref["configuration"][k]), k
        lp = float(ref["log_probability"][k])
        assert abs(got["log_probability"][k] - lp) <= 1e-9 * max(1.0, abs(lp)), k
        np.testing.assert_allclose(got["conditionals"][k], ref["conditionals"][k], rtol=1e-9, atol=1e-9)
        mt = float(ref["max_environment_truncation"][k])
        assert abs(got["max_environment_truncation"][k] - mt) <= max(1e-12, 1e-6 * mt), k
        tln = float(ref["tensor_log_norm"][k])
        assert abs(got["tensor_log_norm"][k] - tln) <= 1e-9 * max(1.0, abs(tln)), k
        assert abs(got["log_amplitude"][k] - ref["log_amplitude"][k]) <= 1e-10 * max(1.0, abs(ref["log_amplitude"][k]))
        say(f"  sample {int(ref['sample_index'][k])}: log p {got['log_probability'][k]:.12f} "
            f"(oracle {lp:.12f}), tensor_log_norm {got['tensor_log_norm'][k]:.10f}")
    say("FP64 energies of samples 1..4")
    m = np.isfinite(ref["energy_per_site"])
    model = spp.TFIM(1.0, 2.0, symmetric=True)
    e = spp.measure_energies(cache, ref["configuration"][m].astype(np.int64), model, CHI)
    for j, k in enumerate(np.nonzero(m)[0]):
        say(f"  sample {int(ref['sample_index'][k])}: E {e['energy_total'][j]:.12f} "
            f"(oracle {ref['energy_total'][k]:.12f})")
        assert abs(e["energy_total"][j] - ref["energy_total"][k]) <= 1e-8 * max(1.0, abs(ref["energy_total"][k]))
        assert abs(e["log_norm"][j] - ref["log_norm"][k]) <= 1e-9 * max(1.0, abs(ref["log_norm"][k]))


@pytest.mark.parametrize("precision", ["fp32", "bf16x6"])
def test_cfg2_fp32_energy_per_site_within_1e6_of_oracle(spp, gpu_ctx, fixture, precision):
    pepo, ref = fixture
    say(f"8x8 TFIM h=2 beta0=0.5 D=6 chi=32: {precision} draws and energies against the FP64 oracle")
    try:
        cache, uni, got = _draws(spp, gpu_ctx, pepo, ref, precision)
        for k in range(len(ref["sample_index"])):
            if np.array_equal(got["configuration"][k], ref["configuration"][k]):
                lp = float(ref["log_probability"][k])
                assert abs(got["log_probability"][k] - lp) <= 1e-4 * max(1.0, abs(lp)), k
                continue
            L = cache.lattice
            cfg_ref = [[int(v) for v in row] for row in ref["configuration"][k]]
            d = boundary_distance(cache.basis, _Ref(ref["conditionals"][k]), cfg_ref, got["configuration"][k],
                                  uni[k], L)
            say(f"  sample {int(ref['sample_index'][k])} differs; uniform {d:.2e} from the boundary")
            assert d <= 1e-5, d
        m = np.isfinite(ref["energy_per_site"])
        model = spp.TFIM(1.0, 2.0, symmetric=True)
        cfg = ref["configuration"][m].astype(np.int64)
        e = spp.measure_energies(cache, cfg, model, CHI)
        de = np.abs(e["energy_per_site"] - ref["energy_per_site"][m])
        # weights of the mixed pipeline: FP32 log norms over FP32 draw probabilities
        same = [np.array_equal(got["configuration"][k], ref["configuration"][k]) for k in np.nonzero(m)[0]]
        assert all(same)
        E = _energy_ensemble(e["energy_per_site"], e["log_norm"], got["log_probability"][m])
        Er = _energy_ensemble(ref["energy_per_site"][m], ref["log_norm"][m], ref["log_probability"][m])
        say(f"  max |dE/site| {np.max(de):.3e}, ensemble E/site {E:.12f} (oracle {Er:.12f})")
        assert abs(E - Er) <= 1e-6
        assert np.max(de) <= 2e-6
    finally:
        gpu_ctx.set_precision("fp64")
