"""The contraction kernel's tile shapes and its autotuner: every tile shape of
a kernel family gives bit-identical products (the same matrix-core
instruction walks K in the same order), so Born draws with the autotuner on
equal those with it off bit for bit, and the tuned launch classes are reused."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILES = {"fp64": 5, "fp32": 6, "bf16x6": 4, "bf16x3": 4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x6", "bf16x3"])
@pytest.mark.parametrize("shape,evens", [((300, 70, 37), None), ((257, 648, 130), (129, 324, 65)),
                                         ((40, 20, 1100), (17, 10, 555))])
def test_every_tile_shape_gives_the_same_product(spp, gpu_ctx, precision, shape, evens):
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    say("tile shapes", precision, shape, evens)
    ref = None
    try:
        for t in range(TILES[precision]):
            gpu_ctx.set_option("gemm_tile", t)
            C = gpu_ctx.gemm(A, B, precision, evens=evens)
            if ref is None:
                ref = C
            else:
                assert np.array_equal(np.isnan(C), np.isnan(ref))
                ok = ~np.isnan(ref)
                assert np.array_equal(C[ok], ref[ok]), (precision, t)
    finally:
        gpu_ctx.set_option("gemm_tile", -1)
    ok = ~np.isnan(ref)
    # a Z2 product only sums the K indices of the row's parity (first e indices even)
    if evens is not None:
        pm, pn, pk = (np.arange(n) >= e for n, e in zip(shape, evens))
        A = A * (pm[:, None] == pk[None, :])
        B = B * (pk[:, None] == pn[None, :])
    bound = 1e-10 if precision == "fp64" else 1e-4
    assert np.max(np.abs(ref[ok] - (A @ B)[ok])) < bound * np.max(np.abs(A) @ np.abs(B))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotuned_draws_are_bit_identical(spp, precision):
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(4, 4), D=(2, 2), eps=0.3, seed=8)
    uni = np.stack([spp.sample_uniforms(2, i, 16) for i in range(1, 9)])
    res = {}
    for mode in (0, 1):
        ctx = spp.Context(0, precision, autotune=mode)
        try:
            cache = spp.SamplingCache(ctx, pepo, 8)
            first = spp.sample_configurations(cache, uniforms=uni)
            again = spp.sample_configurations(cache, uniforms=uni)
            cnt = ctx.counters()
        finally:
            ctx.close()
        assert np.array_equal(first["log_probability"], again["log_probability"])
        res[mode] = (first, cnt)
    say(precision, "tuned classes", res[1][1]["tuned_classes"], "changed", res[1][1]["tuned_changed"],
        "tuning ms", res[1][1]["tune_ms"])
    assert res[0][1]["tuned_classes"] == 0 and res[1][1]["tuned_classes"] > 0
    a, b = res[0][0], res[1][0]
    assert np.array_equal(a["configuration"], b["configuration"])
    assert np.array_equal(a["log_probability"], b["log_probability"])
    assert np.array_equal(a["conditionals"], b["conditionals"])
    assert np.array_equal(a["tensor_log_norm"], b["tensor_log_norm"])
