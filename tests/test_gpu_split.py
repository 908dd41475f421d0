"""GPU checks of the BF16 split-product contraction modes ("bf16x6": three
BF16 pieces per FP32 operand, six piece products; "bf16x3": two pieces, three
products; FP32 storage and accumulation in both) through the library's
matrix-core self-test, against FP64 numpy products.

Error bars, elementwise against |A| |B|: the FP32 accumulation of K terms
contributes at most about (K + 16) 2^-24 (the BF16 instruction sums 16
products per step); the pieces omitted by "bf16x6" are below 3 2^-27 and those
omitted by "bf16x3" below 3 2^-18 relative per product. A factor 2 covers
the order in which the matrix cores add. A kernel that used only the leading
pieces would fail the first bar by orders of magnitude (2^-8 relative)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OMITTED = {"bf16x6": 3 * 2.0 ** -27, "bf16x3": 3 * 2.0 ** -18}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bound(A, B, precision):
    K = A.shape[1]
    scale = np.abs(A) @ np.abs(B)
    return 2 * ((K + 16) * 2.0 ** -24 + OMITTED[precision]) * scale + 1e-300


@pytest.mark.parametrize("precision", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("shape", [(7, 5, 3), (130, 70, 37), (256, 300, 64), (256, 576, 40), (300, 31, 1100)])
@pytest.mark.parametrize("layout", [0, 1, 2, 3])
def test_split_gemm_dense(spp, gpu_ctx, precision, shape, layout):
    """Dense products in all four operand memory orders, over the three block
    tile shapes (N <= 32; N a few multiples of 96; general), ragged edges, K
    below one step of 16 and K above one pass of 1024 offsets."""
    say("split gemm", precision, shape, layout)
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((M, K)) * np.exp(rng.uniform(-3, 3, (M, K)))
    B = rng.standard_normal((K, N))
    C = gpu_ctx.gemm(A, B, precision, layout=layout)
    ref = A @ B
    err = np.abs(C - ref)
    bound = _bound(A, B, precision)
    assert np.all(err <= bound), float(np.max(err / bound))


def test_split_modes_differ_in_accuracy(spp, gpu_ctx):
    """The six-product mode is an FP32-grade product and the three-product mode
    a 2^-16-grade one. With operands exactly representable in FP32 (no input
    rounding) their errors on one product differ by more than 8x (measured:
    fp32 4.7e-7, bf16x6 4.0e-7, bf16x3 4.6e-6 with rounded inputs)."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((256, 128)).astype(np.float32).astype(np.float64)
    B = rng.standard_normal((128, 256)).astype(np.float32).astype(np.float64)
    ref = A @ B
    e = {p: float(np.max(np.abs(gpu_ctx.gemm(A, B, p) - ref)) / np.max(np.abs(ref)))
         for p in ("fp32", "bf16x6", "bf16x3")}
    say("max relative errors", e)
    assert e["bf16x6"] < 1e-6 and e["fp32"] < 1e-6
    assert e["bf16x3"] > 8 * e["bf16x6"]
    assert e["bf16x3"] < 1e-4


@pytest.mark.parametrize("precision", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("layout", [0, 3])
def test_split_z2_gemm(spp, gpu_ctx, precision, layout):
    """Z2 legs: only the parity-matching output blocks are written, and a block
    whose K range is empty (all K indices even) receives exact zeros."""
    M, N, K = 150, 140, 20
    eM, eN, eK = 70, 60, K
    rng = np.random.default_rng(5)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    pm, pn, pk = np.arange(M) >= eM, np.arange(N) >= eN, np.arange(K) >= eK
    A[pm[:, None] != pk[None, :]] = 0.0
    B[pk[:, None] != pn[None, :]] = 0.0
    C = gpu_ctx.gemm(A, B, precision, evens=(eM, eN, eK), layout=layout)
    neutral = pm[:, None] == pn[None, :]
    assert np.all(np.isnan(C[~neutral]))
    assert np.all(C[eM:, eN:] == 0.0)
    bound = _bound(A, B, precision)
    assert np.all(np.abs(C - A @ B)[neutral] <= bound[neutral])


def test_cache_keeps_its_product_mode(spp):
    """A sampling cache computes its draws with the products it was built with,
    whatever the context is set to at draw time: a bf16x3 cache drawn while the
    context is back on FP32 gives the draws of the same cache drawn in bf16x3
    (bit for bit), and its log p differs from the FP32 cache's by about the
    three-product rounding (2^-16 grade), not by zero."""
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(4, 4), D=(2, 2), eps=0.3, seed=5)
    uni = np.stack([spp.sample_uniforms(1, i, 16) for i in range(1, 5)])
    ctx = spp.Context(0, "bf16x3")
    try:
        split = spp.SamplingCache(ctx, pepo, 6)
        a = spp.sample_configurations(split, uniforms=uni)
        ctx.set_precision("fp32")
        b = spp.sample_configurations(split, uniforms=uni)
        plain = spp.SamplingCache(ctx, pepo, 6)
        c = spp.sample_configurations(plain, uniforms=uni)
    finally:
        ctx.close()
    say("bf16x3 cache vs fp32 cache dlogp", np.max(np.abs(a["log_probability"] - c["log_probability"])))
    assert np.array_equal(a["configuration"], b["configuration"])
    assert np.array_equal(a["log_probability"], b["log_probability"])
    assert np.array_equal(a["conditionals"], b["conditionals"])
    assert np.max(np.abs(a["log_probability"] - c["log_probability"])) > 0
