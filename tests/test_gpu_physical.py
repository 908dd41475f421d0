"""GPU parity of physical-kind Born sampling (the resampling proposals,
src/sampling/born.jl:116-134 with _site_hole born.jl:75-83 and
_positive_metric(...; square_root=true) src/states/tensors.jl:65-94) against
the FP64 oracle, and of its eigensolver against numpy.

Bars: configurations identical; log p and conditionals within 1e-9 (FP64);
the PSD reports within 1e-8 absolute for the minimum relative eigenvalue and
the negative spectral fraction, 1e-10 for the hermiticity defect. The count of
negative eigenvalues includes eigenvalues at round-off level (|lambda| of a
few 1e-16 of the spectral scale), whose signs no two eigensolvers share, so it
is held to the oracle's count within the number of such eigenvalues only on
the cases where none are expected (exact chi: the metric is PSD and the count
is a round-off count on both sides, checked not to exceed a few per cent of
the block)."""
import itertools
import sys

import numpy as np
import pytest

from conftest import to_product_state, oracle_cfg, charged_peps, boundary_distance
from oracle.model import Lattice
from oracle import born as ob

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_psd_eigensolver_matches_numpy(spp, gpu_ctx):
    """Tridiagonalisation + QL with projections (spp_selftest_psd): every
    eigenvalue and sum_j f(lambda_j) (u_j^T v)^2 for f = max(., 0) and the
    identity, on symmetric blocks of sizes 1..648 (the D=6 site-metric block)
    with indefinite spectra, and on a block whose triangles differ (averaged)."""
    _lib = spp._lib
    lib = _lib.load()
    rng = np.random.default_rng(5)
    sizes = [1, 2, 3, 40, 200, 648]
    ld = max(sizes)
    mats = np.zeros((len(sizes), ld, ld))
    vecs = np.zeros((len(sizes), 2, ld))
    for i, n in enumerate(sizes):
        G = rng.standard_normal((n, n))
        A = G @ G.T / n - 0.2 * np.eye(n)  # a few negative eigenvalues
        if i == 3:
            A = A + 0.01 * rng.standard_normal((n, n))  # nonsymmetric: the two triangles are averaged
        mats[i, :n, :n] = A
        vecs[i, :, :n] = rng.standard_normal((2, n))
    eig = np.zeros((len(sizes), ld))
    proj = np.zeros((len(sizes), 2, ld))
    fail = np.zeros(len(sizes), dtype=np.int32)
    say("psd selftest", sizes)
    _lib.check(gpu_ctx.handle, lib.spp_selftest_psd(gpu_ctx.handle, len(sizes), ld, _lib.iptr(np.array(sizes, np.int32)),
                                                    _lib.dptr(mats), _lib.dptr(vecs), _lib.dptr(eig),
                                                    _lib.dptr(proj), _lib.iptr(fail)))
    assert not fail.any()
    for i, n in enumerate(sizes):
        A = mats[i, :n, :n]
        S = (A + A.T) / 2
        lam, V = np.linalg.eigh(S)
        scale = max(1.0, np.max(np.abs(lam)))
        assert np.max(np.abs(np.sort(eig[i, :n]) - lam)) < 1e-12 * scale * max(1, n / 50)
        for v in range(2):
            w = vecs[i, v, :n]
            for f in (lambda x: np.maximum(x, 0.0), lambda x: x):
                ref = w @ (V * f(lam)[None, :]) @ V.T @ w
                got = np.sum(f(eig[i, :n]) * proj[i, v, :n] ** 2)
                assert abs(got - ref) < 1e-11 * scale * (w @ w)


def _compare_draws(got, ref_results, L, tol=1e-9):
    for i, ref in enumerate(ref_results):
        assert np.array_equal(got["configuration"][i], oracle_cfg(ref, L)), i
        assert abs(got["log_probability"][i] - ref["log_probability"]) <= tol * max(1, abs(ref["log_probability"]))
        nc = len(ref["conditionals"])
        assert np.max(np.abs(got["conditionals"][i][:nc] - np.array(ref["conditionals"]))) <= tol
        assert abs(got["max_environment_truncation"][i] - ref["max_environment_truncation"]) <= 1e-9


def _compare_repairs(spp, cache, got, ref_results, exact):
    for i, ref in enumerate(ref_results):
        mine = spp.sampling.repair_records(cache, got["repairs"][i])
        assert len(mine) == len(ref["repairs"]), (len(mine), len(ref["repairs"]))
        for a, b in zip(mine, ref["repairs"]):
            assert a["site"] == b["site"]
            if "reason" in b:
                assert a.get("reason") == b["reason"] and a["allowed_outcomes"] == b["allowed_outcomes"]
                continue
            assert abs(a["minimum_relative_eigenvalue"] - b["minimum_relative_eigenvalue"]) < 1e-8
            assert abs(a["negative_spectral_fraction"] - b["negative_spectral_fraction"]) < 1e-8
            assert abs(a["hermiticity_defect"] - b["hermiticity_defect"]) < 1e-10
            if exact:
                # round-off negatives only: both small
                assert a["negative_spectral_fraction"] < 1e-10 and b["negative_spectral_fraction"] < 1e-10


@pytest.mark.parametrize("symmetric,D,chi", [(True, (1, 1), 16), (True, (1, 1), 2), (True, (2, 1), 3),
                                             (False, 2, 16), (False, 3, 3), (True, (2, 2), 16)])
def test_physical_draws_match_oracle(spp, gpu_ctx, symmetric, D, chi):
    """Physical-kind draws of a 3x3 PEPS with mixed auxiliary charges (a
    projected particle's legs) at exact and truncating chi: configurations,
    log p, conditionals, truncation and the per-site PSD reports against the
    oracle on the same uniforms; the truncating cases have indefinite site
    metrics, so the clipping changes the conditionals. D=(2, 2) at exact chi:
    site metrics of large null space (the QL iteration's norm-relative
    deflation)."""
    L = Lattice(3, 3)
    aux = [1, 0, 0, 1, 1, 0, 1, 0, 1]
    st = charged_peps(L, D, 21, aux, symmetric)
    basis = ob.resolve_basis(symmetric)
    ocache = ob.SamplingCache(st, chi, basis=basis, kind="physical")
    pst = to_product_state(spp, st)
    cache = spp.SamplingCache(gpu_ctx, pst, chi, kind="physical")
    uni = np.stack([ob.counter_uniforms(7, i, 9) for i in range(1, 7)])
    say("physical draws", symmetric, D, chi)
    got = spp.sample_configurations(cache, uniforms=uni)
    refs = [ob.sample_configuration(ocache, uniforms=uni[i]) for i in range(len(uni))]
    _compare_draws(got, refs, L)
    _compare_repairs(spp, cache, got, refs, exact=chi >= 16)
    negfrac = max(r["negative_spectral_fraction"] for ref in refs for r in ref["repairs"] if "site" in r
                  and "negative_spectral_fraction" in r)
    say("  largest negative spectral fraction", negfrac)
    if chi < 16:
        assert negfrac > 1e-10  # the truncated metrics are indefinite: the clipping is exercised
    # evaluation mode on the drawn configurations gives the same numbers
    ev = spp.sample_configurations(cache, configurations=got["configuration"])
    assert np.array_equal(ev["log_probability"], got["log_probability"]) or \
        np.max(np.abs(ev["log_probability"] - got["log_probability"])) < 1e-12


def test_physical_evaluation_every_configuration(spp, gpu_ctx):
    """Evaluation mode over all 2^6 configurations of a Z2 2x3 PEPS at exact
    chi: log p against the oracle (which is pinned to the brute-force Born
    distribution), -Inf with no conditionals outside the state's charge,
    probabilities summing to one."""
    L = Lattice(2, 3)
    aux = [1, 0, 1, 1, 0, 0]
    st = charged_peps(L, (1, 1), 11, aux, True)
    basis = ob.resolve_basis(True)
    ocache = ob.SamplingCache(st, 64, basis=basis, kind="physical")
    cache = spp.SamplingCache(gpu_ctx, to_product_state(spp, st), 64, kind="physical")
    cfgs = []
    for ks in itertools.product([0, 1], repeat=6):
        cfgs.append([[basis.labels[ks[x + 2 * y]] for y in range(3)] for x in range(2)])
    cfgs = np.array(cfgs)
    say("physical evaluation, 64 configurations")
    got = spp.sample_configurations(cache, configurations=cfgs)
    total = 0.0
    for i, cfg in enumerate(cfgs):
        ref = ob.sample_configuration(ocache, configuration=cfg)
        if ref["log_probability"] == -np.inf:
            assert got["log_probability"][i] == -np.inf
            one = spp.sample_configuration(cache, configuration=cfg)
            assert one["conditionals"] == [] and one["repairs"] == []
            continue
        assert abs(got["log_probability"][i] - ref["log_probability"]) < 1e-10
        total += np.exp(got["log_probability"][i])
    assert abs(total - 1.0) < 1e-10


def test_physical_fp32_draws(spp):
    """FP32 contractions (FP64 metric eigensolver): the same draws as FP64 on
    a truncating 3x3 Z2 case, log p within 1e-5."""
    L = Lattice(3, 3)
    aux = [0, 1, 1, 0, 0, 1, 0, 1, 0]
    st = charged_peps(L, (2, 1), 33, aux, True)
    uni = np.stack([ob.counter_uniforms(9, i, 9) for i in range(1, 6)])
    res = {}
    for prec in ("fp64", "fp32"):
        ctx = spp.Context(0, prec)
        try:
            cache = spp.SamplingCache(ctx, to_product_state(spp, st), 4, kind="physical")
            res[prec] = spp.sample_configurations(cache, uniforms=uni)
        finally:
            ctx.close()
    assert np.array_equal(res["fp64"]["configuration"], res["fp32"]["configuration"])
    assert np.max(np.abs(res["fp64"]["log_probability"] - res["fp32"]["log_probability"])) < 1e-5


@pytest.mark.parametrize("symmetric,D,chi", [(True, (2, 1), 3), (True, (2, 2), 16), (False, 3, 3)])
def test_physical_draws_fp32_match_oracle(spp, symmetric, D, chi):
    """Physical-kind draws with FP32 contractions (the factorised Gram transfers
    and left steps where the layers allow them) against the FP64 oracle on the
    same uniforms: configurations identical, except a draw whose first differing
    site has its uniform within 1e-4 of a cumulative-probability boundary of the
    oracle's conditional there; log p of the agreeing draws within 1e-4.

    The truncating cases use a chi that does not cut a degenerate pair: a
    projected PEPS row is a ket row times its own bra, so a north boundary's
    Gram spectrum holds the products s_i s_j of the ket boundary's values, each
    pair i != j twice. A chi that keeps one of a pair leaves the kept direction
    to rounding, in any implementation, and with the large truncation of these
    tiny lattices the draws' log p then differ at the 1e-2 level in FP64 as
    well (the same uniforms, from the first site below that boundary on)."""
    L = Lattice(3, 3)
    aux = [1, 0, 0, 1, 1, 0, 1, 0, 1]
    st = charged_peps(L, D, 23, aux, symmetric)
    basis = ob.resolve_basis(symmetric)
    ocache = ob.SamplingCache(st, chi, basis=basis, kind="physical")
    ctx = spp.Context(0, "fp32")
    try:
        cache = spp.SamplingCache(ctx, to_product_state(spp, st), chi, kind="physical")
        uni = np.stack([ob.counter_uniforms(9, i, 9) for i in range(1, 9)])
        say("physical draws fp32", symmetric, D, chi)
        got = spp.sample_configurations(cache, uniforms=uni)
    finally:
        ctx.close()
    near = []
    for i in range(len(uni)):
        ref = ob.sample_configuration(ocache, uniforms=uni[i])
        cfg = oracle_cfg(ref, L)
        if np.array_equal(cfg, got["configuration"][i]):
            assert abs(got["log_probability"][i] - ref["log_probability"]) <= 1e-4 * max(1.0, abs(ref["log_probability"]))
            continue
        near.append(boundary_distance(basis, ref, cfg, got["configuration"][i], uni[i], L))
    assert len(near) <= 1 and all(d < 1e-4 for d in near), near
