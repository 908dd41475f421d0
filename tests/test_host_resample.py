"""Host logic of annealed Born resampling (src/resample/selection.jl,
src/resample/event.jl), pinned by independent routes: linear-space sums
instead of log-sum-exp, a grid minimisation of the risk, the inverse-CDF walk
of a known uniform sequence, closed forms of the fresh-sample count."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def R(spp):
    from stopeps_jl_amd import resample
    return resample


def test_mixture_log_probability_linear_space(R):
    assert R.mixture_log_probability(-3.0, -1.0, 0.0) == -3.0
    for lam in (0.1, 0.5, 0.9):
        ref = math.log((1 - lam) * math.exp(-3.0) + lam * math.exp(-1.0))
        assert abs(R.mixture_log_probability(-3.0, -1.0, lam) - ref) < 1e-14
    assert R.mixture_log_probability(-2.0, -math.inf, 0.25) == pytest.approx(math.log(0.75) - 2.0, abs=1e-14)
    with pytest.raises(ValueError):
        R.mixture_log_probability(0.0, 0.0, 1.0)


def test_loo_log_probabilities_linear_space(R):
    rng = np.random.default_rng(1)
    K = 5
    log_mass = rng.normal(size=K)
    scores = rng.normal(size=(K, K)) - 3
    scores[1, 3] = -math.inf  # a member that cannot propose a configuration
    got = R.loo_log_probabilities(log_mass, scores)
    m = np.exp(log_mass)
    S = np.exp(scores)
    for j in range(K):
        others = [k for k in range(K) if k != j]
        ref = math.log(sum(m[k] * S[k, j] for k in others) / sum(m[k] for k in others))
        assert abs(got[j] - ref) < 1e-13
    # a member's own score never enters its configuration's LOO probability
    s2 = scores.copy()
    np.fill_diagonal(s2, 50.0)
    assert np.allclose(R.loo_log_probabilities(log_mass, s2), got, atol=1e-14)
    assert R.loo_log_probabilities([0.0], [[1.0]])[0] == -math.inf


def test_select_lambda_minimises_the_risk(R, spp):
    """The selected lambda's risk is within the golden-section tolerance of the
    minimum of risk(lambda) = log sum_j a_j / ((1 - lambda) + lambda r_j) on a
    fine grid over [0, lambda_max], and is one of {0, midpoint, lambda_max}."""
    opts = spp.ResamplingOptions()
    rng = np.random.default_rng(3)
    for trial in range(6):
        K = 6
        a2 = rng.normal(size=K)
        lr = rng.normal(scale=2.0, size=K)
        if trial == 5:
            lr[:] = -5.0  # the particles never propose the configurations: lambda = 0
        sel = R.select_lambda(a2, lr, opts)
        a = np.exp(a2 - a2.max())
        r = np.exp(lr)

        def risk(lam):
            return math.log(np.sum(a / ((1 - lam) + lam * r)))
        grid = np.linspace(0, opts.lambda_max, 20001)
        best = min(risk(v) for v in grid)
        assert risk(sel["lambda"]) <= best + 1e-6
        assert abs(sel["log_risk_selected"] - risk(sel["lambda"])) < 1e-12
        assert sel["fallback"] == "none"
        if trial == 5:
            assert sel["lambda"] == 0.0
    one = R.select_lambda([0.3], [1.0], opts)
    assert one["lambda"] == 0.0 and one["fallback"] == "single_lineage"


def test_categorical_and_sir_follow_the_inverse_cdf(R):
    with np.errstate(divide="ignore"):
        lm = np.log([0.2, 0.0, 0.5, 0.3])
    assert [R.categorical(lm, u) for u in (0.0, 0.1999, 0.2001, 0.69, 0.71, 0.9999999)] == [0, 0, 2, 2, 3, 3]
    # SIR: pool within the target is kept whole
    keep = R.sir_select_until_unique([0.1, -0.2], [2, 1], 4, seed=1, event=1)
    assert keep["stop_reason"] == "pool_within_target" and keep["multiplicity"] == [2, 1]
    assert keep["output_relative_log_weight"] == [0.1, -0.2]
    # a known uniform sequence: masses 1:2:1 (weights 1,1,1 with multiplicities 1,2,1)
    us = iter([0.1, 0.6, 0.3, 0.9, 0.5])  # -> 0, 1, 1, 2 (third distinct), then 1
    sel = R.sir_select_until_unique([0.0, 0.0, 0.0], [1, 2, 1], 2, seed=0, event=0,
                                    uniform_for_draw=lambda slot: next(us))
    assert sel["selected_candidate"] == [0, 1, 1]
    assert sel["discarded_candidate"] == 2 and sel["stop_reason"] == "next_distinct"
    assert sel["multiplicity"] == [1, 2, 0] and sel["active"] == [0, 1] and sel["Nv_post"] == 3
    assert abs(sel["common_log_weight"] - (math.log(4.0) - math.log(4.0))) < 1e-15
    # default streams: distinct counts never exceed the target; multiplicities sum to the draws kept
    rng = np.random.default_rng(4)
    lw = rng.normal(size=20)
    sel = R.sir_select_until_unique(lw, [1] * 20, 5, seed=7, event=2)
    assert len(sel["active"]) <= 5 and sum(sel["multiplicity"]) == sel["Nv_post"]
    assert sel["draw_count"] == sel["Nv_post"] + (sel["discarded_candidate"] is not None)


def test_event_uniforms_are_distinct_streams(R):
    a = R.event_uniforms(5, R.DOMAIN_FRESH, 1, 3, 16)
    assert np.all((a >= 0) & (a < 1)) and len(set(a)) == 16
    assert np.array_equal(a, R.event_uniforms(5, R.DOMAIN_FRESH, 1, 3, 16))
    for other in (R.event_uniforms(5, R.DOMAIN_MIXTURE, 1, 3, 16), R.event_uniforms(5, R.DOMAIN_FRESH, 2, 3, 16),
                  R.event_uniforms(5, R.DOMAIN_FRESH, 1, 4, 16), R.event_uniforms(6, R.DOMAIN_FRESH, 1, 3, 16)):
        assert not np.array_equal(a, other)


def test_resampling_plan_fresh_count(R, spp):
    """M = clamp(Ns - floor(eta_p Ns), 0, Ns - 1) (event.jl:10-11), the fresh
    count grows to 1 below lambda_threshold with several sweeps (event.jl:16-19),
    and the risk and LOO terms combine the records as stated there."""
    c = spp.RunConfig(lattice=spp.Lattice(2, 2), Ns=8)
    recs = []
    for i, (rel, cnt, pln, p0) in enumerate([(0.0, 3, -1.0, -2.0), (-1.0, 2, -0.5, -2.5), (-3.0, 3, -2.0, -1.5)]):
        recs.append({"id": i + 1, "count": cnt, "relative_log_weight": rel, "weight_log_scale": 0.0,
                     "physical_log_norm": pln, "log_p0": p0, "energy_per_site": -1.0 - i, "beta": 1.0,
                     "max_state_truncation": 0.0, "max_environment_truncation": 0.0,
                     "max_boundary_entropy": 0.0, "configuration": np.array([[i]])})
    scores = np.array([[-1.0, -2.0, -3.0], [-2.5, -1.0, -2.0], [-3.0, -2.0, -0.5]])
    plan = R.resampling_plan(recs, scores, c)
    w = np.exp([0.0, -1.0, -3.0]) * np.array([3, 2, 3])
    eta_p = min(max(w.sum() ** 2 / (8 * np.sum(w ** 2)), 1 / 8), 3 / 8)
    assert plan["M"] == min(max(8 - math.floor(eta_p * 8), 0), 7)
    mass = np.array([0.0, -1.0, -3.0]) + np.log([3, 2, 3])
    assert np.allclose(plan["parent_log_mass"], mass, atol=1e-15)
    base = np.array([-1.0 + 2.0, -0.5 + 2.5, -2.0 + 1.5])
    assert np.allclose(plan["risk_log_amplitude_squared"], mass + base - base.max(), atol=1e-15)
    opts = spp.ResamplingOptions()
    lam = plan["selection"]["lambda"]
    expect = max(1, plan["M"]) if lam < opts.lambda_threshold else plan["M"]
    assert R.resample_fresh_count(plan, opts) == expect
    assert R.resample_stop(0.5, 1, opts) == "lambda_threshold"
    assert R.resample_stop(0.1, 3, opts) == "sweep_limit" and R.resample_stop(0.1, 2, opts) == "continue"
    assert R.resample_slot(c, 2, 3) == 11
