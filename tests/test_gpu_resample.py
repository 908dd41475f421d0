"""GPU parity of an annealed Born resampling event (src/resample/event.jl,
scores.jl) on particles evolved by the device: the K x K leave-one-out scores
(physical-kind evaluations under every particle's proposal cache), the fresh
draws of every mixture member and their proposal probabilities log q, against
the oracle's physical- and auxiliary-kind sampling of the same states with the
same event streams; and a run of run_stochastic with resampling on.

Bars: FP64 log probabilities within 1e-9 (relative to max(1, |log p|)), the
same configurations, -Inf where the oracle has -Inf."""
import math
import sys

import numpy as np
import pytest

from conftest import to_oracle_state, oracle_cfg
from oracle import born as ob

pytestmark = pytest.mark.gpu

L = (3, 3)
H = 2.0
BETA0, BETA1, DT = 0.2, 0.4, 0.05
CAPS = (1, 1)
CHI = 16
CHI_PROPOSAL = 6
NS = 8


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _config(spp, **kw):
    base = dict(lattice=spp.Lattice(*L), model=spp.TFIM(1.0, H, symmetric=True), beta0=BETA0, beta_log=[],
                beta_measure=[BETA0, BETA1], delta_tau_pur=DT, delta_tau_peps=DT, D_pur=sum(CAPS),
                D_peps=sum(CAPS), bond_space_pur=CAPS, bond_space_peps=CAPS, chi_environment=CHI,
                chi_proposal=CHI_PROPOSAL, Ns=NS, seed=3, record_boundary_entropy=False)
    base.update(kw)
    return spp.RunConfig(**base)


def _close(a, b, tol=1e-9):
    if b == -math.inf or a == -math.inf:
        return a == b
    return abs(a - b) <= tol * max(1.0, abs(b))


def test_event_scores_draws_and_proposals_match_oracle(spp, gpu_ctx):
    from stopeps_jl_amd import resample as R
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = _config(spp, resampling=spp.ResamplingOptions(enabled=False))
    say("resampling event parity: purify 3x3 Z2 to beta0, 8 draws, evolve to beta1")
    cp = spp.run_purification(c, ctx)
    out = spp.run_stochastic(c, checkpoint=cp, ctx=ctx)
    particles = out["particles"]
    records = [p.record() for p in particles]
    cache = spp.SamplingCache(ctx, cp.state, CHI)
    S = R.EventSettings(ctx, cache, NS, c.lattice, c.seed, spp.ResamplingOptions(), CHI_PROPOSAL, None, c.model,
                        CHI, CHI, c.update, 2, sum(CAPS), DT, c.beta_measure, bond_space=CAPS)
    memo = R.ResamplingScores()
    caches = {0: cache}
    say(f"  {len(particles)} particles: leave-one-out scores")
    ev = R.evaluate_resampling(memo, caches, particles, records, S, 0)
    scores = memo.matrix(ev["score_reference"])
    basis = ob.resolve_basis(True)
    ostates = [to_oracle_state(p.state.download()) for p in particles]
    ocaches = [ob.SamplingCache(st, CHI_PROPOSAL, basis=basis, kind="physical") for st in ostates]
    for i, oc in enumerate(ocaches):
        for j, r in enumerate(records):
            ref = ob.sample_configuration(oc, configuration=r["configuration"])["log_probability"]
            assert _close(scores[i, j], ref), (i, j, scores[i, j], ref)
    # fresh draws with a forced mixture coefficient so that particles propose
    plan = dict(ev["plan"])
    plan["selection"] = dict(plan["selection"], **{"lambda": 0.8})
    M = 6
    say("  fresh draws from the mixture (lambda 0.8)")
    batch = R.draw_resampling_batch(memo, caches, records, plan, M, S, event=1, sweep=1)
    opepo = to_oracle_state(cp.state.download())
    obase = ob.SamplingCache(opepo, CHI, basis=basis)
    ns = L[0] * L[1]
    assert any(d["member_id"] != 0 for d in batch["draws"])
    for d in batch["draws"]:
        u = R.event_uniforms(c.seed, R.DOMAIN_FRESH, 1, d["slot"], ns)
        oc = obase if d["member_id"] == 0 else ocaches[[r["id"] for r in records].index(d["member_id"])]
        ref = ob.sample_configuration(oc, uniforms=u)
        assert np.array_equal(d["configuration"], oracle_cfg(ref, c.lattice)), d["index"]
        assert _close(d["selected_log_probability"], ref["log_probability"])
    # proposal probabilities of every distinct configuration
    mass = plan["parent_log_mass"]
    norm = R.rs_logsumexp(mass)
    for cfg, prop in zip(batch["configurations"], batch["proposals"]):
        p0 = ob.sample_configuration(obase, configuration=cfg)["log_probability"]
        pr = R.rs_logsumexp([mass[k] + ob.sample_configuration(ocaches[k], configuration=cfg)["log_probability"]
                             for k in range(len(records))]) - norm
        assert _close(prop["log_p0"], p0) and _close(prop["log_pr"], pr)
        assert _close(prop["log_q"], R.mixture_log_probability(p0, pr, 0.8))


def test_run_stochastic_with_resampling(spp, gpu_ctx):
    """run_stochastic with resampling on and eta_th = 1 (an event at every node
    whose weights are not all equal): events are recorded with their LOO
    evaluations and sweeps, fresh particles carry ids above Ns and the event's
    generation, the post-event ensemble has at most Ns representatives whose
    multiplicities are the SIR counts, and the curve stays finite."""
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = _config(spp, resampling=spp.ResamplingOptions(eta_th=1.0, max_resample_sweep=2))
    say("run_stochastic with resampling, 3x3 Z2, eta_th = 1")
    cp = spp.run_purification(c, ctx)
    out = spp.run_stochastic(c, checkpoint=cp, ctx=ctx)
    events = out["diagnostics"]["events"]
    assert events, "no resampling event"
    for e in events:
        assert e["stop_reason"] in ("lambda_threshold", "sweep_limit", "no_fresh_draws")
        assert 1 <= len(e["sweeps"]) <= 2 and e["lambda_evaluations"]
        sel = e["selection"]
        assert len(sel["active"]) <= NS
        for r in e["fresh"]:
            assert r["id"] > NS and r["generation"] == e["event"]
        post = e["post"]
        assert sum(r["count"] for r in post) == sel["Nv_post"]
        assert all(math.isfinite(r["relative_log_weight"]) for r in post)
        for s in e["sweeps"]:
            for d, prop in zip(s["grouping"]["first_draw"], s["proposals"]):
                assert math.isfinite(prop["log_q"]) and prop["log_q"] <= 1e-12
    for row in out["observables"]["curve"]:
        if row["beta"] >= BETA0:
            assert math.isfinite(row["lnZ"]) and math.isfinite(row["energy_per_site"])
            assert 0 < row["eta"] <= 1
    assert out["observables"]["replay"], "fresh particles were not replayed"
