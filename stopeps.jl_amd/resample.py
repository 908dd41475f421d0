"""Annealed Born resampling of the stochastic ensemble (src/resample/event.jl,
src/resample/scores.jl, src/resample/selection.jl, src/driver/execution.jl:28-31),
FP64 host bookkeeping around the device path.

An event starts when the effective-sample fraction eta falls below eta_th
(src/driver/run.jl:117-122). Every particle gets a physical-kind sampling
cache of its evolved PEPS (``SamplingCache(..., kind="physical")``: the
PSD-projected site metrics on the device); the K x K leave-one-out scores
(each particle's proposal evaluated on every particle's configuration) set the
mixture coefficient lambda of the proposal q = (1 - lambda) p0 + lambda p_r,
p0 the beta0 Born distribution and p_r the weighted mixture of the particles'
physical proposals. Fresh samples are drawn from q, evolved again from beta0
through the recorded step endpoints, and the pool of old and fresh particles is
resampled by SIR to the target count. Every configuration probability is one
batched evaluation-mode draw per proposal member on the device.

Julia's MersenneTwister streams of ``_event_rng`` (src/sampling/born.jl:23-26)
cannot be reproduced without Julia; each stream is replaced by the documented
counter stream of the same identity (seed, domain, event, item), like the
per-sample uniforms of ``sample_uniforms``.
"""
import math
import time

import numpy as np

from .particles import (Particle, aligned_weights, advance_particles, group_configurations, measure_at, statistics)

MASK64 = (1 << 64) - 1
# _event_rng domains (event.jl:99, 107; selection.jl:86)
DOMAIN_MIXTURE = 0x4D49585455524543
DOMAIN_FRESH = 0x46524553485F424F
DOMAIN_SIR = 0x5349525F534C4F54


def _splitmix64(v):
    v = (v + 0x9E3779B97F4A7C15) & MASK64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK64
    return v ^ (v >> 31)


def event_uniforms(seed, domain, event, item, count):
    """The first ``count`` uniforms of the stream ``_event_rng(seed, domain,
    event, item)`` (born.jl:23-26), in the counter form this library uses for
    every reference stream: with s = splitmix64(splitmix64(splitmix64(seed ^
    domain) ^ event) ^ item), u_k = (splitmix64(s ^ k) >> 11) 2^-53."""
    s = _splitmix64(_splitmix64(_splitmix64((int(seed) ^ domain) & MASK64) ^ (int(event) & MASK64))
                    ^ (int(item) & MASK64))
    return np.array([(_splitmix64(s ^ k) >> 11) * 2.0 ** -53 for k in range(count)])


# ----------------------------------------------------------------------------
# selection.jl
# ----------------------------------------------------------------------------

def rs_logsumexp(x):
    """``_rs_logsumexp`` (selection.jl:1-7)."""
    x = np.asarray(x, dtype=np.float64)
    if x.size == 0:
        return -math.inf
    pivot = float(np.max(x))
    if pivot == -math.inf:
        return -math.inf
    if not math.isfinite(pivot):
        raise ValueError("nonfinite positive log mass")
    return pivot + math.log(float(np.sum(np.exp(x - pivot))))


def mixture_log_probability(log_p0, log_pr, lam):
    """``_mixture_log_probability`` (selection.jl:9-13): log((1-lambda) p0 + lambda pr)."""
    if not 0 <= lam < 1:
        raise ValueError("invalid resampling mixture coefficient")
    if lam == 0:
        return float(log_p0)
    return rs_logsumexp([math.log1p(-lam) + log_p0, math.log(lam) + log_pr])


def loo_log_probabilities(log_mass, scores):
    """``_loo_log_probabilities`` (selection.jl:16-25): for each configuration j
    the log probability of the mixture of the other members' proposals,
    weighted by their masses."""
    log_mass = np.asarray(log_mass, dtype=np.float64)
    scores = np.asarray(scores, dtype=np.float64)
    K = len(log_mass)
    if scores.shape != (K, K):
        raise ValueError("LOO scores must be K by K")
    out = np.zeros(K)
    for j in range(K):
        others = [k for k in range(K) if k != j]
        num = rs_logsumexp([log_mass[k] + scores[k, j] for k in others])
        out[j] = num - (rs_logsumexp(log_mass[others]) if others else 0.0)
    return out


def select_lambda(log_amplitude_squared, log_ratio, opts):
    """``_select_lambda`` (selection.jl:28-66): golden-section minimisation of
    the estimated second moment over lambda in [0, lambda_max], then the best
    of {0, the bracket's midpoint, lambda_max} (the smaller on ties)."""
    a2 = np.asarray(log_amplitude_squared, dtype=np.float64)
    lr = np.asarray(log_ratio, dtype=np.float64)
    if not (len(a2) == len(lr) > 0):
        raise ValueError("invalid LOO arrays")
    if not (np.all(np.isfinite(a2)) and all(math.isfinite(v) or v == -math.inf for v in lr)):
        raise ValueError("invalid LOO risk or probability ratio")
    offset = float(np.max(a2))
    amplitude = a2 - offset

    def risk(lam):
        return rs_logsumexp([amplitude[j] - mixture_log_probability(0.0, lr[j], lam) for j in range(len(amplitude))])

    if len(amplitude) == 1 or opts.lambda_max == 0:
        return {"lambda": 0.0, "log_risk_offset": offset, "log_risk_zero": risk(0.0),
                "log_risk_selected": risk(0.0), "log_risk_upper": risk(opts.lambda_max),
                "fallback": "single_lineage" if len(amplitude) == 1 else "zero_interval"}
    left, right = 0.0, opts.lambda_max
    golden = (math.sqrt(5.0) - 1) / 2
    x1, x2 = right - golden * (right - left), left + golden * (right - left)
    f1, f2 = risk(x1), risk(x2)
    while right - left > opts.lambda_tolerance:
        if f1 <= f2:
            right, x2, f2 = x2, x1, f1
            x1 = right - golden * (right - left)
            f1 = risk(x1)
        else:
            left, x1, f1 = x1, x2, f2
            x2 = left + golden * (right - left)
            f2 = risk(x2)
    candidates = sorted([0.0, (left + right) / 2, opts.lambda_max])
    risks = [risk(v) for v in candidates]
    best = min(risks)
    tie = 64 * np.finfo(np.float64).eps * max(1.0, max(abs(r) for r in risks))
    index = next(i for i, r in enumerate(risks) if r <= best + tie)
    return {"lambda": candidates[index], "log_risk_offset": offset, "log_risk_zero": risks[0],
            "log_risk_selected": risks[index], "log_risk_upper": risks[-1], "fallback": "none"}


def categorical(log_mass, u):
    """``_categorical`` (selection.jl:68-80) with the uniform given; 0-based."""
    z = rs_logsumexp(log_mass)
    if not math.isfinite(z):
        raise ValueError("proposal requires positive total probability")
    cumulative = 0.0
    last = -1
    for j, lm in enumerate(log_mass):
        if lm == -math.inf:
            continue
        last = j
        cumulative += math.exp(lm - z)
        if u < cumulative:
            return j
    return last


def sir_select_until_unique(log_weight, input_multiplicity, Ns, seed, event, uniform_for_draw=None):
    """``_sir_select_until_unique`` (selection.jl:84-119): draw pool members with
    probability proportional to weight x multiplicity until Ns distinct ones are
    held and the next draw would add another (or 3 Ns draws); a pool within the
    target is kept whole. Indices are 0-based."""
    if uniform_for_draw is None:
        def uniform_for_draw(slot):
            return float(event_uniforms(seed, DOMAIN_SIR, event, slot, 1)[0])
    if not (isinstance(Ns, (int, np.integer)) and Ns >= 1):
        raise ValueError("invalid representative target")
    log_weight = np.asarray(log_weight, dtype=np.float64)
    mult_in = [int(m) for m in input_multiplicity]
    if not (len(log_weight) == len(mult_in) > 0):
        raise ValueError("empty or inconsistent SIR pool")
    if not np.all(np.isfinite(log_weight)):
        raise ValueError("pool weights must be finite")
    if not all(m > 0 for m in mult_in):
        raise ValueError("pool multiplicities must be positive integers")
    K_pool, Nv_pool = len(log_weight), sum(mult_in)
    if K_pool <= Ns:
        return {"selected_candidate": [], "multiplicity": list(mult_in), "active": list(range(K_pool)),
                "common_log_weight": 0.0, "output_relative_log_weight": list(log_weight), "draw_count": 0,
                "discarded_candidate": None, "stop_reason": "pool_within_target", "K_pool": K_pool,
                "Nv_pool": Nv_pool, "Nv_post": Nv_pool}
    mass = log_weight + np.log(np.asarray(mult_in, dtype=np.float64))
    multiplicity = [0] * K_pool
    ids = []
    distinct = draw_count = 0
    discarded = None
    for slot in range(1, 3 * Ns + 1):
        i = categorical(mass, uniform_for_draw(slot))
        draw_count = slot
        if distinct == Ns and multiplicity[i] == 0:
            discarded = i
            break
        ids.append(i)
        distinct += multiplicity[i] == 0
        multiplicity[i] += 1
    return {"selected_candidate": ids, "multiplicity": multiplicity,
            "active": [i for i in range(K_pool) if multiplicity[i] > 0],
            "common_log_weight": rs_logsumexp(mass) - math.log(Nv_pool),
            "output_relative_log_weight": [0.0] * K_pool, "draw_count": draw_count,
            "discarded_candidate": discarded, "stop_reason": "draw_limit" if discarded is None else "next_distinct",
            "K_pool": K_pool, "Nv_pool": Nv_pool, "Nv_post": len(ids)}


# ----------------------------------------------------------------------------
# scores.jl
# ----------------------------------------------------------------------------

class ResamplingScores:
    """``ResamplingScores`` (scores.jl:1-14): configuration columns and the
    scored (member, column) entries of one event; member 0 is the beta0 Born
    distribution."""

    def __init__(self):
        self.configurations = []
        self.columns = {}
        self.entries = []
        self.index = {}

    def score_columns(self, configurations):
        """``_score_columns!`` (scores.jl:16-21)."""
        out = []
        for x in configurations:
            key = np.asarray(x).tobytes()
            if key not in self.columns:
                self.configurations.append(np.array(x))
                self.columns[key] = len(self.configurations) - 1
            out.append(self.columns[key])
        return out

    def record(self, entry):
        """``_record_score!`` (scores.jl:27-32)."""
        key = (entry["member_id"], entry["configuration_index"])
        if key in self.index:
            return
        self.entries.append(entry)
        self.index[key] = len(self.entries) - 1

    def cached(self, member_id, column):
        """``_cached_score`` (scores.jl:33)."""
        return self.entries[self.index[(member_id, column)]]

    def matrix(self, reference):
        """``_score_matrix`` (scores.jl:34-35): [member, configuration] log probabilities."""
        return np.array([[self.cached(i, c)["log_probability"] for c in reference["configuration_indices"]]
                         for i in reference["member_ids"]])


def _score_entry(member_id, column, log_probability, max_trunc, repairs, source):
    return {"member_id": member_id, "configuration_index": column, "log_probability": float(log_probability),
            "max_environment_truncation": float(max_trunc), "repairs": repairs, "source": source}


def score_configurations(memo, caches, member_ids, configurations, evaluate):
    """``_score_configurations!`` (scores.jl:38-71): every member's probability
    of every (new) configuration column. ``evaluate(cache, configurations)`` is
    the batched evaluation-mode draw; one call per member covers all its
    unscored columns. Returns the reference (member ids, columns)."""
    columns = memo.score_columns(configurations)
    unique_columns = list(dict.fromkeys(columns))
    for mid in member_ids:
        todo = [c for c in unique_columns if (mid, c) not in memo.index]
        if not todo or mid not in caches:
            continue
        r = evaluate(caches[mid], np.array([memo.configurations[c] for c in todo]))
        for k, c in enumerate(todo):
            memo.record(_score_entry(mid, c, r["log_probability"][k], r["max_environment_truncation"][k],
                                     r["repairs"][k], "score"))
    for mid in member_ids:
        for c in unique_columns:
            if (mid, c) not in memo.index:
                raise RuntimeError("missing proposal scores")
    return {"member_ids": list(member_ids), "configuration_indices": columns}


def seed_draw_scores(memo, draws):
    """``_seed_draw_scores!`` (scores.jl:74-87): a draw already evaluated its
    generating member."""
    columns = memo.score_columns([d["configuration"] for d in draws])
    for d, c in zip(draws, columns):
        memo.record(_score_entry(d["member_id"], c, d["selected_log_probability"], d["max_environment_truncation"],
                                 d["repairs"], "draw"))
    return columns


# ----------------------------------------------------------------------------
# event.jl
# ----------------------------------------------------------------------------

def resampling_plan(records, scores, c):
    """``_resampling_plan`` (event.jl:1-14)."""
    relative, _ = aligned_weights(records)
    relative = np.asarray(relative)
    mass = relative - relative.max() + np.log([r["count"] for r in records])
    loo = loo_log_probabilities(mass, scores)
    p0 = np.array([r["log_p0"] for r in records])
    physical_baseline = np.array([r["physical_log_norm"] for r in records]) - p0
    risk = mass + (physical_baseline - physical_baseline.max())
    selection = select_lambda(risk, loo - p0, c.resampling)
    stats = statistics(records, c.Ns, c.lattice.Lx * c.lattice.Ly)
    M = int(min(max(c.Ns - math.floor(stats["eta_p"] * c.Ns), 0), c.Ns - 1))
    return {"selection": selection, "M": M, "loo_log_probability": loo, "loo_log_ratio": loo - p0,
            "risk_log_amplitude_squared": risk, "parent_log_mass": mass, "pre_statistics": stats}


def resample_fresh_count(plan, opts):
    """``_resample_fresh_count`` (event.jl:16-19)."""
    needs_growth = opts.max_resample_sweep > 1 and plan["selection"]["lambda"] < opts.lambda_threshold
    return max(1, plan["M"]) if needs_growth else plan["M"]


def resample_stop(lam, sweep, opts):
    """``_resample_stop`` (event.jl:20-24)."""
    if lam >= opts.lambda_threshold:
        return "lambda_threshold"
    if sweep >= opts.max_resample_sweep:
        return "sweep_limit"
    return "continue"


def resample_slot(c, sweep, d):
    """``_resample_slot`` (event.jl:25): (sweep - 1) Ns + d."""
    return (sweep - 1) * c.Ns + d


# ----------------------------------------------------------------------------
# the event on the device (event.jl:27-234)
# ----------------------------------------------------------------------------

class EventSettings:
    """What an event takes from its run: the context, the beta0 auxiliary
    sampling cache (member 0), RunConfig-like fields (Ns, lattice, seed,
    resampling options) and the evolution / observation settings that the
    replay of fresh particles shares with the ensemble."""

    def __init__(self, ctx, cache, Ns, lattice, seed, resampling, chi_proposal, proposal_space, model,
                 chi_environment, chi_measurement, method, order, D, delta_tau, beta_measure, bond_space=None,
                 pepo_live=None, record_entropy=False, environment_space=None, measurement_space=None):
        self.ctx = ctx
        self.cache = cache
        self.Ns = int(Ns)
        self.lattice = lattice
        self.seed = int(seed)
        self.resampling = resampling
        self.chi_proposal = int(chi_proposal)
        self.proposal_space = proposal_space
        self.model = model
        self.chi_environment = chi_environment
        self.chi_measurement = chi_measurement
        self.method = method
        self.order = order
        self.D = D
        self.delta_tau = delta_tau
        self.beta_measure = beta_measure
        self.bond_space = bond_space
        self.pepo_live = pepo_live
        self.record_entropy = record_entropy
        self.environment_space = environment_space
        self.measurement_space = measurement_space

    @property
    def nsites(self):
        return self.lattice.Lx * self.lattice.Ly


def _evaluate(cache, configurations):
    """Evaluation-mode ``sample_configuration`` of every configuration under
    one member's cache (one batched device call)."""
    from .sampling import sample_configurations, repair_records
    r = sample_configurations(cache, configurations=np.asarray(configurations))
    reps = [repair_records(cache, x) for x in r["repairs"]] if "repairs" in r else [[] for _ in configurations]
    return {"log_probability": r["log_probability"], "max_environment_truncation": r["max_environment_truncation"],
            "repairs": reps}


def _proposal_cache(S, particle):
    """``SamplingCache(p.state, _boundary(c, :proposal); kind=:physical,
    basis=caches[0].basis)`` (event.jl:27-33) on the particle's resident PEPS."""
    from .sampling import SamplingCache
    return SamplingCache(S.ctx, particle.state, S.chi_proposal, basis=S.cache.basis, kind="physical",
                         sector_space=S.proposal_space)


def evaluate_resampling(memo, caches, particles, records, S, after_sweep):
    """``_evaluate_resampling!`` (event.jl:27-47): proposal caches of the
    particles that lack one, the K x K leave-one-out scores and the plan."""
    t0 = time.perf_counter()
    for p in particles:
        if p.id not in caches:
            caches[p.id] = _proposal_cache(S, p)
    reference = score_configurations(memo, caches, [r["id"] for r in records],
                                     [r["configuration"] for r in records], _evaluate)
    plan = resampling_plan(records, memo.matrix(reference), S)
    return {"after_sweep": after_sweep, "plan": plan, "score_reference": reference,
            "loo_seconds": time.perf_counter() - t0}


def draw_resampling_batch(memo, caches, records, plan, M, S, event, sweep):
    """``_draw_resampling_batch!`` (event.jl:96-158): M mixture members (0 =
    the beta0 distribution with probability 1 - lambda, else a particle drawn
    by its mass), one draw from each member's cache, the baseline and mixture
    scores of the distinct configurations, and each draw's proposal log q."""
    lam = plan["selection"]["lambda"]
    mass = plan["parent_log_mass"]
    members = []
    for d in range(1, M + 1):
        u = event_uniforms(S.seed, DOMAIN_MIXTURE, event, resample_slot(S, sweep, d), 2)
        members.append(categorical(mass, u[1]) + 1 if u[0] < lam else 0)
    draws = [None] * M
    by_member = {}
    for d in range(1, M + 1):
        m = members[d - 1]
        mid = 0 if m == 0 else records[m - 1]["id"]
        by_member.setdefault(mid, []).append(d)
    from .sampling import sample_configurations, repair_records
    for mid, ds in by_member.items():
        uni = np.stack([event_uniforms(S.seed, DOMAIN_FRESH, event, resample_slot(S, sweep, d), S.nsites)
                        for d in ds])
        r = sample_configurations(caches[mid], uniforms=uni)
        for k, d in enumerate(ds):
            reps = repair_records(caches[mid], r["repairs"][k]) if "repairs" in r else []
            draws[d - 1] = {"index": d, "slot": resample_slot(S, sweep, d), "member": members[d - 1],
                            "member_id": mid, "configuration": np.array(r["configuration"][k]),
                            "selected_log_probability": float(r["log_probability"][k]),
                            "max_environment_truncation": float(r["max_environment_truncation"][k]),
                            "repairs": reps}
    seed_draw_scores(memo, draws)
    grouping = group_configurations([d["configuration"] for d in draws])
    configurations = [draws[i]["configuration"] for i in grouping["first_draw"]]
    baseline_reference = score_configurations(memo, caches, [0], configurations, _evaluate)
    score_reference = None if lam == 0 else score_configurations(memo, caches, [r["id"] for r in records],
                                                                 configurations, _evaluate)
    normalizer = rs_logsumexp(mass)
    proposals = []
    for column in baseline_reference["configuration_indices"]:
        p0 = memo.cached(0, column)
        scored = [] if score_reference is None else [memo.cached(i, column) for i in score_reference["member_ids"]]
        pr = None if score_reference is None else \
            rs_logsumexp(np.asarray(mass) + np.array([s["log_probability"] for s in scored])) - normalizer
        q = mixture_log_probability(p0["log_probability"], pr, lam)
        err = max(p0["max_environment_truncation"], max((s["max_environment_truncation"] for s in scored),
                                                        default=0.0))
        proposals.append({"log_p0": p0["log_probability"], "log_pr": pr, "log_q": q,
                          "max_environment_truncation": err})
    for i, d in enumerate(draws):
        prop = proposals[grouping["draw_to_group"][i]]
        d["log_probability"] = prop["log_q"]
        d["max_environment_truncation"] = max(d["max_environment_truncation"], prop["max_environment_truncation"])
    return {"members": members, "draws": draws, "grouping": grouping, "configurations": configurations,
            "proposals": proposals, "baseline_reference": baseline_reference, "score_reference": score_reference}


def replay_fresh(fresh, replay, S, event, sweep, out=None):
    """``_replay_fresh!`` (event.jl:60-93): evolve the fresh particles from
    beta0 through the recorded step endpoints of every node, measuring energy
    at the last node when it is a measurement node; one replay statistics row
    per node. Returns the particles' records after the last node."""
    records = []
    for node, segment in enumerate(replay):
        if node > 0:
            advance_particles(S.ctx, fresh, S.model, segment["steps"], S.chi_environment, S.chi_measurement,
                              S.method, node == len(replay) - 1 and measure_at(S.beta_measure, segment["beta"]),
                              S.order, D=S.D, record_entropy=S.record_entropy,
                              environment_space=S.environment_space, measurement_space=S.measurement_space)
        records = [p.record() for p in fresh]
        if records and out is not None:
            stats = statistics(records, sum(r["count"] for r in records), S.nsites)
            stats.update({"event": event, "sweep": sweep, "source": "fresh_replay",
                          "time_seconds": sum(p.time_seconds for p in fresh)})
            out.append({"event": event, "sweep": sweep, "statistics": stats, "samples": records})
    return records


def pool_selection(old, fresh, S, event):
    """``_pool_selection`` (event.jl:49-57): SIR over the pool of old and fresh
    particle records on their common weight scale."""
    pool = list(old) + list(fresh)
    relative, scale = aligned_weights(pool)
    relative = np.asarray(relative)
    pivot = float(relative.max())
    sel = sir_select_until_unique(relative - pivot, [r["count"] for r in pool], S.Ns, S.seed, event)
    sel["pool"] = pool
    sel["common_log_weight"] = scale + pivot + sel["common_log_weight"]
    sel["pool_statistics"] = statistics(pool, S.Ns, S.nsites)
    return sel


def redistribute(particles, selection):
    """``_redistribute`` with ``_set_selected!`` (execution.jl:28-31,
    particles.jl:145-150): the selected particles with their new
    multiplicities and weights on the selection's common scale."""
    pool = {p.id: p for p in particles}
    out = []
    for i in selection["active"]:
        p = pool[selection["pool"][i]["id"]]
        p.count = int(selection["multiplicity"][i])
        p.relative_log_weight = float(selection["output_relative_log_weight"][i])
        p.weight_log_scale = float(selection["common_log_weight"])
        out.append(p)
    return out


def resample(particles, records, S, replay, event, next_sample_id, out=None):
    """``_resample!`` (event.jl:160-234): leave-one-out evaluation, up to
    max_resample_sweep sweeps of fresh proposals (each replayed from beta0),
    re-evaluation after every sweep until lambda reaches lambda_threshold, then
    SIR over old and fresh particles. Returns (particles, records, statistics,
    next_sample_id); ``out`` (a result dictionary) receives the audit record
    under diagnostics.events and the replay rows under observables.replay."""
    from .particles import project_particles
    started = time.perf_counter()
    beta = replay[-1]["beta"]
    opts = S.resampling
    memo = ResamplingScores()
    caches = {0: S.cache}
    evaluations, sweeps = [], []
    replay_rows = out["observables"]["replay"] if out is not None else None
    try:
        evaluation = evaluate_resampling(memo, caches, particles, records, S, 0)
        evaluations.append(evaluation)
        plan = evaluation["plan"]
        M = resample_fresh_count(plan, opts)
        pool_particles = list(particles)
        all_fresh = []
        stop_reason = "continue"
        for sweep in range(1, opts.max_resample_sweep + 1):
            t0 = time.perf_counter()
            batch = draw_resampling_batch(memo, caches, records + all_fresh, plan, M, S, event, sweep)
            proposal_time = time.perf_counter() - t0
            stop_reason = "no_fresh_draws" if M == 0 else resample_stop(plan["selection"]["lambda"], sweep, opts)
            if stop_reason != "continue":
                for k in [k for k in caches if k != 0]:
                    del caches[k]
            first_id = next_sample_id
            next_sample_id += S.Ns
            t0 = time.perf_counter()
            g = batch["grouping"]
            fresh = project_particles(
                S.ctx, S.cache, batch["configurations"], [first_id + d for d in g["first_draw"]],
                g["multiplicity"], [p["log_p0"] for p in batch["proposals"]],
                [p["log_q"] for p in batch["proposals"]],
                [p["max_environment_truncation"] for p in batch["proposals"]], S.model, S.chi_measurement,
                len(replay) == 1 and measure_at(S.beta_measure, beta), bond_space=S.bond_space,
                pepo_live=S.pepo_live, seed=S.seed, generation=event, record_entropy=S.record_entropy,
                measurement_space=S.measurement_space)
            fresh_records = [] if M == 0 else replay_fresh(fresh, replay, S, event, sweep, replay_rows)
            replay_time = time.perf_counter() - t0
            pool_particles += fresh
            all_fresh += fresh_records
            sweeps.append({"sweep": sweep, "evaluation": len(evaluations), "lambda": plan["selection"]["lambda"],
                           "M": M, "members": batch["members"], "draws": batch["draws"], "grouping": g,
                           "proposals": batch["proposals"], "baseline_reference": batch["baseline_reference"],
                           "score_reference": batch["score_reference"],
                           "fresh_ids": [r["id"] for r in fresh_records],
                           "pool_representatives": len(records) + len(all_fresh),
                           "pool_virtual_samples": sum(r["count"] for r in records + all_fresh),
                           "timing": {"proposal_seconds": proposal_time, "replay_seconds": replay_time}})
            if stop_reason != "continue":
                break
            evaluation = evaluate_resampling(memo, caches, pool_particles, records + all_fresh, S, sweep)
            evaluations.append(evaluation)
            plan = evaluation["plan"]
            if plan["selection"]["lambda"] >= opts.lambda_threshold:
                stop_reason = "lambda_threshold"
                break
    finally:
        caches.clear()
    selection = pool_selection(records, all_fresh, S, event)
    nxt = redistribute(pool_particles, selection)
    post = [p.record() for p in nxt]
    stats = statistics(post, S.Ns, S.nsites)
    stats.update({"source": "sampling", "time_seconds": sum(p.time_seconds for p in particles)})
    if out is not None:
        out["diagnostics"]["events"].append({
            "event": event, "beta": beta, "M": M, "lambda_evaluations": evaluations, "sweeps": sweeps,
            "stop_reason": stop_reason,
            "score_table": {"configurations": memo.configurations,
                            "entries": sorted(memo.entries, key=lambda r: (r["member_id"], r["configuration_index"]))},
            "pre": records, "fresh": all_fresh, "selection": selection, "post": post, "post_statistics": stats,
            "accepted": True, "proposal": "finite_chi_psd_sequential_basis",
            "timing": {"loo_seconds": sum(e["loo_seconds"] for e in evaluations),
                       "proposal_seconds": sum(s["timing"]["proposal_seconds"] for s in sweeps),
                       "replay_seconds": sum(s["timing"]["replay_seconds"] for s in sweeps),
                       "total_seconds": time.perf_counter() - started}})
    return nxt, post, stats, next_sample_id
