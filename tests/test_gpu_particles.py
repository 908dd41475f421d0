"""GPU end-to-end parity of the stochastic-PEPS pipeline (stopeps.jl_amd/particles.py
over the C ABI): purification of the identity PEPO by full update to beta0,
Born draws, one particle per distinct configuration with its weight
log_norm - log q, full-update evolution of the particles to the next node,
observation and the ensemble statistics — against the same pipeline composed
from the oracle (run.jl:2-43, 45-66, 87-130; particles.jl:42-102;
ensemble.jl:8-40). The configuration is one whose oracle results move by
< 1e-11 under 1e-13 noise on the initial PEPO."""
import sys

import numpy as np
import pytest

from oracle.model import Lattice, TFIM, identity_pepo, normalize_state
from oracle import born as ob
from oracle import boundary as obm
from oracle import energy as oe
from oracle import ensemble as oens
from oracle import fu as ofu

pytestmark = pytest.mark.gpu

L3 = (3, 3)
H = 2.0
BETA0, BETA1, DT = 0.2, 0.4, 0.05
CAPS = (1, 1)
CHI = 16
NS = 6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_pipeline(bspace=None):
    lat = Lattice(*L3)
    chi = CHI if bspace is None else obm.BoundaryMPS(CHI, bspace)
    m = TFIM(J=1.0, h=H, symmetric=True)
    opts = ofu.FullUpdate(max_iterations=40, linear_max_iterations=100)
    D = sum(CAPS)
    pepo = identity_pepo(lat)
    normalize_state(pepo)
    pur = []
    b = 0.0
    pur_entropy = 0.0
    while b < BETA0 - 1e-12:
        rep = ofu.evolution_step(pepo, m, DT, D, CAPS, opts, chi, record_entropy=True)
        pur_entropy = max(pur_entropy, rep["max_boundary_entropy"])
        b += 2 * DT
    pepo.beta = BETA0
    pur.append(oe.measure_energy(pepo, m, chi, record_entropy=True))
    pur[-1]["fu_entropy"] = pur_entropy
    cache = ob.SamplingCache(pepo, chi)
    ns = lat.Lx * lat.Ly
    draws = [ob.sample_configuration(cache, uniforms=ob.counter_uniforms(1, i, ns)) for i in range(1, NS + 1)]
    cfgs = [d["configuration"] for d in draws]
    grp = oens.group_configurations(cfgs)
    parts = []
    for u, i in enumerate(grp["first_draw"]):
        ps, _ = ob.project_configuration(pepo, cfgs[i])
        r = oe.measure_energy(ps, m, chi, record_entropy=True)
        parts.append({"state": ps, "rel": r["log_norm"] - draws[i]["log_probability"],
                      "tln": r["tensor_log_norm"], "e": r["energy_per_site"], "count": grp["multiplicity"][u],
                      "Se": r["Se"]})
    stats = [oens.ensemble_statistics([p["rel"] for p in parts], [p["count"] for p in parts],
                                      [p["e"] for p in parts], NS, nsites=ns)]
    for p in parts:
        norm = 0.0
        bb = BETA0
        p["fu_entropy"] = 0.0
        while bb < BETA1 - 1e-12:
            rep = ofu.evolution_step(p["state"], m, DT, D, CAPS, opts, chi, record_entropy=True)
            norm += sum(x["log_normalization"] for x in rep["normalizations"])
            p["fu_entropy"] = max(p["fu_entropy"], rep["max_boundary_entropy"])
            bb += 2 * DT
        r = oe.measure_energy(p["state"], m, chi, record_entropy=True)
        p["Se"] = r["Se"]
        p["rel"] += 2 * norm + (r["tensor_log_norm"] - p["tln"])
        p["tln"] = r["tensor_log_norm"]
        p["e"] = r["energy_per_site"]
    stats.append(oens.ensemble_statistics([p["rel"] for p in parts], [p["count"] for p in parts],
                                          [p["e"] for p in parts], NS, nsites=ns))
    return pur, cfgs, grp, parts, stats


@pytest.mark.parametrize("bspace", [None, (2, 1)])
def test_stochastic_pipeline_matches_oracle(spp, gpu_ctx, bspace):
    say(f"stochastic PEPS pipeline 3x3 Z2 D=(1,1) chi=16, boundary space {bspace}: "
        "RunConfig -> run_purification -> run_stochastic")
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = spp.RunConfig(lattice=spp.Lattice(*L3), model=spp.TFIM(1.0, H, symmetric=True), beta0=BETA0, beta_log=[],
                      beta_measure=[BETA0, BETA1], delta_tau_pur=DT, delta_tau_peps=DT, D_pur=sum(CAPS),
                      D_peps=sum(CAPS), bond_space_pur=CAPS, bond_space_peps=CAPS, chi_environment=CHI,
                      update=spp.FullUpdate(max_iterations=40, linear_max_iterations=100), Ns=NS, seed=1,
                      resampling=spp.ResamplingOptions(enabled=False), environment_space=bspace)
    cp = spp.run_purification(c, ctx)
    rows = cp.purification
    pur, cfgs, grp, parts, stats = oracle_pipeline(bspace)
    assert [r["beta"] for r in rows] == [0.0, BETA0]
    assert abs(rows[-1]["energy_total"] - pur[0]["energy_total"]) < 1e-9
    assert abs(rows[-1]["lnZ"] - pur[0]["log_norm"]) < 1e-9
    # beta = 0 (not a measurement node, so only lognorm): the normalised
    # identity PEPO has lnZ = log Tr 1 = Lx*Ly*log 2
    assert np.isnan(rows[0]["energy_total"]) and abs(rows[0]["lnZ"] - 9 * np.log(2.0)) < 1e-12
    # boundary entropy (record_boundary_entropy defaults to true): no steps
    # before beta = 0, the FU maximum of the steps to beta0, and the measured Se
    assert np.isnan(rows[0]["max_boundary_entropy"]) and rows[0]["Se"] is None
    assert abs(rows[-1]["max_boundary_entropy"] - pur[0]["fu_entropy"]) < 1e-8
    for side in ("north", "south"):
        np.testing.assert_allclose(rows[-1]["Se"][side], pur[0]["Se"][side], atol=1e-9)
    result = spp.run_stochastic(c, cp, ctx)
    assert result["schema_version"] == "2.7.0" and result["metadata"]["complete"]
    curve = result["observables"]["curve"]
    assert [r["beta"] for r in curve] == spp.temperature_nodes(c)
    particles = result["particles"]
    assert len(particles) == len(grp["first_draw"])
    assert len(result["observables"]["samples"]) == 2 * len(particles)
    for p, i, mult in zip(particles, grp["first_draw"], grp["multiplicity"]):
        assert np.array_equal(p.configuration, cfgs[i]) and p.count == mult
    for p, q in zip(particles, parts):
        say(f"  particle {p.id}: weight {p.relative_log_weight:.10f} vs {q['rel']:.10f}, "
            f"E/site {p.energy_per_site:.10f} vs {q['e']:.10f}")
        assert abs(p.relative_log_weight - q["rel"]) < 1e-8
        assert abs(p.energy_per_site - q["e"]) < 1e-9
        assert abs(p.beta - BETA1) < 1e-15
        assert abs(p.max_boundary_entropy - q["fu_entropy"]) < 1e-8
        for side in ("north", "south"):
            np.testing.assert_allclose(p.Se[side], q["Se"][side], atol=1e-9)
    srows = [r for r in curve if r.get("source") == "sampling"]
    assert [r["beta"] for r in srows] == [BETA0, BETA1]
    for r, s in zip(srows, stats):
        for k in ("energy_per_site", "energy_per_site_error", "lnZ", "lnZ_error", "eta", "ESS"):
            assert abs(r[k] - s[k]) <= 1e-8 * max(1.0, abs(s[k])), (k, r[k], s[k])
        assert r["virtual_samples"] == NS and r["representatives"] == len(parts)
    # new particles carry no FU entropy (NaN); after the steps the maximum over particles
    assert np.isnan(srows[0]["max_boundary_entropy"])
    assert abs(srows[1]["max_boundary_entropy"] - max(q["fu_entropy"] for q in parts)) < 1e-8
