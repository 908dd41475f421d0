"""The reference's plain CPU-runnable workload as its configuration states it,
through the GPU entry points: 4x4 open TFIM J=1 h=0.7, Z2 bond space
Z2Space(0=>2,1=>2) (D=4), second-order Trotter with delta_tau=0.05, PEPO
purification from beta=0 to beta0=1.0 by full update, then 64 Born samples
(seed 1, sample indices 1..64) evolved by full update to beta in {1.5, 2, 3, 4},
boundary chi=16, FP64, FullUpdate(40, 100), resampling off
(src/driver/run.jl:2-43, 76-134).

Checks are those of the reference's reading guide
(docs/src/guide/results.md:7-20) plus the physics the result must obey: the
run is complete; the curve's betas are temperature_nodes(c); every energy,
error and lnZ is finite; lnZ(0) = N log 2 and E(0) = 0; the energy falls
monotonically with beta; eta lies in [1/Nv, 1]; and the same workload on the
dense model (ComplexSpace(2), D=4, XBasis) agrees with the Z2 run within the
combined sampling error plus a truncation allowance."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BETAS = [1.5, 2.0, 3.0, 4.0]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(spp, ctx, symmetric):
    c = spp.RunConfig(lattice=spp.Lattice(4, 4), model=spp.TFIM(J=1.0, h=0.7, symmetric=symmetric), beta0=1.0,
                      beta_log=[], beta_measure=[0.0, 1.0] + BETAS, delta_tau_pur=0.05, delta_tau_peps=0.05,
                      D_pur=4, D_peps=4, bond_space_pur=(2, 2) if symmetric else None,
                      bond_space_peps=(2, 2) if symmetric else None, chi_environment=16,
                      update=spp.FullUpdate(max_iterations=40, linear_max_iterations=100), Ns=64, seed=1,
                      resampling=spp.ResamplingOptions(enabled=False))
    assert spp.validate_config(c) is c
    cp = spp.run_purification(c, ctx)
    r = spp.run_stochastic(c, cp, ctx)
    return c, r


def _check_run(spp, c, r, label):
    assert r["metadata"]["complete"] and r["metadata"]["stage_complete"]
    assert r["schema_version"] == "2.7.0"
    rows = r["observables"]["curve"]
    assert [row["beta"] for row in rows] == spp.temperature_nodes(c) == [0.0, 1.0] + BETAS
    for row in rows:
        say(f"  {label} beta {row['beta']}: E/site {row['energy_per_site']:.6f} +- "
            f"{row['energy_per_site_error']:.6f}, lnZ {row['lnZ']:.6f}, eta {row['eta']}")
        for k in ("energy_per_site", "energy_per_site_error", "lnZ", "lnZ_error"):
            assert np.isfinite(row[k]), (row["beta"], k)
    assert abs(rows[0]["lnZ"] - 16 * np.log(2.0)) < 1e-10 and abs(rows[0]["energy_per_site"]) < 1e-12
    e = [row["energy_per_site"] for row in rows]
    assert all(b < a for a, b in zip(e, e[1:])), e
    for row in rows[1:]:
        Nv = row["virtual_samples"]
        assert Nv == 64 and 1.0 / Nv <= row["eta"] <= 1.0
    assert len(r["observables"]["samples"]) > 0
    return {row["beta"]: row for row in rows}


def test_cfg1_z2_purification_and_ensemble_to_beta4(spp, gpu_ctx):
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    say("4x4 Z2 TFIM h=0.7 D=4 (2+2) chi=16: purify to 1.0, 64 samples to beta 1.5, 2, 3, 4")
    c, rz = _run(spp, ctx, True)
    z = _check_run(spp, c, rz, "Z2")
    say("4x4 dense TFIM h=0.7 D=4 chi=16, same workload")
    c, rd = _run(spp, ctx, False)
    d = _check_run(spp, c, rd, "dense")
    for b in [1.0] + BETAS:
        err = np.hypot(z[b]["energy_per_site_error"], d[b]["energy_per_site_error"])
        say(f"  beta {b}: Z2 {z[b]['energy_per_site']:.6f}  dense {d[b]['energy_per_site']:.6f}  (sigma {err:.6f})")
        assert abs(z[b]["energy_per_site"] - d[b]["energy_per_site"]) < 4 * err + 5e-3
