"""The reference's quick-start configuration through the GPU entry points
(examples/transverse_Ising_model/config.quickstart.jl: 4x4 dense TFIM J=1
h=0.7, XBasis, beta0=0.1, beta_log = beta_measure = [0, 0.1, 0.2, 0.3],
dtau 0.025, D=4, chi=16, FullUpdate(40, 100), Ns=32, resampling off), with the
result checks of the reference's reading guide (docs/src/guide/results.md:7-20):
the run is complete, the curve's betas are temperature_nodes(c), every energy
and error is finite. Also: lnZ(0) = N log 2 and the energy falls with beta."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_quickstart_configuration(spp, gpu_ctx):
    print("quick-start configuration 4x4 dense TFIM, 32 samples", flush=True)
    sys.stdout.flush()
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = spp.RunConfig(lattice=spp.Lattice(4, 4), model=spp.TFIM(J=1.0, h=0.7), sampling_basis="XBasis", beta0=0.1,
                      beta_log=[0.0, 0.1, 0.2, 0.3], beta_measure=[0.0, 0.1, 0.2, 0.3], delta_tau_pur=0.025,
                      delta_tau_peps=0.025, D_pur=4, D_peps=4, chi_environment=16, chi_sampling=16,
                      chi_measurement=16, chi_proposal=16,
                      update=spp.FullUpdate(max_iterations=40, linear_max_iterations=100), contraction="BoundaryMPSRow",
                      Ns=32, seed=20260924, resampling=spp.ResamplingOptions(enabled=False))
    assert spp.validate_config(c) is c
    cp = spp.run_purification(c, ctx)
    r = spp.run_stochastic(c, cp, ctx)
    assert r["metadata"]["complete"]
    rows = r["observables"]["curve"]
    assert [row["beta"] for row in rows] == spp.temperature_nodes(c)
    assert all(np.isfinite(row["energy_per_site"]) and np.isfinite(row["energy_per_site_error"]) for row in rows)
    for row in rows:
        print(f"  beta {row['beta']}: E/site {row['energy_per_site']:.6f} +- {row['energy_per_site_error']:.6f}, "
              f"lnZ {row['lnZ']:.6f}", flush=True)
    assert abs(rows[0]["lnZ"] - 16 * np.log(2.0)) < 1e-10 and abs(rows[0]["energy_per_site"]) < 1e-12
    e = [row["energy_per_site"] for row in rows]
    assert all(b < a for a, b in zip(e, e[1:]))
    assert len(r["observables"]["samples"]) > 0 and r["schema_version"] == "2.7.0"
