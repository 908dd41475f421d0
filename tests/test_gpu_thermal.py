"""Thermal physics of the whole GPU pipeline (RunConfig -> run_purification ->
run_stochastic) against exact diagonalization of a 3x3 open TFIM: the
purified PEPO's energy and lnZ at beta0, and the stochastic-PEPS estimates at
beta0 and after full-update evolution, within their sampling errors plus a
Trotter/truncation allowance. The draws are deterministic (sample indices 1..Ns
of seed 1), so the test is reproducible."""
import sys

import numpy as np
import pytest

from oracle.model import Lattice, TFIM, exact_thermal_observables

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_thermal_energy_3x3_matches_exact_diagonalization(spp, gpu_ctx):
    say("thermal 3x3 TFIM J=1 h=0.7 Z2, D=4 (2+2), chi=16: purification to 0.5, 64 samples evolved to 1.0")
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = spp.RunConfig(lattice=spp.Lattice(3, 3), model=spp.TFIM(1.0, 0.7, symmetric=True), beta0=0.5, beta_log=[],
                      beta_measure=[0.5, 1.0], delta_tau_pur=0.05, delta_tau_peps=0.05, D_pur=4, D_peps=4,
                      bond_space_pur=(2, 2), bond_space_peps=(2, 2), chi_environment=16,
                      update=spp.FullUpdate(max_iterations=40, linear_max_iterations=100), Ns=64, seed=1,
                      resampling=spp.ResamplingOptions(enabled=False))
    cp = spp.run_purification(c, ctx)
    r = spp.run_stochastic(c, cp, ctx)
    L = Lattice(3, 3)
    m = TFIM(J=1.0, h=0.7, symmetric=False)  # ED in the Z basis: the spectrum is basis independent
    pur = cp.purification[-1]
    ex0 = exact_thermal_observables(L, m, 0.5)
    say(f"  purification beta 0.5: E/site {pur['energy_per_site']:.6f} (exact {ex0['energy_per_site']:.6f}), "
        f"lnZ {pur['lnZ']:.6f} (exact {ex0['lnZ']:.6f})")
    assert abs(pur["energy_per_site"] - ex0["energy_per_site"]) < 2e-3
    assert abs(pur["lnZ"] - ex0["lnZ"]) < 2e-3 * 9
    for row in [q for q in r["observables"]["curve"] if q.get("source") == "sampling"]:
        ex = exact_thermal_observables(L, m, row["beta"])
        say(f"  stochastic beta {row['beta']}: E/site {row['energy_per_site']:.6f} +- {row['energy_per_site_error']:.6f} "
            f"(exact {ex['energy_per_site']:.6f}), lnZ {row['lnZ']:.6f} +- {row['lnZ_error']:.6f} "
            f"(exact {ex['lnZ']:.6f}), eta {row['eta']:.3f}")
        assert abs(row["energy_per_site"] - ex["energy_per_site"]) < 4 * row["energy_per_site_error"] + 5e-3
        assert abs(row["lnZ"] - ex["lnZ"]) < 4 * row["lnZ_error"] + 2e-3 * 9
