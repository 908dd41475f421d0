"""run_purification writing results/checkpoint.npz and run_stochastic resuming
from it (src/driver/run.jl:33-36, 76-88) give the same ensemble as the
in-memory checkpoint; the result file holds the returned record."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stochastic_run_from_the_checkpoint_file(spp, gpu_ctx, tmp_path):
    print("checkpoint file round trip, 3x3 Z2 TFIM", flush=True)
    sys.stdout.flush()
    ctx = gpu_ctx
    ctx.set_precision("fp64")
    c = spp.RunConfig(lattice=spp.Lattice(3, 3), model=spp.TFIM(1.0, 0.7, symmetric=True), beta0=0.2,
                      beta_log=[], beta_measure=[0.2, 0.3], delta_tau_pur=0.05, delta_tau_peps=0.05, D_pur=4,
                      D_peps=4, bond_space_pur=(2, 2), bond_space_peps=(2, 2), chi_environment=16, Ns=8, seed=1,
                      update=spp.FullUpdate(max_iterations=10, linear_max_iterations=50),
                      resampling=spp.ResamplingOptions(enabled=False), output_directory=str(tmp_path))
    cp = spp.run_purification(c, ctx)
    path = tmp_path / "results" / "checkpoint.npz"
    assert path.exists()
    with pytest.raises(FileExistsError):
        spp.run_purification(c, ctx)
    r_mem = spp.run_stochastic(c, cp, ctx)
    r_file = spp.run_stochastic(c, ctx=ctx)  # loads results/checkpoint.npz
    for a, b in zip(r_mem["observables"]["curve"], r_file["observables"]["curve"]):
        assert a["beta"] == b["beta"]
        for k in ("energy_per_site", "lnZ"):  # NaN where the node is not measured
            assert np.isclose(a[k], b[k], rtol=1e-12, atol=0, equal_nan=True), (a["beta"], k)
    saved = spp.load_result(str(tmp_path / "results" / "result.npz"))
    assert saved["metadata"]["complete"] and saved["schema_version"] == "2.7.0"
    assert [row["beta"] for row in saved["observables"]["curve"]] == [row["beta"] for row in r_file["observables"]["curve"]]
    np.testing.assert_array_equal([row["energy_per_site"] for row in saved["observables"]["curve"]],
                                  [row["energy_per_site"] for row in r_file["observables"]["curve"]])
