"""Result and checkpoint files (src/io/results.jl:8-80, src/io/run_output.jl:1-39):
round trips, atomic publication and the checks of load_checkpoint (CPU only)."""
import math
import os

import numpy as np
import pytest


def test_result_round_trip(spp, tmp_path):
    rng = np.random.default_rng(1)
    obs = {"curve": [{"beta": 0.5, "energy_per_site": -1.9, "eta": float("nan"), "lnZ_error": float("inf"),
                      "Se": {"north": rng.standard_normal((7, 7)), "south": np.zeros((2, 2))}}],
           "samples": [{"configuration": np.array([[1, 2], [2, 1]]), "count": 3, "bond": ((0, 0), (1, 0))}],
           "big": rng.standard_normal((10, 20)), 7: "integer key"}
    path = tmp_path / "results" / "result.npz"
    out = spp.save_result(str(path), method="stoPEPS", model={"type": "TFIM", "J": 1.0, "h": 2.0},
                          parameters={"Ns": 16, "bond_space_peps": (3, 3)}, observables=obs,
                          diagnostics={"history": [{"phase": "initialization", "beta": 0.0}]},
                          metadata={"run_id": "r1", "complete": True})
    assert out == os.path.abspath(path)
    d = spp.load_result(str(path))
    assert d["schema_version"] == spp.RESULT_SCHEMA_VERSION == "2.7.0" and d["method"] == "stoPEPS"
    assert d["parameters"]["bond_space_peps"] == (3, 3) and d["metadata"]["run_id"] == "r1"
    row = d["observables"]["curve"][0]
    assert math.isnan(row["eta"]) and row["lnZ_error"] == float("inf")
    np.testing.assert_array_equal(row["Se"]["north"], obs["curve"][0]["Se"]["north"])
    np.testing.assert_array_equal(d["observables"]["big"], obs["big"])
    assert d["observables"]["samples"][0]["bond"] == ((0, 0), (1, 0)) and d["observables"][7] == "integer key"
    np.testing.assert_array_equal(d["observables"]["samples"][0]["configuration"], [[1, 2], [2, 1]])


def test_atomic_publication_and_overwrite(spp, tmp_path):
    path = str(tmp_path / "r.npz")
    spp.save_result(path, method="a", model={}, observables={"x": 1})
    with pytest.raises(FileExistsError):
        spp.save_result(path, method="b", model={}, observables={"x": 2}, overwrite=False)
    assert spp.load_result(path)["method"] == "a"
    spp.save_result(path, method="c", model={}, observables={"x": 3}, overwrite=True)
    assert spp.load_result(path)["observables"]["x"] == 3
    assert sorted(os.listdir(tmp_path)) == ["r.npz"]  # no temporary files left behind
    with pytest.raises(ValueError):
        spp.save_result(str(tmp_path / "r.jld2"), method="a", model={})


def _config(spp, beta0=0.5):
    return spp.RunConfig(lattice=spp.Lattice(3, 2), model=spp.TFIM(1.0, 2.0, symmetric=True), beta0=beta0,
                         beta_log=[], beta_measure=[beta0], bond_space_pur=(2, 1), Ns=4, seed=1,
                         resampling=spp.ResamplingOptions(enabled=False))


def test_checkpoint_round_trip_and_checks(spp, tmp_path):
    c = _config(spp)
    st = spp.random_peps(spp.Lattice(3, 2), (2, 1), seed=3, kind="pepo")
    st.beta, st.log_amplitude = 0.5, 1.25
    cp = spp.ThermalCheckpoint(st, spp.parameter_dict(c), [{"beta": 0.5, "energy_per_site": -1.0, "Se": None}],
                               [{"phase": "initialization", "beta": 0.0}], "run-7")
    path = str(tmp_path / "results" / "checkpoint.npz")
    spp.save_checkpoint(path, cp)
    with pytest.raises(FileExistsError):
        spp.save_checkpoint(path, cp)
    back = spp.load_checkpoint(path, c)
    assert back.run_id == "run-7" and back.purification[0]["energy_per_site"] == -1.0
    assert back.state.beta == 0.5 and back.state.log_amplitude == 1.25
    for xy in st.lattice.site_order():
        np.testing.assert_array_equal(back.state[xy].data, st[xy].data)
        for a, b in zip(back.state[xy].parity, st[xy].parity):
            np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError, match="checkpoint model, lattice or beta0 mismatch"):
        spp.load_checkpoint(path, _config(spp, beta0=0.4))
    d = spp.load_result(path)
    d["parameters"]["checkpoint_format"] = "other"
    bad = str(tmp_path / "bad.npz")
    spp.save_result(bad, method=d["method"], model=d["model"], parameters=d["parameters"],
                    observables=d["observables"], diagnostics=d["diagnostics"], metadata=d["metadata"])
    with pytest.raises(ValueError, match="unsupported checkpoint format"):
        spp.load_checkpoint(bad, c)
