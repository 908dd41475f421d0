"""Host bookkeeping of the particles (stopeps.jl_amd/particles.py) on the CPU:
ensemble statistics, beta grids and configuration grouping against the
oracle's restatement and against closed forms."""
import math

import numpy as np
import pytest

from oracle import ensemble as oens


@pytest.fixture(scope="module")
def P(spp):
    from stopeps_jl_amd import particles
    return particles


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_ensemble_statistics_matches_oracle(P, seed):
    rng = np.random.default_rng(seed)
    K = int(rng.integers(1, 40))
    rel = rng.normal(scale=3.0, size=K) - 50.0
    mult = rng.integers(1, 5, size=K)
    e = rng.normal(size=K)
    Ns = int(mult.sum()) + int(rng.integers(0, 3))
    got = P.ensemble_statistics(rel, mult, e, Ns, log_scale=1.5, nsites=9)
    ref = oens.ensemble_statistics(rel, mult, e, Ns, log_scale=1.5, nsites=9)
    for k, v in ref.items():
        if isinstance(v, float) and math.isnan(v):
            assert math.isnan(got[k]), k
        else:
            assert got[k] == pytest.approx(v, rel=1e-13, abs=1e-13), k


def test_ensemble_statistics_closed_forms(P):
    e = np.array([-1.0, -2.0, -3.0])
    r = P.ensemble_statistics(np.zeros(3), [1, 1, 1], e, 3)
    assert r["eta"] == pytest.approx(1.0, abs=1e-15)
    assert r["energy_per_site"] == pytest.approx(-2.0, abs=1e-15)
    assert r["energy_per_site_error"] == pytest.approx(np.std(e, ddof=1) / np.sqrt(3), abs=1e-14)
    assert r["lnZ"] == pytest.approx(0.0, abs=1e-15) and r["lnZ_error"] == pytest.approx(0.0, abs=1e-15)
    # one dominant weight: eta -> 1/Nv, the estimate is that particle's energy
    r = P.ensemble_statistics(np.array([0.0, -800.0, -800.0]), [1, 1, 1], e, 3)
    assert r["eta"] == pytest.approx(1 / 3, abs=1e-12) and r["energy_per_site"] == pytest.approx(-1.0, abs=1e-12)
    # multiplicities count as repeated samples
    r = P.ensemble_statistics(np.zeros(2), [2, 1], np.array([-1.0, -4.0]), 3, nsites=4)
    assert r["energy_per_site"] == pytest.approx(-2.0, abs=1e-15) and r["virtual_samples"] == 3
    assert r["energy_total"] == pytest.approx(-8.0, abs=1e-14) and r["representatives"] == 2
    # lnZ of equal weights exp(c): log_scale + c
    r = P.ensemble_statistics(np.full(4, 2.5), [1, 1, 1, 1], np.zeros(4), 4, log_scale=-1.0)
    assert r["lnZ"] == pytest.approx(1.5, abs=1e-15)
    # unmeasured energies propagate as NaN, weights still give lnZ
    r = P.ensemble_statistics(np.zeros(2), [1, 1], np.array([np.nan, 1.0]), 2)
    assert math.isnan(r["energy_per_site"]) and r["lnZ"] == pytest.approx(0.0)
    with pytest.raises(ValueError):
        P.ensemble_statistics([], [], [], 1)
    with pytest.raises(ValueError):
        P.ensemble_statistics([np.inf], [1], [0.0], 1)


def test_beta_grids_match_oracle(P):
    for beta0, blog, bmeas, dt in [(0.5, [0.0, 0.5, 1.0, 2.0], [0.5, 2.0], 0.05),
                                   (0.1, list(np.arange(0.0, 0.31, 0.05)), [0.0, 0.1, 0.2, 0.3], 0.025),
                                   (1.0, [1.5, 2.0, 3.0, 4.0], [1.5, 2.0, 3.0, 4.0], 0.05)]:
        nodes = P.temperature_nodes(beta0, blog, bmeas)
        assert nodes == oens.temperature_nodes(beta0, blog, bmeas)
        for a, b in zip(nodes[:-1], nodes[1:]):
            s = P.step_schedule(a, b, dt)
            assert s == oens.step_schedule(a, b, dt)
            assert s[-1] == b and all(y > x for x, y in zip([a] + s[:-1], s))
            assert all(abs((y - x) - 2 * dt) < 1e-12 for x, y in zip([a] + s[:-2], s[:-1]))
    # beta 0.5 -> 2 with 2 dtau = 0.1 is fifteen steps ending exactly on 2
    s = P.step_schedule(0.5, 2.0, 0.05)
    assert len(s) == 15 and s[-1] == 2.0
    assert P.step_schedule(1.0, 1.0, 0.05) == []
    with pytest.raises(ValueError):
        P.step_schedule(1.0, 0.5, 0.05)


def test_group_configurations_matches_oracle(P):
    rng = np.random.default_rng(4)
    cfgs = rng.integers(1, 3, size=(30, 2, 2))
    got = P.group_configurations(cfgs)
    ref = oens.group_configurations(cfgs)
    assert got == ref
    assert sum(got["multiplicity"]) == 30
    for d, g in enumerate(got["draw_to_group"]):
        assert np.array_equal(cfgs[d], cfgs[got["first_draw"][g]])


def test_normalize_state(spp):
    from stopeps_jl_amd.states import normalize_state, random_peps, Lattice
    st = random_peps(Lattice(2, 2), (1, 1), seed=3)
    for (x, y) in st.lattice.site_order():
        st.sites[x][y].data *= 3.0
    rec = normalize_state(st)
    assert len(rec) == 4 and all(abs(r["log_normalization"] - np.log(3.0)) < 1e-14 for r in rec)
    assert abs(st.log_amplitude - 4 * np.log(3.0)) < 1e-13
    assert all(abs(np.linalg.norm(st[xy].data) - 1.0) < 1e-14 for xy in st.lattice.site_order())


def test_run_config_mirror(spp):
    """RunConfig defaults, derived chis, validation and the beta grid
    (config.jl:20-115)."""
    c = spp.RunConfig()
    assert c.chi_sampling == c.chi_measurement == c.chi_proposal == c.chi_environment == 32
    assert c.beta0 == 0.1 and c.Ns == 16 and c.seed == 20260924 and c.resampling.enabled
    assert spp.temperature_nodes(c) == [0.0, 0.05, 0.1, 0.15, 0.2, 0.25, 0.3]
    spp.validate_config(c)
    for bad in (dict(delta_tau_pur=0.0), dict(Ns=0), dict(trotter_order=3), dict(beta0=-1.0),
                dict(bond_space_pur=(0, 0), model=spp.TFIM(1.0, 1.0, symmetric=True))):
        with pytest.raises(ValueError):
            spp.validate_config(spp.RunConfig(**bad))
    p = spp.parameter_dict(spp.RunConfig(model=spp.TFIM(1.0, 3.04, symmetric=True), lattice=spp.Lattice(4, 4)))
    assert p["model"] == {"type": "TFIM", "J": 1.0, "h": 3.04, "physical_space": "Z2Space(0=>1,1=>1)"}
    assert p["lattice"] == {"type": "Lattice", "Lx": 4, "Ly": 4} and p["contraction"] == "BoundaryMPS-Row"
    assert p["sampling_basis"]["labels"] == [1, 2]
    from stopeps_jl_amd.config import bond_capacity, device_support
    z2 = spp.RunConfig(model=spp.TFIM(1.0, 1.0, symmetric=True))
    assert bond_capacity(z2, 6, (3, 3)) == (3, 3) and bond_capacity(z2, 4, None) == (4, 4)
    device_support(spp.RunConfig(), stochastic=True)  # resampling (enabled by default) runs on the device
    with pytest.raises(NotImplementedError):
        device_support(spp.RunConfig(contraction="BoundaryMPSDiagonal"), stochastic=False)


def test_nan_max_follows_julia(P):
    """``max``/``maximum`` on floats in the reference propagate NaN (a new
    particle's max_boundary_entropy is NaN, so is the node's maximum)."""
    assert P.nan_max([0.1, 0.3, 0.2]) == 0.3
    assert math.isnan(P.nan_max([0.1, math.nan, 0.2]))
    assert math.isnan(P.nan_max([0.5], init=math.nan))
    assert P.nan_max([], init=0.0) == 0.0
    with pytest.raises(ValueError):
        P.nan_max([])


def test_boundary_space_parsing(spp):
    """Boundary sector budgets of ``BoundaryMPS(chi, sector_space)``: an
    (even, odd) pair for Z2, one trivial sector for a dense space."""
    from stopeps_jl_amd.sampling import boundary_caps
    assert boundary_caps(None) == (-1, -1)
    assert boundary_caps((4, 2)) == (4, 2) and boundary_caps([0, 3]) == (0, 3)
    assert boundary_caps(6) == (6, -1)
    for bad in ((0, 0), (-1, 2), 0, -3):
        with pytest.raises(ValueError):
            boundary_caps(bad)


def test_run_config_boundary_spaces_and_entropy_convention(spp):
    """Sampling and measurement spaces default to the environment space
    (config.jl:44-47); the entropy conventions are the reference's strings
    (config.jl:147-151)."""
    c = spp.RunConfig(model=spp.TFIM(1.0, 1.0, symmetric=True), environment_space=(4, 2))
    assert c.sampling_space == c.measurement_space == c.proposal_space == (4, 2)
    spp.validate_config(c)
    from stopeps_jl_amd.config import device_support
    device_support(spp.RunConfig(environment_space=(4, 2), resampling=spp.ResamplingOptions(enabled=False)),
                   stochastic=True)
    with pytest.raises(ValueError):
        spp.validate_config(spp.RunConfig(model=spp.TFIM(1.0, 1.0, symmetric=True), measurement_space=(0, 0)))
    p = spp.parameter_dict(c)
    assert p["boundary_entropy_convention"] == "pretruncation_Schmidt_entropy; nats; sector_dimension_weighted"
    assert p["boundary_entropy_log_scope"] == "maximum_over_FU_boundary_cuts_in_output_interval"
    assert p["boundary_entropy_profile_scope"] == ("energy_measurement_boundaries; Se.north[y,b] and Se.south[y,b]; "
                                                   "cut_after_row_y_and_column_b")
