"""Fixtures of the 8x8 TFIM Born-draw workload (J=1, h=2.0, Z2 bonds 3+3,
D=6, chi=32, beta0=0.5), the configuration the headline throughput is quoted on.

Two stages:

``--stage pepo`` (needs the GPU library): ``run_purification`` of the
configuration (identity PEPO, second-order Trotter, delta_tau=0.05, five full-
update steps to beta0=0.5, FullUpdate(max_iterations=40,
linear_max_iterations=100), bond space Z2Space(0=>3,1=>3), chi=32, FP64;
src/driver/run.jl:2-43). The beta0 PEPO is trimmed to its live bond sectors
(the reference's truncated spaces) and written with its parities, beta and
log_amplitude to ``cfg2_pepo_beta05.npz``. The frozen PEPO is the input of
every check below; how it was produced does not enter the parity claim.

``--stage oracle`` (CPU only): the FP64 oracle's Born draws of sample indices
1..8 on that PEPO (uniforms ``sample_uniforms(1, i, 64)``, one per site in
row-major order), their log probabilities, conditionals and
max_environment_truncation (src/sampling/born.jl:86-163), the projected
samples' tensor log norms at chi=32 (src/contraction/boundary_mps.jl:180-191)
and ``measure_energy`` (src/measurement/energy.jl:31-78) of samples 1..4, into
``cfg2_oracle.npz``. One process per sample with one BLAS thread each.

Usage:
  python tests/golden/make_cfg2_fixture.py --stage pepo --out gpurun_out/cfg2_pepo_beta05.npz
  python tests/golden/make_cfg2_fixture.py --stage oracle --jobs 8
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

PEPO_FILE = os.path.join(HERE, "cfg2_pepo_beta05.npz")
ORACLE_FILE = os.path.join(HERE, "cfg2_oracle.npz")
L, J, H, CHI, BETA0, DTAU, SEED = 8, 1.0, 2.0, 32, 0.5, 0.05, 1
DRAWS, ENERGIES = 8, 4


def trim_live(state, live):
    """Keep the live slots of every virtual leg (the first live[site, leg, q]
    indices of parity q): the bond spaces the reference holds after the
    full-update truncations."""
    from stopeps_jl_amd.states import FiniteState, Site
    lat = state.lattice
    sites = [[None] * lat.Ly for _ in range(lat.Lx)]
    for (x, y) in lat.site_order():
        s = state[(x, y)]
        a, par = s.data, list(s.parity)
        for v, k in enumerate((0, 1, 4, 5)):
            p = par[k]
            e, o = int(live[x + y * lat.Lx, v, 0]), int(live[x + y * lat.Lx, v, 1])
            keep = np.concatenate([np.nonzero(p == 0)[0][:e], np.nonzero(p == 1)[0][:o]])
            keep.sort(kind="stable")
            a = np.take(a, keep, axis=k)
            par[k] = p[keep]
        sites[x][y] = Site(np.ascontiguousarray(a), par)
    return FiniteState(lat, sites, state.kind, beta=state.beta, log_amplitude=state.log_amplitude)


def load_arrays(path=PEPO_FILE):
    """(Lx, Ly, {(x, y): (data, parities)}, beta, log_amplitude) of a state file."""
    f = np.load(path, allow_pickle=False)
    Lx, Ly = int(f["Lx"]), int(f["Ly"])
    sites = {}
    for i in range(Lx * Ly):
        x, y = i % Lx, i // Lx
        sites[(x, y)] = (f[f"site{i}"], [f[f"par{i}_{k}"].astype(np.int64) for k in range(6)])
    return Lx, Ly, sites, float(f["beta"]), float(f["log_amplitude"])


def oracle_pepo(path=PEPO_FILE):
    from oracle.model import Lattice, State
    from oracle.z2 import Z2
    Lx, Ly, sites, beta, la = load_arrays(path)
    return State(Lattice(Lx, Ly), [[Z2(*sites[(x, y)]) for y in range(Ly)] for x in range(Lx)], "pepo", beta, la)


def stage_pepo(out):
    import __graft_entry__ as ge
    spp = ge.load_package()
    c = spp.RunConfig(lattice=spp.Lattice(L, L), model=spp.TFIM(J, H, symmetric=True), beta0=BETA0, beta_log=[],
                      beta_measure=[BETA0], delta_tau_pur=DTAU, delta_tau_peps=DTAU, D_pur=6, D_peps=6,
                      bond_space_pur=(3, 3), bond_space_peps=(3, 3), chi_environment=CHI, trotter_order=2,
                      update=spp.FullUpdate(max_iterations=40, linear_max_iterations=100),
                      record_boundary_entropy=False, Ns=1024, seed=SEED,
                      resampling=spp.ResamplingOptions(enabled=False))
    ctx = spp.Context(0, "fp64")
    t0 = time.perf_counter()
    def progress(info):
        print(f"  full-update step to beta={info['beta']:.3f}: {info['time_seconds']:.1f} s, max env truncation "
              f"{float(np.max(info['max_environment_truncation'])):.3e}", flush=True)
    cp = spp.run_purification(c, ctx, on_step=progress)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    row = cp.purification[-1]
    print(f"purified to beta={cp.state.state.beta} in {dt:.1f} s; E/site={row['energy_per_site']:.12f}, "
          f"max state truncation {row['max_state_truncation']:.3e}", flush=True)
    live = cp.state.live()
    st = trim_live(cp.state.download(), live)
    spp.save_state_npz(out, st, gpu_energy_per_site=row["energy_per_site"], gpu_lnZ=row["lnZ"], live=live,
                       purification_seconds=dt)
    print("wrote", out, flush=True)


def _oracle_sample(i):
    """Draw of sample index i, its projected lognorm or energy (worker)."""
    from oracle import born as ob
    from oracle import boundary as obm
    from oracle import energy as oe
    from oracle.model import TFIM
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)
    pepo = oracle_pepo()
    cache = ob.SamplingCache(pepo, CHI)
    u = ob.counter_uniforms(SEED, i, L * L)
    t0 = time.perf_counter()
    r = ob.sample_configuration(cache, uniforms=u)
    cfg = np.array([[r["configuration"][x][y] for y in range(L)] for x in range(L)])
    ps, _ = ob.project_configuration(pepo, cfg)
    out = {"configuration": cfg, "log_probability": r["log_probability"],
           "conditionals": np.array(r["conditionals"]), "max_environment_truncation": r["max_environment_truncation"],
           "log_amplitude": ps.log_amplitude, "south_max_truncation": [e.max_truncation for e in cache.south]}
    if i <= ENERGIES:
        e = oe.measure_energy(ps, TFIM(J, H, symmetric=True), CHI)
        out.update({"energy_total": e["energy_total"], "energy_per_site": e["energy_per_site"],
                    "log_norm": e["log_norm"], "tensor_log_norm": e["tensor_log_norm"]})
    else:
        n = obm.lognorm(ps, CHI)
        out.update({"energy_total": np.nan, "energy_per_site": np.nan, "log_norm": n["log_norm"],
                    "tensor_log_norm": n["tensor_log_norm"]})
    print(f"sample {i}: {time.perf_counter() - t0:.0f} s", flush=True)
    return out


def stage_oracle(jobs):
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=jobs) as ex:
        res = list(ex.map(_oracle_sample, range(1, DRAWS + 1)))
    keys = ["log_probability", "max_environment_truncation", "log_amplitude", "energy_total", "energy_per_site",
            "log_norm", "tensor_log_norm"]
    np.savez_compressed(ORACLE_FILE, sample_index=np.arange(1, DRAWS + 1), seed=SEED, chi=CHI,
                        configuration=np.stack([r["configuration"] for r in res]).astype(np.int8),
                        conditionals=np.stack([r["conditionals"] for r in res]),
                        south_max_truncation=np.array(res[0]["south_max_truncation"]),
                        **{k: np.array([r[k] for r in res], dtype=np.float64) for k in keys})
    print("wrote", ORACLE_FILE, flush=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--stage", choices=["pepo", "oracle"], required=True)
    ap.add_argument("--out", default=PEPO_FILE)
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    if a.stage == "pepo":
        stage_pepo(a.out)
    else:
        stage_oracle(a.jobs)
