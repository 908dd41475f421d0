#!/usr/bin/env python3
"""Born samples/s of the batched sampler on one MI355X.

Workload (one "step"): one batched `sample_configuration` of B samples on the
beta0 sampling cache of the 8x8 open TFIM (J=1, h=2.0) PEPO purified to
beta0=0.5 (Z2 bond space Z2Space(0=>3,1=>3), D=6, physical/auxiliary d=2; the
frozen fixture tests/golden/cfg2_pepo_beta05.npz), boundary chi=32, row-major
draws with one injected uniform per site (sample indices 1, 2, ... of seed 1).
The headline run computes the contraction GEMMs in FP32 on MFMA (FP64
eigensolvers, probabilities and log weights); the plain path is FP64
everywhere at batch 1, and the batched FP64 point runs the same batch in FP64.
The cache build is timed separately and amortised over the whole ensemble.

The roofline fraction uses the algorithmic work of a draw: the Z2 block-sparse
GEMM flops of the reference's contraction sequence (stopeps.jl_amd/flops.py),
per launch of the dominant kernel, over that kernel's average launch time.

Prints one JSON line (rank 0).
"""
import argparse
import gc
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Born samples/s, 8×8 TFIM D=6 χ=32: batched MFMA vs plain FP64, "
          "ΔE/site≤1e-6")
# MI355X dense matrix TFLOP/s; the BF16 split modes: 2500 (BF16) over their
# piece products per FP32 product
PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16x6": 2500.0 / 6, "bf16x3": 2500.0 / 3}
# the contraction kernels of each precision; FP32 includes the factorised Gram
# transfer kernel of the draw's absorbs (counted with them by the library)
KERNEL = {"fp32": "z2gemm_kernel<float> + kron_rows_kernel", "fp64": "z2gemm_kernel<double>",
          "bf16x6": "z2gemm_split_kernel<..., 3> + kron_rows_kernel",
          "bf16x3": "z2gemm_split_kernel<..., 2> + kron_rows_kernel"}
PEPO_FIXTURE = os.path.join(ROOT, "tests", "golden", "cfg2_pepo_beta05.npz")
ORACLE_FIXTURE = os.path.join(ROOT, "tests", "golden", "cfg2_oracle.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_row(args):
    """One row of a Born draw by the FP64 oracle on one core (worker process):
    right row environments, both candidate left steps and the drawn left step
    per site, and the compressed north absorb (src/sampling/born.jl:105-161),
    on the fixture PEPO's double layers; boundaries are seeded random Z2
    tensors of the row's production shapes (the cost depends on shapes only)."""
    y, chi, seed = args
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)
    sys.path.insert(0, ROOT)
    from oracle import boundary as bm
    from oracle import born as ob
    from oracle.model import Lattice, State
    from oracle.z2 import Z2, symmetrize
    f = np.load(PEPO_FIXTURE, allow_pickle=False)
    L = int(f["Lx"])
    sites = [[Z2(f[f"site{x + yy * L}"], [f[f"par{x + yy * L}_{k}"].astype(np.int64) for k in range(6)])
              for yy in range(L)] for x in range(L)]
    pepo = State(Lattice(L, L), sites, "pepo")
    basis = ob.resolve_basis(True)
    cands = [[bm.double_layer(ob.project(pepo[(x, y)], basis, k)) for k in range(2)] for x in range(L)]
    rowl = [c[0] for c in cands]
    rng = np.random.default_rng(seed + y)

    def boundary(axis, trivial):
        w = [E.p[0] for E in rowl]
        e = [E.p[2] for E in rowl]
        out = []
        for x in range(L):
            def bond(b):
                if trivial or b in (0, L):
                    return np.array([0])
                c = min(chi, int(np.prod([len(p) for p in e[:b]])), int(np.prod([len(p) for p in w[b:]])))
                return np.array([0] * ((c + 1) // 2) + [1] * (c // 2))
            lp, rp = bond(x), bond(x + 1)
            t = Z2(rng.standard_normal((len(lp), len(rowl[x].p[axis]), len(rp))), [lp, rowl[x].p[axis], rp])
            out.append(symmetrize(t))
        return out
    north = boundary(1, y == 0)
    south = boundary(3, y == L - 1)
    t0 = time.perf_counter()
    _, right = bm.row_environments(north, rowl, south, side="right")
    left = bm.row_edge(north, rowl, south, "left")
    for x in range(L):
        steps = [bm.left_step(left, north[x], E, south[x]) for E in cands[x]]
        _ = [bm.join_scalars(s, right[x + 1]) for s in steps]
        if x < L - 1:
            left = bm.safe_scale(bm.left_step(left, north[x], rowl[x], south[x]))
    bm.absorb(bm.Boundary(north), rowl, chi, "north")
    return y, time.perf_counter() - t0


def host_info():
    """Host CPU model, core count, ROCm version and BLAS library of this box."""
    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/opt/rocm/.info/version") as fh:
            rocm = fh.read().strip()
    except OSError:
        rocm = None
    blas = None
    try:
        from threadpoolctl import threadpool_info
        libs = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if libs:
            blas = f"{libs[0].get('internal_api')} {libs[0].get('version')}"
    except Exception:
        pass
    return {"cpu": model, "nproc": os.cpu_count(), "rocm": rocm, "blas": blas, "numpy": np.__version__}


def cpu_baseline(chi=32, seed=5, procs=None):
    """The FP64 oracle (CPU restatement of the reference, not the reference,
    which needs Julia) on the box's host cores: ``procs`` single-thread
    processes, each timing one row of a Born draw; rows 0..7 are spread over
    the processes and a draw's single-core time is the sum of the mean row
    times. Throughput = procs / that time (each core drawing whole samples)."""
    import multiprocessing as mp
    procs = procs or min(16, os.cpu_count() or 1)
    f = np.load(PEPO_FIXTURE, allow_pickle=False)
    L = int(f["Ly"])
    old = {k: os.environ.get(k) for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS")}
    for k in old:
        os.environ[k] = "1"
    try:
        t0 = time.perf_counter()
        with mp.get_context("spawn").Pool(procs) as pool:
            res = pool.map(_cpu_row, [(k % L, chi, seed) for k in range(max(procs, L))])
        wall = time.perf_counter() - t0
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    per_row = {y: float(np.mean([t for (r, t) in res if r == y])) for y in range(L)}
    t_draw = sum(per_row.values())
    return {"value": procs / t_draw, "unit": "Born samples/s", "cores": procs, "kind": "port",
            "sample": (f"FP64 NumPy oracle (CPU restatement, not the reference), {procs} single-thread processes, "
                       f"each timing one row of an 8x8 D=6 chi={chi} draw on the fixture PEPO (rows 0..{L - 1} "
                       f"spread over the processes; random boundaries of the production shapes); one draw = "
                       f"{t_draw:.1f} core-seconds, {wall:.1f} s wall"),
            "row_seconds": per_row, "host": host_info()}


def measured_traffic(precision, batch, L, chi, kron):
    """HBM bytes per launch of the contraction kernels for this workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent or
    for another workload)."""
    path = os.path.join(ROOT, "profiles", f"contraction_{precision}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if (t.get("precision"), t.get("batch"), t.get("L"), t.get("chi"), t.get("kron")) != (precision, batch, L, chi,
                                                                                       kron):
        return None, None
    return t["traffic_bytes_per_launch"], t.get("source")


def uneven_timing(spp, L=16, chi=32, batch=64, spread=(2, 3, 4), seed=11):
    """Uneven batch at the 16x16 lattice: batch random Z2 PEPS whose bond
    dimensions spread over D in `spread` (seeded; D split as (ceil(D/2),
    floor(D/2))), embedded in the capacity of the largest D and packed into one
    batched log-norm call (the per-sample parity splits drive the ragged
    contraction kernel, tile shapes tuned per launch class), FP32 contractions.
    Against the same states in FP64 (accuracy) and against a batch whose states
    all have the largest D (what padding to one capacity would cost)."""
    rng = np.random.default_rng(seed)
    lat = spp.Lattice(L, L)
    dmax = max(spread)
    cap = ((dmax + 1) // 2, dmax // 2)
    Ds = rng.choice(spread, size=batch)
    states, lives = [], []
    for k, D in enumerate(Ds):
        st = spp.random_peps(lat, ((D + 1) // 2, D // 2), seed=1000 + k)
        padded, live = spp.pad_bonds(st, cap)
        states.append(padded)
        lives.append(live)
    full = [spp.random_peps(lat, cap, seed=2000 + k) for k in range(batch)]
    out = {"workload": f"log-norm of {batch} random {L}x{L} Z2 PEPS, bond D spread over {list(spread)} "
                       f"(counts {[int(np.sum(Ds == d)) for d in spread]}), capacity D={dmax}, chi={chi}",
           "batch": batch}
    for prec in ("fp32", "fp64"):
        ctx = spp.Context(0, prec)
        try:
            dev = [spp.DeviceState(ctx, s, live=lv) for s, lv in zip(states, lives)]
            spp.lognorms(ctx, dev, chi)  # warm-up: tunes the launch classes
            ctx.synchronize()
            t0 = time.perf_counter()
            r = spp.lognorms(ctx, dev, chi)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            out[prec] = {"value": batch / dt, "unit": "states/s", "log_norm": r["log_norm"]}
            if prec == "fp32":
                devf = [spp.DeviceState(ctx, s) for s in full]
                spp.lognorms(ctx, devf, chi)
                ctx.synchronize()
                t0 = time.perf_counter()
                spp.lognorms(ctx, devf, chi)
                ctx.synchronize()
                out["fp32_all_at_capacity"] = {"value": batch / (time.perf_counter() - t0), "unit": "states/s"}
            log(f"uneven {prec}: {out[prec]['value']:.3f} states/s")
        finally:
            ctx.close()
    a, b = out["fp32"].pop("log_norm"), out["fp64"].pop("log_norm")
    out["fp32"]["max_rel_dlog_norm_vs_fp64"] = float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))
    out["uneven_gain_over_capacity"] = out["fp32"]["value"] / out["fp32_all_at_capacity"]["value"]
    return out


def adjoint_timing(spp, reps=2, batch=8, chi=36):
    """The log-norm gradient workload: 8x8 Z2 PEPS, D=6 (3+3), d=2, i.i.d.
    N(0,1) blocks scaled to unit norm (seed 7), chi=36, FP64 contractions: all
    64 site gradients of one state (and, as a throughput figure, of a batch of
    states with seeds 7, 8, ...)."""
    lat = spp.Lattice(8, 8)
    states = [spp.random_peps(lat, (3, 3), seed=7 + k) for k in range(batch)]
    ctx = spp.Context(0, "fp64", autotune=0)  # timed on first calls: heuristic tile shapes
    dev = [spp.DeviceState(ctx, s) for s in states]
    spp.lognorm_gradients(ctx, dev[:1], chi)  # warm-up
    ctx.synchronize()
    ctx.reset_counters()
    t0 = time.perf_counter()
    for _ in range(reps):
        g, info = spp.lognorm_gradients(ctx, dev[:1], chi)
    ctx.synchronize()
    t1 = (time.perf_counter() - t0) / reps
    cnt = ctx.counters()
    # log<psi|psi> is homogeneous of degree 2 in each site: <A_i, grad_i> = 2
    euler = max(abs(float(np.sum(states[0][xy].data * g[0][xy])) - 2.0) for xy in lat.site_order())
    t0 = time.perf_counter()
    spp.lognorm_gradients(ctx, dev, chi)
    ctx.synchronize()
    tb = time.perf_counter() - t0
    return {"value": 1.0 / t1, "unit": "gradients/s (64 site gradients of one state)", "dtype": "fp64",
            "workload": f"gradient of log<psi|psi>, 8x8 Z2 PEPS D=6 (3+3), d=2, seed 7, chi={chi}",
            "s_per_gradient": t1, "log_norm": float(info["log_norm"][0]),
            "max_environment_truncation": float(info["max_environment_truncation"][0]),
            "max_euler_residual": euler,
            "eigensolver_iterations_per_solve": cnt["eig_iterations"] / max(cnt["eig_calls"], 1),
            "batch": batch, "batch_s_per_state": tb / batch}


def entropy_timing(spp, n=16, chi=32, J=1.0, h=2.0):
    """The boundary-entropy observer's cost: FP64 energy measurement of n
    projected samples of the synthetic 8x8 D=6 PEPO (eps 1.0, seed 1) at
    chi=32, without and with record_entropy (every south and north cut's
    pre-truncation spectrum), each timed once after a warm-up."""
    lat = spp.Lattice(8, 8)
    pepo = spp.synthetic_thermal_pepo(lat, (3, 3), 1.0, seed=1)
    ctx = spp.Context(0, "fp64", autotune=0)  # timed on first calls: heuristic tile shapes
    cache = spp.SamplingCache(ctx, pepo, chi)
    model = spp.TFIM(J, h, symmetric=True)
    labels = np.asarray(cache.basis.labels)
    cfg = labels[np.random.default_rng(0).integers(0, 2, size=(n, 8, 8))]
    spp.measure_energies(cache, cfg[:2], model)
    out = {}
    for rec in (False, True):
        ctx.synchronize()
        t0 = time.perf_counter()
        r = spp.measure_energies(cache, cfg, model, record_entropy=rec)
        ctx.synchronize()
        out[rec] = (time.perf_counter() - t0, r)
    return {"workload": f"FP64 energy of {n} projected samples, 8x8 Z2 PEPO D=6, chi={chi}",
            "s_without_entropy": out[False][0], "s_with_entropy": out[True][0],
            "overhead": out[True][0] / out[False][0] - 1.0,
            "max_boundary_entropy": float(np.max(out[True][1]["max_boundary_entropy"]))}


def energy_timing(spp, n=32, chi=32, J=1.0, h=2.0, krons=(1, 0)):
    """FP32 energy measurement of n projected samples of the frozen beta0 PEPO
    (random labels, seed 0) at chi, one call after a 4-sample warm-up call
    (which tunes the tile shapes), per value of the ``kron`` option: the sweeps'
    absorbs through the single-layer tables (1) or the dense double layers (0)."""
    pepo, _ = spp.load_state_npz(PEPO_FIXTURE)
    model = spp.TFIM(J, h, symmetric=True)
    out, res = {}, {}
    for kron in krons:
        ctx = spp.Context(0, "fp32", kron=kron)
        try:
            cache = spp.SamplingCache(ctx, pepo, chi)
            labels = np.asarray(cache.basis.labels)
            cfg = labels[np.random.default_rng(0).integers(0, 2, size=(n, 8, 8))]
            spp.measure_energies(cache, cfg[:4], model)
            ctx.synchronize()
            ctx.reset_counters()
            t0 = time.perf_counter()
            r = spp.measure_energies(cache, cfg, model)
            ctx.synchronize()
            t = time.perf_counter() - t0
            res[kron] = r
            out[f"kron{kron}"] = {"value": n / t, "unit": "energy measurements/s", "s": t,
                                  "kron_launches": ctx.counters()["kron_launches"]}
            log(f"energy kron={kron}: {n / t:.3f} measurements/s")
        finally:
            ctx.close()
    if 1 in res and 0 in res:
        out["max_abs_dE_per_site_kron_vs_dense"] = float(np.max(np.abs(res[1]["energy_per_site"] -
                                                                      res[0]["energy_per_site"])))
    out["workload"] = f"FP32 energy of {n} projected samples (random labels), cfg2 beta0 PEPO 8x8 D=6, chi={chi}"
    return out


def fu_timing(spp, batch=16, chi=32, J=1.0, h=2.0, delta=0.05, warmup=False):
    """The full-update evolution workload: one second-order step (sweeps H->,
    V->, V<-, H<-; 224 bond updates) of a batch of 8x8 Z2 PEPS, D=6 (3+3),
    d=2, i.i.d. N(0,1) blocks scaled to unit norm (seeds 100, 101, ...), TFIM
    gates, boundary chi=32, FullUpdate with 40 ALS and 100 PCG iterations
    (the options of the reference's example configurations), FP64. One timed
    step (a step is tens of seconds); with ``warmup`` an untimed step first
    tunes the tile shapes, otherwise the heuristic shapes are used."""
    lat = spp.Lattice(8, 8)
    ctx = spp.Context(0, "fp64", autotune=1 if warmup else 0)
    dev = [spp.DeviceState(ctx, spp.random_peps(lat, (3, 3), seed=100 + k)) for k in range(batch)]
    model = spp.TFIM(J, h, symmetric=True)
    method = spp.FullUpdate(max_iterations=40, error_tolerance=1e-8, diagonal_shift=1e-10,
                            linear_max_iterations=100, linear_tolerance=1e-9)
    if warmup:
        spp.evolution_steps(ctx, dev, model, delta, chi, method)
        ctx.synchronize()
        log(f"full-update warm-up step done (batch {batch})")
    ctx.reset_counters()
    t0 = time.perf_counter()
    r = spp.evolution_steps(ctx, dev, model, delta, chi, method)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    cnt = ctx.counters()
    rec = r["records"]
    return {"value": batch / dt, "unit": "FU sample-steps/s", "dtype": "fp64",
            "workload": f"full-update step (2nd order, delta={delta}), 8x8 Z2 PEPS D=6 (3+3), d=2, chi={chi}, "
                        f"TFIM J={J} h={h}, 40 ALS x 100 PCG iterations, {batch} states (seeds 100..)",
            "batch": batch, "s_per_step": dt, "s_per_sample_step": dt / batch,
            "bond_updates_per_step": int(rec.shape[1]), "mean_als_iterations": float(rec[:, :, 3].mean()),
            "max_final_cost": float(rec[:, :, 1].max()),
            "max_environment_truncation": float(np.max(r["max_environment_truncation"])),
            "contraction_tflops_over_step": cnt["gemm_flops"] / dt / 1e12}


def physical_timing(spp, L=8, chi=32, batch=256, precision="fp64", D=(3, 3)):
    """Physical-kind Born draws (the resampling proposals of one particle,
    src/sampling/born.jl:116-134): a seeded random 8x8 Z2 PEPS with D=6 (3+3),
    physical-kind cache at chi=32, one batch of 256 draws (the size of a
    leave-one-out scoring batch: one member's cache over K configurations) after
    a warm-up batch of the same size. Per site and sample the device forms the site metric, makes
    it PSD (two 648x648 parity blocks: Householder tridiagonalisation and QL)
    and draws."""
    peps = spp.random_peps(spp.Lattice(L, L), D, seed=7)
    uni = [np.stack([spp.sample_uniforms(11, k * batch + i, L * L) for i in range(1, batch + 1)]) for k in range(2)]
    ctx = spp.Context(0, precision)
    try:
        t0 = time.perf_counter()
        cache = spp.SamplingCache(ctx, peps, chi, kind="physical")
        ctx.synchronize()
        tc = time.perf_counter() - t0
        spp.sample_configurations(cache, uniforms=uni[0])
        ctx.synchronize()
        t0 = time.perf_counter()
        r = spp.sample_configurations(cache, uniforms=uni[1])
        ctx.synchronize()
        dt = time.perf_counter() - t0
        rep = r["repairs"]
        out = {"value": batch / dt, "unit": "physical-kind Born samples/s", "dtype": precision,
               "workload": f"physical-kind draws (PSD site metric), random {L}x{L} Z2 PEPS D={sum(D)} ({D[0]}+{D[1]}), "
                           f"chi={chi}, {batch} samples per call",
               "batch": batch, "cache_build_s": tc, "s_per_call": dt,
               "metric_blocks_per_call": 2 * batch * L * L,
               "largest_negative_spectral_fraction": float(np.nanmax(rep[:, :, 3])),
               "min_relative_eigenvalue": float(np.nanmin(rep[:, :, 0]))}
        del cache
    finally:
        ctx.close()
    return out


def precision_sweep(spp, flops, L=6, chi=64, batch=8, D=(4, 4)):
    """Born draws at the bond and boundary dimensions of the 12x12 workload
    (D=8 as Z2Space(0=>4, 1=>4), chi=64) on a reduced LxL lattice, in every
    contraction precision: FP64, FP32 on the FP32 matrix cores, and FP32 from
    six or three BF16 piece products. The same uniforms (seed 1, indices
    1..batch) in each; accuracy is stated against the FP64 draws. Input: the
    synthetic thermal PEPO (beta=0 identity plus a seeded 0.3 perturbation,
    unit-norm sites; no purified beta0=0.4 PEPO at these sizes is committed)."""
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(L, L), D=D, eps=0.3, seed=4)
    uni = np.stack([spp.sample_uniforms(1, i, L * L) for i in range(1, batch + 1)])
    scope = "the 12x12 D=8 chi=64 workload" if L == 12 else "reduced lattice of the 12x12 D=8 chi=64 workload"
    out = {"workload": f"batched Born draws, {L}x{L} synthetic Z2 PEPO, D={sum(D)} ({D[0]}+{D[1]}), d=2, chi={chi}, "
                       f"{batch} samples per step ({scope})",
           "algorithmic_flops_per_draw": flops.draw_flops(pepo, chi), "samples": batch}
    ctx = spp.Context(0, "fp64", autotune=0)  # timed on first calls: heuristic tile shapes
    ref = None
    try:
        for prec in ("fp64", "fp32", "bf16x6", "bf16x3"):
            ctx.set_precision(prec)
            t0 = time.perf_counter()
            cache = spp.SamplingCache(ctx, pepo, chi)
            ctx.synchronize()
            tc = time.perf_counter() - t0
            t0 = time.perf_counter()
            d = spp.sample_configurations(cache, uniforms=uni)
            ctx.synchronize()
            dt = time.perf_counter() - t0
            row = {"value": batch / dt, "unit": "Born samples/s", "cache_build_s": tc}
            if ref is None:
                ref = d
            else:
                same = [np.array_equal(d["configuration"][k], ref["configuration"][k]) for k in range(batch)]
                dl = [abs(d["log_probability"][k] - ref["log_probability"][k]) for k in range(batch) if same[k]]
                row.update(identical_configurations=int(sum(same)),
                           max_abs_dlog_probability_vs_fp64=float(max(dl)) if dl else None,
                           max_rel_dtensor_log_norm_vs_fp64=float(np.max(
                               np.abs(d["tensor_log_norm"] - ref["tensor_log_norm"])
                               / np.maximum(1.0, np.abs(ref["tensor_log_norm"])))))
            out[prec] = row
            log(f"sweep {prec}: {row['value']:.3f} samples/s")
            del cache, d
        for prec in ("bf16x6", "bf16x3"):
            out[prec]["gain_over_fp32"] = out[prec]["value"] / out["fp32"]["value"]
    finally:
        ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128, help="Born samples per step")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "bf16x6", "bf16x3"])
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--kron", type=int, default=1, choices=[0, 1, 2],
                    help="1: factorised Gram transfers in the FP32 draws on the vector ALUs (default); "
                         "2: the same on the matrix cores; 0: dense double layers")
    ap.add_argument("--plain-draws", type=int, default=3, help="batch-1 FP64 draws for the plain path")
    ap.add_argument("--check-samples", type=int, default=8,
                    help="drawn samples measured in the energy-throughput figure")
    ap.add_argument("--J", type=float, default=1.0)
    ap.add_argument("--h", type=float, default=2.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="timed loop only (for rocprofv3)")
    ap.add_argument("--entropy-only", action="store_true",
                    help="time only the boundary-entropy observer (energy with and without record_entropy)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the log-norm gradient workload (for rocprofv3)")
    ap.add_argument("--fu-only", action="store_true", help="time only the full-update workload")
    ap.add_argument("--energy-only", type=str, default=None, metavar="KRONS",
                    help="time only the FP32 energy measurement for these comma-separated kron values (e.g. 1,0)")
    ap.add_argument("--physical-only", action="store_true", help="time only the physical-kind draws")
    ap.add_argument("--fu-warmup", action="store_true", help="an untimed full-update step first (tunes the tiles)")
    ap.add_argument("--fu-batch", type=int, default=16, help="states per full-update step (0: skip)")
    ap.add_argument("--sweep-lattice", type=int, default=6,
                    help="lattice edge of the D=8 chi=64 precision sweep (12: the full lattice; 0: skip)")
    ap.add_argument("--sweep-batch", type=int, default=8, help="samples per precision in the sweep")
    ap.add_argument("--sweep-only", action="store_true", help="run only the precision sweep")
    ap.add_argument("--uneven-only", action="store_true",
                    help="time only the uneven-bond batch at the 16x16 lattice (log-norms)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed loop, copy 256 MiB device to device (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the sampler shards samples across GPUs by design, "
              "but multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    spp = ge.load_package()
    from stopeps_jl_amd import flops
    if args.entropy_only:
        print(json.dumps({"boundary_entropy": entropy_timing(spp)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_timing(spp)}), flush=True)
        return
    if args.energy_only:
        krons = tuple(int(v) for v in args.energy_only.split(","))
        print(json.dumps({"energy_measurement": energy_timing(spp, krons=krons)}), flush=True)
        return
    if args.fu_only:
        print(json.dumps({"full_update": fu_timing(spp, batch=max(args.fu_batch, 1), warmup=args.fu_warmup)}),
              flush=True)
        return
    if args.uneven_only:
        print(json.dumps({"uneven_batch": uneven_timing(spp)}), flush=True)
        return
    if args.sweep_only:
        print(json.dumps({"precision_sweep": precision_sweep(spp, flops, L=args.sweep_lattice,
                                                             batch=args.sweep_batch)}), flush=True)
        return
    if args.physical_only:
        print(json.dumps({"physical_draws": {p: physical_timing(spp, precision=p) for p in ("fp64", "fp32")}}),
              flush=True)
        return
    pepo, _ = spp.load_state_npz(PEPO_FIXTURE)
    L = pepo.lattice.Lx
    ns = L * L
    B = args.batch
    w_draw = flops.draw_flops(pepo, args.chi)  # algorithmic flops of one draw

    ctx = spp.Context(0, args.precision, kron=args.kron)
    t0 = time.perf_counter()
    cache = spp.SamplingCache(ctx, pepo, args.chi)
    ctx.synchronize()
    t_cache = time.perf_counter() - t0
    log(f"cache build {t_cache:.2f} s")

    def uniforms(first, n):
        return np.stack([spp.sample_uniforms(1, i, ns) for i in range(first, first + n)])

    index = 1
    for _ in range(args.warmup):
        spp.sample_configurations(cache, uniforms=uniforms(index, B))
        index += B
    ctx.synchronize()
    tuned = ctx.counters()  # the warm-up tunes the contraction tile shapes of every launch class
    ctx.reset_counters()
    ctx.set_option("profile", 1)
    draws = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        draws.append(spp.sample_configurations(cache, uniforms=uniforms(index, B)))
        index += B
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    cnt = ctx.counters()
    ctx.set_option("profile", 0)
    ndraws = B * args.steps
    value = ndraws / elapsed
    log(f"timed {args.steps} x {B} draws in {elapsed:.2f} s -> {value:.3f} samples/s")
    traffic, traffic_src = measured_traffic(args.precision, B, L, args.chi, args.kron)
    kc = cnt["fp64" if args.precision == "fp64" else "fp32"]  # launches of this precision's contraction kernel
    avg_ms = kc["ms"] / max(kc["launches"], 1)
    draws_per_launch = ndraws / max(kc["launches"], 1)
    achieved = w_draw * draws_per_launch / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    executed_tf = kc["flops"] / (kc["ms"] * 1e-3) / 1e12 if kc["ms"] > 0 else 0.0
    peak = PEAK[args.precision]  # split modes: the BF16 peak over the piece products per FP32 product
    result = {
        "metric": METRIC, "value": value, "unit": "Born samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"batched Born draws, {L}x{L} TFIM J=1 h=2.0 Z2 PEPO at beta0=0.5, D=6 (3+3), d=2, "
                               f"chi={args.chi}, {B} samples per step", "lattice": [L, L], "D": 6, "chi": args.chi,
                   "samples_per_step": B, "cache_build_s": t_cache,
                   "gram_transfers": {0: "dense double layers", 1: "factorised (ket and bra single layers)",
                                      2: "factorised, matrix-core kernel"}[args.kron],
                   "input": "tests/golden/cfg2_pepo_beta05.npz: the beta0=0.5 PEPO of full-update purification "
                            "(5 steps of dtau=0.05, FP64), frozen; uniforms of sample indices 1.. of seed 1"},
        "roofline": {"bound": "mfma", "kernel": KERNEL[args.precision],
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                     "traffic_source": traffic_src,
                     "algorithmic_flops_per_draw": w_draw, "draws_per_launch": draws_per_launch,
                     "avg_launch_ms": avg_ms, "launches": kc["launches"],
                     "step_frac": w_draw * value / (peak * 1e12),
                     "executed_tflops": executed_tf, "executed_frac": executed_tf / peak,
                     "executed_flops_per_draw": cnt["gemm_flops"] / ndraws,
                     "kernel_share_of_step": kc["ms"] / (1e3 * elapsed),
                     "all_contractions_share_of_step": cnt["gemm_ms"] / (1e3 * elapsed)},
    }
    result["autotune"] = {"tuned_classes": tuned["tuned_classes"], "classes_off_heuristic": tuned["tuned_changed"],
                          "tuning_s": tuned["tune_ms"] / 1e3,
                          "note": "tile shapes tuned per launch class during the warm-up; bit-identical results"}
    if args.profile_only:
        if args.copy_calibration:
            result["copy_calibration_bytes"] = ctx.copy_calibration(64 << 20)
        print(json.dumps(result), flush=True)
        return

    # the headline batch with the products on the BF16 matrix cores (the
    # context's workspace is already sized, so one timed step each)
    # the same batch launched kernel by kernel (no captured row sweeps), then
    # also with the heuristic tile shapes instead of the tuned ones
    ctx.set_option("graphs", 0)
    t0 = time.perf_counter()
    spp.sample_configurations(cache, uniforms=uniforms(index, B))
    ctx.synchronize()
    eager = B / (time.perf_counter() - t0)
    ctx.set_option("autotune", 0)
    t0 = time.perf_counter()
    spp.sample_configurations(cache, uniforms=uniforms(index, B))
    ctx.synchronize()
    result["autotune"]["heuristic_value"] = B / (time.perf_counter() - t0)
    result["autotune"]["gain"] = eager / result["autotune"]["heuristic_value"]
    result["graphs"] = {"eager_value": eager, "gain": value / eager,
                        "note": "row sweeps of the draw captured as HIP graphs and replayed (value) vs launched "
                                "kernel by kernel (eager_value), tuned tiles in both"}
    ctx.set_option("autotune", 1)
    ctx.set_option("graphs", 1)
    log(f"eager launches: {eager:.3f} samples/s; heuristic tiles: {result['autotune']['heuristic_value']:.3f}")
    split_caches = {}
    if args.precision == "fp32":
        split = {}
        for prec in ("bf16x6", "bf16x3"):
            ctx.set_precision(prec)
            scache = spp.SamplingCache(ctx, pepo, args.chi)
            spp.sample_configurations(scache, uniforms=uniforms(10 ** 6, B))  # warm-up (tunes this mode's tiles)
            ctx.synchronize()
            t0 = time.perf_counter()
            spp.sample_configurations(scache, uniforms=uniforms(1, B))
            ctx.synchronize()
            v = B / (time.perf_counter() - t0)
            split[prec] = {"value": v, "unit": "Born samples/s", "gain_over_fp32": v / value,
                           "accuracy": "precision_check." + prec}
            log(f"{prec}: {v:.3f} samples/s")
            split_caches[prec] = scache
        ctx.set_precision(args.precision)
        result["split_modes"] = split
    # batched FP64: the same batch, FP64 everywhere
    pctx = spp.Context(0, "fp64")
    pcache = spp.SamplingCache(pctx, pepo, args.chi)
    spp.sample_configurations(pcache, uniforms=uniforms(10 ** 6, B))  # warm-up (tunes the FP64 tiles)
    spp.sample_configurations(pcache, uniforms=uniforms(10 ** 6, 1))  # and those of the batch-1 path
    spp.sample_configurations(pcache, uniforms=uniforms(10 ** 6 + 1, 1))  # which captures its row sweeps
    pctx.synchronize()
    t0 = time.perf_counter()
    spp.sample_configurations(pcache, uniforms=uniforms(index, B))
    pctx.synchronize()
    batched64 = B / (time.perf_counter() - t0)
    log(f"batched FP64: {batched64:.3f} samples/s")
    # plain FP64 path: the same kernels, FP64, one sample at a time
    t0 = time.perf_counter()
    for i in range(args.plain_draws):
        spp.sample_configurations(pcache, uniforms=uniforms(index + i, 1))
    pctx.synchronize()
    plain = args.plain_draws / (time.perf_counter() - t0)
    log(f"plain FP64: {plain:.3f} samples/s")
    pctx.set_option("graphs", 0)
    t0 = time.perf_counter()
    for i in range(args.plain_draws):
        spp.sample_configurations(pcache, uniforms=uniforms(index + i, 1))
    pctx.synchronize()
    plain_eager = args.plain_draws / (time.perf_counter() - t0)
    pctx.set_option("graphs", 1)
    log(f"plain FP64 without graphs: {plain_eager:.3f} samples/s")
    result["plain_fp64"] = {"value": plain, "unit": "Born samples/s", "batch": 1,
                            "speedup_of_value": value / plain, "value_without_graphs": plain_eager}
    result["batched_fp64"] = {"value": batched64, "unit": "Born samples/s", "batch": B,
                              "batching_gain": batched64 / plain, "precision_gain": value / batched64}
    model = spp.TFIM(args.J, args.h, symmetric=True)
    result["precision_check"] = oracle_check(spp, [(args.precision, cache), *split_caches.items(), ("fp64", pcache)],
                                             model, args)
    # energy throughput of the headline's drawn samples
    k = min(args.check_samples, B)
    cfg = draws[0]["configuration"][:k]
    ctx.set_option("autotune", 0)  # timed on its first call: heuristic tile shapes
    ctx.synchronize()
    t0 = time.perf_counter()
    spp.measure_energies(cache, cfg, model)
    ctx.synchronize()
    t_energy = time.perf_counter() - t0
    ctx.set_option("autotune", 1)
    result["energy_measurement"] = {"value": k / t_energy, "unit": "energy measurements/s",
                                    "dtype": args.precision, "samples": int(k), "chi": args.chi}
    # release the draw workspaces (the batched FP64 one holds ~0.1 TB) so the
    # later workloads size their chunks from the whole device
    del cache, pcache, draws, split_caches
    ctx.close()
    pctx.close()
    gc.collect()
    if args.sweep_lattice > 0:
        log("precision sweep ...")
        result["precision_sweep"] = precision_sweep(spp, flops, L=args.sweep_lattice, batch=args.sweep_batch)
    result["adjoint"] = adjoint_timing(spp)
    log("physical-kind draws ...")
    result["physical_draws"] = physical_timing(spp)
    p32 = physical_timing(spp, precision="fp32")
    result["physical_draws"]["fp32"] = {k: p32[k] for k in ("value", "unit", "dtype", "s_per_call",
                                                           "largest_negative_spectral_fraction")}
    log(f"physical-kind draws: {result['physical_draws']['value']:.3f} samples/s (fp32: {p32['value']:.3f})")
    if args.fu_batch > 0:
        log("full-update step ...")
        result["full_update"] = fu_timing(spp, batch=args.fu_batch, J=args.J, h=args.h, warmup=True)
    if not args.no_cpu:
        log("CPU baseline ...")
        try:
            result["cpu_baseline"] = cpu_baseline(chi=args.chi)
        except Exception as exc:  # the baseline must not hide the GPU result
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result), flush=True)


def oracle_check(spp, caches, model, args):
    """Each precision's caches (name, cache) against the FP64 oracle's committed
    vectors on the same PEPO (tests/golden/cfg2_oracle.npz): draws of the
    fixture's sample indices (configurations, log p), and the energy per site
    of the oracle's measured samples, per sample and as the weighted ensemble
    estimate exp(log_norm - log q) (src/driver/particles.jl:43-52,
    ensemble.jl:17)."""
    try:
        ref = np.load(ORACLE_FIXTURE, allow_pickle=False)
    except OSError:
        return {"error": "oracle vectors not found"}
    ns = caches[0][1].lattice.Lx * caches[0][1].lattice.Ly
    idx = [int(i) for i in ref["sample_index"]]
    uni = np.stack([spp.sample_uniforms(int(ref["seed"]), i, ns) for i in idx])
    out = {"reference": "FP64 oracle, tests/golden/cfg2_oracle.npz", "samples": len(idx)}
    for name, c in caches:
        d = spp.sample_configurations(c, uniforms=uni)
        same = [np.array_equal(d["configuration"][k], ref["configuration"][k]) for k in range(len(idx))]
        dl = [abs(d["log_probability"][k] - ref["log_probability"][k]) for k in range(len(idx)) if same[k]]
        # the discarded weight of the truncations: FP32 products put noise of
        # about 1e-7 relative on it, so it is compared with the oracle's value
        dmt = np.abs(d["max_environment_truncation"] - ref["max_environment_truncation"])
        out[name] = {"identical_configurations": int(sum(same)),
                     "max_abs_dlog_probability": float(max(dl)) if dl else None,
                     "max_environment_truncation": float(np.max(d["max_environment_truncation"])),
                     "max_abs_dmax_environment_truncation": float(np.max(dmt))}
    m = np.isfinite(ref["energy_per_site"])
    cfg = ref["configuration"][m].astype(np.int64)
    for name, c in caches:
        e = spp.measure_energies(c, cfg, model)
        lw = e["log_norm"] - ref["log_probability"][m]
        w = np.exp(lw - np.max(lw))
        E = float(np.sum(w * e["energy_per_site"]) / np.sum(w))
        lwr = ref["log_norm"][m] - ref["log_probability"][m]
        wr = np.exp(lwr - np.max(lwr))
        Er = float(np.sum(wr * ref["energy_per_site"][m]) / np.sum(wr))
        out[name]["max_abs_dE_per_site"] = float(np.max(np.abs(e["energy_per_site"] - ref["energy_per_site"][m])))
        out[name]["ensemble_dE_per_site"] = abs(E - Er)
        out[name]["meets_dE_per_site_1e-6"] = bool(abs(E - Er) <= 1e-6)
        out[name]["ensemble_E_per_site"] = E
    out["max_environment_truncation"] = float(np.max(ref["max_environment_truncation"]))
    out["model"] = f"TFIM J={args.J} h={args.h} (Z2)"
    return out

if __name__ == "__main__":
    main()
