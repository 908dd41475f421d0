#!/usr/bin/env python3
"""Born samples/s of the batched sampler on one MI355X.

Workload (one "step"): one batched `sample_configuration` of B samples on the
beta0 sampling cache of an 8x8 open-lattice Z2 PEPO with bond space
Z2Space(0=>3,1=>3) (D=6), physical/auxiliary d=2, boundary chi=32, row-major
draws with one injected uniform per site. The headline run computes the
contraction GEMMs in FP32 on MFMA (FP64 eigensolvers, probabilities and
log weights); the plain path is FP64 everywhere at batch 1. The cache build
is timed separately and amortised over the whole ensemble.

Prints one JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Born samples/s, 8×8 TFIM D=6 χ=32: batched MFMA vs plain FP64, "
          "ΔE/site≤1e-6")
PEAK = {"fp32": 157.3, "fp64": 78.6}  # MI355X dense matrix TFLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(L=8, D=(3, 3), chi=32, seed=5, threads=None):
    """Time the FP64 oracle on one row of a Born draw at the same shapes and
    extrapolate to a draw (8 rows). One row = right row environments, the two
    candidate left steps and the drawn left step per site, and the compressed
    north absorb (src/sampling/born.jl:105-161). Boundaries are seeded random
    tensors of the production shapes (cost depends on shapes only)."""
    from threadpoolctl import threadpool_limits
    from oracle.model import Lattice, random_state
    from oracle.z2 import Z2, symmetrize
    from oracle import boundary as bm
    threads = threads or min(16, os.cpu_count() or 1)
    rng = np.random.default_rng(seed)
    lat = Lattice(L, L)
    pepo = random_state(lat, D, "pepo", seed)
    y = 2
    rowl = [bm.double_layer(pepo[(x, y)]) for x in range(L)]
    cands = [rowl[x] for x in range(L)]

    def bnd(phys_p):
        half = np.array([0] * (chi // 2) + [1] * (chi - chi // 2))
        out = []
        for x in range(L):
            lp = np.array([0]) if x == 0 else half
            rp = np.array([0]) if x == L - 1 else half
            t = Z2(rng.standard_normal((len(lp), len(phys_p[x]), len(rp))), [lp, phys_p[x], rp])
            out.append(symmetrize(t))
        return out
    north = bnd([E.p[1] for E in rowl])
    south = bnd([E.p[3] for E in rowl])
    with threadpool_limits(threads):
        t0 = time.perf_counter()
        _, right = bm.row_environments(north, rowl, south, side="right")
        left = bm.row_edge(north, rowl, south, "left")
        for x in range(L):
            steps = [bm.left_step(left, north[x], E, south[x]) for E in (cands[x], cands[x])]
            _ = [bm.join_scalars(s, right[x + 1]) for s in steps]
            left = bm.safe_scale(bm.left_step(left, north[x], cands[x], south[x]))
        t_env = time.perf_counter() - t0
        t1 = time.perf_counter()
        bm.absorb(bm.Boundary(north), rowl, chi, "north")
        t_abs = time.perf_counter() - t1
    t_row = t_env + t_abs
    return {"value": 1.0 / (L * t_row), "unit": "Born samples/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 NumPy oracle (CPU restatement, not the reference), one row of an {L}x{L} D=6 chi={chi} "
                       f"draw timed ({t_env:.1f} s environments + {t_abs:.1f} s compressed absorb), x{L} rows "
                       f"per draw; OpenBLAS {threads} threads")}


def measured_traffic(precision, batch, L, chi):
    """HBM bytes per launch of the dominant kernel for this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (None if absent or for
    another workload)."""
    path = os.path.join(ROOT, "profiles", f"z2gemm_{precision}_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None, None
    if (t.get("precision"), t.get("batch"), t.get("L"), t.get("chi")) != (precision, batch, L, chi):
        return None, None
    return t["traffic_bytes_per_launch"], t.get("source")


def adjoint_timing(spp, reps=2, batch=8, chi=36):
    """The log-norm gradient workload: 8x8 Z2 PEPS, D=6 (3+3), d=2, i.i.d.
    N(0,1) blocks scaled to unit norm (seed 7), chi=36, FP64 contractions: all
    64 site gradients of one state (and, as a throughput figure, of a batch of
    states with seeds 7, 8, ...)."""
    lat = spp.Lattice(8, 8)
    states = [spp.random_peps(lat, (3, 3), seed=7 + k) for k in range(batch)]
    ctx = spp.Context(0, "fp64")
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
    ctx = spp.Context(0, "fp64")
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


def fu_timing(spp, batch=16, chi=32, J=1.0, h=2.0, delta=0.05):
    """The full-update evolution workload: one second-order step (sweeps H->,
    V->, V<-, H<-; 224 bond updates) of a batch of 8x8 Z2 PEPS, D=6 (3+3),
    d=2, i.i.d. N(0,1) blocks scaled to unit norm (seeds 100, 101, ...), TFIM
    gates, boundary chi=32, FullUpdate with 40 ALS and 100 PCG iterations
    (the options of the reference's example configurations), FP64. One timed
    step: a step is tens of seconds and allocates only its scratch."""
    lat = spp.Lattice(8, 8)
    ctx = spp.Context(0, "fp64")
    dev = [spp.DeviceState(ctx, spp.random_peps(lat, (3, 3), seed=100 + k)) for k in range(batch)]
    model = spp.TFIM(J, h, symmetric=True)
    method = spp.FullUpdate(max_iterations=40, error_tolerance=1e-8, diagonal_shift=1e-10,
                            linear_max_iterations=100, linear_tolerance=1e-9)
    ctx.synchronize()
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=128, help="Born samples per step")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--L", type=int, default=8)
    ap.add_argument("--chi", type=int, default=32)
    ap.add_argument("--eps", type=float, default=1.0, help="perturbation size of the synthetic PEPO")
    ap.add_argument("--plain-draws", type=int, default=3, help="batch-1 FP64 draws for the plain path")
    ap.add_argument("--check-samples", type=int, default=8,
                    help="drawn samples re-evaluated in FP64 (log p) and measured in both precisions (energy)")
    ap.add_argument("--J", type=float, default=1.0)
    ap.add_argument("--h", type=float, default=2.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="timed loop only (for rocprofv3)")
    ap.add_argument("--entropy-only", action="store_true",
                    help="time only the boundary-entropy observer (energy with and without record_entropy)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="time only the log-norm gradient workload (for rocprofv3)")
    ap.add_argument("--fu-only", action="store_true", help="time only the full-update workload")
    ap.add_argument("--fu-batch", type=int, default=16, help="states per full-update step (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the timed loop, copy 256 MiB device to device (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the sampler shards samples across GPUs by design, "
              "but multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import __graft_entry__ as ge
    spp = ge.load_package()
    if args.entropy_only:
        print(json.dumps({"boundary_entropy": entropy_timing(spp)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_timing(spp)}), flush=True)
        return
    if args.fu_only:
        print(json.dumps({"full_update": fu_timing(spp, batch=max(args.fu_batch, 1))}), flush=True)
        return
    L = args.L
    lat = spp.Lattice(L, L)
    pepo = spp.synthetic_thermal_pepo(lat, (3, 3), args.eps, seed=1)
    ns = L * L
    B = args.batch

    ctx = spp.Context(0, args.precision)
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
    value = B * args.steps / elapsed
    log(f"timed {args.steps} x {B} draws in {elapsed:.2f} s -> {value:.3f} samples/s")
    traffic, traffic_src = measured_traffic(args.precision, B, L, args.chi)
    kc = cnt[args.precision]  # launches of the contraction kernel of this precision
    gemm_tf = kc["flops"] / (kc["ms"] * 1e-3) / 1e12 if kc["ms"] > 0 else 0.0
    avg_ms = kc["ms"] / max(kc["launches"], 1)
    result = {
        "metric": METRIC, "value": value, "unit": "Born samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
        "config": {"workload": f"batched Born draws, {L}x{L} Z2 PEPO D=6 (3+3), d=2, chi={args.chi}, "
                               f"{B} samples per step", "lattice": [L, L], "D": 6, "chi": args.chi,
                   "samples_per_step": B, "cache_build_s": t_cache,
                   "input": f"synthetic PEPO: identity embedded in D=6 bonds + Gaussian perturbation eps={args.eps}, seed 1"},
        "roofline": {"bound": "mfma", "kernel": f"z2gemm_kernel<{'float' if args.precision == 'fp32' else 'double'}>",
                     "achieved": gemm_tf, "peak": PEAK[args.precision], "unit": "TFLOP/s",
                     "frac": gemm_tf / PEAK[args.precision], "traffic": traffic,
                     "traffic_unit": "bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                     "traffic_source": traffic_src,
                     "avg_launch_ms": avg_ms, "launches": kc["launches"],
                     "flops_per_launch": kc["flops"] / max(kc["launches"], 1),
                     "kernel_share_of_step": kc["ms"] / (1e3 * elapsed),
                     "all_contractions_share_of_step": cnt["gemm_ms"] / (1e3 * elapsed),
                     "flops_per_draw": cnt["gemm_flops"] / (B * args.steps)},
    }
    if args.profile_only:
        if args.copy_calibration:
            result["copy_calibration_bytes"] = ctx.copy_calibration(64 << 20)
        print(json.dumps(result), flush=True)
        return

    # plain FP64 path: the same kernels, FP64, one sample at a time
    pctx = spp.Context(0, "fp64")
    pcache = spp.SamplingCache(pctx, pepo, args.chi)
    spp.sample_configurations(pcache, uniforms=uniforms(10 ** 6, 1))  # warm-up
    pctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.plain_draws):
        spp.sample_configurations(pcache, uniforms=uniforms(index + i, 1))
    pctx.synchronize()
    plain = args.plain_draws / (time.perf_counter() - t0)
    result["plain_fp64"] = {"value": plain, "unit": "Born samples/s", "batch": 1,
                            "speedup_of_value": value / plain}
    # precision check: FP64 evaluation of FP32-drawn configurations
    k = min(args.check_samples, B)
    d0 = draws[0]
    ev = spp.sample_configurations(pcache, configurations=d0["configuration"][:k])
    dlogp = np.max(np.abs(ev["log_probability"] - d0["log_probability"][:k]))
    dtln = np.max(np.abs(ev["tensor_log_norm"] - d0["tensor_log_norm"][:k]))
    # energy of the same samples (TFIM J, h at the sampling chi), batched
    # FP32-MFMA contractions vs the FP64 path; the ensemble estimate weighs
    # each sample by exp(log_norm - log q) (particles.jl:43-52, ensemble.jl:17)
    model = spp.TFIM(args.J, args.h, symmetric=True)
    cfg = d0["configuration"][:k]
    ctx.synchronize()
    t0 = time.perf_counter()
    e_lo = spp.measure_energies(cache, cfg, model)
    ctx.synchronize()
    t_energy = time.perf_counter() - t0
    e_hi = spp.measure_energies(pcache, cfg, model)

    def ensemble(e, lw):
        w = np.exp(lw - np.max(lw))
        return float(np.sum(w * e) / np.sum(w))
    E_lo = ensemble(e_lo["energy_per_site"], e_lo["log_norm"] - d0["log_probability"][:k])
    E_hi = ensemble(e_hi["energy_per_site"], e_hi["log_norm"] - ev["log_probability"])
    result["precision_check"] = {
        "samples": int(k), "max_abs_dlog_probability": float(dlogp), "max_abs_dtensor_log_norm": float(dtln),
        "max_environment_truncation": float(np.max(d0["max_environment_truncation"])),
        "max_abs_dE_per_site": float(np.max(np.abs(e_lo["energy_per_site"] - e_hi["energy_per_site"]))),
        "ensemble_E_per_site": E_hi, "ensemble_dE_per_site": abs(E_lo - E_hi),
        "model": f"TFIM J={args.J} h={args.h} (Z2)"}
    result["energy_measurement"] = {"value": k / t_energy, "unit": "energy measurements/s",
                                    "dtype": args.precision, "samples": int(k), "chi": args.chi}
    result["adjoint"] = adjoint_timing(spp)
    if args.fu_batch > 0:
        log("full-update step ...")
        result["full_update"] = fu_timing(spp, batch=args.fu_batch, J=args.J, h=args.h)
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(L=L, chi=args.chi)
        except Exception as exc:  # the baseline must not hide the GPU result
            result["cpu_baseline"] = {"value": None, "error": repr(exc)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
