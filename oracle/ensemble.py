"""Weights, beta grids and ensemble statistics (oracle restatement of
src/statistics/ensemble.jl, src/driver/config.jl:63-76, src/driver/particles.jl:42-64,86-102)."""
import numpy as np

EPS = np.finfo(np.float64).eps


def same_beta(a, b):
    """``_same_beta`` (config.jl:63)."""
    return abs(a - b) <= 64 * EPS * max(1.0, abs(a), abs(b))


def temperature_nodes(beta0, beta_log, beta_measure):
    """``temperature_nodes`` (config.jl:68-76)."""
    nodes = sorted([0.0, beta0] + list(beta_log) + list(beta_measure))
    out = []
    for b in nodes:
        b = beta0 if same_beta(b, beta0) else b
        if out and same_beta(out[-1], b):
            continue
        out.append(b)
    return out


def step_schedule(start, target, delta_tau):
    """``_step_schedule`` (particles.jl:55-64)."""
    steps = []
    b = float(start)
    while not same_beta(b, target):
        if not target > b:
            raise ValueError("evolution requires increasing beta endpoints")
        b = min(target, b + 2 * delta_tau)
        if same_beta(b, target):
            b = float(target)
        steps.append(b)
    return steps


def ensemble_statistics(relative, multiplicity, energy, Ns, log_scale=0.0, nsites=1):
    """``ensemble_statistics`` (ensemble.jl:8-25)."""
    relative = np.asarray(relative, dtype=np.float64)
    multiplicity = np.asarray(multiplicity, dtype=np.float64)
    energy = np.asarray(energy, dtype=np.float64)
    if not (len(relative) == len(multiplicity) == len(energy) > 0):
        raise ValueError("empty or inconsistent ensemble")
    if not (np.all(np.isfinite(relative)) and np.all(multiplicity > 0)):
        raise ValueError("invalid particle weights/counts")
    Nv = multiplicity.sum()
    pivot = relative.max()
    w = np.exp(relative - pivot)
    mass = multiplicity * w
    total = mass.sum()
    mu1 = total / Nv
    mu2 = np.sum(multiplicity * w ** 2) / Nv
    eta = min(max(mu1 ** 2 / mu2, 1 / Nv), 1.0)

    def sem(f):
        return np.nan if Nv <= 1 else np.sqrt(max(0.0, np.sum(multiplicity * np.abs(f) ** 2) / (Nv * (Nv - 1))))

    measured = bool(np.all(np.isfinite(energy)))
    E = float(np.sum(mass * energy) / total) if measured else np.nan
    Eerror = sem(w * (energy - E) / mu1) if measured else np.nan
    lnZ_error = sem((w - mu1) / mu1)
    eta_error = sem((2 * mu1 / mu2) * (w - mu1) - (mu1 ** 2 / mu2 ** 2) * (w ** 2 - mu2))
    eta_p = min(max(total ** 2 / (Ns * np.sum(mass ** 2)), 1 / Ns), len(relative) / Ns)
    return {"energy_total": nsites * E, "energy_per_site": E, "energy_per_site_error": Eerror,
            "lnZ": float(log_scale + pivot + np.log(mu1)), "lnZ_error": lnZ_error, "eta": eta,
            "eta_error": eta_error, "ESS": Nv * eta, "eta_p": eta_p, "target_samples": Ns,
            "virtual_samples": Nv, "representatives": len(relative)}


def new_particle_weight(log_norm, log_q):
    """``_new_particle`` weight (particles.jl:42-52): relative_log_weight = log_norm - log_q."""
    return log_norm - log_q


def advance_weight(relative, normalizations, new_tensor_log_norm, old_tensor_log_norm):
    """``_advance_particle!`` weight update (particles.jl:94):
    relative += 2*sum(log_normalization) + (tensor_log_norm_new - old)."""
    return relative + 2 * float(np.sum(normalizations)) + (new_tensor_log_norm - old_tensor_log_norm)


def group_configurations(configurations):
    """``_group_configurations`` (selection.jl:122-140): first draw, multiplicity, map."""
    first, mult, mapping, lookup = [], [], [], {}
    for d, cfg in enumerate(configurations):
        key = tuple(np.asarray(cfg).T.ravel().tolist())
        j = lookup.get(key)
        if j is None:
            first.append(d)
            mult.append(0)
            j = len(first) - 1
            lookup[key] = j
        mult[j] += 1
        mapping.append(j)
    return {"first_draw": first, "multiplicity": mult, "draw_to_group": mapping}
