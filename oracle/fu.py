"""Full-update (FU) imaginary-time evolution (oracle restatement of
src/evolution/full_update.jl; test infrastructure).

Dense Z2 tensors (``oracle.z2``); real FP64 arithmetic, so conjugation is the
identity and every ``_out`` / ``dual`` of the reference is a no-op. Leg order of
a site is (W, N, P, Aux, E, S). The PSD repair of the bond metric
(``_positive_reduced_metric``, full_update.jl:177-340) is restated on the
complete bond hole as a dense matrix per Z2 block; ``update_bond`` runs it
where the reference does (full_update.jl:491-501).
"""
import numpy as np

from .z2 import Z2, EPS, ein, permute, left_orth, right_orth, tsvd, eigh_blocks, as_matrix, fuse
from .boundary import (act, double_layer, double_layers, boundary_environments, absorb, row_environments,
                       left_step, right_step, safe_scale, Slab, row)
from .boundary import record_entropy as _record_entropy
from .model import normalize_state

RESIDUAL_ROUNDOFF_TOLERANCE = 1e-8  # full_update.jl:18


class FullUpdate:
    """``FullUpdate`` options (full_update.jl:11-17)."""

    def __init__(self, max_iterations=4, error_tolerance=1e-8, diagonal_shift=1e-10,
                 linear_max_iterations=24, linear_tolerance=1e-9):
        self.max_iterations = int(max_iterations)
        self.error_tolerance = float(error_tolerance)
        self.diagonal_shift = float(diagonal_shift)
        self.linear_max_iterations = int(linear_max_iterations)
        self.linear_tolerance = float(linear_tolerance)


class RepairNeeded(RuntimeError):
    """The reference would run ``_positive_reduced_metric`` here."""


# ---------------------------------------------------------------------------
# reduced pair (full_update.jl:21-29) and its restoration (:128-133)
# ---------------------------------------------------------------------------

PA = {"horizontal": (0, 1, 5, 3, 2, 4), "vertical": (0, 1, 4, 3, 2, 5)}
PB = {"horizontal": (1, 4, 5, 3, 2, 0), "vertical": (0, 4, 5, 3, 2, 1)}


def reduced_pair(A, B, gate, direction):
    """QR of each site into Q (four exterior legs | r) and a core (r, P, bond);
    target[r,p,q,s] = a[r,u,k] b[k,v,s] gate[p,q,u,v] (full_update.jl:21-29)."""
    pa, pb = PA[direction], PB[direction]
    Qa, Ca = left_orth(permute(A, pa), 4)          # Ca[r, P, bond]
    Qb, Cb = left_orth(permute(B, pb), 4)          # Cb[s, P, bond]
    b = permute(Cb, (2, 1, 0))                     # b[bond, P, s]
    target = ein("ruk,kvs,pquv->rpqs", Ca, b, gate)
    return {"Qa": Qa, "Qb": Qb, "target": target, "pa": pa, "pb": pb}


def restore_pair(pair, A, B):
    """``_restore_pair`` (full_update.jl:128-133)."""
    first = ein("uvwar,rpb->uvwapb", pair["Qa"], A)
    second = ein("uvwar,bpr->uvwapb", pair["Qb"], B)
    return permute(first, np.argsort(pair["pa"])), permute(second, np.argsort(pair["pb"]))


# ---------------------------------------------------------------------------
# bond environment and reduced metric (boundary_mps.jl:290-353, full_update.jl:135-172)
# ---------------------------------------------------------------------------

class BondEnvironment:
    """``BondEnvironment`` (boundary_mps.jl:290-298): left/right row messages
    (three legs; four for a vertical slab), the top and bottom boundary tensors."""

    def __init__(self, left, right, top, bottom, direction):
        self.left, self.right, self.top, self.bottom, self.direction = left, right, top, bottom, direction


def bond_environment(top, bottom, x, direction, left, right):
    """``bond_environment`` (boundary_mps.jl:300-324) with given messages."""
    if direction == "horizontal":
        return BondEnvironment(left, right, (top[x], top[x + 1]), (bottom[x], bottom[x + 1]), direction)
    return BondEnvironment(left, right, (top[x],), (bottom[x],), direction)


def reduced_layer(Q):
    """``_reduced_layer`` (full_update.jl:142-148): A[W2, N2, E2, rb, rk] with the
    (ket, bra) pairs of the first three legs fused."""
    A = ein("wneur,WNEuR->wWnNeERr", Q, Q)
    for k in (4, 2, 0):
        A = fuse(A, k, 2)
    return A


def reduced_metric(env, Qa, Qb):
    """N[rb, sb, rk, sk] (full_update.jl:150-172)."""
    A = reduced_layer(Qa)
    B = reduced_layer(Qb)
    left, right = env.left, env.right
    if env.direction == "horizontal":
        T1, T2 = env.top
        B1, B2 = env.bottom
        X = ein("awb,anc->bncw", left, T1)
        Y = ein("bncw,wnsRr->bcsRr", X, A)
        L = ein("bcsRr,bsf->cfRr", Y, B1)
        U = ein("deg,cvd->cveg", right, T2)
        V = ein("cveg,vetSs->cgtSs", U, B)
        R = ein("cgtSs,ftg->cfSs", V, B2)
        return ein("cfRr,cfSs->RSrs", L, R)
    T = env.top[0]
    bottom = env.bottom[0]
    X = ein("awvb,anc->wvbnc", left, T)
    Y = ein("wvbnc,wneRr->vbceRr", X, A)
    L = ein("vbceRr,cefd->vbfdRr", Y, right)
    R = ein("vftSs,btd->vfSsbd", B, bottom)
    return ein("vbfdRr,vfSsbd->RSrs", L, R)


def bond_hole_metric(env, Qa, Qb):
    """The same metric from the complete six-leg hole (``bond_hole``,
    boundary_mps.jl:326-345) contracted with the reduced layers: an independent
    contraction order that checks ``reduced_metric``."""
    A = reduced_layer(Qa)
    B = reduced_layer(Qb)
    left, right = env.left, env.right
    if env.direction == "horizontal":
        T1, T2 = env.top
        B1, B2 = env.bottom
        X = ein("awb,anc->wbnc", left, T1)
        Lh = ein("wbnc,bsf->cwnfs", X, B1)
        Rh = ein("deg,cvd,ftg->cveft", right, T2, B2)
        H = ein("cwnfs,cveft->wnsvet", Lh, Rh)
        return ein("wnsvet,wnsRr,vetQq->RQrq", H, A, B)
    T = env.top[0]
    Bt = env.bottom[0]
    Lh = ein("awvb,anc->wvbnc", left, T)
    Rh = ein("bsd,cefd->bscef", Bt, right)
    H = ein("wvbnc,bscef->wnevfs", Lh, Rh)
    return ein("wnevfs,wneRr,vfsQq->RQrq", H, A, B)


def bond_hole(env):
    """The complete bond hole (``bond_hole`` / ``_grouped_bond_hole``,
    boundary_mps.jl:326-353, full_update.jl:228-257) over the six exterior
    double legs: H[w,n,s,v,e,t] (horizontal: W, N, S of the first site, N, E,
    S of the second) or H[w,n,e,v,f,s] (vertical: W, N, E of the upper site,
    W, E, S of the lower)."""
    left, right = env.left, env.right
    if env.direction == "horizontal":
        T1, T2 = env.top
        B1, B2 = env.bottom
        X = ein("awb,anc->wbnc", left, T1)
        Lh = ein("wbnc,bsf->cwnfs", X, B1)
        Rh = ein("deg,cvd,ftg->cveft", right, T2, B2)
        return ein("cwnfs,cveft->wnsvet", Lh, Rh)
    T = env.top[0]
    Bt = env.bottom[0]
    Lh = ein("awvb,anc->wvbnc", left, T)
    Rh = ein("bsd,cefd->bscef", Bt, right)
    return ein("wvbnc,bscef->wnevfs", Lh, Rh)


def positive_reduced_metric(env, Qa, Qb):
    """``_positive_reduced_metric`` (full_update.jl:289-331): the grouped hole
    as a matrix (rows: the ket copies of the first site's three exterior legs,
    then the second's; columns: the bra copies), scaled by its norm, made
    symmetric, eigendecomposed per Z2 block; every positive eigenpair gives
    W = sqrt(lambda) v' and the piece W'W contracted with the Q factors
    (``_grouped_repair_piece``, :277-287). Returns (N[rb, sb, rk, sk], report)."""
    H = bond_hole(env)
    if not np.all(np.isfinite(H.a)):
        raise ValueError("nonfinite norm metric")
    dims = [int(np.sqrt(n)) for n in H.shape]
    legpar = []  # parity of each virtual (ket or bra) index of each exterior leg
    for k in range(6):
        pk = H.p[k].reshape(dims[k], dims[k])  # fused pair parity = ket + bra
        legpar.append(pk[:, 0] ^ pk[0, 0])     # ket parities (bra index 0 fixed)
    A = H.a.reshape([d for m in dims for d in (m, m)])
    A = np.transpose(A, [2 * k for k in range(6)] + [2 * k + 1 for k in range(6)])
    n = int(np.prod(dims))
    M = A.reshape(n, n)
    rowpar = np.zeros(1, dtype=np.int64)
    for k in range(6):
        rowpar = ((rowpar[:, None] + legpar[k][None, :]) % 2).reshape(-1)
    scale = float(np.linalg.norm(M))
    if scale > 0:
        M = M / scale
    defect = float(np.sqrt(np.sum((M - M.T) ** 2)))
    M = (M + M.T) / 2
    # Q factors as [exterior (3 legs fused), aux, r]
    qa = Qa.a.reshape(-1, Qa.shape[3], Qa.shape[4])
    qb = Qb.a.reshape(-1, Qb.shape[3], Qb.shape[4])
    na = qa.shape[0]
    result = np.zeros((qa.shape[2], qb.shape[2], qa.shape[2], qb.shape[2]))  # [rb, sb, rk, sk]
    minimum = spectral = negw = neg2 = spec2 = 0.0
    nneg = 0
    for q in (0, 1):
        idx = np.nonzero(rowpar == q)[0]
        if len(idx) == 0:
            continue
        w, V = np.linalg.eigh(M[np.ix_(idx, idx)])
        minimum = min(minimum, float(np.min(w)))
        spectral = max(spectral, float(np.max(np.abs(w))))
        nneg += int(np.sum(w < 0))
        negw += float(np.sum(w < 0))
        neg2 += float(np.sum(np.minimum(w, 0) ** 2))
        spec2 += float(np.sum(w ** 2))
        for k in np.nonzero(w > 0)[0]:
            v = np.zeros(n)
            v[idx] = np.sqrt(w[k]) * V[:, k]             # row k of W (ket side)
            Wk = v.reshape(na, -1)                       # [first site's legs, second site's legs]
            ket = np.einsum("ab,axr,bys->xrys", Wk, qa, qb)
            result += np.einsum("xRyS,xrys->RSrs", ket, ket)  # W'W with both Q factors, aux traced
    tiny = np.finfo(np.float64).tiny
    report = {"minimum_relative_eigenvalue": minimum / max(spectral, tiny), "negative_eigenvalues": negw,
              "negative_multiplets": nneg, "negative_spectral_fraction": np.sqrt(neg2) / max(np.sqrt(spec2), tiny),
              "hermiticity_defect": defect if scale > 0 else defect / tiny}
    rp = Qa.p[4]
    sp = Qb.p[4]
    return Z2(result, [rp, sp, rp, sp]), report


# ---------------------------------------------------------------------------
# environment gauge (full_update.jl:55-117)
# ---------------------------------------------------------------------------

def transform_metric(N, Li, Ri):
    """``_fu_transform_metric`` (full_update.jl:55-62)."""
    first = ein("ijkl,ix->jklx", N, Li)
    second = ein("jklx,jy->klxy", first, Ri)
    third = ein("klxy,kz->lxyz", second, Li)
    return ein("lxyz,lw->xyzw", third, Ri)


def environment_gauge(N):
    """``_fu_environment_gauge`` (full_update.jl:64-104)."""
    if not np.all(np.isfinite(N.a)):
        raise ValueError("nonfinite FU norm metric")
    scale = N.norm()
    if not scale > 0:
        raise ValueError("zero FU norm metric")
    ra, sa = N.shape[0], N.shape[1]
    M, pr, pc = as_matrix(N.scale(1.0 / scale), 2)
    H = (M + M.T) / 2
    blocks = eigh_blocks(H, pr)
    allv = np.concatenate([b[2] for b in blocks]) if blocks else np.zeros(0)
    spectral = float(np.max(np.abs(allv))) if allv.size else 0.0
    spectral = 1.0 if spectral == 0.0 else spectral
    relative_floor = max(np.sqrt(EPS), ra * sa * EPS)
    floor = relative_floor * spectral
    # Z = root * vectors' : rows indexed by eigenvector k (parity of its block)
    rows, kpar = [], []
    for q, idx, w, V in blocks:
        root = np.sqrt(np.maximum(w, floor))
        for k in range(len(w)):
            z = np.zeros(M.shape[1])
            z[idx] = root[k] * V[:, k]
            rows.append(z)
            kpar.append(q)
    Z = Z2(np.asarray(rows).reshape(len(rows), ra, sa), [np.asarray(kpar), N.p[2], N.p[3]])
    _, left = left_orth(permute(Z, (0, 2, 1)), 2)      # ((k, sk), rk) -> left[x, rk]
    _, right = left_orth(permute(Z, (0, 1, 2)), 2)     # ((k, rk), sk) -> right[y, sk]
    Li = Z2(np.linalg.inv(left.a), [left.p[1], left.p[0]])
    Ri = Z2(np.linalg.inv(right.a), [right.p[1], right.p[0]])
    metric = transform_metric(N, Li, Ri)
    diag = {"relative_floor": relative_floor,
            "floored_eigenvalues": int(np.sum(allv < floor)),
            "minimum_relative_eigenvalue": float(np.min(allv) / spectral) if allv.size else 0.0,
            "hermiticity_defect": float(np.linalg.norm(M - M.T))}
    return {"left": left, "right": right, "left_inverse": Li, "right_inverse": Ri, "metric": metric,
            "diagnostic": diag}


def svd_initial(target, D, caps):
    """``_fu_svd_initial`` (full_update.jl:31-42): truncrank(D) & truncspace & trunctol."""
    cutoff = EPS * max(target.shape) * target.norm()
    U, S, V, bp, _, _ = tsvd(target, 2, chi=D, caps=caps, atol=cutoff)
    if len(S) == 0:
        raise ValueError("gate produced a zero target")
    root = np.sqrt(S)
    return Z2(U.a * root, U.p), Z2(V.a * root[:, None, None], V.p)


def initialize_pair(pair, N, D, caps):
    """``_initialize_pair`` (full_update.jl:106-113)."""
    g = environment_gauge(N)
    wt = act(act(pair["target"], g["left"], 0), g["right"], 3)
    left, right = svd_initial(wt, D, caps)
    out = dict(pair)
    out.update(left=left, right=right, working_target=wt, working_metric=g["metric"],
               left_inverse=g["left_inverse"], right_inverse=g["right_inverse"], gauge=g["diagnostic"])
    return out


def original_cores(pair, A, B):
    """``_fu_original_cores`` (full_update.jl:115)."""
    return act(A, pair["left_inverse"], 0), act(B, pair["right_inverse"], 2)


# ---------------------------------------------------------------------------
# ALS fit (full_update.jl:342-482)
# ---------------------------------------------------------------------------

def apply_pair_metric(N, theta):
    return ein("rstu,tpqu->rpqs", N, theta)


def pair_tensor(A, B):
    return ein("rpb,bqs->rpqs", A, B)


def residual_cost(N, A, B, target):
    res = pair_tensor(A, B) - target
    return float(np.sum(res.a * apply_pair_metric(N, res).a))


def repair_trigger(raw, target):
    scale = max(abs(target), np.finfo(np.float64).tiny)
    if not (np.isfinite(raw) and np.isfinite(target)):
        raise ValueError("nonfinite FU residual/target")
    if raw / scale < -RESIDUAL_ROUNDOFF_TOLERANCE:
        return "negative_residual"
    if target < 0:
        return "negative_target_norm"
    return None


def cost_evaluation(raw, target):
    relative = raw / max(abs(target), np.finfo(np.float64).tiny)
    if not np.isfinite(relative):
        raise ValueError("nonfinite relative FU residual")
    if relative < -RESIDUAL_ROUNDOFF_TOLERANCE:
        raise ValueError("negative FU residual needs PSD repair")
    return abs(relative), relative < 0


def solve_core(M, pr, rhs, x0, opts):
    """``_solve_core`` (full_update.jl:371-400): Jacobi-preconditioned CG with a
    diagonal shift, on the matrix M (rows with parities pr) and a right-hand
    side with one column per physical index."""
    diag = np.maximum(np.diag(M), 0.0)
    scale = float(np.max(diag)) if diag.size else 0.0
    tiny = np.finfo(np.float64).tiny
    shift = opts.diagonal_shift * max(scale, tiny)
    floor = max(shift, EPS * scale, tiny)
    pre = np.maximum(diag + shift, floor)[:, None]

    def action(v):
        return M @ v + shift * v

    x = x0.copy()
    r = rhs - action(x)
    rhsnorm = max(np.linalg.norm(rhs), tiny)
    residual = np.linalg.norm(r) / rhsnorm
    if residual <= opts.linear_tolerance:
        return x, residual
    z = r / pre
    p = z.copy()
    rz = float(np.sum(r * z))
    for _ in range(opts.linear_max_iterations):
        Ap = action(p)
        curvature = float(np.sum(p * Ap))
        if not curvature > EPS * max(np.linalg.norm(p) * np.linalg.norm(Ap), tiny):
            break
        alpha = rz / curvature
        x = x + alpha * p
        r = r - alpha * Ap
        residual = np.linalg.norm(r) / rhsnorm
        if residual <= opts.linear_tolerance:
            break
        z = r / pre
        nxt = float(np.sum(r * z))
        if not nxt >= 0:
            break
        p = z + (nxt / max(rz, tiny)) * p
        rz = nxt
    return x, float(residual)


def left_normal_equation(metric, right, wt):
    """``_left_normal_equation`` (full_update.jl:402-408): normal[(r,k),(r',k')]
    and rhs[(r,k), p]."""
    normal = ein("fsgn,bps,opn->fbgo", metric, right, right)
    rhs = ein("fpqs,bqs->fbp", wt, right)
    return normal, rhs


def right_normal_equation(metric, left, wt):
    """``_right_normal_equation`` (full_update.jl:410-416): normal[(k,s),(k',s')]."""
    normal = ein("wseN,wpb,epo->bsoN", metric, left, left)
    rhs = ein("wqps,wqb->bsp", wt, left)
    return normal, rhs


def optimize_pair(pair, N, opts):
    """``_optimize_pair`` (full_update.jl:418-482)."""
    A, B = pair["left"], pair["right"]
    target = pair["working_target"]
    metric = pair["working_metric"]
    target_norm = float(np.sum(pair["target"].a * apply_pair_metric(N, pair["target"]).a))
    oA, oB = original_cores(pair, A, B)
    raw = residual_cost(N, oA, oB, pair["target"])
    trig = repair_trigger(raw, target_norm)
    if trig is not None:
        raise RepairNeeded(trig)
    initial_cost, fallback = cost_evaluation(raw, target_norm)
    bestA, bestB, best = oA, oB, initial_cost
    best_factors = None
    residual = 0.0
    iterations = 0
    limit = False
    stop = "disabled" if opts.max_iterations == 0 else "iteration_limit"
    if fallback:
        stop = "residual_fallback"
    elif opts.max_iterations > 0 and best <= opts.error_tolerance:
        stop = "error_tolerance"
    elif opts.max_iterations > 0:
        wt = apply_pair_metric(metric, target)
        for it in range(1, opts.max_iterations + 1):
            MA, ya = left_normal_equation(metric, B, wt)
            Mm = MA.a.reshape(MA.shape[0] * MA.shape[1], -1)
            x0 = permute(A, (0, 2, 1)).a.reshape(Mm.shape[0], -1)
            sol, ra = solve_core(Mm, None, ya.a.reshape(Mm.shape[0], -1), x0, opts)
            solt = Z2(sol.reshape(A.shape[0], A.shape[2], A.shape[1]), [A.p[0], A.p[2], A.p[1]])
            QA, R = left_orth(permute(solt, (0, 2, 1)), 2)       # QA[r,p,k], R[k,bond]
            right_initial = ein("kb,bqs->kqs", R, B)
            A = QA
            MB, yb = right_normal_equation(metric, A, wt)
            Mm = MB.a.reshape(MB.shape[0] * MB.shape[1], -1)
            x0 = permute(right_initial, (0, 2, 1)).a.reshape(Mm.shape[0], -1)
            sol, rb = solve_core(Mm, None, yb.a.reshape(Mm.shape[0], -1), x0, opts)
            ri = right_initial
            solt = Z2(sol.reshape(ri.shape[0], ri.shape[2], ri.shape[1]), [ri.p[0], ri.p[2], ri.p[1]])
            C, QB = right_orth(permute(solt, (0, 2, 1)), 1)        # C[bond,k'], QB[k',q,s]
            A = ein("rpk,kc->rpc", QA, C)
            B = QB
            oA, oB = original_cores(pair, A, B)
            raw = residual_cost(N, oA, oB, pair["target"])
            trig = repair_trigger(raw, target_norm)
            if trig is not None:
                raise RepairNeeded(trig)
            cost, fb = cost_evaluation(raw, target_norm)
            residual = max(ra, rb)
            iterations = it
            if fb:
                stop = "residual_fallback"
                break
            if cost <= best:
                best, bestA, bestB = cost, oA, oB
                best_factors = (QA, C, QB)
            if best <= opts.error_tolerance:
                stop = "error_tolerance"
                break
            limit = it == opts.max_iterations
    if best_factors is not None:
        A, B = balance_bond(*best_factors)
        bestA, bestB = original_cores(pair, A, B)
    return {"left": bestA, "right": bestB, "initial_cost": initial_cost, "final_cost": best,
            "residual": residual, "iterations": iterations, "iteration_limit_reached": limit, "stop_reason": stop}


def balance_bond(QA, C, QB):
    """``_fu_balance_bond`` (full_update.jl:44-52): compact SVD of the bond centre."""
    U, S, V, _, _, _ = tsvd(C, 1)
    root = np.sqrt(S)
    Us = Z2(U.a * root, U.p)
    Vs = Z2(V.a * root[:, None], V.p)
    return ein("rpk,kc->rpc", QA, Us), ein("ck,kqs->cqs", Vs, QB)


def update_bond(state, i, j, gate, D, caps, opts, env, direction, force_repair=False):
    """``update_bond!`` (full_update.jl:484-508). When the residual check
    requests it (or ``force_repair``, a test hook taking that branch at the
    initial stage), the reduced metric of the PSD-projected complete hole
    replaces N and the pair is initialised and fitted again; a second request
    raises the reference's error."""
    pair = reduced_pair(state[i], state[j], gate, direction)
    N = reduced_metric(env, pair["Qa"], pair["Qb"])
    init = initialize_pair(pair, N, D, caps)
    repairs = []
    try:
        if force_repair:
            raise RepairNeeded("forced")
        fit = optimize_pair(init, N, opts)
    except RepairNeeded as trig:
        repaired, report = positive_reduced_metric(env, pair["Qa"], pair["Qb"])
        report["reason"] = str(trig)
        repairs.append(report)
        init = initialize_pair(pair, repaired, D, caps)
        try:
            fit = optimize_pair(init, repaired, opts)
        except RepairNeeded:
            raise ValueError("PSD repair did not remove the FU residual failure")
    state[i], state[j] = restore_pair(pair, fit["left"], fit["right"])
    rec = {k: fit[k] for k in ("initial_cost", "final_cost", "residual", "iterations", "iteration_limit_reached",
                               "stop_reason")}
    rec.update(bond=(i, j), direction=direction, repairs=repairs)
    return rec


# ---------------------------------------------------------------------------
# sweeps and the evolution step (full_update.jl:510-597)
# ---------------------------------------------------------------------------

def sweep(state, model, delta, D, caps, opts, chi, direction, reverse, entropy=None, force=(), first=0):
    """``_sweep!`` (full_update.jl:510-561) for the row boundary-MPS contraction.
    ``entropy``: the ``BoundaryEntropy`` dict ({"maximum": ...}) every absorb
    reports to, or None."""
    L = state.lattice
    Lx, Ly = L.Lx, L.Ly
    layers = double_layers(state)
    north, south, env_err = boundary_environments(layers, chi, side="north" if reverse else "south",
                                                  entropy=entropy)

    def obs(side, row_cut):
        if entropy is None:
            return None
        return lambda bond, value: _record_entropy(entropy, side, row_cut, bond, value)
    ymax = Ly if direction == "horizontal" else Ly - 1
    xmax = Lx - 1 if direction == "horizontal" else Lx
    records = []
    ys = range(ymax - 1, -1, -1) if reverse else range(ymax)
    for y in ys:
        top = north[y].tensors
        lastrow = y + (0 if direction == "horizontal" else 1)
        bottom = south[lastrow + 1].tensors

        def row_of():
            if direction == "horizontal":
                return row(layers, y)
            return [Slab(layers[x][y], layers[x][y + 1]) for x in range(Lx)]
        edges_l, edges_r = row_environments(top, row_of(), bottom, side="left" if reverse else "right")
        rolling = edges_r[Lx] if reverse else edges_l[0]
        xs = range(xmax - 1, -1, -1) if reverse else range(xmax)
        for x in xs:
            i = (x, y)
            j = (x + 1, y) if direction == "horizontal" else (x, y + 1)
            finalx = x + 1 if direction == "horizontal" else x
            left = edges_l[x] if reverse else rolling
            right = rolling if reverse else edges_r[finalx + 1]
            env = bond_environment(top, bottom, x, direction, left, right)
            gate = model.trotter_gate(L, i, j, delta)
            gate = Z2(gate, [state[i].p[2], state[j].p[2], state[i].p[2], state[j].p[2]])
            rec = update_bond(state, i, j, gate, D, caps, opts, env, direction,
                              force_repair=(first + len(records)) in force)
            rec["max_environment_truncation"] = max(north[y].max_truncation, south[lastrow + 1].max_truncation)
            records.append(rec)
            layers[i[0]][i[1]] = double_layer(state[i])
            layers[j[0]][j[1]] = double_layer(state[j])
            column = finalx if reverse else x
            if x != (0 if reverse else xmax - 1):
                E = layers[column][y] if direction == "horizontal" else Slab(layers[column][y], layers[column][y + 1])
                if reverse:
                    rolling = safe_scale(right_step(top[column], E, bottom[column], rolling))
                else:
                    rolling = safe_scale(left_step(rolling, top[column], E, bottom[column]))
        if reverse:
            south[lastrow] = absorb(south[lastrow + 1], row(layers, lastrow), chi, "south", obs("south", lastrow - 1))
            env_err = max(env_err, south[lastrow].max_truncation)
        else:
            north[y + 1] = absorb(north[y], row(layers, y), chi, "north", obs("north", y + 1))
            env_err = max(env_err, north[y + 1].max_truncation)
    return records, env_err


def evolution_step(state, model, delta, D, caps, opts, chi, order=2, record_entropy=False, force_repair=()):
    """``evolution_step!`` (full_update.jl:563-597) with FullUpdate and
    BoundaryMPS(chi). ``caps`` = (even, odd) sector budget of the bond space, or None.
    ``force_repair``: bond updates (sweep order) that take the PSD-repair branch."""
    records = []
    env_err = 0.0
    entropy = {"maximum": 0.0} if record_entropy else None
    gate_delta = delta / 2 if order == 2 else delta
    sweeps = ((("horizontal", False), ("vertical", False), ("vertical", True), ("horizontal", True))
              if order == 2 else (("horizontal", False), ("vertical", False)))
    L = state.lattice
    if len(L) == 1:
        state[(0, 0)] = act(state[(0, 0)], Z2(model.onsite_gate(delta), [state[(0, 0)].p[2]] * 2), 2)
    else:
        for direction, reverse in sweeps:
            if not L.bonds(direction):
                continue
            r, e = sweep(state, model, gate_delta, D, caps, opts, chi, direction, reverse, entropy,
                         force=force_repair, first=len(records))
            records += r
            env_err = max(env_err, e)
    state.beta += 2 * delta
    norms = normalize_state(state)
    return {"beta": state.beta, "records": records, "normalizations": norms,
            "max_environment_truncation": env_err,
            "max_boundary_entropy": float("nan") if entropy is None else entropy["maximum"],
            "max_state_truncation": max((r["final_cost"] for r in records), default=0.0)}
