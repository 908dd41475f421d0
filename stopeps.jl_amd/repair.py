"""PSD repair of the full-update bond metric on the host: the callback that
``spp_fu_evolution_step`` calls for a bond whose residual check requests the
repair (``_positive_reduced_metric`` and the repair branch of ``update_bond!``,
src/evolution/full_update.jl:177-340, 484-508).

The library contracts the sample's complete bond hole H on the device and
passes it with the Q factors of the reduced pair. Here H is read as a matrix
over the ket and bra copies of its six exterior legs, scaled by its norm,
made symmetric, eigendecomposed per Z2 block (LAPACK through numpy, as the
reference uses ``eigen!``), its negative eigenvalues dropped, and the projected
hole contracted with the Q factors into the repaired reduced metric. The
library then runs the bond solve of that sample again on it.
"""
import ctypes

import numpy as np

from . import _lib


class RepairProblem(ctypes.Structure):
    """``spp_fu_repair_problem`` (include/stopeps_hip.h)."""
    _fields_ = [("sample", ctypes.c_int), ("bond", ctypes.c_int), ("horizontal", ctypes.c_int),
                ("hole_dims", ctypes.c_int * 6), ("hole_evens", ctypes.c_int * 6),
                ("hole_q", ctypes.c_int * 6), ("hole_slot", ctypes.c_int * 6),
                ("hole", ctypes.POINTER(ctypes.c_double)),
                ("qa_dims", ctypes.c_int * 6), ("qa_evens", ctypes.c_int * 6),
                ("qb_dims", ctypes.c_int * 6), ("qb_evens", ctypes.c_int * 6),
                ("qa_core", ctypes.c_int), ("qb_core", ctypes.c_int),
                ("qa", ctypes.POINTER(ctypes.c_double)), ("qb", ctypes.POINTER(ctypes.c_double)),
                ("target_norm", ctypes.c_double)]


REPAIR_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(RepairProblem),
                             ctypes.POINTER(ctypes.c_double))


def pair_order(m, e):
    """Index of the (ket i, bra j) pair in the library's parity-sorted pair
    leg of a virtual leg of dimension m with e even indices: even pairs
    (even-even, odd-odd) first, then odd pairs (even-odd, odd-even)."""
    o = m - e
    P = np.empty((m, m), dtype=np.int64)
    for i in range(m):
        for j in range(m):
            if i < e and j < e:
                P[i, j] = i * e + j
            elif i >= e and j >= e:
                P[i, j] = e * e + (i - e) * o + (j - e)
            elif i < e:
                P[i, j] = e * e + o * o + i * o + (j - e)
            else:
                P[i, j] = e * e + o * o + e * o + (i - e) * e + j
    return P


def psd_projection(M, parity):
    """PSD projection of a real matrix in its Z2 blocks, as
    ``_positive_reduced_metric`` forms it (full_update.jl:289-331): M is scaled
    by its Frobenius norm, symmetrised, each block (rows and columns of one
    parity) eigendecomposed and its negative eigenvalues dropped. Returns
    (H+ of the scaled matrix, report) with the reference's diagnostics."""
    scale = float(np.linalg.norm(M))
    S = M / scale if scale > 0 else M.copy()
    defect = float(np.linalg.norm(S - S.T))
    S = (S + S.T) / 2
    P = np.zeros_like(S)
    minimum, spectral, nneg, neg2, spec2 = 0.0, 0.0, 0, 0.0, 0.0
    for q in (0, 1):
        idx = np.nonzero(parity == q)[0]
        if len(idx) == 0:
            continue
        w, V = np.linalg.eigh(S[np.ix_(idx, idx)])
        minimum = min(minimum, float(w.min()))
        spectral = max(spectral, float(np.abs(w).max()))
        nneg += int(np.sum(w < 0))
        neg2 += float(np.sum(np.minimum(w, 0) ** 2))
        spec2 += float(np.sum(w ** 2))
        keep = w > 0
        Vk = V[:, keep]
        P[np.ix_(idx, idx)] = (Vk * w[keep]) @ Vk.T
    tiny = np.finfo(np.float64).tiny
    report = {"minimum_relative_eigenvalue": minimum / max(spectral, tiny), "negative_eigenvalues": float(nneg),
              "negative_multiplets": nneg, "negative_spectral_fraction": np.sqrt(neg2) / max(np.sqrt(spec2), tiny),
              "hermiticity_defect": defect if scale > 0 else defect / tiny}
    return P, report


def repaired_metric(hole, hole_dims, hole_evens, hole_q, hole_slot, qa, qa_dims, qa_evens, qa_core, qb, qb_dims,
                    qb_evens, qb_core):
    """Reduced metric N[pair(rk, rb)][pair(sk, sb)] of the PSD-projected bond
    hole (the layouts of ``spp_fu_repair_problem``). Returns (metric, report)."""
    dims = [int(v) for v in hole_dims]
    H = np.asarray(hole, dtype=np.float64).reshape([m * m for m in dims])
    # pair legs -> (ket, bra) index pairs
    for k in reversed(range(6)):
        P = pair_order(dims[k], int(hole_evens[k]))
        shape = list(H.shape)
        H = np.take(H, P.reshape(-1), axis=k).reshape(shape[:k] + [dims[k], dims[k]] + shape[k + 1:])
    # order the exterior legs as (legs of qa, legs of qb), ket legs then bra legs
    order = [k for k in range(6) if hole_q[k] == 0] + [k for k in range(6) if hole_q[k] == 1]
    H = np.transpose(H, [2 * k for k in order] + [2 * k + 1 for k in order])
    n = int(np.prod(dims))
    M = H.reshape(n, n)
    par = np.zeros(1, dtype=np.int64)
    for k in order:
        p = (np.arange(dims[k]) >= int(hole_evens[k])).astype(np.int64)
        par = ((par[:, None] + p[None, :]) % 2).reshape(-1)
    Hp, report = psd_projection(M, par)

    def factor(q, qd, core, owner):
        Q = np.asarray(q, dtype=np.float64).reshape([int(v) for v in qd])
        slots = [int(hole_slot[k]) for k in order if hole_q[k] == owner]
        Q = np.transpose(Q, slots + [2, 3, int(core)])
        ext = int(np.prod([Q.shape[i] for i in range(len(slots))]))
        return Q.reshape(ext, Q.shape[len(slots)] * Q.shape[len(slots) + 1], Q.shape[-1])

    Qa = factor(qa, qa_dims, qa_core, 0)
    Qb = factor(qb, qb_dims, qb_core, 1)
    na, nb = Qa.shape[0], Qb.shape[0]
    Hp = Hp.reshape(na, nb, na, nb)                      # [ket a, ket b, bra a, bra b]
    Na = np.einsum("kxr,bxs->kbrs", Qa, Qa)              # [ket a, bra a, rk, rb]
    Nb = np.einsum("kxr,bxs->kbrs", Qb, Qb)              # [ket b, bra b, sk, sb]
    T = np.einsum("klbc,lcst->kbst", Hp, Nb)
    N = np.einsum("kbst,kbru->rust", T, Na)              # [rk, rb, sk, sb]
    ra, rb = Qa.shape[2], Qb.shape[2]
    Pa = pair_order(ra, int(qa_evens[qa_core]))
    Pb = pair_order(rb, int(qb_evens[qb_core]))
    out = np.zeros((ra * ra, rb * rb))
    out[Pa[:, :, None, None], Pb[None, None, :, :]] = N
    return out, report


class RepairCallback:
    """The host callback of one context: computes the repaired metric and
    keeps the reference's repair records (``repairs`` of ``update_bond!``)."""

    def __init__(self):
        self.records = []
        self.error = None
        self._fn = REPAIR_FN(self._call)

    def _call(self, user, problem, metric):
        try:
            p = problem.contents
            hn = int(np.prod([int(v) ** 2 for v in p.hole_dims]))
            qan = int(np.prod(list(p.qa_dims)))
            qbn = int(np.prod(list(p.qb_dims)))
            hole = np.ctypeslib.as_array(p.hole, shape=(hn,))
            qa = np.ctypeslib.as_array(p.qa, shape=(qan,))
            qb = np.ctypeslib.as_array(p.qb, shape=(qbn,))
            out, report = repaired_metric(hole, list(p.hole_dims), list(p.hole_evens), list(p.hole_q),
                                          list(p.hole_slot), qa, list(p.qa_dims), list(p.qa_evens), p.qa_core, qb,
                                          list(p.qb_dims), list(p.qb_evens), p.qb_core)
            self.transform(out)
            n = out.size
            np.ctypeslib.as_array(metric, shape=(n,))[:] = out.reshape(-1)
            report.update(sample=int(p.sample), bond=int(p.bond), direction="horizontal" if p.horizontal else "vertical",
                          reason="negative_residual", target_norm=float(p.target_norm))
            self.records.append(report)
            return 0
        except Exception as exc:  # reported by the library as a failed call
            self.error = exc
            return -1

    def transform(self, metric):
        """Hook on the repaired metric before it goes back (identity)."""

    def install(self, ctx):
        _lib.check(ctx.handle, ctx._lib.spp_fu_set_repair(ctx.handle, ctypes.cast(self._fn, ctypes.c_void_p), None))

    @staticmethod
    def uninstall(ctx):
        _lib.check(ctx.handle, ctx._lib.spp_fu_set_repair(ctx.handle, None, None))
