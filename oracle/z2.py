"""Z2-graded dense tensors and sector-aware factorizations (oracle).

Restates the parts of TensorKit 0.17 / MatrixAlgebraKit 0.6 that the
reference's hot path uses (not vendored in the reference; their call sites are
cited at each user). For Z2 (bosonic, abelian, self-dual) a block-sparse
TensorMap is equivalent to a dense array whose entries vanish unless the index
parities sum to zero; ``permute`` is a transpose, ``_fuse`` a reshape with
parities added, and matrix factorizations split into two parity blocks.
"""
import numpy as np

EPS = np.finfo(np.float64).eps


class Z2:
    """Dense array ``a`` with a parity vector per leg ``p``."""

    __slots__ = ("a", "p")

    def __init__(self, a, p):
        self.a = np.asarray(a)
        self.p = [np.asarray(x, dtype=np.int64) % 2 for x in p]
        if len(self.p) != self.a.ndim:
            raise ValueError("one parity vector per leg")
        for pp, n in zip(self.p, self.a.shape):
            if len(pp) != n:
                raise ValueError("parity vector length mismatch")

    @property
    def ndim(self):
        return self.a.ndim

    @property
    def shape(self):
        return self.a.shape

    def copy(self):
        return Z2(self.a.copy(), [x.copy() for x in self.p])

    def conj(self):
        return Z2(self.a.conj(), self.p)

    def scale(self, s):
        return Z2(self.a * s, self.p)

    def __add__(self, other):
        _check_same_parities(self, other)
        return Z2(self.a + other.a, self.p)

    def __sub__(self, other):
        _check_same_parities(self, other)
        return Z2(self.a - other.a, self.p)

    def norm(self):
        return float(np.linalg.norm(self.a.ravel()))


def _check_same_parities(x, y):
    if x.shape != y.shape or any(not np.array_equal(a, b) for a, b in zip(x.p, y.p)):
        raise ValueError("space mismatch")


def fused_parity(plist):
    """Parity of each fused index, C order (last leg fastest)."""
    out = np.zeros(1, dtype=np.int64)
    for p in plist:
        out = ((out[:, None] + np.asarray(p)[None, :]) % 2).ravel()
    return out


def total_parity_grid(t):
    grid = np.zeros(t.shape, dtype=np.int64)
    for k, p in enumerate(t.p):
        shape = [1] * t.ndim
        shape[k] = len(p)
        grid = grid + p.reshape(shape)
    return grid % 2


def charge_defect(t):
    """Largest |entry| that violates neutrality (0 for a symmetric tensor)."""
    if t.a.size == 0:
        return 0.0
    bad = total_parity_grid(t) != 0
    return float(np.max(np.abs(t.a[bad]), initial=0.0))


def symmetrize(t):
    """Zero every entry whose index parities do not sum to 0."""
    a = t.a.copy()
    a[total_parity_grid(t) != 0] = 0
    return Z2(a, t.p)


def ein(spec, *ts):
    """``np.einsum`` with parity bookkeeping; letters shared by operands must
    carry identical parity vectors (the Z2 analogue of matching spaces)."""
    lhs, rhs = spec.replace(" ", "").split("->")
    terms = lhs.split(",")
    if len(terms) != len(ts):
        raise ValueError("operand count")
    par = {}
    for term, t in zip(terms, ts):
        if len(term) != t.ndim:
            raise ValueError(f"rank mismatch in {spec}")
        for ch, p in zip(term, t.p):
            if ch in par:
                if not np.array_equal(par[ch], p):
                    raise ValueError(f"space mismatch on index {ch} in {spec}")
            else:
                par[ch] = p
    a = np.einsum(spec, *[t.a for t in ts], optimize=True)
    return Z2(a, [par[ch] for ch in rhs])


def permute(t, perm):
    return Z2(np.transpose(t.a, perm), [t.p[k] for k in perm])


def fuse(t, first, count):
    """Fuse legs first..first+count-1 (0-based) into one leg, C order.
    Restates ``_fuse`` (src/states/tensors.jl:122-139); the order of the fused
    basis differs from TensorKit's fusion isomorphism by a permutation that no
    scalar depends on."""
    if count == 1:
        return t
    shape = list(t.shape)
    newdim = int(np.prod(shape[first:first + count]))
    newshape = shape[:first] + [newdim] + shape[first + count:]
    p = t.p[:first] + [fused_parity(t.p[first:first + count])] + t.p[first + count:]
    return Z2(t.a.reshape(newshape), p)


def split(t, axis, dims, parities):
    """Inverse of :func:`fuse` (``_split``, src/states/tensors.jl:141-155)."""
    if fused_parity(parities).tolist() != t.p[axis].tolist():
        raise ValueError("tensor splitting space mismatch")
    shape = list(t.shape)
    newshape = shape[:axis] + list(dims) + shape[axis + 1:]
    p = t.p[:axis] + [np.asarray(x) for x in parities] + t.p[axis + 1:]
    return Z2(t.a.reshape(newshape), p)


def as_matrix(t, nrows):
    """(rows legs | columns legs) view: matrix, row parities, column parities."""
    rs = int(np.prod(t.shape[:nrows], dtype=np.int64))
    cs = int(np.prod(t.shape[nrows:], dtype=np.int64))
    return t.a.reshape(rs, cs), fused_parity(t.p[:nrows]), fused_parity(t.p[nrows:])


# ----------------------------------------------------------------------------
# Truncation selection (MatrixAlgebraKit truncrank / truncspace / trunctol and
# their intersection ``&``, used at src/contraction/boundary_mps.jl:18 and
# src/evolution/full_update.jl:33-34).
# ----------------------------------------------------------------------------

def select_kept(values_by_sector, chi=None, caps=None, atol=None):
    """Return the number of values kept in each sector.

    ``values_by_sector[q]`` is sorted in descending order. ``chi``: keep the
    ``chi`` largest values over all sectors (truncrank). ``caps[q]``: keep at
    most ``caps[q]`` in sector q (truncspace). ``atol``: keep values > atol
    (trunctol). Several strategies combine by intersection of the kept index
    sets. Ties at the truncrank cut are broken towards sector 0 and towards
    earlier values, which is parity unpinned against MatrixAlgebraKit.
    """
    nsec = len(values_by_sector)
    keep = [len(v) for v in values_by_sector]
    if chi is not None:
        allv = []
        for q, v in enumerate(values_by_sector):
            for i, s in enumerate(v):
                allv.append((-float(s), q, i))
        allv.sort()
        count = [0] * nsec
        for (_, q, _) in allv[:max(int(chi), 0)]:
            count[q] += 1
        keep = [min(a, b) for a, b in zip(keep, count)]
    if caps is not None:
        keep = [min(k, int(c)) for k, c in zip(keep, caps)]
    if atol is not None:
        keep = [min(k, int(np.sum(np.asarray(v) > atol))) for k, v in zip(keep, values_by_sector)]
    return keep


def _blocks(pr, pc):
    for q in (0, 1):
        rows = np.nonzero(pr == q)[0]
        cols = np.nonzero(pc == q)[0]
        yield q, rows, cols


def svd_blocks(M, pr, pc):
    """Compact SVD per parity block: list of (q, rows, cols, U, S, Vh)."""
    out = []
    for q, rows, cols in _blocks(pr, pc):
        if len(rows) == 0 or len(cols) == 0:
            out.append((q, rows, cols, np.zeros((len(rows), 0)), np.zeros(0), np.zeros((0, len(cols)))))
            continue
        B = M[np.ix_(rows, cols)]
        U, S, Vh = np.linalg.svd(B, full_matrices=False)
        out.append((q, rows, cols, U, S, Vh))
    return out


def svd_trunc(M, pr, pc, chi=None, caps=None, atol=None):
    """Truncated SVD with sector-aware selection.

    Returns (U, S, Vh, bond_parity, err) where U: rows x k, Vh: k x cols and
    ``err`` = 2-norm of the discarded singular values (MAK truncation_error!).
    Kept columns are ordered sector 0 then sector 1, descending within each.
    """
    blocks = svd_blocks(M, pr, pc)
    keep = select_kept([b[4] for b in blocks], chi=chi, caps=caps, atol=atol)
    k = sum(keep)
    U = np.zeros((M.shape[0], k), dtype=M.dtype)
    Vh = np.zeros((k, M.shape[1]), dtype=M.dtype)
    S = np.zeros(k)
    bp = np.zeros(k, dtype=np.int64)
    off = 0
    err2 = 0.0
    spectra = []
    for (q, rows, cols, Ub, Sb, Vb), kq in zip(blocks, keep):
        spectra.append(Sb)
        U[rows, off:off + kq] = Ub[:, :kq]
        Vh[off:off + kq, cols] = Vb[:kq, :]
        S[off:off + kq] = Sb[:kq]
        bp[off:off + kq] = q
        err2 += float(np.sum(Sb[kq:] ** 2))
        off += kq
    return U, S, Vh, bp, np.sqrt(err2), spectra


def qr_blocks(M, pr, pc):
    """left_orth: M = Q R per parity block, Q rows x k with k_q = min(m_q, n_q).
    R gets a nonnegative diagonal (a gauge choice; scalars do not depend on it)."""
    kq = []
    parts = []
    for q, rows, cols in _blocks(pr, pc):
        if len(rows) == 0 or len(cols) == 0:
            parts.append((q, rows, cols, np.zeros((len(rows), 0)), np.zeros((0, len(cols)))))
            kq.append(0)
            continue
        Q, R = np.linalg.qr(M[np.ix_(rows, cols)], mode="reduced")
        d = np.sign(np.diag(R))
        d[d == 0] = 1
        Q = Q * d[None, :]
        R = R * d[:, None]
        parts.append((q, rows, cols, Q, R))
        kq.append(Q.shape[1])
    k = sum(kq)
    Qf = np.zeros((M.shape[0], k), dtype=M.dtype)
    Rf = np.zeros((k, M.shape[1]), dtype=M.dtype)
    bp = np.zeros(k, dtype=np.int64)
    off = 0
    for (q, rows, cols, Q, R), n in zip(parts, kq):
        Qf[rows, off:off + n] = Q
        Rf[off:off + n, cols] = R
        bp[off:off + n] = q
        off += n
    return Qf, Rf, bp


def lq_blocks(M, pr, pc):
    """right_orth: M = L Q with Q having orthonormal rows, per parity block."""
    Qt, Rt, bp = qr_blocks(M.T.conj(), pc, pr)
    return Rt.T.conj(), Qt.T.conj(), bp


def eigh_blocks(H, p):
    """Hermitian eigendecomposition per parity block of a square matrix."""
    out = []
    for q in (0, 1):
        idx = np.nonzero(p == q)[0]
        if len(idx) == 0:
            out.append((q, idx, np.zeros(0), np.zeros((0, 0))))
            continue
        w, V = np.linalg.eigh(H[np.ix_(idx, idx)])
        out.append((q, idx, w, V))
    return out


def left_orth(t, nrows):
    """QR of ``t`` viewed as (first nrows legs | rest): returns Q (legs of the
    rows + new bond) and R (new bond + column legs)."""
    M, pr, pc = as_matrix(t, nrows)
    Q, R, bp = qr_blocks(M, pr, pc)
    Qt = Z2(Q.reshape(t.shape[:nrows] + (Q.shape[1],)), t.p[:nrows] + [bp])
    Rt = Z2(R.reshape((R.shape[0],) + t.shape[nrows:]), [bp] + t.p[nrows:])
    return Qt, Rt


def right_orth(t, nrows):
    M, pr, pc = as_matrix(t, nrows)
    L, Q, bp = lq_blocks(M, pr, pc)
    Lt = Z2(L.reshape(t.shape[:nrows] + (L.shape[1],)), t.p[:nrows] + [bp])
    Qt = Z2(Q.reshape((Q.shape[0],) + t.shape[nrows:]), [bp] + t.p[nrows:])
    return Lt, Qt


def tsvd(t, nrows, chi=None, caps=None, atol=None):
    """Truncated SVD of ``t`` as (rows | cols): U, S (1-D), V, bond parity, err, spectra."""
    M, pr, pc = as_matrix(t, nrows)
    U, S, Vh, bp, err, spectra = svd_trunc(M, pr, pc, chi=chi, caps=caps, atol=atol)
    Ut = Z2(U.reshape(t.shape[:nrows] + (U.shape[1],)), t.p[:nrows] + [bp])
    Vt = Z2(Vh.reshape((Vh.shape[0],) + t.shape[nrows:]), [bp] + t.p[nrows:])
    return Ut, S, Vt, bp, err, spectra


def diag(S, bp):
    return Z2(np.diag(S), [bp, bp])


def unit_vector(parity=0):
    """Unit OBC leg tensor (``_unit_tensor``, src/states/tensors.jl:116-119)."""
    return np.ones(1), np.asarray([parity])
