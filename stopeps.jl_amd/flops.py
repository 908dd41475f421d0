"""Algorithmic work of the Born draw: the Z2 block-sparse GEMM flops of the
reference's contraction sequence, the yardstick of the roofline fraction.

One draw (``sample_configuration``, src/sampling/born.jl:86-163) runs, per row
y and column x: one ``_right_step`` (row environments,
src/contraction/boundary_mps.jl:240-246), one ``_left_step`` per basis
candidate (born.jl:120-129), the selected ``_left_step`` for x < Lx
(born.jl:148-150) and the site's ``_absorb`` product (boundary_mps.jl:136-151).
Each step is a chain of pairwise contractions; a contraction of Z2-neutral
tensors C[M,N] = A[M,K] B[K,N] splits into one GEMM per total parity pi of the
index groups, so it costs sum_pi |M_pi| |K_pi| |N_pi| multiply-adds, with
|G_pi| the number of index tuples of group G with parity pi. The compression
factorizations (the reference's LQ and SVD) and their absorb-back products
are not counted.

Legs are (even, odd) sector dimensions. A double-layer leg pairs a ket and a
bra copy of a virtual leg (e, o): (e^2 + o^2, 2 e o). Boundary bonds of
dimension c = min(chi, exact) are counted at the even split (ceil(c/2),
floor(c/2)) unless a per-bond split is given; the exact bond of cut b (after
column b) is the product of the pair legs left or right of it, whichever is
smaller.
"""


def pair(leg):
    """Double-layer (ket, bra) pair leg of a virtual leg (even, odd)."""
    e, o = leg
    return (e * e + o * o, 2 * e * o)


def fuse(*legs):
    """Sector dimensions of the product of legs."""
    e, o = 1, 0
    for (a, b) in legs:
        e, o = e * a + o * b, e * b + o * a
    return (e, o)


def dim(leg):
    return leg[0] + leg[1]


def macs(M, K, N):
    """Multiply-adds of C[M,N] = A[M,K] B[K,N] over neutral tensors; M, K, N
    are lists of legs."""
    m, k, n = fuse(*M), fuse(*K), fuse(*N)
    return m[0] * k[0] * n[0] + m[1] * k[1] * n[1]


def even_split(c):
    return ((c + 1) // 2, c // 2)


def _bond_dims(pairs_w, pairs_e, chi):
    """Boundary bonds c(0..Lx) of a row whose double layers have W legs
    pairs_w and E legs pairs_e."""
    Lx = len(pairs_w)
    bonds = [(1, 0)]
    for b in range(1, Lx):
        left = dim(fuse(*pairs_e[:b]))
        right = dim(fuse(*pairs_w[b:]))
        bonds.append(even_split(min(chi, left, right)))
    bonds.append((1, 0))
    return bonds


def draw_macs(layers, Lx, Ly, chi, candidates=2):
    """Multiply-adds of one Born draw. ``layers[(x, y)]`` = double-layer legs
    (w, n, e, s) as (even, odd) pairs (edge legs (1, 0))."""
    total = 0
    for y in range(Ly):
        w = [layers[(x, y)][0] for x in range(Lx)]
        e = [layers[(x, y)][2] for x in range(Lx)]
        top = _bond_dims(w, e, chi) if y > 0 else [(1, 0)] * (Lx + 1)
        bot = _bond_dims(w, e, chi) if y < Ly - 1 else [(1, 0)] * (Lx + 1)
        for x in range(Lx):
            lw, ln, le, ls = layers[(x, y)]
            tl, tr, bl, br = top[x], top[x + 1], bot[x], bot[x + 1]
            # _right_step: T[tl,n,tr] R[tr,e,br] -> X[tl,n,e,br]; X E over (n,e) -> Y[tl,br,w,s];
            # Y B over (s,br) -> R'[tl,w,bl]
            right = macs([tl, ln], [tr], [le, br]) + macs([tl, br], [ln, le], [lw, ls]) + \
                macs([tl, lw], [ls, br], [bl])
            # _left_step: L[tl,w,bl] T[tl,n,tr] -> X[w,bl,n,tr]; X E over (w,n) -> Y[bl,tr,e,s];
            # Y B over (bl,s) -> L'[tr,e,br]
            left = macs([lw, bl], [tl], [ln, tr]) + macs([bl, tr], [lw, ln], [le, ls]) + \
                macs([tr, le], [bl, ls], [br])
            # _absorb: A[l,n,r] E[w,n,e,s] over n
            absorb = macs([tl, tr], [ln], [lw, le, ls])
            total += right + candidates * left + (x < Lx - 1) * left + absorb
    return total


def lognorm_macs(layers, Lx, Ly, chi):
    """Multiply-adds of the absorb products of ``lognorm`` (Ly north absorbs)."""
    total = 0
    for y in range(Ly):
        w = [layers[(x, y)][0] for x in range(Lx)]
        e = [layers[(x, y)][2] for x in range(Lx)]
        top = _bond_dims(w, e, chi) if y > 0 else [(1, 0)] * (Lx + 1)
        for x in range(Lx):
            lw, ln, le, ls = layers[(x, y)]
            total += macs([top[x], top[x + 1]], [ln], [lw, le, ls])
    return total


def state_layers(state):
    """Double-layer legs (w, n, e, s) of every site of a ``FiniteState``."""
    out = {}
    for (x, y) in state.lattice.site_order():
        p = state[(x, y)].parity
        legs = [(int((q == 0).sum()), int((q == 1).sum())) for q in (p[0], p[1], p[4], p[5])]
        out[(x, y)] = tuple(pair(v) for v in legs)
    return out


def draw_flops(state, chi):
    """W_draw: 2 x multiply-adds of one Born draw from the PEPO ``state``."""
    L = state.lattice
    return 2 * draw_macs(state_layers(state), L.Lx, L.Ly, chi)


def lognorm_flops(state, chi):
    L = state.lattice
    return 2 * lognorm_macs(state_layers(state), L.Lx, L.Ly, chi)
