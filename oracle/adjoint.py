"""Gradient of the log norm with respect to every site tensor (oracle
restatement; test infrastructure).

The reference has no adjoint. The object it builds for one site is
``_site_hole`` (src/sampling/born.jl:75-83): the single-site norm matrix N_i
from the boundary environments and the row messages around site i, so that
<psi|psi> = sum A_i[ket] N_i[ket, bra] A_i[bra] over the virtual legs (the
physical and auxiliary legs are traced). For real tensors

    d log <psi|psi> / d A_i = (N_i A_i + N_i^T A_i) / (A_i N_i A_i),

exact when chi is at least the exact boundary bond (every environment exact);
with truncation it is the environment gradient, not the derivative of the
truncation itself.
"""
import numpy as np

from .z2 import Z2, ein, split
from . import boundary as bm


def site_hole(L, T, B, R):
    """``_site_hole`` before its split (born.jl:76-78): H[w,n,e,s] over the
    fused (ket, bra) pair legs."""
    X = ein("awb,anc->wbnc", L, T)
    Y = ein("bsd,ced->bsce", B, R)
    return ein("wbnc,bsce->wnes", X, Y)


def split_hole(H, A):
    """Split H's pair legs into (ket, bra) with A's spaces (the _split loop of born.jl:79-81)."""
    out = H
    for k, axis in reversed(list(enumerate((0, 1, 4, 5)))):
        d = A.shape[axis]
        p = A.p[axis]
        out = split(out, k, (d, d), (p, p))
    return out  # [wk, wb, nk, nb, ek, eb, sk, sb]


def site_gradient(H, A):
    """(N A + N^T A) / (A N A) for one site."""
    H8 = split_hole(H, A)
    G1 = ein("WwNnEeSs,wnpaes->WNpaES", H8, A)   # bra contracted: ket-shaped
    G2 = ein("wWnNeEsS,wnpaes->WNpaES", H8, A)   # ket contracted: bra-shaped
    Z = float(np.sum(A.a * G1.a))
    return Z2((G1.a + G2.a) / Z, A.p), Z


def lognorm_gradient(state, chi):
    """Environment gradient of log <psi|psi> for every site: dict (x, y) -> Z2
    shaped like the site, plus the log norm."""
    Lx, Ly = state.lattice.Lx, state.lattice.Ly
    layers = bm.double_layers(state)
    north, south, _ = bm.boundary_environments(layers, chi)
    grads = {}
    for y in range(Ly):
        rowl = bm.row(layers, y)
        top = north[y].tensors
        bottom = south[y + 1].tensors
        left, right = bm.row_environments(top, rowl, bottom)
        for x in range(Lx):
            H = site_hole(left[x], top[x], bottom[x], right[x + 1])
            grads[(x, y)], _ = site_gradient(H, state[(x, y)])
    return grads, bm.lognorm(state, chi)
