"""Factorised double-layer products (csrc/kron.hip) on the GPU.

The kernel applies a double layer E = sum_phi S (x) S through its single
layer S. Its reference here is the dense contraction with the double layer
built from S in numpy (FP64), in the parity-sorted pair order and folded
layouts of the absorb's intermediates: stage 0 is
z2[l,R,F,w,p] = sum_{n,e} z1[l,R,F,n,e] E[w,n,e,p], stage 1 is
z3[l,R,w,W,N] = sum_{F,p} z2[l,R,F,w,p] E[W,N,F,p] (boundary_mps.jl:136-151
contracts the same double layer), stage 2, the draw's candidate left step,
Y[b,c,e,s] = sum_{w,n} X[w,b,c,n] E[w,n,e,s] (boundary_mps.jl:233-239), and
stage 3, the first transfer of an absorb in the rotated frame (the energy's
south sweep, boundary_mps.jl:193-209), z2[l,R,F,e,n] = sum_{s,w} z1[l,R,F,s,w]
E[w,n,e,s]; the rotated frame's second transfer is stage 2. Both kernels run every case: the vector one (the default) and the matrix-core
one (16x16x4 FP32 MFMA tiles). The draws through the factorised route are then compared with
the dense route and with the FP64 oracle."""
import ctypes

import numpy as np
import pytest

from conftest import to_product_state
from oracle import born as ob

pytestmark = pytest.mark.gpu


def pair_table(D, e):
    """Pair index of (ket i, bra j): ee, oo, eo, oe blocks, bra fastest."""
    o = D - e
    t = np.zeros((D, D), dtype=np.int64)
    for i in range(D):
        for j in range(D):
            pi, pj = int(i >= e), int(j >= e)
            ri, rj = i - e * pi, j - e * pj
            cnt = (e, o)
            base = {(0, 0): 0, (1, 1): e * e, (0, 1): e * e + o * o, (1, 0): e * e + o * o + e * o}[(pi, pj)]
            t[i, j] = base + ri * cnt[pj] + rj
    return t


def parities(dim, even):
    return (np.arange(dim) >= even).astype(np.int64)


def random_layer(rng, d, e, dphi, ephi, q):
    S = rng.normal(size=(d[0], d[1], dphi, d[2], d[3]))
    par = (parities(d[0], e[0])[:, None, None, None, None] + parities(d[1], e[1])[None, :, None, None, None]
           + parities(dphi, ephi)[None, None, :, None, None] + parities(d[2], e[2])[None, None, None, :, None]
           + parities(d[3], e[3])[None, None, None, None, :]) % 2
    S[par != q] = 0.0
    return S


def double_layer(S, d, e):
    """E[w, n, e, s] over pair legs."""
    t8 = np.einsum("anfcs,ANfCS->aAnNcCsS", S, S)
    P = [pair_table(d[k], e[k]) for k in range(4)]
    E = np.zeros(tuple(d[k] * d[k] for k in range(4)))
    idx = np.ix_(*[np.arange(d[k]) for k in range(4) for _ in range(2)])
    a, A, n, N, c, C, s, S_ = idx
    E[P[0][a, A], P[1][n, N], P[2][c, C], P[3][s, S_]] = t8
    return E


def pair_leg(D, e):
    return D * D, e * e + (D - e) * (D - e)


def neutral_random(rng, legs):
    """Dense random tensor over (dim, even) legs, zero unless the parities sum to 0."""
    x = rng.normal(size=[dm for dm, _ in legs])
    par = np.zeros(x.shape, dtype=np.int64)
    for k, (dm, ev) in enumerate(legs):
        shape = [1] * len(legs)
        shape[k] = dm
        par = par + parities(dm, ev).reshape(shape)
    x[par % 2 != 0] = 0.0
    return x


def fold_last(x, legs):
    """Folded storage: the last leg keeps the parity half its neighbours fix."""
    dm, ev = legs[-1]
    f = max(ev, dm - ev)
    rest = x.shape[:-1]
    par = np.zeros(rest, dtype=np.int64)
    for k, (d_, e_) in enumerate(legs[:-1]):
        shape = [1] * len(rest)
        shape[k] = d_
        par = par + parities(d_, e_).reshape(shape)
    par %= 2
    out = np.zeros(rest + (f,))
    valid = np.zeros(rest + (f,), dtype=bool)
    for p in (0, 1):
        cnt = ev if p == 0 else dm - ev
        m = par == p
        out[m, :cnt] = x[m][:, p * ev:p * ev + cnt]
        valid[m, :cnt] = True
    return out, valid


def run_kron(spp, ctx, S, d, e, dphi, ephi, stage, nl, el, nr, er, x, rows=0):
    lib = spp._lib.load()
    di = np.array(d, dtype=np.int32)
    ei = np.array(e, dtype=np.int32)
    W, N, Ep, Sp = [pair_leg(d[k], e[k]) for k in range(4)]
    fold = lambda leg: max(leg[1], leg[0] - leg[1])
    if stage == 0:
        ny = nl * nr * Ep[0] * W[0] * fold(Sp)
    elif stage == 3:
        ny = nl * nr * W[0] * Ep[0] * fold(N)
    elif stage == 1:
        ny = nl * nr * W[0] * W[0] * fold(N)
    else:
        ny = nl * nr * Ep[0] * fold(Sp)
    xf = np.ascontiguousarray(x, dtype=np.float32).ravel()
    y = np.zeros(ny, dtype=np.float32)
    Sd = np.ascontiguousarray(S, dtype=np.float64)
    rc = lib.spp_selftest_kron(ctx.handle, di.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                               ei.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), dphi, ephi,
                               Sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), stage, nl, el, nr, er,
                               xf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               y.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), rows)
    if rc != 0:
        raise AssertionError(f"spp_selftest_kron returned {rc}: {lib.spp_last_error(ctx.handle).decode()}")
    return y


CASES = [
    ((2, 2, 2, 2), (1, 1, 1, 1), 2, 1, 0),
    ((3, 2, 3, 2), (2, 1, 1, 1), 2, 1, 1),
    ((2, 3, 2, 3), (1, 2, 2, 1), 1, 1, 0),
    ((2, 2, 2, 2), (2, 2, 2, 2), 2, 2, 0),   # ungraded legs (dense model)
    ((4, 4, 4, 4), (2, 2, 2, 2), 2, 1, 1),
    ((6, 6, 6, 6), (3, 3, 3, 3), 2, 1, 0),   # 8x8 D=6 workload, auxiliary kind
    ((6, 6, 6, 6), (3, 3, 3, 3), 1, 1, 1),   # physical kind
    ((8, 8, 8, 8), (4, 4, 4, 4), 2, 1, 1),   # 12x12 D=8 workload (widest register tile)
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"d{c[0]}e{c[1]}phi{c[2]}q{c[4]}")
@pytest.mark.parametrize("stage", [0, 1, 2, 3])
@pytest.mark.parametrize("kernel", [1, 2], ids=["vector", "matrix_cores"])
def test_kron_kernel_matches_dense_double_layer(spp, gpu_ctx, case, stage, kernel):
    gpu_ctx.set_option("kron", kernel)
    try:
        check_kron_case(spp, gpu_ctx, case, stage)
    finally:
        gpu_ctx.set_option("kron", 1)


def check_kron_case(spp, gpu_ctx, case, stage):
    d, e, dphi, ephi, q = case
    rng = np.random.default_rng(7 + stage + 3 * sum(d))
    S = random_layer(rng, d, e, dphi, ephi, q)
    E = double_layer(S, d, e)
    W, N, Ep, Sp = [pair_leg(d[k], e[k]) for k in range(4)]
    nl, el, nr, er = (3, 2, 4, 1) if max(d) < 8 else (1, 1, 2, 1)
    lleg, rleg = (nl, el), (nr, er)
    if stage == 0:
        legs_in = [lleg, rleg, Ep, N, Ep]          # z1[l, R, F, n, e]
        z = neutral_random(rng, legs_in)
        ref = np.tensordot(z, E, axes=([3, 4], [1, 2]))          # [l, R, F, w, p]
        legs_out = [lleg, rleg, Ep, W, Sp]
    elif stage == 1:
        legs_in = [lleg, rleg, Ep, W, Sp]          # z2[l, R, F, w, p]
        z = neutral_random(rng, legs_in)
        ref = np.tensordot(z, E, axes=([2, 4], [2, 3]))          # [l, R, w, W, N]
        legs_out = [lleg, rleg, W, W, N]
    elif stage == 2:
        legs_in = [W, lleg, rleg, N]               # X[w, b, c, n] of a candidate left step
        z = neutral_random(rng, legs_in)
        ref = np.tensordot(z, E, axes=([0, 3], [0, 1]))          # [b, c, e, s]
        legs_out = [lleg, rleg, Ep, Sp]
    else:
        # the rotated frame's first transfer, frame legs (W', N', E', S') = (E, S, W, N):
        # z2[l, R, F, w', p'] = sum_{n', e'} z1[l, R, F, n', e'] E[w=e', n=p', e=w', s=n']
        legs_in = [lleg, rleg, W, Sp, W]           # z1[l, R, F, n', e']
        z = neutral_random(rng, legs_in)
        ref = np.tensordot(z, E, axes=([3, 4], [3, 0])).swapaxes(3, 4)   # [l, R, F, w', p']
        legs_out = [lleg, rleg, W, Ep, N]
    x, _ = fold_last(z, legs_in)
    want, valid = fold_last(ref, legs_out)
    scale = max(np.abs(want).max(), 1e-30)
    for rows in ((0, 1, 7, 16) if max(d) < 8 else (0, 3)):
        got = run_kron(spp, gpu_ctx, S, d, e, dphi, ephi, stage, nl, el, nr, er, x, rows).reshape(want.shape)
        err = np.abs(got - want)[valid].max()
        assert err <= 2e-6 * scale, (rows, err, scale)


@pytest.mark.parametrize("D,chi", [(2, 6), (3, 8)])
def test_kron_draws_match_dense_route(spp, D, chi):
    """FP32 draws through the factorised Gram transfers (vector and matrix-core
    kernels) against the dense double-layer route on the same uniforms: the
    same configurations (up to a uniform within FP32 rounding of a boundary)
    and log p to FP32 rounding."""
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(4, 4), D=(D - D // 2, D // 2), eps=0.3, seed=11)
    uni = np.stack([spp.sample_uniforms(3, i, 16) for i in range(1, 17)])
    out = {}
    for kron in (1, 2, 0):
        ctx = spp.Context(0, "fp32", kron=kron, graphs=0)
        try:
            cache = spp.SamplingCache(ctx, pepo, chi)
            ctx.reset_counters()
            out[kron] = (spp.sample_configurations(cache, uniforms=uni), ctx.counters()["kron_launches"])
        finally:
            ctx.close()
    assert out[1][1] > 0 and out[2][1] > 0 and out[0][1] == 0
    b = out[0][0]
    for kernel in (1, 2):
        a = out[kernel][0]
        same = [np.array_equal(a["configuration"][i], b["configuration"][i]) for i in range(len(uni))]
        assert sum(same) >= len(uni) - 1
        for i in range(len(uni)):
            if same[i]:
                assert abs(a["log_probability"][i] - b["log_probability"][i]) <= 1e-4 * max(1.0, abs(b["log_probability"][i]))
                np.testing.assert_allclose(a["conditionals"][i], b["conditionals"][i], atol=1e-4)


def test_kron_draws_match_oracle(spp):
    """The factorised route against the FP64 oracle on a 4x4 D=2 PEPO."""
    from oracle.model import Lattice, embed_identity_pepo
    L = Lattice(4, 4)
    pepo = embed_identity_pepo(L, (1, 1), 0.4, seed=5)
    state = to_product_state(spp, pepo)
    ctx = spp.Context(0, "fp32")
    try:
        cache = spp.SamplingCache(ctx, state, 6)
        uni = np.stack([spp.sample_uniforms(2, i, 16) for i in range(1, 9)])
        got = spp.sample_configurations(cache, uniforms=uni)
        assert ctx.counters()["kron_launches"] > 0
    finally:
        ctx.close()
    ocache = ob.SamplingCache(pepo, 6)
    agree = 0
    for i in range(len(uni)):
        ref = ob.sample_configuration(ocache, uniforms=uni[i])
        cfg = np.array([[ref["configuration"][x][y] for y in range(4)] for x in range(4)])
        if np.array_equal(cfg, got["configuration"][i]):
            agree += 1
            assert abs(ref["log_probability"] - got["log_probability"][i]) <= 1e-4 * max(1.0, abs(ref["log_probability"]))
    assert agree >= len(uni) - 1


@pytest.mark.parametrize("D,chi", [(3, 8)])
def test_kron_energies_match_dense_route(spp, D, chi):
    """FP32 energies of projected samples with the sweeps' absorbs (north, and
    south in the rotated frame) and the row and slab terms' plain-layer
    products through the single-layer tables (vector and matrix-core kernels)
    against the dense double layers: energies, log norms and the recorded
    boundary entropies to FP32 rounding."""
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(4, 4), D=(D - D // 2, D // 2), eps=0.3, seed=13)
    rng = np.random.default_rng(5)
    out = {}
    for kron in (1, 2, 0):
        ctx = spp.Context(0, "fp32", kron=kron, graphs=0)
        try:
            cache = spp.SamplingCache(ctx, pepo, chi)
            cfgs = np.asarray(cache.basis.labels)[rng.integers(0, 2, size=(6, 4, 4))] if kron == 1 else out[1][2]
            ctx.reset_counters()
            got = spp.measure_energies(cache, cfgs, spp.TFIM(1.0, 2.0, symmetric=True), chi=chi, record_entropy=True)
            out[kron] = (got, ctx.counters()["kron_launches"], cfgs)
        finally:
            ctx.close()
    assert out[1][1] > 0 and out[2][1] > 0 and out[0][1] == 0
    b = out[0][0]
    assert np.all(b["status"] == 0)
    for kernel in (1, 2):
        a = out[kernel][0]
        assert np.all(a["status"] == 0)
        for k in ("energy_total", "energy_per_site", "log_norm", "tensor_log_norm"):
            np.testing.assert_allclose(a[k], b[k], rtol=2e-5, atol=2e-5, err_msg=k)
        for i in range(len(a["Se"])):
            for side in ("north", "south"):
                np.testing.assert_allclose(a["Se"][i][side], b["Se"][i][side], atol=2e-4)
