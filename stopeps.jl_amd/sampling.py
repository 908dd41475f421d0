"""Born sampling API on the MI355X library, mirroring StoPEPS.jl
(src/sampling/born.jl, src/contraction/boundary_mps.jl).

``SamplingCache`` and ``sample_configuration`` keep the reference's names,
arguments and error messages; ``sample_configurations`` is the batched form
that replaces the per-sample loop of ``_initial_particles``
(src/driver/run.jl:45-66). All contraction work runs in the HIP library;
this module only marshals arrays.
"""
import contextlib
import ctypes

import numpy as np

from . import _lib
from .states import parity_sorted, resolve_basis, normalize_state, FiniteState, Site

MASK64 = (1 << 64) - 1
# Contraction arithmetic: FP64; FP32 on the FP32 matrix cores; FP32 operands
# and accumulation with the products formed from three BF16 pieces per operand
# (six piece products, FP32-equivalent) or two (three products, about 2^-16
# relative) on the BF16 matrix cores.
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x6": 2, "bf16x3": 3}


class BornError(ValueError):
    """Numerical failure of a Born draw (reference messages, born.jl:137-142)."""


class Context:
    """One device context (one HIP stream) of the library."""

    def __init__(self, device=0, precision="fp64", **options):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.spp_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise _lib.HIPLibraryError(f"spp_create failed ({rc}): no usable HIP device")
        self.handle = h
        self.set_precision(precision)
        for k, v in options.items():
            self.set_option(k, v)

    def set_precision(self, precision):
        self.precision = precision
        _lib.check(self.handle, self._lib.spp_set_precision(self.handle, PRECISIONS[precision]))

    def set_option(self, name, value):
        _lib.check(self.handle, self._lib.spp_set_option(self.handle, name.encode(), float(value)))

    def counters(self):
        out = np.zeros(23)
        rc = self._lib.spp_get_counters(self.handle, _lib.dptr(out), 23)
        if rc < 0:
            _lib.check(self.handle, rc)
        return {"gemm_flops": out[0], "gemm_launches": out[1], "subspace_iterations": out[2],
                "workspace_peak_bytes": out[3], "gemm_ms": out[4], "gemm_flops_profiled": out[5],
                "fp64": {"launches": out[6], "ms": out[7], "flops": out[8]},
                "fp32": {"launches": out[9], "ms": out[10], "flops": out[11]},
                "eig_calls": out[12], "eig_iterations": out[13], "eig_filter_gemms": out[14],
                "fu_repairs": out[15], "tuned_classes": out[16], "tuned_changed": out[17], "tune_ms": out[18],
                "graph_captures": out[19], "graph_replays": out[20], "eig_unconverged": out[21],
                "kron_launches": out[22]}

    def reset_counters(self):
        self._lib.spp_reset_counters(self.handle)

    @contextlib.contextmanager
    def boundary_space(self, space):
        """Run the enclosed calls with the boundary sector budget of
        ``BoundaryMPS(chi, space)`` (boundary_mps.jl:11-18): an (even, odd)
        pair for a Z2 state, an int for a dense one, or None."""
        caps = boundary_caps(space)
        self.set_option("boundary_cap_even", caps[0])
        self.set_option("boundary_cap_odd", caps[1])
        try:
            yield
        finally:
            self.set_option("boundary_cap_even", -1)
            self.set_option("boundary_cap_odd", -1)

    def synchronize(self):
        _lib.check(self.handle, self._lib.spp_synchronize(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self._lib.spp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def gemm(self, A, B, precision="fp64", evens=None, layout=0):
        """Self-test: C = A @ B through the library's matrix-core kernel. With
        ``evens = (eM, eN, eK)`` the indices are Z2 legs (first e entries even)
        and only parity-matching entries of C are computed (others NaN).
        ``layout`` bit 0 / bit 1 hand A / B to the device transposed in memory."""
        A = np.asarray(A, dtype=np.float64)
        B = np.asarray(B, dtype=np.float64)
        M, K = A.shape
        K2, N = B.shape
        if K != K2:
            raise ValueError("shape mismatch")
        As = np.ascontiguousarray(A.T if layout & 1 else A)
        Bs = np.ascontiguousarray(B.T if layout & 2 else B)
        C = np.zeros((M, N))
        eM, eN, eK = evens if evens is not None else (M, N, K)
        _lib.check(self.handle, self._lib.spp_selftest_gemm(self.handle, PRECISIONS[precision], M, N, K, eM, eN, eK,
                                                             int(layout), _lib.dptr(As), _lib.dptr(Bs),
                                                             _lib.dptr(C)))
        return C

    def copy_calibration(self, n):
        """One launch of the library's plain copy kernel over n FP32 values
        (8 n bytes of HBM traffic), for scaling hardware traffic counters."""
        _lib.check(self.handle, self._lib.spp_selftest_copy(self.handle, int(n)))
        return 8 * int(n)


class DeviceState:
    """A ``FiniteState`` uploaded to the device (``spp_state_upload``), or a
    state the device created (``adopt``)."""

    @classmethod
    def adopt(cls, ctx, handle, shell):
        """Wrap a device-created state handle; ``shell`` is a ``FiniteState``
        with the state's lattice, kind, beta, log amplitude and parity-sorted
        leg spaces (its arrays are zero-stride placeholders)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.state = shell
        self._arrays = [shell[(x, y)].data for (x, y) in shell.lattice.site_order()]
        self.handle = handle
        return self

    def __init__(self, ctx, state, live=None):
        self.ctx = ctx
        self.state = state
        L = state.lattice
        dims, evens, arrays = [], [], []
        for (x, y) in L.site_order():
            a, ev = parity_sorted(state[(x, y)])
            arrays.append(a)
            dims += list(a.shape)
            evens += ev
        self._arrays = arrays
        dims = np.asarray(dims, dtype=np.int32)
        evens = np.asarray(evens, dtype=np.int32)
        ptrs = (_lib.c_double_p * len(arrays))(*[_lib.dptr(a) for a in arrays])
        h = ctypes.c_void_p()
        kind = 0 if state.kind == "pepo" else 1
        _lib.check(ctx.handle, ctx._lib.spp_state_upload(ctx.handle, L.Lx, L.Ly, kind, _lib.iptr(dims),
                                                         _lib.iptr(evens), ptrs, state.beta,
                                                         state.log_amplitude, ctypes.byref(h)))
        self.handle = h
        if live is not None:
            lv = np.ascontiguousarray(live, dtype=np.int32).reshape(-1)
            _lib.check(ctx.handle, ctx._lib.spp_state_set_live(ctx.handle, h, _lib.iptr(lv)))

    def live(self):
        """Live (even, odd) counts per site and virtual leg W, N, E, S: [Lx*Ly, 4, 2]."""
        L = self.state.lattice
        lv = np.zeros(L.Lx * L.Ly * 8, dtype=np.int32)
        bufs = [np.empty_like(a) for a in self._arrays]
        ptrs = (_lib.c_double_p * len(bufs))(*[_lib.dptr(b) for b in bufs])
        _lib.check(self.ctx.handle, self.ctx._lib.spp_state_download(self.ctx.handle, self.handle, ptrs, None, None,
                                                                     _lib.iptr(lv)))
        return lv.reshape(L.Lx * L.Ly, 4, 2)

    def download(self):
        """The device state as a ``FiniteState`` in the leg order of the uploaded one."""
        L = self.state.lattice
        bufs = [np.empty_like(a) for a in self._arrays]
        ptrs = (_lib.c_double_p * len(bufs))(*[_lib.dptr(b) for b in bufs])
        beta = ctypes.c_double()
        la = ctypes.c_double()
        _lib.check(self.ctx.handle, self.ctx._lib.spp_state_download(self.ctx.handle, self.handle, ptrs,
                                                                     ctypes.byref(beta), ctypes.byref(la), None))
        sites = [[None] * L.Ly for _ in range(L.Lx)]
        for b, (x, y) in zip(bufs, L.site_order()):
            src = self.state[(x, y)]
            a = b
            for k, p in enumerate(src.parity):
                order = np.argsort(p, kind="stable")
                a = np.take(a, np.argsort(order), axis=k)
            sites[x][y] = Site(np.ascontiguousarray(a), src.parity)
        return FiniteState(L, sites, self.state.kind, beta=beta.value, log_amplitude=la.value)

    def __del__(self):
        try:
            if self.handle:
                self.ctx._lib.spp_state_free(self.handle)
        except Exception:
            pass


def boundary_caps(space):
    """(even, odd) budgets of a boundary sector space; -1 = none. A dense
    space n is one trivial sector, which the device holds as the even one."""
    if space is None:
        return (-1, -1)
    if isinstance(space, (tuple, list)):
        e, o = (int(v) for v in space)
        if min(e, o) < 0 or e + o <= 0:
            raise ValueError("a boundary sector space needs nonnegative budgets and a nonzero dimension")
        return (e, o)
    n = int(space)
    if n <= 0:
        raise ValueError("a boundary sector space needs a positive dimension")
    return (n, -1)


KINDS = {"auxiliary": 0, "physical": 1}


class SamplingCache:
    """``SamplingCache(state, BoundaryMPS(chi, sector_space); kind, basis)``
    (src/sampling/born.jl:45-73). ``kind="auxiliary"`` (the default) samples a
    PEPO's auxiliary indices; ``kind="physical"`` samples a PEPS's physical
    indices with the PSD-projected site metric (the resampling proposals,
    src/resample/event.jl:27-33). ``state`` is a host ``FiniteState`` or a
    resident ``DeviceState`` (e.g. a particle). The draws from the cache
    compress with the same boundary method."""

    def __init__(self, ctx, state, chi, basis=None, kind="auxiliary", sector_space=None):
        if kind not in KINDS:
            raise ValueError("unknown sampling kind")
        dev = state if isinstance(state, DeviceState) else None
        host = dev.state if dev is not None else state
        if kind == "auxiliary" and host.kind != "pepo":
            raise ValueError("the auxiliary-kind sampling cache needs a PEPO")
        if kind == "physical" and host.kind != "peps":
            raise ValueError("proposal requires PEPS")
        self.ctx = ctx
        self.state = host
        self.kind = kind
        self.chi = int(chi)
        self.sector_space = sector_space
        boundary_caps(sector_space)
        self.basis = resolve_basis(host, basis)
        self.lattice = host.lattice
        self._dev = dev if dev is not None else DeviceState(ctx, host)
        U = np.ascontiguousarray(self.basis.U, dtype=np.float64)
        h = ctypes.c_void_p()
        with ctx.boundary_space(sector_space):
            _lib.check(ctx.handle, ctx._lib.spp_sampling_cache_build_kind(ctx.handle, self._dev.handle, self.chi,
                                                                          _lib.dptr(U), KINDS[kind],
                                                                          ctypes.byref(h)))
        self.handle = h
        smt = np.zeros(self.lattice.Ly + 1)
        ctx._lib.spp_sampling_cache_info(self.handle, _lib.dptr(smt))
        self.south_max_truncation = smt

    def __del__(self):
        try:
            if self.handle:
                self.ctx._lib.spp_sampling_cache_free(self.handle)
        except Exception:
            pass


def splitmix64(v):
    """``_splitmix64`` (src/sampling/born.jl:1-6)."""
    v = (v + 0x9E3779B97F4A7C15) & MASK64
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK64
    return v ^ (v >> 31)


def sample_uniforms(seed, index, nsites):
    """Per-site uniforms of sample ``index`` (1-based) for ``seed``: the
    documented counter-based stream that stands in for ``sample_rng(seed, index)``
    (born.jl:13-22) where Julia's MersenneTwister is unavailable:
    u_k = (splitmix64(splitmix64(splitmix64(seed) ^ index) ^ k) >> 11) * 2^-53."""
    if seed < 0 or index < 1:
        raise ValueError("invalid random stream identity")
    base = splitmix64(splitmix64(seed & MASK64) ^ (index & MASK64))
    return np.array([(splitmix64(base ^ k) >> 11) * 2.0 ** -53 for k in range(nsites)])


def _raise_status(cache, status, fail):
    bad = np.nonzero(status)[0]
    if len(bad) == 0:
        return
    i = int(bad[0])
    site = int(fail[i])
    x, y = site % cache.lattice.Lx + 1, site // cache.lattice.Lx + 1
    if status[i] == 1:
        raise BornError(f"invalid Born conditional at ({x},{y}); increase chi_sampling")
    if status[i] == 2:
        raise BornError(f"Born conditional has zero total mass at ({x},{y})")
    if status[i] == 4:
        raise BornError("nonfinite norm metric")
    if status[i] == 5:
        raise BornError("sampling prefix violates the total symmetry sector")
    if status[i] == 6:
        raise BornError(f"site metric eigensolver did not converge at ({x},{y})")
    raise BornError("nonfinite row environment")


REPAIR_FIELDS = ("minimum_relative_eigenvalue", "negative_eigenvalues", "negative_multiplets",
                 "negative_spectral_fraction", "hermiticity_defect")


def repair_records(cache, rep):
    """The reference's ``repairs`` list of one physical-kind draw (born.jl:118-131)
    from its [Lx*Ly][6] array: one PSD report per site reached, in draw order,
    each followed by a ``zero_proposal_weights`` record where the proposal fell
    back to the uniform distribution."""
    out = []
    Lx = cache.lattice.Lx
    for i, r in enumerate(np.asarray(rep)):
        if np.isnan(r[0]):
            break
        site = (i % Lx + 1, i // Lx + 1)
        d = {k: float(v) for k, v in zip(REPAIR_FIELDS, r[:5])}
        d["negative_multiplets"] = int(r[2])
        d["site"] = site
        out.append(d)
        if r[5] > 0:
            out.append({"site": site, "reason": "zero_proposal_weights", "allowed_outcomes": int(r[5])})
    return out


def sample_configurations(cache, uniforms=None, configurations=None, check=True):
    """Batched ``sample_configuration`` over n samples.

    ``uniforms``: array [n, Lx*Ly] (row-major sites, x fastest). ``configurations``:
    evaluation mode, array [n, Lx, Ly] of basis labels. Returns a dict of arrays:
    configuration [n, Lx, Ly] (labels), log_probability [n], conditionals
    [n, Lx*Ly], max_environment_truncation [n], tensor_log_norm [n] and
    log_amplitude [n] of the projected normalised PEPS (at the sampling chi),
    status [n], fail_site [n].
    """
    L = cache.lattice
    ns = L.Lx * L.Ly
    if uniforms is None and configurations is None:
        raise ValueError("an explicit RNG is required")
    labels_in = None
    if configurations is not None:
        cfg = np.asarray(configurations)
        n = cfg.shape[0]
        if cfg.shape[1:] != (L.Lx, L.Ly):
            raise ValueError("configuration lattice")
        idx = np.vectorize(cache.basis.index)(cfg)
        labels_in = np.ascontiguousarray(np.transpose(idx, (0, 2, 1)).reshape(n, ns), dtype=np.int32)
        uni = None
    else:
        uni = np.ascontiguousarray(uniforms, dtype=np.float64)
        n = uni.shape[0]
        if uni.shape != (n, ns):
            raise ValueError("uniforms must have shape [n, Lx*Ly]")
    labels = np.zeros((n, ns), dtype=np.int32)
    logp = np.zeros(n)
    cond = np.zeros((n, ns))
    mt = np.zeros(n)
    tln = np.zeros(n)
    lamp = np.zeros(n)
    status = np.zeros(n, dtype=np.int32)
    fail = np.zeros(n, dtype=np.int32)
    rep = np.zeros((n, ns, 6)) if cache.kind == "physical" else None
    ctx = cache.ctx
    with ctx.boundary_space(cache.sector_space):
        _lib.check(ctx.handle, ctx._lib.spp_born_sample_batch_repairs(
            ctx.handle, cache.handle, n, _lib.dptr(uni), _lib.iptr(labels_in), _lib.iptr(labels), _lib.dptr(logp),
            _lib.dptr(cond), _lib.dptr(mt), _lib.dptr(tln), _lib.dptr(lamp), _lib.iptr(status), _lib.iptr(fail),
            _lib.dptr(rep)))
    if check:
        _raise_status(cache, status, fail)
    lab = np.asarray(cache.basis.labels)[labels].reshape(n, L.Ly, L.Lx).transpose(0, 2, 1)
    if labels_in is not None:
        lab = np.array(configurations)
    out = {"configuration": lab, "log_probability": logp, "conditionals": cond,
           "max_environment_truncation": mt, "tensor_log_norm": tln, "log_amplitude": lamp,
           "status": status, "fail_site": fail}
    if rep is not None:
        out["repairs"] = rep
    return out


def sample_configuration(cache, uniforms=None, configuration=None):
    """``sample_configuration(cache; rng | configuration)`` (born.jl:86-163) for
    one sample; ``uniforms`` replaces ``rng`` (one value per site, row-major)."""
    if configuration is not None:
        r = sample_configurations(cache, configurations=np.asarray(configuration)[None])
    else:
        if uniforms is None:
            raise ValueError("an explicit RNG is required")
        r = sample_configurations(cache, uniforms=np.asarray(uniforms)[None])
    lp = float(r["log_probability"][0])
    cond = list(r["conditionals"][0])
    if lp == -np.inf:
        # the draw stopped at its first zero conditional, or before any site
        # when a physical configuration missed the state's charge
        charge_rejected = "repairs" in r and np.isnan(r["repairs"][0][0][0])
        cond = [] if charge_rejected else cond[:int(np.argmin(np.asarray(cond) > 0)) + 1]
    repairs = repair_records(cache, r["repairs"][0]) if "repairs" in r else []
    return {"configuration": r["configuration"][0], "log_probability": lp, "conditionals": cond,
            "max_environment_truncation": float(r["max_environment_truncation"][0]), "repairs": repairs}


def project_configurations(ctx, pepo, configurations, basis=None, bond_space=None):
    """``project_configuration`` (born.jl:27-36) of n configurations of a
    resident PEPO (``DeviceState``) on the device (``spp_project_batch``):
    returns the normalised projected PEPS as resident ``DeviceState`` objects
    (parity-sorted legs, interior bonds in ``bond_space`` = (even, odd) when
    given, the PEPO's live sectors kept) and the normalisation logs [n, Lx*Ly]."""
    host = pepo.state
    L = host.lattice
    ns = L.Lx * L.Ly
    basis = resolve_basis(host, basis)
    cfg = np.asarray(configurations)
    n = cfg.shape[0]
    idx = np.vectorize(basis.index)(cfg) if n else np.zeros((0, L.Lx, L.Ly), dtype=np.int64)
    labels = np.ascontiguousarray(np.transpose(idx, (0, 2, 1)).reshape(n, ns), dtype=np.int32)
    U = np.ascontiguousarray(basis.U, dtype=np.float64)
    cap = None if bond_space is None else np.ascontiguousarray([int(v) for v in bond_space], dtype=np.int32)
    handles = (ctypes.c_void_p * max(n, 1))()
    lamp = np.zeros(n)
    norms = np.zeros((n, ns))
    _lib.check(ctx.handle, ctx._lib.spp_project_batch(ctx.handle, pepo.handle, n, _lib.iptr(labels), _lib.dptr(U),
                                                      _lib.iptr(cap), handles, _lib.dptr(lamp), _lib.dptr(norms)))
    out = []
    aux_graded = bool(np.any(host[(0, 0)].parity[3]))
    for s_ in range(n):
        sites = [[None] * L.Ly for _ in range(L.Lx)]
        for (x, y) in L.site_order():
            A = host[(x, y)]
            par = []
            for k in range(6):
                e = int(np.sum(A.parity[k] == 0))
                o = len(A.parity[k]) - e
                edge = (k == 0 and x == 0) or (k == 1 and y == 0) or (k == 4 and x == L.Lx - 1) or \
                    (k == 5 and y == L.Ly - 1)
                if k in (0, 1, 4, 5) and cap is not None and not edge:
                    e, o = int(cap[0]), int(cap[1])
                if k == 3:
                    q = basis.charges[labels[s_, x + y * L.Lx]] if aux_graded else 0
                    e, o = (0, 1) if q else (1, 0)
                par.append(np.array([0] * e + [1] * o))
            shape = tuple(len(p) for p in par)
            sites[x][y] = Site(np.broadcast_to(np.zeros(1), shape), par)
        shell = FiniteState(L, sites, "peps", beta=host.beta, log_amplitude=float(lamp[s_]))
        out.append(DeviceState.adopt(ctx, ctypes.c_void_p(handles[s_]), shell))
    return out, norms


def project_configuration(pepo, configuration, basis=None):
    """``project_configuration`` (born.jl:27-36): aux projection on the host
    (a slice per site) followed by ``normalize_state!`` (finite.jl:89-100)."""
    basis = resolve_basis(pepo, basis)
    L = pepo.lattice
    sites = [[None] * L.Ly for _ in range(L.Lx)]
    for x in range(L.Lx):
        for y in range(L.Ly):
            A = pepo[(x, y)]
            k = basis.index(configuration[x][y])
            a = np.tensordot(A.data, basis.U[:, k], axes=([3], [0]))[:, :, :, None, :, :]
            a = np.moveaxis(a, 3, 3)
            charge = basis.charges[k] if np.any(A.parity[3]) else 0
            par = list(A.parity)
            par[3] = np.array([charge])
            sites[x][y] = Site(np.ascontiguousarray(a), par)
    st = FiniteState(L, sites, "peps", beta=pepo.beta, log_amplitude=pepo.log_amplitude)
    return st, normalize_state(st)
