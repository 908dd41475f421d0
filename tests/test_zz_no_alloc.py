"""The no-allocation variant of the batched lognorm (spp_lognorm_batch_noalloc)
against the allocating entry point (spp_lognorm_batch), both through the C ABI:
two calls in a row (the second on a smaller lattice and batch), results equal
to 1e-12 relative, and a state too large for the static scratch rejected with
a message before any launch."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _lognorm(spp, ctx, name, devs, chi):
    n = len(devs)
    handles = (ctypes.c_void_p * n)(*[d.handle.value for d in devs])
    out = {k: np.full(n, np.nan) for k in ("log_norm", "tensor_log_norm", "max_trunc")}
    status = np.full(n, -1, dtype=np.int32)
    rc = getattr(ctx._lib, name)(ctx.handle, n, handles, int(chi), spp._lib.dptr(out["log_norm"]),
                                 spp._lib.dptr(out["tensor_log_norm"]), spp._lib.dptr(out["max_trunc"]),
                                 spp._lib.iptr(status))
    out["status"] = status
    return rc, out


def _states(spp, ctx, L, D, count, seed):
    return [spp.DeviceState(ctx, spp.random_peps(spp.Lattice(L, L), D, seed=seed + k)) for k in range(count)]


def test_noalloc_lognorm_matches_allocating_entry_point(spp):
    ctx = spp.Context(0, "fp64")
    try:
        for (L, D, count, chi, seed) in ((4, (2, 1), 12, 12, 40), (3, (1, 1), 3, 6, 60)):
            say(f"lognorm without allocation: {count} states {L}x{L} D={D} chi={chi}")
            devs = _states(spp, ctx, L, D, count, seed)
            rc, ref = _lognorm(spp, ctx, "spp_lognorm_batch", devs, chi)
            assert rc == 0, ctx._lib.spp_last_error(ctx.handle)
            rc, got = _lognorm(spp, ctx, "spp_lognorm_batch_noalloc", devs, chi)
            assert rc == 0, ctx._lib.spp_last_error(ctx.handle)
            assert np.array_equal(got["status"], ref["status"]) and np.all(ref["status"] == 0)
            for k in ("log_norm", "tensor_log_norm"):
                np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=0)
            np.testing.assert_allclose(got["max_trunc"], ref["max_trunc"], rtol=1e-9, atol=1e-15)
        say("a state beyond the static scratch is rejected")
        big = _states(spp, ctx, 8, (3, 3), 1, 7)
        rc, got = _lognorm(spp, ctx, "spp_lognorm_batch_noalloc", big, 32)
        assert rc != 0
        msg = ctx._lib.spp_last_error(ctx.handle)
        msg = msg.decode() if isinstance(msg, bytes) else str(msg)
        assert "static scratch" in msg, msg
        assert np.all(got["status"] == -1)  # nothing was written
    finally:
        ctx.close()
