"""Captured row sweeps of the batched draw (HIP graphs replayed on later chunks
and calls of a sampling cache): every replay gives the draws of the
kernel-by-kernel run bit for bit, with the program's own eager work
(other draws, a log norm) and synchronisations between the replays; several
chunks per call, evaluation mode, the physical kind, FP64 and FP32."""
import sys

import numpy as np
import pytest

from conftest import to_product_state, charged_peps
from oracle.model import Lattice

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b):
    for k in ("configuration", "log_probability", "conditionals", "max_environment_truncation"):
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_replays_equal_eager_draws(spp, precision):
    pepo = spp.synthetic_thermal_pepo(spp.Lattice(4, 4), D=(2, 2), eps=0.3, seed=9)
    uni = [np.stack([spp.sample_uniforms(3, 100 * r + i, 16) for i in range(1, 9)]) for r in range(4)]
    eager_ctx = spp.Context(0, precision, graphs=0)
    try:
        ecache = spp.SamplingCache(eager_ctx, pepo, 8)
        eager = [spp.sample_configurations(ecache, uniforms=u) for u in uni]
        eager_chunks = spp.sample_configurations(ecache, uniforms=uni[0][:7])
        eager_eval = spp.sample_configurations(ecache, configurations=eager[1]["configuration"])
    finally:
        eager_ctx.close()
    ctx = spp.Context(0, precision, graphs=1)
    other = spp.DeviceState(ctx, spp.random_peps(spp.Lattice(4, 4), (2, 2), seed=3))
    try:
        cache = spp.SamplingCache(ctx, pepo, 8)
        for rep in range(3):
            for r, u in enumerate(uni):
                got = spp.sample_configurations(cache, uniforms=u)
                _same(got, eager[r])
                # eager work between replays: a draw of another size and a log norm
                spp.sample_configurations(cache, uniforms=uni[2][:3])
                spp.lognorms(ctx, [other], 8)
                ctx.synchronize()
        cnt = ctx.counters()
        say(precision, "captures", cnt["graph_captures"], "replays", cnt["graph_replays"])
        assert cnt["graph_captures"] > 0 and cnt["graph_replays"] > cnt["graph_captures"]
        # several chunks per call (3 + 3 + 1), replayed across calls
        ctx.set_option("chunk", 3)
        for rep in range(3):
            _same(spp.sample_configurations(cache, uniforms=uni[0][:7]), eager_chunks)
        ctx.set_option("chunk", 0)
        for rep in range(3):
            _same(spp.sample_configurations(cache, configurations=eager[1]["configuration"]), eager_eval)
    finally:
        ctx.close()


def test_graph_replays_physical_kind(spp):
    L = Lattice(3, 3)
    st = to_product_state(spp, charged_peps(L, (2, 1), 4, [1, 0, 0, 1, 1, 0, 1, 0, 1], True))
    uni = np.stack([spp.sample_uniforms(5, i, 9) for i in range(1, 6)])
    res = {}
    for g in (0, 1):
        ctx = spp.Context(0, "fp64", graphs=g)
        try:
            cache = spp.SamplingCache(ctx, st, 4, kind="physical")
            runs = [spp.sample_configurations(cache, uniforms=uni) for _ in range(4)]
            res[g] = runs
            if g:
                assert ctx.counters()["graph_replays"] > 0
        finally:
            ctx.close()
    for r in res[1]:
        _same(r, res[0][0])
        assert np.array_equal(np.nan_to_num(r["repairs"], nan=-7.0), np.nan_to_num(res[0][0]["repairs"], nan=-7.0))
