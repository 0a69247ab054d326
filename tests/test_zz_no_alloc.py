"""twg_smooth_batch_noalloc (the smoothing batch with its work-queue counter
in a file-scope device array, no device allocation) against twg_smooth_batch
through the C-ABI: two calls in a row, the second on a smaller batch, each
bit-exact with the allocating entry point on the same inputs (statuses,
positions, new qualities)."""
import ctypes

import numpy as np
import pytest
import torch

from tetwild_amd import _capi as C
from tetwild_amd import workloads as W
from tetwild_amd.ops import SmoothBatch

pytestmark = pytest.mark.gpu


def run_both(ctx, B, surf, tile, chunk):
    """the same batch through both entry points; returns both results"""
    outs = []
    for name in ("twg_smooth_batch", "twg_smooth_batch_noalloc"):
        batch = SmoothBatch(ctx, B, sf=surf, bt=surf, eps2=1e-6, sampling_dist=1e-3,
                            old_q=np.full(len(B["ring_tets"]), 1e3))
        batch.p_out.fill_(float("nan"))
        batch.status.fill_(255)
        batch.new_q.fill_(float("nan"))
        torch.cuda.synchronize()
        ctx._bind()
        print(f"{name}: n={batch.n} tile={tile} chunk={chunk}", flush=True)
        C.check(getattr(ctx.L, name)(ctx.h, ctypes.byref(batch.desc), C.TWG_FP64, tile, chunk),
                name)
        torch.cuda.synchronize()
        outs.append(tuple(t.cpu().numpy() for t in (batch.status, batch.p_out, batch.new_q)))
    return outs


def test_noalloc_matches_smooth_batch(ctx):
    V, F = W.bumpy_sphere(freq=40, seed=0)
    surf = ctx.surface(V, F)
    B = W.smoothing_batch(3000, seed=5, surface_frac=0.2, V=V, F=F)
    # second call: the first 700 problems (a smaller batch, same arrays)
    k = 700
    B2 = dict(B)
    B2["center"] = B["center"][:k]
    B2["kind"] = B["kind"][:k]
    B2["ring_off"] = B["ring_off"][:k + 1]
    B2["tri_off"] = B["tri_off"][:k + 1]
    for arrays, tile, chunk in ((B, 64, 4), (B2, 32, 1), (B, C.TWG_SMOOTH_TILE_SPLIT, 2),
                                (B2, C.TWG_SMOOTH_TILE_SPLIT, 1)):
        (s0, p0, q0), (s1, p1, q1) = run_both(ctx, arrays, surf, tile, chunk)
        assert (s0 != 255).all()
        assert np.array_equal(s0, s1)
        assert np.array_equal(p0.view(np.int64), p1.view(np.int64))
        assert np.array_equal(q0.view(np.int64), q1.view(np.int64))
        assert (s0 == 0).sum() > 0


def test_noalloc_rejects_bad_arguments(ctx):
    d = C.SmoothDesc()
    d.n = -1
    r = ctx.L.twg_smooth_batch_noalloc(ctx.h, ctypes.byref(d), C.TWG_FP64, 0, 0)
    assert r == C.TWG_EINVAL
    r = ctx.L.twg_smooth_batch_noalloc(ctx.h, ctypes.byref(d), C.TWG_FP64, 48, 0)
    assert r == C.TWG_EINVAL
