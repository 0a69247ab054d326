"""GPU parity at the benchmark sizes on the 1,003,520-facet bumpy sphere
(eps = sampling_dist = 1e-3), against the multi-threaded CPU oracle:

* envelope faces: 2M query triangles (edge 0.002), every decision, FP64 and
  mixed;
* envelope points: 2M samples in a +-1 % band around eps, every decision;
* smoothing: the 200k one-ring batch (20 % surface) in full -- status and
  final position bit-exact, new ring qualities bit-exact for moved surface
  vertices, FP64 and mixed, 64-lane tiles and the 64/32 split.

Bar: identical decisions and bit-identical positions / qualities."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tetwild_amd import _capi as C
from tetwild_amd import workloads as W
from tetwild_amd.ops import SmoothBatch

pytestmark = pytest.mark.gpu

EPS = 1e-3


@pytest.fixture(scope="module")
def sphere(ctx):
    V, F = W.bumpy_sphere(freq=224, seed=0)
    surf = ctx.surface(V, F)
    # the oracle's tree over the facets in the order the device tree holds
    S = O.Surface(V, F[surf.order()])
    yield V, F, surf, S
    surf.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def test_fullsize_faces(ctx, dev, sphere):
    V, F, surf, S = sphere
    n = 2_000_000
    Q = W.query_triangles(V, F, n, eps=EPS, edge=0.002, seed=0)
    tq = torch.from_numpy(Q).to(dev)
    ref = S.face_out_mt(Q, EPS * EPS, EPS)
    for prec in (0, 1):
        print(f"faces n={n} precision={prec}", flush=True)
        out = ctx.envelope_faces(surf, tq, EPS * EPS, EPS, precision=prec).cpu().numpy()
        mism = np.nonzero(out != ref)[0]
        assert mism.size == 0, f"precision {prec}: {mism.size} mismatches, first {mism[:10]}"
    assert 0.2 < ref.mean() < 0.8


def test_fullsize_points(ctx, dev, sphere):
    V, F, surf, S = sphere
    n = 2_000_000
    P = W.envelope_samples(V, F, n, eps=EPS, band=0.01, seed=0)
    tp = torch.from_numpy(P).to(dev)
    ref = S.point_out_mt(P, EPS * EPS)
    for prec in (0, 1):
        print(f"points n={n} precision={prec}", flush=True)
        out = ctx.envelope_points(surf, tp, EPS * EPS, precision=prec).cpu().numpy()
        mism = np.nonzero(out != ref)[0]
        assert mism.size == 0, f"precision {prec}: {mism.size} mismatches, first {mism[:10]}"
    assert 0.2 < ref.mean() < 0.8


def test_fullsize_smoothing(ctx, dev, sphere):
    V, F, surf, S = sphere
    n = 200_000
    B = W.smoothing_batch(n, seed=0, surface_frac=0.2, V=V, F=F)
    q, _ = ctx.tet_qualities(*(torch.from_numpy(a).to(dev) for a in
                               (B["ring_tets"], B["x"], B["y"], B["z"])))
    old_q = q.cpu().numpy()
    rp, rs, rq, _ = O.smooth_batch_mt(B, old_q, S, S, EPS * EPS, EPS)
    batch = SmoothBatch(ctx, B, sf=surf, bt=surf, eps2=EPS * EPS, sampling_dist=EPS, old_q=old_q)
    moved_surf = np.nonzero((rs == 0) & (B["kind"] > 0))[0]
    assert (B["kind"] > 0).mean() > 0.15 and moved_surf.size > 0.1 * n
    for prec, tile in ((0, 64), (1, 64), (0, C.TWG_SMOOTH_TILE_SPLIT)):
        print(f"smoothing n={n} precision={prec} tile={tile}", flush=True)
        batch.status.fill_(255)
        batch.run(precision=prec, tile=tile, chunk=1)
        torch.cuda.synchronize()
        p, s, nq = (t.cpu().numpy() for t in (batch.p_out, batch.status, batch.new_q))
        mism = np.nonzero(s != rs)[0]
        assert mism.size == 0, f"precision {prec}: status mismatch at {mism[:10]}"
        assert np.array_equal(bits(p), bits(rp)), f"precision {prec}: positions differ"
        ro = B["ring_off"]
        sel = np.concatenate([np.arange(ro[i], ro[i + 1]) for i in moved_surf])
        assert np.array_equal(bits(nq[sel]), bits(rq[sel])), f"precision {prec}: qualities differ"
