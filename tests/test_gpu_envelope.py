"""GPU parity of (b): sampleTriangle, isPointOutEnvelop, isFaceOutEnvelop and
nearest_facet through the C-ABI against the CPU oracle.

Bar: sample coordinates bit-exact; in/out decisions identical (FP64 and the
mixed FP32-box traversal); nearest facet, point and distance bit-exact when
the oracle's tree uses the same facet order."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tetwild_amd import workloads as W

pytestmark = pytest.mark.gpu


def dev_t(dev, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.fixture(scope="module")
def bumpy():
    V, F = W.bumpy_sphere(freq=40, seed=0)
    return V, F


def test_sample_triangles_bitexact(ctx, dev):
    st = W.Stream(30, 0)
    tris = []
    for k in range(3000):
        t = st.normal(9).reshape(3, 3)
        t = t / np.linalg.norm(t[1] - t[0]) * 1e-3 * st.uniform(1, 0.3, 30)[0]
        tris.append((t + st.normal(3)).reshape(9))
    tris.append(np.array([0, 0, 0, 1.0, 0, 0, 0.5, 0.1, 0]))  # integer ratio N
    tris = np.array(tris)
    counts, offs, pts = ctx.sample_triangles(dev_t(dev, tris), 1e-3)
    counts = counts.cpu().numpy()
    pts = pts.cpu().numpy()
    k = 0
    for i, t in enumerate(tris):
        d = 1e-3 if i < 3000 else 0.25
        if i == 3000:
            c2, o2, p2 = ctx.sample_triangles(dev_t(dev, t[None]), 0.25)
            ref = O.sample_triangle(t[:3], t[3:6], t[6:], 0.25)
            assert np.array_equal(p2.cpu().numpy(), ref)
            continue
        ref = O.sample_triangle(t[:3], t[3:6], t[6:], d)
        assert counts[i] == len(ref)
        assert np.array_equal(pts[k:k + counts[i]], ref), i
        k += counts[i]


@pytest.mark.parametrize("precision", [0, 1])
def test_envelope_points_match_oracle(ctx, dev, bumpy, precision):
    V, F = bumpy
    eps = 1e-3
    S = O.Surface(V, F)
    surf = ctx.surface(V, F)
    P = W.envelope_samples(V, F, 20000, eps=eps, band=0.01, seed=1)
    far = W.Stream(2, 0).normal(3000).reshape(1000, 3) * 1.5
    P = np.concatenate([P, far, V[:500] * 0.999])
    out, st = ctx.envelope_points(surf, dev_t(dev, P), eps * eps, precision=precision, stats=True)
    out = out.cpu().numpy()
    ref = S.point_out(P, eps * eps)
    assert np.array_equal(out, ref)
    assert 0.2 < ref.mean() < 0.8
    st = st.cpu().numpy()
    print(f"precision={precision} samples={st[0]} traversals={st[1]} boxes={st[2]} leaves={st[3]}")


@pytest.mark.parametrize("precision", [0, 1])
def test_envelope_faces_match_oracle(ctx, dev, bumpy, precision):
    V, F = bumpy
    eps = 1e-3
    S = O.Surface(V, F)
    surf = ctx.surface(V, F)
    Q = W.query_triangles(V, F, 15000, eps=eps, edge=0.004, seed=3)
    # degenerate queries: collinear and repeated vertices
    deg = np.stack([Q[:20, :3], Q[:20, :3] * 0.5 + Q[:20, 3:6] * 0.5, Q[:20, 3:6]], 1).reshape(20, 9)
    deg2 = np.concatenate([Q[20:30, :3], Q[20:30, :3], Q[20:30, 6:]], 1)
    Q = np.concatenate([Q, deg, deg2])
    out, st = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision,
                                 stats=True)
    out = out.cpu().numpy()
    ref, tot = S.face_out(Q, eps * eps, eps)
    assert np.array_equal(out, ref)
    assert 0.2 < ref.mean() < 0.8
    # without statistics a batch this small runs a wave per triangle
    out_w = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision)
    assert np.array_equal(out_w.cpu().numpy(), ref)
    st = st.cpu().numpy()
    print(f"precision={precision} out-rate={ref.mean():.3f} samples={st[0]} (all {tot}) "
          f"traversals={st[1]} boxes={st[2]} leaves={st[3]}")


def test_nearest_bitexact(ctx, dev, bumpy):
    V, F = bumpy
    surf = ctx.surface(V, F)
    order = surf.order()
    S = O.Surface(V, F[order])  # same facet order as the device tree
    P = W.envelope_samples(V, F, 5000, eps=3e-3, band=0.5, seed=7)
    P = np.concatenate([P, V[:300]])  # on vertices: ties between facets
    cp, d2, fid = ctx.nearest(surf, dev_t(dev, P))
    cpo, d2o, fido = S.nearest(P)
    assert np.array_equal(d2.cpu().numpy(), d2o)
    assert np.array_equal(cp.cpu().numpy(), cpo)
    assert np.array_equal(fid.cpu().numpy().astype(np.int64), fido)


def test_degenerate_surface_facets_and_tiny_surfaces(ctx, dev):
    # boundary-tree style facets (v0, v1, v1) and a single-facet surface
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]])
    F = np.array([[0, 1, 1], [1, 3, 3], [0, 2, 2]], dtype=np.int32)
    P = W.Stream(5, 0).uniform(3000, -0.2, 1.2).reshape(1000, 3)
    for FF in (F, F[:1]):
        surf = ctx.surface(V, FF)
        S = O.Surface(V, FF)
        out = ctx.envelope_points(surf, dev_t(dev, P), 0.01).cpu().numpy()
        assert np.array_equal(out, S.point_out(P, 0.01))
        empty = torch.empty((0, 3), dtype=torch.float64, device=dev)
        assert ctx.envelope_points(surf, empty, 0.01).numel() == 0


@pytest.mark.parametrize("nf", [255, 256, 257, 511, 512, 513, 1024, 1025])
def test_tree_top_in_lds_boundaries(ctx, dev, bumpy, nf):
    """surfaces whose trees end just below, at and above the levels the
    kernels copy into LDS (512 FP64 / 1024 FP32 node slots)"""
    V, F = bumpy
    F = F[:nf]
    eps = 1e-3
    S = O.Surface(V, F)
    surf = ctx.surface(V, F)
    P = W.envelope_samples(V, F, 3000, eps=eps, band=0.01, seed=nf)
    Q = W.query_triangles(V, F, 2000, eps=eps, edge=0.004, seed=nf)
    ref_p = S.point_out(P, eps * eps)
    ref_f, _ = S.face_out(Q, eps * eps, eps)
    for precision in (0, 1):
        out = ctx.envelope_points(surf, dev_t(dev, P), eps * eps, precision=precision)
        assert np.array_equal(out.cpu().numpy(), ref_p), precision
        out = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision)
        assert np.array_equal(out.cpu().numpy(), ref_f), precision
    assert 0.05 < ref_p.mean() < 0.95 and 0.05 < ref_f.mean() < 0.95


def test_bad_surface_rejected(ctx):
    V = np.zeros((3, 3))
    F = np.array([[0, 1, 5]], dtype=np.int32)
    with pytest.raises(Exception):
        ctx.surface(V, F)


def test_dense_faces_wave_lists(ctx, dev):
    """A Morton-ordered batch with >= 8 queries per facet runs faces_list_kernel
    (a wave's facets listed in LDS): decisions vs the oracle and vs the
    per-sample tree walk (the statistics path)"""
    V, F = W.bumpy_sphere(freq=16, seed=0)
    eps = 1e-3
    n = 70000
    assert n >= 65536 and n >= 8 * len(F)
    surf = ctx.surface(V, F)
    Q = W.query_triangles(V, F, n, eps=eps, edge=0.002, seed=21)
    ref, _ = O.Surface(V, F).face_out(Q, eps * eps, eps)
    assert 0.2 < ref.mean() < 0.8
    for precision in (0, 1):
        ctx.list_stats(True)
        out = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision).cpu().numpy()
        listed, walked = ctx.list_stats(False)
        walk, _ = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision, stats=True)
        assert np.array_equal(out, ref), precision
        assert np.array_equal(walk.cpu().numpy(), ref), precision
        # the list kernel ran, and most of its waves tested their lists
        print(f"precision {precision}: {listed} waves listed, {walked} walked", flush=True)
        assert listed + walked == (n + 255) // 256 * 4 and listed > 0.9 * (listed + walked)


def test_dense_faces_mixed_sizes(ctx, dev):
    """faces_list_kernel's wave walk with uneven waves: every 97th triangle
    exactly collinear (left to the exact pass: a lane with no triangle that
    still walks others' samples) and every 89th scaled five times about its
    centroid (~25x the samples of its neighbours, spread over the wave);
    decisions vs the oracle"""
    V, F = W.bumpy_sphere(freq=16, seed=0)
    eps = 1e-3
    n = 70000
    surf = ctx.surface(V, F)
    Q = W.query_triangles(V, F, n, eps=eps, edge=0.002, seed=23).reshape(n, 3, 3)
    c = Q.mean(axis=1, keepdims=True)
    Q[::89] = c[::89] + 5.0 * (Q[::89] - c[::89])
    Q[::97, 2] = 0.5 * (Q[::97, 0] + Q[::97, 1])
    Q = np.ascontiguousarray(Q.reshape(n, 9))
    ref, _ = O.Surface(V, F).face_out(Q, eps * eps, eps)
    assert 0.2 < ref.mean() < 0.8 and 0.0 < ref[::89].mean() < 1.0
    for precision in (0, 1):
        ctx.list_stats(True)
        out = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=precision).cpu().numpy()
        listed, walked = ctx.list_stats(False)
        print(f"precision {precision}: {listed} waves listed, {walked} walked", flush=True)
        assert listed > 0.9 * (listed + walked)
        assert np.array_equal(out, ref), (precision, np.nonzero(out != ref)[0][:10])


def test_dense_points_wave_lists(ctx, dev):
    """A Morton-ordered batch with >= 8 points per facet runs points_list_kernel
    (a wave's facets listed in LDS): decisions vs the oracle and vs the
    per-point tree walk (the statistics path); a few points with NaN
    coordinates send their waves to the tree walk"""
    V, F = W.bumpy_sphere(freq=16, seed=0)
    eps = 1e-3
    n = 70000
    assert n >= 65536 and n >= 8 * len(F)
    surf = ctx.surface(V, F)
    P = W.envelope_samples(V, F, n, eps=eps, band=0.01, seed=22)
    ref = O.Surface(V, F).point_out(P, eps * eps)
    assert 0.2 < ref.mean() < 0.8
    Pn = P.copy()
    Pn[::9973, 1] = np.nan
    for precision in (0, 1):
        for Q, nan in ((P, False), (Pn, True)):
            ctx.list_stats(True)
            out = ctx.envelope_points(surf, dev_t(dev, Q), eps * eps, precision=precision).cpu().numpy()
            listed, walked = ctx.list_stats(False)
            walk, _ = ctx.envelope_points(surf, dev_t(dev, Q), eps * eps, precision=precision, stats=True)
            assert np.array_equal(out, walk.cpu().numpy()), precision
            # most waves tested their lists; the NaN points' waves walked the tree
            print(f"precision {precision} nan {nan}: {listed} listed, {walked} walked", flush=True)
            assert listed + walked == (n + 255) // 256 * 4 and listed > 0.9 * (listed + walked)
            if nan:
                assert walked > 0
        assert np.array_equal(ctx.envelope_points(surf, dev_t(dev, P), eps * eps,
                                                  precision=precision).cpu().numpy(), ref), precision


def test_large_surface_faces(ctx, dev):
    """1M-facet bumpy sphere, 100k queries: decisions vs the oracle's tree"""
    V, F = W.bumpy_sphere(freq=224, seed=0)
    eps = 1e-3
    surf = ctx.surface(V, F)
    Q = W.query_triangles(V, F, 100000, eps=eps, edge=0.002, seed=11)
    out64 = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps).cpu().numpy()
    outm = ctx.envelope_faces(surf, dev_t(dev, Q), eps * eps, eps, precision=1).cpu().numpy()
    S = O.Surface(V, F)
    sel = np.arange(0, 100000, 10)
    ref, _ = S.face_out(Q[sel], eps * eps, eps)
    assert np.array_equal(out64[sel], ref)
    assert np.array_equal(outm, out64)
