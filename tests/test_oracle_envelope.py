"""CPU pins of the oracle's envelope pieces (no GPU).

  * point-triangle distance: exact rational brute force (plane projection
    against the three edges) on random, degenerate and point triangles;
  * sampleTriangle (src/tetwild/Common.cpp:143-228): invariants -- the three
    vertices are samples, every sample lies in the triangle, spacing d along
    the sampled lines, the longest-edge count formula;
  * envelope decisions: tree traversal == brute force over all facets, for any
    facet order; and == exact rational distances away from the eps^2 band;
  * nearest facet: squared distance == brute-force minimum.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as O
from tetwild_amd import workloads as W


def exact_d2(p, a, b, c):
    P = [Fr(float(v)) for v in p]
    A = [Fr(float(v)) for v in a]
    B = [Fr(float(v)) for v in b]
    C = [Fr(float(v)) for v in c]

    def sub(u, v): return [u[i] - v[i] for i in range(3)]
    def dot(u, v): return sum(u[i] * v[i] for i in range(3))

    def seg(u, v):
        w = sub(v, u)
        l2 = dot(w, w)
        if l2 == 0:
            return dot(sub(p_, u), sub(p_, u))
        t = dot(sub(p_, u), w) / l2
        t = min(max(t, Fr(0)), Fr(1))
        q = [u[i] + t * w[i] for i in range(3)]
        return dot(sub(q, p_), sub(q, p_))
    p_ = P
    best = min(seg(A, B), seg(A, C), seg(B, C))
    e0, e1 = sub(B, A), sub(C, A)
    a00, a01, a11 = dot(e0, e0), dot(e0, e1), dot(e1, e1)
    det = a00 * a11 - a01 * a01
    if det != 0:
        d = sub(A, P)
        b0, b1 = dot(d, e0), dot(d, e1)
        s = (a01 * b1 - a11 * b0) / det
        t = (a01 * b0 - a00 * b1) / det
        if s >= 0 and t >= 0 and s + t <= 1:
            q = [A[i] + s * e0[i] + t * e1[i] for i in range(3)]
            best = min(best, dot(sub(q, P), sub(q, P)))
    return best


def test_point_triangle_distance_exact():
    # Eberly's region method (as in geogram) treats a triangle as a segment
    # set only when det = |e0|^2 |e1|^2 - (e0.e1)^2 < 1e-30; a nearly (not
    # exactly) collinear facet is ill-conditioned for it in the reference too,
    # so degenerate cases here are exactly degenerate: integer coordinates.
    st = W.Stream(21, 0)
    for k in range(400):
        tri = st.normal(9).reshape(3, 3)
        if k % 5 == 0:
            tri = st.integers(9, -8, 8).reshape(3, 3).astype(float)
            tri[2] = tri[0] + 0.5 * (tri[1] - tri[0])          # exact segment
        if k % 17 == 0:
            tri[1] = tri[0].copy(); tri[2] = tri[0].copy()    # a point
        p = st.normal(3) * (0.3 if k % 2 else 2.0) + tri.mean(0)
        d2, cp = O.point_tri_d2(p, tri[0], tri[1], tri[2])
        ex = float(exact_d2(p, tri[0], tri[1], tri[2]))
        scale = np.sum((p - tri[0]) ** 2) + np.sum((tri[1] - tri[0]) ** 2) + \
            np.sum((tri[2] - tri[0]) ** 2)
        assert abs(d2 - ex) <= 1e-14 * scale, (k, d2, ex)
        # the returned closest point realises the distance
        assert abs(np.sum((cp - p) ** 2) - d2) <= 1e-13 * scale


def test_sample_triangle_invariants():
    st = W.Stream(22, 0)
    d = 1e-3
    for k in range(300):
        tri = st.normal(9).reshape(3, 3)
        tri = tri / np.linalg.norm(tri[1] - tri[0]) * d * st.uniform(1, 0.5, 25)[0]
        tri += st.normal(3)
        ps = O.sample_triangle(tri[0], tri[1], tri[2], d)
        # the vertices are samples, bit for bit
        for v in tri:
            assert np.any(np.all(ps == v, axis=1))
        # inside the triangle
        e0, e1 = tri[1] - tri[0], tri[2] - tri[0]
        M = np.stack([e0, e1], 1)
        st_, *_ = np.linalg.lstsq(M, (ps - tri[0]).T, rcond=None)
        resid = ps - tri[0] - (M @ st_).T
        scale = max(np.linalg.norm(e0), np.linalg.norm(e1))
        assert np.max(np.abs(resid)) < 1e-9 * scale
        assert st_.min() > -1e-9 and (st_.sum(0)).max() < 1 + 1e-9
        # longest-edge count: samples at n*d for n = 0..N' (N' = floor(N) or
        # N - 1 when N is an integer), then v1
        ls = [np.sum((tri[i] - tri[(i + 1) % 3]) ** 2) for i in range(3)]
        mi = max(i for i in range(3) if ls[i] == max(ls))
        N = np.sqrt(ls[mi]) / d
        if N <= 1:
            assert len(ps) == 3
            continue
        Np = N - 1 if N == int(N) else np.floor(N)
        first = ps[:int(Np) + 1]
        gaps = np.linalg.norm(np.diff(first, axis=0), axis=1)
        assert np.allclose(gaps, d, rtol=1e-9)
        assert np.array_equal(ps[int(Np) + 1], tri[(mi + 1) % 3])


def test_sample_triangle_small_and_integer_ratio():
    d = 0.25
    # longest edge exactly 4 d: N is an integer, so 4 points along it then v1
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, 0.1, 0]])
    ps = O.sample_triangle(tri[0], tri[1], tri[2], d)
    assert np.allclose(ps[:4, 0], [0, 0.25, 0.5, 0.75]) and np.array_equal(ps[4], tri[1])
    tiny = np.array([[0.0, 0, 0], [0.1, 0, 0], [0, 0.1, 0]])
    assert len(O.sample_triangle(*tiny, 1.0)) == 3


@pytest.fixture(scope="module")
def small_surface():
    V, F = W.icosphere_jittered(levels=3, sigma=0.002, seed=1)
    return V, F


def near_points(V, F, n, eps, seed):
    P, f, N = W.surface_points(V, F, n, seed=seed, tag=40)
    st = W.Stream(seed, 41)
    d = eps * (1 + st.uniform(n, -0.02, 0.02)) * np.where(st.uniform(n) < 0.5, 1, -1)
    Q = P + d[:, None] * N
    far = st.normal(3 * 50).reshape(50, 3) * 2
    return np.concatenate([Q, far, P[:50] * 0.5])


def test_tree_equals_brute_force(small_surface):
    V, F = small_surface
    eps = 0.02
    eps2 = eps * eps
    pts = near_points(V, F, 1500, eps, 3)
    S = O.Surface(V, F)
    tree = S.point_out(pts, eps2)
    brute = S.point_out_brute(pts, eps2)
    assert np.array_equal(tree, brute)
    assert 0.2 < tree.mean() < 0.8
    # facet order does not matter
    perm = np.argsort(W.Stream(9, 9).uniform(len(F)))
    S2 = O.Surface(V, F[perm])
    assert np.array_equal(S2.point_out(pts, eps2), tree)


def test_decisions_vs_exact_rationals():
    V, F = W.icosphere_jittered(levels=1, sigma=0.01, seed=2)
    eps = 0.05
    eps2 = eps * eps
    pts = near_points(V, F, 150, eps, 4)
    S = O.Surface(V, F)
    dec = S.point_out(pts, eps2)
    E2 = Fr(eps2)
    checked = 0
    for p, o in zip(pts, dec):
        m = min(exact_d2(p, V[f[0]], V[f[1]], V[f[2]]) for f in F)
        if abs(m - E2) <= Fr(1, 10**12) * E2:
            continue  # rounding band of eps_2: reported, not asserted
        assert o == (1 if m > E2 else 0)
        checked += 1
    assert checked > 140


def test_face_out_is_any_sample_out(small_surface):
    V, F = small_surface
    eps, d = 0.01, 0.01
    Q = W.query_triangles(V, F, 300, eps=eps, edge=0.03, seed=5)
    S = O.Surface(V, F)
    out, tot = S.face_out(Q, eps * eps, d)
    for i in range(300):
        ps = O.sample_triangle(Q[i, :3], Q[i, 3:6], Q[i, 6:], d)
        assert out[i] == int(S.point_out_brute(ps, eps * eps).any())
    assert 0.1 < out.mean() < 0.9 and tot > 300 * 3
    # degenerate (collinear) triangles are "in"
    deg = np.array([[0, 0, 0, 1, 1, 1, 2, 2, 2.0], [5, 5, 5, 5, 5, 5, 7, 7, 7.0]])
    assert np.all(S.face_out(deg, eps * eps, d)[0] == 0)


def test_nearest_is_minimum(small_surface):
    V, F = small_surface
    pts = near_points(V, F, 400, 0.05, 6)
    S = O.Surface(V, F)
    cp, d2, fid = S.nearest(pts)
    assert np.array_equal(d2, S.min_d2_brute(pts))
    # the closest point lies on the reported facet
    for i in range(0, 400, 37):
        f = F[fid[i]]
        dd, _ = O.point_tri_d2(pts[i], V[f[0]], V[f[1]], V[f[2]])
        assert dd == d2[i]
