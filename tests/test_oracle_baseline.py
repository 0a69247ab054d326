"""The multi-core CPU baseline (oracle/baseline.c): the batch loops over host
threads give the sequential loops' results bit for bit, and the near-tie
accounting counts what it says it counts."""
import numpy as np

from oracle import oracle as O
from tetwild_amd import workloads as W


def test_amips_and_qualities_mt_match_sequential():
    x, y, z, tets, _ = W.kuhn_lattice(9, seed=2)
    for mode in (0, 1, 2):
        E, G, H = O.amips_batch(tets, x, y, z, mode)
        Em, Gm, Hm = O.amips_batch_mt(tets, x, y, z, mode, nthreads=4)
        assert np.array_equal(E.view(np.int64), Em.view(np.int64))
        if mode >= 1:
            assert np.array_equal(G.view(np.int64), Gm.view(np.int64))
        if mode >= 2:
            assert np.array_equal(H.view(np.int64), Hm.view(np.int64))
    q = O.tet_qualities(tets, x, y, z)
    qm = np.empty_like(q)
    O.lib().orc_tet_qualities_mt(O._p(np.ascontiguousarray(tets)), len(tets), O._p(x), O._p(y),
                                 O._p(z), O._p(qm), 3)
    assert np.array_equal(q.view(np.int64), qm.view(np.int64))


def test_envelope_mt_match_sequential():
    V, F = W.icosphere_jittered(levels=3, sigma=0.002, seed=1)
    S = O.Surface(V, F)
    Q = W.query_triangles(V, F, 3000, eps=0.01, edge=0.03, seed=5)
    fo, _ = S.face_out(Q, 1e-4, 0.01)
    assert np.array_equal(S.face_out_mt(Q, 1e-4, 0.01, nthreads=4), fo)
    pts = np.ascontiguousarray(Q[:, :3])
    assert np.array_equal(S.point_out_mt(pts, 1e-4, nthreads=4), S.point_out(pts, 1e-4))
    assert 0 < fo.sum() < len(fo)


def test_smoothing_and_collapse_mt_match_sequential():
    V, F = W.icosphere_jittered(levels=3, sigma=0.002, seed=1)
    B = W.smoothing_batch(300, seed=7, surface_frac=0.3, V=V, F=F)
    S = O.Surface(V, F)
    q = O.tet_qualities(B["ring_tets"], B["x"], B["y"], B["z"])
    eps = 0.01
    p, s, nq = O.smooth_batch(B, q, S, S, eps * eps, eps)
    pm, sm, nqm, risk = O.smooth_batch_mt(B, q, S, S, eps * eps, eps, nthreads=4)
    assert np.array_equal(s, sm)
    assert np.array_equal(p.view(np.int64), pm.view(np.int64))
    assert np.array_equal(nq.view(np.int64), nqm.view(np.int64))
    assert risk[0] > 0 and 0 <= risk[1] <= risk[0] and 0 < risk[2] and 0 <= risk[3] <= risk[2]

    M = W.sheet_mesh(5, seed=4)
    qc = O.tet_qualities(M["tets"], M["x"], M["y"], M["z"])
    C = W.collapse_batch(M, qc)
    sf, bt = O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"])
    e = 0.05 / 5
    for soft in (None, 40.0):
        st, nq = O.collapse_batch(C, sf, bt, e * e, e, soft_energy=soft)
        stm, nqm = O.collapse_batch_mt(C, sf, bt, e * e, e, soft_energy=soft, nthreads=4)
        assert np.array_equal(st, stm)
        assert np.array_equal(np.nan_to_num(nq).view(np.int64), np.nan_to_num(nqm).view(np.int64))


def test_face_risk_counts_band_samples():
    """a triangle parallel to a flat surface at height h has every sample at
    d2 = h^2 up to the rounding of the distance (a few ulps of the squared
    lengths involved, so the band here is 1e-9 eps2 wide): every sample is a
    band sample when eps2 = h^2, none when eps2 is far from it"""
    # flat square surface z = 0 (two facets)
    V = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=np.float64)
    F = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    S = O.Surface(V, F)
    h = 0.25  # h^2 = 1/16 exactly
    tri = np.array([[0, 0, h, 0.3, 0, h, 0, 0.3, h]], dtype=np.float64)
    d = 0.1
    ns = len(O.sample_triangle(tri[0, :3], tri[0, 3:6], tri[0, 6:], d))
    eps2 = h * h
    band = 1e-9 * eps2
    assert S.face_risk(tri, eps2, d, band, nthreads=2) == (ns, ns, 1, 1)
    # eps2 just below: still in the band, the decision "out" rests on it
    assert S.face_risk(tri, np.nextafter(eps2, 0), d, band, nthreads=2) == (ns, ns, 1, 1)
    # far from the band: samples counted, none at risk
    assert S.face_risk(tri, eps2 * 4, d, band, nthreads=2) == (ns, 0, 0, 0)
    assert S.face_risk(tri, eps2 / 4, d, band, nthreads=2) == (ns, 0, 0, 0)
    # a second, clearly-out triangle settles its own decision whatever the band says
    tri2 = np.concatenate([tri, tri + np.array([0, 0, 0, 0, 0, 0.2, 0, 0, 0])])
    c = S.face_risk(tri2, eps2, d, band, nthreads=2)
    assert c[2] == 2 and c[3] == 1


def test_point_risk_counts_band_points():
    V = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], dtype=np.float64)
    F = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    S = O.Surface(V, F)
    h = 0.25
    P = np.array([[0.1, 0.2, h], [0.3, -0.2, -h], [0.0, 0.0, 2 * h]])
    assert S.point_risk(P, h * h, 1e-9 * h * h, nthreads=2) == (3, 2)
    assert S.point_risk(P, 4 * h * h, 1e-9 * h * h, nthreads=2) == (3, 1)
    assert S.point_risk(P, 2 * h * h, 1e-9 * h * h, nthreads=2) == (3, 0)
