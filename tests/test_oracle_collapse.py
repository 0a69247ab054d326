"""Pins the oracle's restatement of collapseAnEdge's checks
(src/tetwild/EdgeCollapser.cpp:284-446, 727-775) with constructed candidates
whose outcome geometry decides: an octahedral one-ring around v1 at the
origin, flat surfaces, and targets chosen inside, outside or behind it.

Each case pins one branch of the control flow and the order of the checks
(point envelope before isFlip before quality before the epsilon faces)."""
import numpy as np
import pytest

from oracle import oracle as O

SUCCESS, FLIP, QUALITY, ENVELOP, ENVELOP_SUC = 0, 1, 2, 3, 4
V1S, V1B, V2S, V2B, V1U = 1, 2, 4, 8, 16

# vertex 0: v1 at the origin; 1..6: +x -x +y -y +z -z
OCT = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1],
                [0, 0, -1]], dtype=np.float64)


def octa_ring():
    tets = []
    for sx in (1, 2):
        for sy in (3, 4):
            for sz in (5, 6):
                t = [0, sx, sy, sz]
                P = OCT[t]
                if O.orient3d(*P) < 0:
                    t = [0, sx, sz, sy]
                tets.append(t)
    return np.array(tets, dtype=np.int32)


RING = octa_ring()
BIG = np.array([[-10, -10, 0], [10, -10, 0], [10, 10, 0], [-10, 10, 0]], dtype=np.float64)
PLANE = O.Surface(BIG, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32))  # z = 0
LINE = O.Surface(np.array([[-10, 0, 0], [10, 0, 0]], dtype=np.float64),
                 np.array([[0, 1, 1]], dtype=np.int32))                    # the x axis


def run(v2pos=None, v2=None, flags=0, old_q=1e50, tris=(), eps=0.01, check_quality=True,
        check_eps=True, soft=None, extra=()):
    P = np.concatenate([OCT, np.array(extra, dtype=np.float64).reshape(-1, 3)])
    if v2pos is not None:
        P = np.concatenate([P, np.array([v2pos], dtype=np.float64)])
        v2 = len(P) - 1
    B = {"x": P[:, 0].copy(), "y": P[:, 1].copy(), "z": P[:, 2].copy(),
         "v1": np.array([0], np.int32), "v2": np.array([v2], np.int32),
         "flags": np.array([flags], np.uint8),
         "ring_off": np.array([0, len(RING)], np.int32), "ring_tets": RING,
         "old_q": np.full(len(RING), float(old_q)),
         "tri_off": np.array([0, len(tris)], np.int32),
         "tri_ids": np.array(tris, dtype=np.int32).reshape(-1, 3)}
    s, nq = O.collapse_batch(B, PLANE, LINE, eps * eps, eps, check_quality=check_quality,
                             check_eps=check_eps, soft_energy=soft)
    return int(s[0]), nq, B, v2


def test_convex_ring_to_a_ring_vertex():
    """collapsing the centre onto +x: the four tets around +x vanish, the four
    around -x are re-coned from +x -- convex, so no flip; their qualities are
    calTetQualities of the explicit new tets"""
    s, nq, B, v2 = run(v2=1)
    assert s == SUCCESS
    removed = (RING == 1).any(1)
    assert removed.sum() == 4 and np.isnan(nq[removed]).all()
    new = np.where(RING == 0, 1, RING)[~removed]
    q = O.tet_qualities(new, B["x"], B["y"], B["z"])
    assert np.array_equal(nq[~removed], q)
    assert (q > 3.0).all() and (q < 1e50).all()


def test_target_behind_the_ring_flips():
    """v2 at (-5, 0, 0): the faces through -x see it from outside"""
    s, _, _, _ = run(v2pos=(-5.0, 0.0, 0.0))
    assert s == FLIP


@pytest.mark.parametrize("old_q,soft,flags,cq,expect", [
    (0.0, None, 0, True, QUALITY),        # any new energy > 0 = old max
    (1e50, None, 0, True, SUCCESS),
    (0.0, 1e50, 0, True, SUCCESS),        # is_soft replaces the old max
    (1e50, 0.0, 0, True, QUALITY),
    (0.0, None, V1U, True, SUCCESS),      # unrounded v1: new max counts as 0
    (0.0, None, 0, False, SUCCESS),       # quality check off
])
def test_quality_rule(old_q, soft, flags, cq, expect):
    s, nq, _, _ = run(v2pos=(0.1, 0.0, 0.0), old_q=old_q, soft=soft, flags=flags,
                      check_quality=cq)
    assert s == expect
    assert np.isfinite(nq).all()


def test_degenerate_edge_skips_quality_and_envelope():
    """v2 at v1's position: quality and epsilon checks are skipped, but a
    surface v1 still reports ENVELOP_SUC"""
    s, _, _, _ = run(v2pos=(0.0, 0.0, 0.0), old_q=0.0)
    assert s == SUCCESS
    # triangle standing out of the plane z = 0, v1 on the surface
    s, _, _, v2 = run(v2pos=(0.0, 0.0, 0.0), old_q=0.0, flags=V1S | V2S, tris=[(7, 3, 5)])
    assert s == ENVELOP_SUC


def test_point_envelope_precedes_flip():
    # v2 behind the ring but on the plane: the point check passes, then FLIP
    assert run(v2pos=(-5.0, 0.0, 0.0), flags=V1S)[0] == FLIP
    # lifted off the plane: ENVELOP before the flip test
    assert run(v2pos=(-5.0, 0.0, 3.0), flags=V1S)[0] == ENVELOP
    # a surface v2 is not point-checked
    assert run(v2pos=(-5.0, 0.0, 3.0), flags=V1S | V2S)[0] == FLIP
    # v2 inside the ring, 0.5 above the plane
    assert run(v2pos=(0.0, 0.0, 0.5), flags=V1S)[0] == ENVELOP
    assert run(v2pos=(0.0, 0.0, 0.005), flags=V1S, eps=0.01)[0] == ENVELOP_SUC


def test_boundary_point_envelope():
    assert run(v2pos=(0.0, 0.5, 0.0), flags=V1B)[0] == ENVELOP
    assert run(v2pos=(0.3, 0.0, 0.0), flags=V1B)[0] == SUCCESS
    assert run(v2pos=(0.0, 0.5, 0.0), flags=V1B | V2B)[0] == SUCCESS
    # boundary check comes before the surface one; both in -> continue
    assert run(v2pos=(0.3, 0.0, 0.0), flags=V1B | V1S)[0] == ENVELOP_SUC


def test_epsilon_faces():
    # flat triangle (v2, +y, -y) on z = 0: in
    assert run(v2=1, flags=V1S | V2S, tris=[(1, 3, 4)])[0] == ENVELOP_SUC
    # (v2, +y, +z) rises to z = 1: out
    assert run(v2=1, flags=V1S | V2S, tris=[(1, 3, 4), (1, 3, 5)])[0] == ENVELOP
    # epsilon check off
    assert run(v2=1, flags=V1S | V2S, tris=[(1, 3, 5)], check_eps=False)[0] == SUCCESS
    # quality precedes the faces
    assert run(v2=1, flags=V1S | V2S, tris=[(1, 3, 5)], old_q=0.0)[0] == QUALITY
    # flip precedes the faces
    assert run(v2pos=(-5.0, 0.0, 0.0), flags=V1S | V2S, tris=[(7, 3, 5)])[0] == FLIP
    # a collinear triangle is never out (isFaceOutEnvelop on a degenerate face)
    s = run(v2=1, flags=V1S | V2S, tris=[(1, 7, 8)], extra=[(3.0, 0, 5.0), (5.0, 0, 10.0)])[0]
    assert s == ENVELOP_SUC
    # no triangles at all (v1 on the rim of an open surface)
    assert run(v2=1, flags=V1S | V2S)[0] == ENVELOP_SUC
