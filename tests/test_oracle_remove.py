"""Pins the oracle's EdgeRemover decision (src/tetwild/EdgeRemover.cpp:97-617)
with rings of 3, 4 and 5 tets around a vertical edge v1-v2 built from
geometry.  The new tets are checked against the re-triangulation constructed
independently here (the shared face for 3-2, a diagonal of the square for
4-4, a fan of the pentagon for 5-6, each coned to v1 and v2); the decision
against the quality rule on those independently built tets."""
import numpy as np
import pytest

from oracle import oracle as O

V1, V2 = 0, 1


def ring_mesh(k, r, h=1.0, twist=0.0, aspect=1.0):
    """v1 = (0, 0, -h), v2 = (0, 0, h), k points on an ellipse (radius r,
    axes stretched by aspect); tets (v1, v2, p_i, p_i+1), each positively
    oriented"""
    ang = 2 * np.pi * np.arange(k) / k + twist + 0.05 * np.sin(np.arange(k) * 1.7)
    P = np.concatenate([[[0, 0, -h], [0, 0, h]],
                        np.stack([r * aspect * np.cos(ang), r / aspect * np.sin(ang),
                                  0.03 * np.cos(3 * ang)], 1)])
    tets = []
    for i in range(k):
        t = [V1, V2, 2 + i, 2 + (i + 1) % k]
        if O.orient3d(*P[t]) < 0:
            t = [V1, V2, 2 + (i + 1) % k, 2 + i]
        tets.append(t)
    return P, np.array(tets, np.int32)


def oriented(P, vs):
    t = list(vs)
    if O.orient3d(*P[t]) < 0:
        t[2], t[3] = t[3], t[2]
    return t


def quality(P, ts):
    return O.tet_qualities(np.array(ts, np.int32), *(np.ascontiguousarray(P[:, c]) for c in range(3)))


def xyz(P):
    return [np.ascontiguousarray(P[:, c]) for c in range(3)]


@pytest.mark.parametrize("r", [0.2, 0.6, 1.5, 4.0])
def test_32_face_and_rule(r):
    P, ring = ring_mesh(3, r)
    old = quality(P, ring)
    new = [oriented(P, [2, 3, 4, V1]), oriented(P, [2, 3, 4, V2])]
    qn = quality(P, new)
    code, nt, nq, eb = O.remove_edge(V1, V2, ring, old, *xyz(P))
    expect = 3 if (O.tet_flip(np.array(new, np.int32), *xyz(P)).sum() == 0 and qn.max() < old.max()) else 0
    assert code == expect
    if code:
        assert sorted(map(sorted, nt.tolist())) == sorted(map(sorted, new))
        # bits in the oracle's vertex order; the independent vertex order is an
        # even permutation, equal up to rounding
        assert np.array_equal(nq, quality(P, nt))
        assert np.allclose(np.sort(nq), np.sort(qn), rtol=1e-12, atol=0)
    assert eb == 0


def test_32_equal_buget():
    P, ring = ring_mesh(3, 0.6)
    old = quality(P, ring)
    code0 = O.remove_edge(V1, V2, ring, old, *xyz(P))[0]
    # with stored qualities equal to the new max, only equal_buget lets it through
    code, nt, nq, _ = O.remove_edge(V1, V2, ring, np.full(3, 1e50), *xyz(P))
    assert code == 3
    same = np.full(3, nq.max())
    assert O.remove_edge(V1, V2, ring, same, *xyz(P), equal_buget=0)[0] == 0
    code, _, _, eb = O.remove_edge(V1, V2, ring, same, *xyz(P), equal_buget=2)
    assert code == 3 and eb == 1
    assert code0 in (0, 3)


@pytest.mark.parametrize("h,aspect,r,twist", [(1.0, 1.0, 0.5, 0.0), (1.0, 1.5, 0.3, 0.0),
                                               (3.0, 2.5, 1.0, 0.4), (3.0, 4.0, 0.3, 0.0),
                                               (0.3, 1.0, 1.0, 0.3)])
def test_44_diagonals(h, aspect, r, twist):
    P, ring = ring_mesh(4, r, h=h, twist=twist, aspect=aspect)
    old = quality(P, ring)
    code, nt, nq, eb = O.remove_edge(V1, V2, ring, old, *xyz(P))
    # the two re-triangulations, built here: a diagonal of the square coned to
    # v1 and v2; the reference tries first the diagonal from the first
    # non-edge vertex of the ring's first tet
    a0 = int(ring[0][2])
    first = (a0, 2 + (a0 - 2 + 2) % 4)
    second = tuple(v for v in (2, 3, 4, 5) if v not in first)
    cands = []
    for a, c in (first, second):
        b, d = [v for v in (2, 3, 4, 5) if v not in (a, c)]
        ts = [oriented(P, [a, c, b, w]) for w in (V1, V2)] + [oriented(P, [a, c, d, w]) for w in (V1, V2)]
        cands.append(ts)
    # removeAnEdge_44's rule (:327-368): a rotation that does not flip must
    # beat the best so far, else the whole removal fails
    best, expect, limit = None, 4, old.max()
    for ts in cands:
        if O.tet_flip(np.array(ts, np.int32), *xyz(P)).sum():
            continue
        qm = quality(P, ts).max()
        if not qm < limit * (1 + 1e-12):
            expect = 0
            break
        best, limit = ts, qm
    if best is None:
        expect = 0
    assert code == expect
    if code == 4:
        assert sorted(map(sorted, nt.tolist())) == sorted(map(sorted, best))
        qb = quality(P, best)
        assert np.array_equal(nq, quality(P, nt)) and nq.max() < old.max()
        assert np.allclose(np.sort(nq), np.sort(qb), rtol=1e-12, atol=0)
    else:
        assert len(nt) == 0


@pytest.mark.parametrize("r", [0.4, 1.0, 2.0])
def test_56_fan(r):
    P, ring = ring_mesh(5, r)
    old = quality(P, ring)
    code, nt, nq, eb = O.remove_edge(V1, V2, ring, old, *xyz(P))
    if code == 0:
        # nothing better: the fans from every pentagon vertex are flipped or worse
        for a in range(5):
            tris = [(a, (a + 1) % 5, (a + 2) % 5), (a, (a + 2) % 5, (a + 3) % 5),
                    (a, (a + 3) % 5, (a + 4) % 5)]
            ts = [oriented(P, [2 + i, 2 + j, 2 + k, w]) for (i, j, k) in tris for w in (V1, V2)]
            fl = O.tet_flip(np.array(ts, np.int32), *xyz(P)).sum()
            assert fl > 0 or quality(P, ts).max() >= old.max()
        return
    assert code == 5 and len(nt) == 6
    got = sorted(map(sorted, nt.tolist()))
    fans = []
    for a in range(5):
        tris = [(a, (a + 1) % 5, (a + 2) % 5), (a, (a + 2) % 5, (a + 3) % 5),
                (a, (a + 3) % 5, (a + 4) % 5)]
        fans.append(sorted(sorted([2 + i, 2 + j, 2 + k, w]) for (i, j, k) in tris for w in (V1, V2)))
    assert got in fans
    assert O.tet_flip(nt, *xyz(P)).sum() == 0
    assert np.array_equal(nq, quality(P, nt)) and nq.max() < old.max()


def test_other_ring_sizes_and_worse_rings():
    P, ring = ring_mesh(6, 1.0)
    assert O.remove_edge(V1, V2, ring, quality(P, ring), *xyz(P))[0] == 0
    for k in (3, 4, 5):
        P, ring = ring_mesh(k, 1.0)
        code, nt, nq, eb = O.remove_edge(V1, V2, ring, np.zeros(k), *xyz(P), equal_buget=0)
        assert code == 0 and len(nt) == 0
