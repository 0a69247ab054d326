"""CPU pins of the control constants of the oracle's NewtonsMethod
(src/tetwild/VertexSmoother.cpp:465-542: MAX_STEP = 15 Newton steps, MAX_IT
= 20 step halvings per step, a step that lowers the energy by less than
1e-5 ends the loop) with rings built so that geometry fixes the outcome.

  * Needle.  One tet whose opposite face is a tiny equilateral triangle (size
    s = 1e-12) at height h below the vertex.  While h >> s its energy grows as
    h^(4/3), so the full Newton step lands at -2h (across the face plane: a
    flip), a = 1/2 at -h/2 (a flip), and a = 1/4 at h/4 is accepted: every
    step takes exactly two halvings and divides h by 4, lowering the energy by
    far more than 1e-5.  The loop is therefore cut by MAX_STEP: the final
    height is 4^-15 h.
  * Blocked needle.  The vertex at the origin, the needle below it, and a
    small well-shaped tet whose opposite face crosses the Newton ray at
    a = 1.75 * 2^-19 of the full step X1.  With the vertex at the origin the
    right-hand side is -a J exactly, so every candidate is a X1 exactly (a is
    a power of two).  Exact rational orientations show that every a >= 2^-18
    flips that tet and a = 2^-19 does not: the first step is accepted at the
    20th try (it = 19).  The second step would need a = 2^-20, a 21st try.
    So MAX_IT = 20 gives exactly p = 2^-19 X1, is_moved = true.
  * Octahedron.  The vertex displaced from the centre of a regular octahedron
    by t; the full step is accepted and lowers the ring energy (measured with
    the reference's own compiled energy function where present) by 3.4e-6
    at t = 4e-4 and 3.1e-5 at t = 1.2e-3: below and above 1e-5, so is_moved
    is false (with the vertex left at the full step, as the reference leaves
    it) and true.

Each pin is also run against copies of the oracle with a planted mistake in
these constants, the halving or the flip test, and must fail there.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import oracle as O

from oracle_mutants import fails


def xyz(P):
    return tuple(np.ascontiguousarray(P[:, k]) for k in range(3))


def positive(tets, P):
    """orders each tet so that its exact orientation is positive"""
    t = np.array(tets, np.int32)
    f = O.tet_flip(t, *xyz(P)).astype(bool)
    t[f] = t[f][:, [0, 1, 3, 2]]
    assert not O.tet_flip(t, *xyz(P)).any()
    return t


def triangle(zc, r, rot=0.0):
    ang = np.array([0, 2 * np.pi / 3, 4 * np.pi / 3]) + rot
    return np.stack([r * np.cos(ang), r * np.sin(ang), np.full(3, zc)], 1)


def exact_orient(a, b, c, d):
    A = [[Fr(float(u)) - Fr(float(w)) for u, w in zip(p, a)] for p in (b, c, d)]
    det = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
           - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
           + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]))
    return (det > 0) - (det < 0)


def ring_energy_ref(t, P):
    """sum of the reference's compiled energy over the ring (the oracle's
    energy where the reference library is not built)"""
    if O.ref() is None:
        return O.new_energy(t, *xyz(P))
    T = np.concatenate([P[tt] for tt in t]).reshape(-1, 12)
    return float(np.sum(O.ref_ejh_T(T)[0]))


def check_needle():
    P = np.vstack([[0.0, 0.0, 1.0], triangle(0.0, 1e-12)])
    t = positive([[0, 1, 2, 3]], P)
    moved, p = O.newtons_method(t, 0, *xyz(P))
    assert moved
    assert abs(p[0]) < 1e-20 and abs(p[1]) < 1e-20
    r = p[2] / 4.0 ** -15
    assert 0.98 < r < 1.02, r


def blocked_ring():
    X1z = -3.0
    for _ in range(3):  # place the blocker against the ring's own full step
        zb = 1.75 * 2.0 ** -19 * abs(X1z)
        P = np.vstack([[0.0, 0.0, 0.0], triangle(-1.0, 1e-12), triangle(-zb, zb, 0.3)])
        t = positive([[0, 1, 2, 3], [0, 4, 5, 6]], P)
        ok, E, J, H, X0 = O.newtons_update(t, 0, *xyz(P))
        assert ok and np.array_equal(X0, [0, 0, 0])
        full, X1 = O.solve3(H, -J)
        X1z = X1[2]
    return P, t, X1


def check_blocked():
    P, t, X1 = blocked_ring()
    blocker = t[1]
    # geometry: every a >= 2^-18 flips the blocker, a = 2^-19 does not
    for k in range(0, 19):
        q = X1 * 2.0 ** -k
        Q = P.copy()
        Q[0] = q
        assert exact_orient(*Q[blocker]) != 1, k
    Q = P.copy()
    Q[0] = X1 * 2.0 ** -19
    assert exact_orient(*Q[blocker]) == 1 and exact_orient(*Q[t[0]]) == 1
    assert ring_energy_ref(t, P) - ring_energy_ref(t, Q) > 1.0
    moved, p = O.newtons_method(t, 0, *xyz(P))
    assert moved
    assert np.array_equal(p, X1 * 2.0 ** -19), (p, X1 * 2.0 ** -19)


def octahedron(tt):
    c = tt * np.array([0.6, 0.48, 0.64])
    V = [np.zeros(3)]
    for i in range(3):
        V += [c + np.eye(3)[i], c - np.eye(3)[i]]
    P = np.array(V)
    t = positive([[0, a, b, cc] for a in (1, 2) for b in (3, 4) for cc in (5, 6)], P)
    return P, t


def check_threshold():
    for tt, want in ((4e-4, False), (1.2e-3, True)):
        P, t = octahedron(tt)
        ok, E, J, H, X0 = O.newtons_update(t, 0, *xyz(P))
        full, X = O.solve3(H, H.reshape(3, 3) @ X0 - J)  # the a = 1 candidate
        Q = P.copy()
        Q[0] = X
        dE = ring_energy_ref(t, P) - ring_energy_ref(t, Q)
        assert (1e-6 < dE < 1e-5) if not want else (1e-5 < dE < 1e-4), dE
        moved, p = O.newtons_method(t, 0, *xyz(P))
        assert moved == want, (tt, moved)
        if not want:
            assert np.array_equal(p, X)  # the accepted position is kept


def test_newton_max_step_needle():
    check_needle()


def test_newton_max_it_blocked():
    check_blocked()


def test_newton_energy_threshold():
    check_threshold()


MUTATIONS = [
    ("MAX_STEP 14", "#define MAX_STEP 15", "#define MAX_STEP 14", check_needle),
    ("MAX_STEP 16", "#define MAX_STEP 15", "#define MAX_STEP 16", check_needle),
    ("MAX_IT 19", "#define MAX_IT 20", "#define MAX_IT 19", check_blocked),
    ("MAX_IT 21", "#define MAX_IT 20", "#define MAX_IT 21", check_blocked),
    ("threshold 1e-6", "fabs(new_energy - old_energy) < 1e-5", "fabs(new_energy - old_energy) < 1e-6",
     check_threshold),
    ("threshold 1e-4", "fabs(new_energy - old_energy) < 1e-5", "fabs(new_energy - old_energy) < 1e-4",
     check_threshold),
    ("no flip test", "if (M ? e_is_flip(M, rt, m) : ring_flip(rt, m, &w)) {",
     "if (0 && (M ? e_is_flip(M, rt, m) : ring_flip(rt, m, &w))) {", check_needle),
    ("halving by 3",
     "                if (M) moved_back(M, v_id, old_p, old_pos);\n                a /= 2.0;\n                continue;\n            }\n            new_energy",
     "                if (M) moved_back(M, v_id, old_p, old_pos);\n                a /= 3.0;\n                continue;\n            }\n            new_energy",
     check_blocked),
    ("is_moved on step 0", "is_moved = step != 0;", "is_moved = 1;", check_threshold),
]


@pytest.mark.parametrize("name,old,new,check", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_newton_pins_catch_mutation(tmp_path, name, old, new, check):
    assert fails(check, tmp_path, "smooth.c", old, new), name
