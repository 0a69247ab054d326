"""CPU pin of the oracle's sampleTriangle (src/tetwild/Common.cpp:143-228)
against an independent model of the sample set it defines.

The reference samples a triangle on a hexagonal lattice of spacing d laid
out in the triangle's own plane.  With v0v1 the longest edge (the LAST
longest of the three, std::minmax_element) put on the x axis (v0 at the
origin, length L) and v2 at (a, h), the model is, in 50-digit arithmetic:

  * N = L / d; N <= 1: the three vertices and nothing else.  An integer N is
    decremented.  Base row: (n d, 0) for n = 0 .. floor(N), then v1.
  * M = floor(h / (sqrt(3)/2 d)); M < 1: v2 and nothing else.
  * row m = 1 .. M lies at y_m = m sqrt(3)/2 d.  Its left and right ends
    on the side edges are x0 = y_m a / h and x1 = L - y_m (L - a) / h; the
    rows stop at the first m with x1 - x0 <= d.  With u = x0 / d,
    n = floor(u + 1/2), bumped by one on even rows when floor(u) == n; the
    row's samples are x = (n + (m mod 2)/2 + i) d, i = 0 .. floor(|x1 - x_start| / d)
    (x_start - x0 <= d < x1 - x0, so the row never starts past its right end).
  * then v2, then the side edges v1 -> v2 and v2 -> v0 at spacing d from
    their first vertex, n = 1 .. floor(N_e) (N_e decremented when integer),
    for edges longer than d.

The oracle's samples must equal the model's in number and order, and in
position to 1e-12 of the coordinates' scale.  Triangles whose floor()
arguments fall within 1e-9 of an integer are skipped (FP64 may round either
way there), except the constructed cases whose ratios are exact in binary.
Each rule of the model is also checked to catch a planted mistake.
"""
from fractions import Fraction as Fr

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O
from tetwild_amd import workloads as W

from oracle_mutants import fails

mp.mp.dps = 50
NEAR = mp.mpf("1e-9")


class Ambiguous(Exception):
    """a floor() argument lies within the FP64 rounding band of an integer"""


def _floor(u, exact=False):
    f = mp.floor(u)
    if not exact and (u - f < NEAR or f + 1 - u < NEAR):
        raise Ambiguous()
    return int(f)


def model(tri, d, exact=False):
    """Expected samples of sampleTriangle (list of 3-vectors, mpf) and the
    branches taken.  exact=True: the caller guarantees that the ratios at
    integer boundaries are exact (no Ambiguous check)."""
    V = [[mp.mpf(float(c)) for c in v] for v in tri]
    d = mp.mpf(float(d))
    # the last longest edge, decided on exact squared lengths
    ls = []
    for i in range(3):
        a, b = tri[i], tri[(i + 1) % 3]
        ls.append(sum((Fr(float(a[k])) - Fr(float(b[k]))) ** 2 for k in range(3)))
    mx = max(ls)
    mi = max(i for i in range(3) if ls[i] == mx)
    srt = sorted(ls)
    if not exact and srt[1] != srt[2] and (srt[2] - srt[1]) < Fr(1, 10**9) * srt[2]:
        raise Ambiguous()
    info = {}

    def sub(u, v):
        return [u[k] - v[k] for k in range(3)]

    def nrm(u):
        return mp.sqrt(sum(c * c for c in u))

    def integer(u):
        return u == mp.floor(u)

    v0, v1, v2 = V[mi], V[(mi + 1) % 3], V[(mi + 2) % 3]
    e = sub(v1, v0)
    L = nrm(e)
    N = L / d
    if N <= 1:
        info["tiny"] = 1
        return [V[0], V[1], V[2]], info
    if integer(N):
        N -= 1
        info["int_ratio"] = 1
    ex = [c / L for c in e]
    w = sub(v2, v0)
    a = sum(w[k] * ex[k] for k in range(3))
    fy = [w[k] - a * ex[k] for k in range(3)]
    h = nrm(fy)
    ey = [c / h for c in fy]

    def at(x, y):
        return [v0[k] + x * ex[k] + y * ey[k] for k in range(3)]

    out = [at(n * d, 0) for n in range(_floor(N, exact) + 1)]
    out.append(v1)
    s32 = mp.sqrt(3) / 2
    M = _floor(h / (s32 * d), exact)
    if M < 1:
        info["flat"] = 1
        out.append(v2)
        return out, info
    for m in range(1, M + 1):
        y = m * s32 * d
        x0 = y * a / h
        x1 = L - y * (L - a) / h
        if abs(x1 - x0 - d) < NEAR * d and not exact:
            raise Ambiguous()
        if x1 - x0 <= d:
            info["break"] = info.get("break", 0) + 1
            break
        u = x0 / d
        n = _floor(u + mp.mpf(1) / 2, exact)
        if m % 2 == 0 and n == _floor(u, exact):
            n += 1
            info["even_bump"] = info.get("even_bump", 0) + 1
        xs = (n + mp.mpf(m % 2) / 2) * d
        N1 = _floor(abs(x1 - xs) / d, exact)
        out += [at(xs + i * d, y) for i in range(N1 + 1)]
        info["rows"] = info.get("rows", 0) + 1
    out.append(v2)
    for p, q in ((v1, v2), (v2, v0)):
        g = sub(q, p)
        le = nrm(g)
        Ne = le / d
        if Ne > 1:
            if integer(Ne):
                Ne -= 1
                info["int_edge"] = info.get("int_edge", 0) + 1
            u = [c / le for c in g]
            out += [[p[k] + n * d * u[k] for k in range(3)] for n in range(1, _floor(Ne, exact) + 1)]
            info["edges"] = info.get("edges", 0) + 1
    return out, info


def agree(tri, d, exact=False):
    """oracle == model; returns the model's branch info (None if skipped)"""
    try:
        exp, info = model(tri, d, exact)
    except Ambiguous:
        return None
    got = O.sample_triangle(tri[0], tri[1], tri[2], d)
    assert len(got) == len(exp), (len(got), len(exp))
    scale = max(1.0, float(np.max(np.abs(tri))))
    E = np.array([[float(c) for c in p] for p in exp])
    assert np.max(np.abs(got - E)) <= 1e-12 * scale, np.max(np.abs(got - E))
    return info


def random_triangles(n, seed):
    """triangles of 1 .. 40 d on the longest edge: acute, obtuse, thin and
    nearly equilateral, at an offset from the origin"""
    st = W.Stream(seed, 0)
    d = 1e-3
    out = []
    for k in range(n):
        kind = k % 4
        L = d * float(st.uniform(1, 0.6, 40)[0])
        if kind == 0:    # general
            apex = [float(st.uniform(1, 0.1, 0.9)[0]) * L, float(st.uniform(1, 0.05, 1.0)[0]) * L]
        elif kind == 1:  # thin (rows cut short or no row at all)
            apex = [float(st.uniform(1, 0.2, 0.8)[0]) * L, float(st.uniform(1, 0.0, 0.12)[0]) * L]
        elif kind == 2:  # near equilateral
            apex = [L / 2 * (1 + 0.1 * float(st.normal(1)[0])), L * 0.866 * (1 + 0.05 * float(st.normal(1)[0]))]
        else:            # obtuse at the apex
            apex = [float(st.uniform(1, 0.3, 0.7)[0]) * L, float(st.uniform(1, 0.05, 0.3)[0]) * L]
        R = np.linalg.qr(st.normal(9).reshape(3, 3))[0]
        o = st.normal(3)
        P = np.array([[0, 0, 0], [L, 0, 0], [apex[0], apex[1], 0.0]]) @ R.T + o
        rot = k % 3  # which vertex comes first
        out.append((np.roll(P, rot, axis=0), d))
    return out


def check_random():
    seen = {}
    checked = 0
    for tri, d in random_triangles(400, 31):
        info = agree(tri, d)
        if info is None:
            continue
        checked += 1
        for key, v in info.items():
            seen[key] = seen.get(key, 0) + 1
    assert checked > 380
    return seen


def check_constructed():
    # 3-4-5 right triangle with d = 1/2: every edge ratio is an exact integer
    # (6, 8, 10), so the decrement rule fires on all three edges
    tri = np.array([[0.0, 0, 0], [4.0, 0, 0], [0.0, 3.0, 0]])
    info = agree(tri, 0.5, exact=True)
    assert info.get("int_ratio") and info.get("int_edge") == 2
    # isosceles with two equal longest edges: the LAST of them is v0v1
    tri = np.array([[0.0, 0, 0], [2.0, 0, 0], [1.0, 3.0, 0]])
    info = agree(tri, 0.3)
    assert info is not None
    ps = O.sample_triangle(tri[0], tri[1], tri[2], 0.3)
    assert np.array_equal(ps[0], tri[2])  # v0 = vs[2] (the last longest edge starts there)
    # shorter than d: the vertices only
    tri = np.array([[0.0, 0, 0], [0.5, 0, 0], [0.2, 0.3, 0]])
    assert agree(tri, 1.0).get("tiny")
    # lower than one row: base row, v1, v2 and no side edges
    # (L / d = 8 exactly: base samples n = 0 .. 7, then v1 and v2)
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, 0.05, 0]])
    info = agree(tri, 0.125, exact=True)
    assert info.get("flat")
    assert len(O.sample_triangle(tri[0], tri[1], tri[2], 0.125)) == 8 + 1 + 1


def test_sample_triangle_matches_lattice_model():
    seen = check_random()
    check_constructed()
    # every rule of the model was exercised
    for key in ("rows", "break", "even_bump", "edges", "flat"):
        assert seen.get(key, 0) > 0, (key, seen)
    print("branches", seen)


# planted mistakes in the oracle's sampleTriangle, each of which the model
# check must catch (tests/oracle_mutants.py builds the mutated copy)
MUTATIONS = [
    ("row parity", "((n + (m % 2) / 2.0) - m * sqrt3_2 / tan_v0) * d",
     "((n + ((m + 1) % 2) / 2.0) - m * sqrt3_2 / tan_v0) * d"),
    ("even-row bump", "if (m % 2 == 0 && n == n1) n += 1;", "if (m % 2 == 1 && n == n1) n += 1;"),
    ("no even-row bump", "if (m % 2 == 0 && n == n1) n += 1;", "(void)n1;"),
    ("row count M", "int M = (int)(h / (sqrt3_2 * d));", "int M = (int)(h / d);"),
    ("row length N1", "int N1 = (int)(dist(v, v1_m) / d);", "int N1 = (int)(dist(v, v1_m) / d) - 1;"),
    ("break rule", "if (dist(v0_m, v1_m) <= d) break;", "if (dist(v0_m, v1_m) <= 0.5 * d) break;"),
    ("row height", "v3 v0_m = add(v0, mulv(n_v0v2, m * sqrt3_2 * d / sin_v0));",
     "v3 v0_m = add(v0, mulv(n_v0v2, m * sqrt3_2 * d / sin_v1));"),
    ("side edge v1v2 start", "for (int n = 1; n <= N; n++) push(&k, add(v1, mulv(mulv(nn, d), (double)n)));",
     "for (int n = 0; n <= N; n++) push(&k, add(v1, mulv(mulv(nn, d), (double)n)));"),
    ("side edge v2v0 spacing", "for (int n = 1; n <= N; n++) push(&k, add(v2, mulv(mulv(nn, d), (double)n)));",
     "for (int n = 1; n <= N; n++) push(&k, add(v2, mulv(mulv(nn, d), (double)n * 0.5)));"),
    ("edge integer rule", "    N = sqrt(ls[(max_i + 2) % 3]) / d;\n    if (N > 1) {\n        if (N == (int)N) N -= 1;",
     "    N = sqrt(ls[(max_i + 2) % 3]) / d;\n    if (N > 1) {\n        if (N == (int)N) N -= 0;"),
    ("longest-edge tie", "if (!(ls[i] < ls[max_i])) max_i = i;", "if (ls[i] > ls[max_i]) max_i = i;"),
    ("base integer rule", "if (N == (int)N) N -= 1;\n\n    v3 v0", "if (N == (int)N) N -= 0;\n\n    v3 v0"),
]


@pytest.mark.parametrize("name,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_lattice_model_catches_mutation(tmp_path, name, old, new):
    def check():
        check_random()
        check_constructed()
    assert fails(check, tmp_path, "envelope.c", old, new), name
