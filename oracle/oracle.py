"""ctypes bindings for the CPU oracle (liboracle.so) and the reference AMIPS
build (_ref/libref_amips.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_int = ctypes.c_int
P = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = load(path)
    return _LIB


@contextlib.contextmanager
def swapped(path):
    """Every binding of this module calls into the library at `path` (a build
    of the oracle's sources, e.g. one with a planted mistake) inside the
    block."""
    global _LIB
    old = lib()
    _LIB = load(path)
    try:
        yield _LIB
    finally:
        _LIB = old


def load(path):
    """ctypes handle on an oracle build, with every signature declared."""
    L = ctypes.CDLL(path)
    sig = {
        "orc_amips_energy": (c_dbl, [P]),
        "orc_amips_jacobian": (None, [P, P]),
        "orc_amips_hessian": (None, [P, P]),
        "orc_amips_ejh": (c_dbl, [P, P, P]),
        "orc_rcbrt": (c_dbl, [c_dbl]),
        "orc_amips_batch": (None, [P, c_i64, P, P, P, c_int, P, P, P]),
        "orc_orient3d": (c_int, [P, P, P, P]),
        "orc_collinear3": (c_int, [P, P, P]),
        "orc_tet_qualities": (None, [P, c_i64, P, P, P, P]),
        "orc_tet_flip": (None, [P, c_i64, P, P, P, P]),
        "orc_sample_triangle": (c_i64, [P, P, P, c_dbl, P, c_i64]),
        "orc_point_tri_d2": (c_dbl, [P, P, P, P, P]),
        "orc_surface_create": (P, [P, c_i64, P, c_i64]),
        "orc_surface_destroy": (None, [P]),
        "orc_point_out": (c_int, [P, P, c_dbl]),
        "orc_face_out": (c_int, [P, P, c_dbl, c_dbl, P]),
        "orc_face_out_batch": (None, [P, P, c_i64, c_dbl, c_dbl, P, P]),
        "orc_point_out_batch": (None, [P, P, c_i64, c_dbl, P]),
        "orc_nearest": (c_i64, [P, P, P, P]),
        "orc_point_out_brute": (c_int, [P, P, c_dbl]),
        "orc_min_d2_brute": (c_dbl, [P, P]),
        "orc_newtons_method": (c_int, [P, c_int, ctypes.c_int32, P, P, P, P]),
        "orc_newtons_update": (c_int, [P, c_int, ctypes.c_int32, P, P, P, P, P, P, P]),
        "orc_new_energy": (c_dbl, [P, c_int, P, P, P]),
        "orc_solve3_qrcp": (c_int, [P, P, P]),
        "orc_smooth_problem": (c_int, [c_int, ctypes.c_int32, P, c_int, P, P, P, P, P, c_int,
                                       P, P, c_dbl, c_dbl, P, P]),
        "orc_energy_stats": (None, [P, c_i64, P, P, P, c_dbl, P, P, P, P, P]),
        "orc_filter_energy": (c_dbl, [P, c_dbl, P]),
        "orc_tet_groups": (None, [P, P, c_i64, P, P, P, P, P, P]),
        "orc_remove_edge": (c_int, [ctypes.c_int32, ctypes.c_int32, P, c_int, P, P, P, P, P,
                                    P, P, P]),
        "orc_collapse_check": (c_int, [ctypes.c_int32, ctypes.c_int32, c_int, P, c_int, P, P,
                                       P, P, P, c_int, P, P, c_dbl, c_dbl, c_int, c_int,
                                       c_int, c_dbl, P]),
        "orc_smooth_problem_risk": (c_int, [c_int, ctypes.c_int32, P, c_int, P, P, P, P, P,
                                            c_int, P, P, c_dbl, c_dbl, P, P, P]),
        "orc_smooth_vertex": (c_int, [c_int, ctypes.c_int32, P, c_int, P, P, P, P, P, P, P, P,
                                      P, c_int, P, P, c_dbl, c_dbl, P]),
        "orc_max_threads": (c_int, []),
        "orc_amips_batch_mt": (None, [P, c_i64, P, P, P, c_int, P, P, P, c_int]),
        "orc_tet_qualities_mt": (None, [P, c_i64, P, P, P, P, c_int]),
        "orc_face_out_batch_mt": (None, [P, P, c_i64, c_dbl, c_dbl, P, c_int]),
        "orc_point_out_batch_mt": (None, [P, P, c_i64, c_dbl, P, c_int]),
        "orc_smooth_batch_mt": (None, [c_i64, P, P, P, P, P, P, P, P, P, P, P, P, c_dbl, c_dbl,
                                       P, P, P, P, c_int]),
        "orc_collapse_batch_mt": (None, [c_i64, P, P, P, P, P, P, P, P, P, P, P, P, P, c_dbl,
                                         c_dbl, c_int, c_int, c_int, c_dbl, P, P, c_int]),
        "orc_face_risk_batch_mt": (None, [P, P, c_i64, c_dbl, c_dbl, c_dbl, P, c_int]),
        "orc_point_risk_batch_mt": (None, [P, P, c_i64, c_dbl, c_dbl, P, c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def ref():
    """The reference's own AMIPS functions, compiled by ref_build.sh.  Returns
    None when the library is absent (it is built only where the reference
    sources exist)."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libref_amips.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_amips_energy.restype = c_dbl
        L.ref_amips_energy.argtypes = [P]
        L.ref_amips_jacobian.argtypes = [P, P]
        L.ref_amips_hessian.argtypes = [P, P]
        L.ref_amips_batch_ejh.argtypes = [P, ctypes.c_long, P, P, P]
        L.ref_amips_batch_e.argtypes = [P, ctypes.c_long, P]
        L.ref_amips_batch_ejh_mt.argtypes = [P, ctypes.c_long, P, P, P, c_int]
        _REF = L
    return _REF


def _p(a):
    return a.ctypes.data_as(P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------- AMIPS ----
def amips_batch(tets, x, y, z, mode=2):
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    x, y, z = _f64(x), _f64(y), _f64(z)
    n = tets.shape[0]
    E = np.empty(n)
    G = np.empty((n, 3))
    H = np.empty((n, 9))
    lib().orc_amips_batch(_p(tets), n, _p(x), _p(y), _p(z), mode, _p(E), _p(G), _p(H))
    return E, G, H


def amips_ejh_T(T):
    """T: (n, 12) gathered tets -> E (n), G (n,3), H (n,9)."""
    T = _f64(T).reshape(-1, 12)
    n = T.shape[0]
    E = np.empty(n)
    G = np.empty((n, 3))
    H = np.empty((n, 9))
    L = lib()
    for i in range(n):
        E[i] = L.orc_amips_ejh(_p(T[i]), _p(G[i]), _p(H[i]))
    return E, G, H


def ref_ejh_T(T):
    R = ref()
    if R is None:
        return None
    T = _f64(T).reshape(-1, 12)
    n = T.shape[0]
    E = np.empty(n)
    G = np.empty((n, 3))
    H = np.empty((n, 9))
    R.ref_amips_batch_ejh(_p(T), n, _p(E), _p(G), _p(H))
    return E, G, H


def rcbrt(x):
    return lib().orc_rcbrt(float(x))


def tet_qualities(tets, x, y, z):
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.empty(tets.shape[0])
    lib().orc_tet_qualities(_p(tets), tets.shape[0], _p(x), _p(y), _p(z), _p(out))
    return out


def tet_flip(tets, x, y, z):
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.empty(tets.shape[0], dtype=np.uint8)
    lib().orc_tet_flip(_p(tets), tets.shape[0], _p(x), _p(y), _p(z), _p(out))
    return out


def orient3d(p, q, r, s):
    a = [_f64(v) for v in (p, q, r, s)]
    return lib().orc_orient3d(*[_p(v) for v in a])


def collinear3(p, q, r):
    a = [_f64(v) for v in (p, q, r)]
    return lib().orc_collinear3(*[_p(v) for v in a])


# ------------------------------------------------------------- envelope ----
def sample_triangle(v0, v1, v2, d):
    a = [_f64(v) for v in (v0, v1, v2)]
    L = lib()
    n = L.orc_sample_triangle(_p(a[0]), _p(a[1]), _p(a[2]), float(d), None, 0)
    ps = np.empty((max(n, 1), 3))
    L.orc_sample_triangle(_p(a[0]), _p(a[1]), _p(a[2]), float(d), _p(ps), n)
    return ps[:n]


def point_tri_d2(p, a, b, c):
    v = [_f64(t) for t in (p, a, b, c)]
    cp = np.empty(3)
    d2 = lib().orc_point_tri_d2(_p(v[0]), _p(v[1]), _p(v[2]), _p(v[3]), _p(cp))
    return d2, cp


class Surface:
    """Envelope surface with the oracle's implicit AABB tree over the facets
    in the order given."""

    def __init__(self, V, F):
        self.V = _f64(V).reshape(-1, 3)
        self.F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        self.h = lib().orc_surface_create(_p(self.V), self.V.shape[0], _p(self.F),
                                          self.F.shape[0])
        if not self.h:
            raise ValueError("empty surface")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_surface_destroy(self.h)
            self.h = None

    def face_out(self, tris, eps2, d):
        tris = _f64(tris).reshape(-1, 9)
        out = np.empty(tris.shape[0], dtype=np.uint8)
        tot = c_i64(0)
        lib().orc_face_out_batch(self.h, _p(tris), tris.shape[0], float(eps2), float(d),
                                 _p(out), ctypes.byref(tot))
        return out, tot.value

    def point_out(self, pts, eps2):
        pts = _f64(pts).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.uint8)
        lib().orc_point_out_batch(self.h, _p(pts), pts.shape[0], float(eps2), _p(out))
        return out

    def face_out_mt(self, tris, eps2, d, nthreads=0):
        """face_out over host threads (0: all OpenMP offers)"""
        tris = _f64(tris).reshape(-1, 9)
        out = np.empty(tris.shape[0], dtype=np.uint8)
        lib().orc_face_out_batch_mt(self.h, _p(tris), tris.shape[0], float(eps2), float(d),
                                    _p(out), int(nthreads))
        return out

    def point_out_mt(self, pts, eps2, nthreads=0):
        pts = _f64(pts).reshape(-1, 3)
        out = np.empty(pts.shape[0], dtype=np.uint8)
        lib().orc_point_out_batch_mt(self.h, _p(pts), pts.shape[0], float(eps2), _p(out),
                                     int(nthreads))
        return out

    def face_risk(self, tris, eps2, d, band, nthreads=0):
        """near-tie counts of isFaceOutEnvelop: (samples, samples with
        |min d2 - eps2| <= band, triangles with such a sample, triangles whose
        decision rests on the band)"""
        tris = _f64(tris).reshape(-1, 9)
        c = np.zeros(4, dtype=np.int64)
        lib().orc_face_risk_batch_mt(self.h, _p(tris), tris.shape[0], float(eps2), float(d),
                                     float(band), _p(c), int(nthreads))
        return tuple(int(v) for v in c)

    def point_risk(self, pts, eps2, band, nthreads=0):
        """(points, points with |min d2 - eps2| <= band)"""
        pts = _f64(pts).reshape(-1, 3)
        c = np.zeros(2, dtype=np.int64)
        lib().orc_point_risk_batch_mt(self.h, _p(pts), pts.shape[0], float(eps2), float(band),
                                      _p(c), int(nthreads))
        return int(c[0]), int(c[1])

    def point_out_brute(self, pts, eps2):
        pts = _f64(pts).reshape(-1, 3)
        L = lib()
        return np.array([L.orc_point_out_brute(self.h, _p(np.ascontiguousarray(p)), float(eps2))
                         for p in pts], dtype=np.uint8)

    def min_d2_brute(self, pts):
        pts = _f64(pts).reshape(-1, 3)
        L = lib()
        return np.array([L.orc_min_d2_brute(self.h, _p(np.ascontiguousarray(p))) for p in pts])

    def nearest(self, pts):
        pts = _f64(pts).reshape(-1, 3)
        n = pts.shape[0]
        cp = np.empty((n, 3))
        d2 = np.empty(n)
        fid = np.empty(n, dtype=np.int64)
        L = lib()
        for i in range(n):
            dd = c_dbl(0.0)
            fid[i] = L.orc_nearest(self.h, _p(np.ascontiguousarray(pts[i])), _p(cp[i]),
                                   ctypes.byref(dd))
            d2[i] = dd.value
        return cp, d2, fid


# ------------------------------------------------------------ smoothing ----
def newtons_method(ring_tets, v_id, x, y, z):
    rt = np.ascontiguousarray(ring_tets, dtype=np.int32).reshape(-1, 4)
    x, y, z = _f64(x), _f64(y), _f64(z)
    p = np.empty(3)
    moved = lib().orc_newtons_method(_p(rt), rt.shape[0], int(v_id), _p(x), _p(y), _p(z), _p(p))
    return bool(moved), p


def newtons_update(ring_tets, v_id, x, y, z):
    rt = np.ascontiguousarray(ring_tets, dtype=np.int32).reshape(-1, 4)
    x, y, z = _f64(x), _f64(y), _f64(z)
    E = c_dbl(0.0)
    J = np.empty(3)
    H = np.empty(9)
    X0 = np.empty(3)
    ok = lib().orc_newtons_update(_p(rt), rt.shape[0], int(v_id), _p(x), _p(y), _p(z),
                                  ctypes.byref(E), _p(J), _p(H), _p(X0))
    return bool(ok), E.value, J, H.reshape(3, 3), X0


def new_energy(ring_tets, x, y, z):
    rt = np.ascontiguousarray(ring_tets, dtype=np.int32).reshape(-1, 4)
    x, y, z = _f64(x), _f64(y), _f64(z)
    return lib().orc_new_energy(_p(rt), rt.shape[0], _p(x), _p(y), _p(z))


def solve3(A, b):
    A = _f64(A).reshape(9)
    b = _f64(b).reshape(3)
    x = np.empty(3)
    full = lib().orc_solve3_qrcp(_p(A), _p(b), _p(x))
    return bool(full), x


def smooth_batch(B, old_q=None, sf=None, bt=None, eps2=0.0, d=1.0):
    """Per-problem oracle over a batch dict (workloads.smoothing_batch layout):
    returns p_out (n,3), status (n), new_q (n_ring; 0 where not written)."""
    L = lib()
    x, y, z = _f64(B["x"]), _f64(B["y"]), _f64(B["z"])
    ro, to = B["ring_off"], B["tri_off"]
    rt_all = np.ascontiguousarray(B["ring_tets"], dtype=np.int32)
    ab_all = np.ascontiguousarray(B["tri_ab"], dtype=np.int32).reshape(-1, 2)
    n = len(B["center"])
    oq = np.zeros(len(rt_all)) if old_q is None else _f64(old_q)
    p_out = np.empty((n, 3))
    status = np.empty(n, dtype=np.uint8)
    new_q = np.zeros(max(len(rt_all), 1))
    for p in range(n):
        r0, r1 = int(ro[p]), int(ro[p + 1])
        t0, t1 = int(to[p]), int(to[p + 1])
        rt = np.ascontiguousarray(rt_all[r0:r1])
        ab = np.ascontiguousarray(ab_all[t0:t1]) if t1 > t0 else np.zeros((1, 2), np.int32)
        q = np.ascontiguousarray(oq[r0:r1]) if r1 > r0 else np.zeros(1)
        po = np.empty(3)
        nq = np.empty(max(r1 - r0, 1))
        s = L.orc_smooth_problem(int(B["kind"][p]), int(B["center"][p]), _p(rt), r1 - r0,
                                 _p(x), _p(y), _p(z), _p(q), _p(ab), t1 - t0,
                                 sf.h if sf is not None else None,
                                 bt.h if bt is not None else None, float(eps2), float(d),
                                 _p(po), _p(nq))
        status[p] = s
        p_out[p] = po
        if s == 0 and int(B["kind"][p]) in (1, 2):
            new_q[r0:r1] = nq[:r1 - r0]
    return p_out, status, new_q


class ExactMesh:
    """posf, exact positions (doubles) and is_rounded flags of a tet mesh, for
    the reference's per-vertex smoothing bodies with its rounding rules
    (orc_smooth_vertex); the arrays are updated in place"""

    def __init__(self, P, Pexact, rounded):
        self.x, self.y, self.z = (np.ascontiguousarray(P[:, k], dtype=np.float64) for k in range(3))
        self.px, self.py, self.pz = (np.ascontiguousarray(Pexact[:, k], dtype=np.float64)
                                     for k in range(3))
        self.r = np.array(rounded, dtype=np.uint8)

    def smooth_vertex(self, kind, v, ring_tets, old_q, tri_ab, sf=None, bt=None, eps2=0.0, d=1.0):
        rt = np.ascontiguousarray(ring_tets, dtype=np.int32)
        m = len(rt)
        oq = _f64(old_q) if m else np.zeros(1)
        ab = np.ascontiguousarray(tri_ab, dtype=np.int32).reshape(-1)
        nt = len(ab) // 2
        if nt == 0:
            ab = np.zeros(2, np.int32)
        nq = np.zeros(max(m, 1))
        s = lib().orc_smooth_vertex(int(kind), int(v), _p(rt), m, _p(self.x), _p(self.y),
                                    _p(self.z), _p(self.px), _p(self.py), _p(self.pz), _p(self.r),
                                    _p(oq), _p(ab), nt, sf.h if sf is not None else None,
                                    bt.h if bt is not None else None, float(eps2), float(d), _p(nq))
        return int(s), nq[:m]

    def posf(self):
        return np.stack([self.x, self.y, self.z], 1)


def max_threads():
    return int(lib().orc_max_threads())


def amips_batch_mt(tets, x, y, z, mode=2, nthreads=0):
    """amips_batch over host threads"""
    t = np.ascontiguousarray(tets, dtype=np.int32)
    n = len(t)
    E = np.empty(n)
    G = np.empty((n, 3)) if mode >= 1 else None
    H = np.empty((n, 9)) if mode >= 2 else None
    lib().orc_amips_batch_mt(_p(t), n, _p(_f64(x)), _p(_f64(y)), _p(_f64(z)), mode, _p(E),
                             _p(G) if G is not None else None, _p(H) if H is not None else None,
                             int(nthreads))
    return E, G, H


def smooth_batch_mt(B, old_q=None, sf=None, bt=None, eps2=0.0, d=1.0, nthreads=0):
    """smooth_batch over host threads; also returns the line-search near-tie
    counts (comparisons, within band, |dE| < 1e-5 tests, within band)"""
    x, y, z = _f64(B["x"]), _f64(B["y"]), _f64(B["z"])
    n = len(B["center"])
    rt = np.ascontiguousarray(B["ring_tets"], dtype=np.int32)
    nr = len(rt)
    oq = np.zeros(max(nr, 1)) if old_q is None else _f64(old_q)
    ab = np.ascontiguousarray(B["tri_ab"], dtype=np.int32).reshape(-1)
    if len(ab) == 0:
        ab = np.zeros(2, np.int32)
    p_out = np.empty((n, 3))
    status = np.empty(n, dtype=np.uint8)
    new_q = np.zeros(max(nr, 1))
    risk = np.zeros(4, dtype=np.int64)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    kind = np.ascontiguousarray(B["kind"], dtype=np.uint8)
    center, ro, to = i32(B["center"]), i32(B["ring_off"]), i32(B["tri_off"])
    lib().orc_smooth_batch_mt(n, _p(kind), _p(center), _p(ro), _p(rt), _p(x), _p(y), _p(z),
                              _p(oq), _p(to), _p(ab), sf.h if sf is not None else None,
                              bt.h if bt is not None else None, float(eps2), float(d),
                              _p(p_out), _p(status), _p(new_q), _p(risk), int(nthreads))
    return p_out, status, new_q, tuple(int(v) for v in risk)


def collapse_batch_mt(B, sf=None, bt=None, eps2=0.0, d=1.0, check_quality=True, check_eps=True,
                      soft_energy=None, nthreads=0):
    """collapse_batch over host threads"""
    x, y, z = _f64(B["x"]), _f64(B["y"]), _f64(B["z"])
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    rt = i32(B["ring_tets"]).reshape(-1)
    tri = i32(B["tri_ids"]).reshape(-1)
    nr = len(rt) // 4
    oq = _f64(B["old_q"]) if nr else np.zeros(1)
    n = len(B["v1"])
    status = np.empty(n, dtype=np.uint8)
    new_q = np.zeros(max(nr, 1))
    lib().orc_collapse_batch_mt(
        n, _p(i32(B["v1"])), _p(i32(B["v2"])), _p(np.ascontiguousarray(B["flags"], np.uint8)),
        _p(i32(B["ring_off"])), _p(rt if nr else np.zeros(4, np.int32)), _p(x), _p(y), _p(z),
        _p(oq), _p(i32(B["tri_off"])), _p(tri if len(tri) else np.zeros(3, np.int32)),
        sf.h if sf is not None else None, bt.h if bt is not None else None, float(eps2),
        float(d), int(bool(check_quality)), int(bool(check_eps)), int(soft_energy is not None),
        float(soft_energy) if soft_energy is not None else 0.0, _p(status), _p(new_q),
        int(nthreads))
    return status, new_q


def collapse_batch(B, sf=None, bt=None, eps2=0.0, d=1.0, check_quality=True, check_eps=True,
                   soft_energy=None):
    """Per-candidate oracle over a batch dict (workloads.collapse_batch layout):
    returns status (n) and new_q (n_ring; NaN for removed tets, 0 where the
    quality stage was not reached)."""
    L = lib()
    x, y, z = _f64(B["x"]), _f64(B["y"]), _f64(B["z"])
    ro, to = B["ring_off"], B["tri_off"]
    rt_all = np.ascontiguousarray(B["ring_tets"], dtype=np.int32).reshape(-1, 4)
    tri_all = np.ascontiguousarray(B["tri_ids"], dtype=np.int32).reshape(-1, 3)
    oq = _f64(B["old_q"]) if len(rt_all) else np.zeros(1)
    fl = B["flags"]
    n = len(B["v1"])
    status = np.empty(n, dtype=np.uint8)
    new_q = np.zeros(max(len(rt_all), 1))
    for c in range(n):
        r0, r1 = int(ro[c]), int(ro[c + 1])
        t0, t1 = int(to[c]), int(to[c + 1])
        rt = np.ascontiguousarray(rt_all[r0:r1]) if r1 > r0 else np.zeros((1, 4), np.int32)
        tr = np.ascontiguousarray(tri_all[t0:t1]) if t1 > t0 else np.zeros((1, 3), np.int32)
        q = np.ascontiguousarray(oq[r0:r1]) if r1 > r0 else np.zeros(1)
        nq = np.zeros(max(r1 - r0, 1))
        status[c] = L.orc_collapse_check(
            int(B["v1"][c]), int(B["v2"][c]), int(fl[c]), _p(rt), r1 - r0, _p(x), _p(y), _p(z),
            _p(q), _p(tr), t1 - t0, sf.h if sf is not None else None,
            bt.h if bt is not None else None, float(eps2), float(d), int(bool(check_quality)),
            int(bool(check_eps)), int(soft_energy is not None),
            float(soft_energy) if soft_energy is not None else 0.0, _p(nq))
        if r1 > r0:
            new_q[r0:r1] = nq[:r1 - r0]
    return status, new_q


def energy_stats(e, removed=None, tets=None, locked=None, thres=10.0):
    """(sum, max, second_max, count, buckets[11]) as the reference's loops"""
    e = _f64(e)
    out = [np.zeros(1) for _ in range(3)]
    cnt = np.zeros(1, np.int64)
    b = np.zeros(11, np.int64)
    rm = None if removed is None else np.ascontiguousarray(removed, dtype=np.uint8)
    tt = None if tets is None else np.ascontiguousarray(tets, dtype=np.int32)
    lk = None if locked is None else np.ascontiguousarray(locked, dtype=np.uint8)
    lib().orc_energy_stats(_p(e), len(e), None if rm is None else _p(rm),
                           None if tt is None else _p(tt), None if lk is None else _p(lk),
                           float(thres), _p(out[0]), _p(out[1]), _p(out[2]), _p(cnt), _p(b))
    return out[0][0], out[1][0], out[2][0], int(cnt[0]), b


def filter_energy(buckets, thres=10.0):
    """getFilterEnergy's return value and is_clean_up (None where not set)"""
    b = np.ascontiguousarray(buckets, dtype=np.int64)
    flag = np.full(1, -1, dtype=np.int32)
    v = lib().orc_filter_energy(_p(b), float(thres), _p(flag))
    return v, (None if flag[0] < 0 else bool(flag[0]))


def tet_groups(tets, off, x, y, z):
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    off = np.ascontiguousarray(off, dtype=np.int32)
    n = len(off) - 1
    fl = np.zeros(max(n, 1), np.uint8)
    mq = np.zeros(max(n, 1))
    q = np.zeros(max(len(tets), 1))
    lib().orc_tet_groups(_p(tets), _p(off), n, _p(_f64(x)), _p(_f64(y)), _p(_f64(z)), _p(fl),
                         _p(mq), _p(q))
    return fl[:n], mq[:n], q[:len(tets)]


def remove_edge(v1, v2, ring, old_q, x, y, z, equal_buget=0):
    """(code 0/3/4/5, new tets (k, 4), new qualities (k,), equal_buget after)"""
    ring = np.ascontiguousarray(ring, dtype=np.int32).reshape(-1, 4)
    oq = _f64(old_q)
    nt = np.zeros((6, 4), np.int32)
    nq = np.zeros(6)
    k = np.zeros(1, np.int32)
    eb = np.array([equal_buget], np.int32)
    code = lib().orc_remove_edge(int(v1), int(v2), _p(ring), len(ring), _p(oq), _p(_f64(x)),
                                 _p(_f64(y)), _p(_f64(z)), _p(eb), _p(nt), _p(nq), _p(k))
    return code, nt[:k[0]].copy(), nq[:k[0]].copy(), int(eb[0])
