"""Seeded synthetic workloads for the mesh-optimisation inner loop.

All random streams come from SplitMix64 counters (so they do not depend on
numpy's generator versions) with Box-Muller normals.

  kuhn_lattice(n, jitter)   jittered Kuhn lattice (6 tets per cube) in the unit
                            cube: n cubes per axis -> (n+1)^3 vertices, 6 n^3 tets
  aspect_tets(n)            separate tets: a regular tet stretched along one
                            (pancake) or two (needle) axes by 10^U[0,4], then a
                            random rotation and translation
  bumpy_sphere(freq)        geodesic icosphere of the given frequency displaced
                            radially by 0.05 * (sum of 8 random real spherical
                            harmonics)
  query_triangles(...)      near-equilateral query triangles offset from a
                            surface along the facet normal
  icosphere_jittered(...)   subdivided icosphere with radial Gaussian jitter
"""
import math

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based SplitMix64 stream: draw k values at a time."""

    def __init__(self, seed, tag=0):
        self.base = np.uint64((seed * 0x100000001B3 + tag * 0x51ED27) & 0xFFFFFFFFFFFFFFFF)
        self.pos = 0

    def u64(self, k):
        with np.errstate(over="ignore"):
            idx = np.arange(self.pos + 1, self.pos + 1 + k, dtype=np.uint64)
            z = self.base + idx * _G
            self.pos += k
            return _mix(z)

    def uniform(self, k, lo=0.0, hi=1.0):
        """U(lo, hi), never exactly lo."""
        u = ((self.u64(k) >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0**-53
        return lo + (hi - lo) * u

    def normal(self, k):
        m = (k + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:k]

    def integers(self, k, lo, hi):
        """uniform integers in [lo, hi]"""
        span = np.uint64(hi - lo + 1)
        return (self.u64(k) % span).astype(np.int64) + lo


# ------------------------------------------------------------------- tets --
_KUHN = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]


def _orient_filter(T):
    """sign of det[q-p; r-p; s-p] for gathered (n,4,3) tets when certain, 0 when
    the floating-point filter cannot decide."""
    a = T[:, 1] - T[:, 0]
    b = T[:, 2] - T[:, 0]
    c = T[:, 3] - T[:, 0]
    m1 = b[:, 1] * c[:, 2] - b[:, 2] * c[:, 1]
    m2 = b[:, 2] * c[:, 0] - b[:, 0] * c[:, 2]
    m3 = b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0]
    det = a[:, 0] * m1 + a[:, 1] * m2 + a[:, 2] * m3
    perm = (np.abs(a[:, 0]) * (np.abs(b[:, 1] * c[:, 2]) + np.abs(b[:, 2] * c[:, 1])) +
            np.abs(a[:, 1]) * (np.abs(b[:, 2] * c[:, 0]) + np.abs(b[:, 0] * c[:, 2])) +
            np.abs(a[:, 2]) * (np.abs(b[:, 0] * c[:, 1]) + np.abs(b[:, 1] * c[:, 0])))
    bound = 2.0**-50 * perm
    s = np.zeros(len(det), dtype=np.int8)
    s[det > bound] = 1
    s[-det > bound] = -1
    return s


def _orient_exact(t):
    from fractions import Fraction as F
    p = [[F(float(v)) for v in row] for row in t]
    a = [p[1][i] - p[0][i] for i in range(3)]
    b = [p[2][i] - p[0][i] for i in range(3)]
    c = [p[3][i] - p[0][i] for i in range(3)]
    det = (a[0] * (b[1] * c[2] - b[2] * c[1]) + a[1] * (b[2] * c[0] - b[0] * c[2]) +
           a[2] * (b[0] * c[1] - b[1] * c[0]))
    return (det > 0) - (det < 0)


def orientation(tets, x, y, z, chunk=1 << 22):
    out = np.empty(len(tets), dtype=np.int8)
    for s in range(0, len(tets), chunk):
        t = tets[s:s + chunk]
        T = np.stack([np.stack([x[t[:, j]], y[t[:, j]], z[t[:, j]]], axis=1) for j in range(4)],
                     axis=1)
        o = _orient_filter(T)
        for i in np.nonzero(o == 0)[0]:
            o[i] = _orient_exact(T[i])
        out[s:s + chunk] = o
    return out


def kuhn_lattice(n, jitter=0.15, seed=0):
    """Returns x, y, z (float64 SoA), tets (int32, (6 n^3, 4)), n_swapped."""
    n1 = n + 1
    h = 1.0 / n
    g = np.arange(n1, dtype=np.float64) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    # vertex id = i + n1*(j + n1*k) -> use Fortran-like order of (i, j, k)
    x = np.ascontiguousarray(X.transpose(2, 1, 0).reshape(-1))
    y = np.ascontiguousarray(Y.transpose(2, 1, 0).reshape(-1))
    z = np.ascontiguousarray(Z.transpose(2, 1, 0).reshape(-1))
    nv = n1 ** 3
    st = Stream(seed, 2)
    x += st.normal(nv) * (jitter * h)
    y += st.normal(nv) * (jitter * h)
    z += st.normal(nv) * (jitter * h)
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ii = ii.transpose(2, 1, 0).reshape(-1)
    jj = jj.transpose(2, 1, 0).reshape(-1)
    kk = kk.transpose(2, 1, 0).reshape(-1)
    base = (ii + n1 * (jj + n1 * kk)).astype(np.int64)
    step = np.array([1, n1, n1 * n1], dtype=np.int64)
    tets = np.empty((len(base), 6, 4), dtype=np.int64)
    for t, perm in enumerate(_KUHN):
        v = base.copy()
        tets[:, t, 0] = v
        v = v + step[perm[0]]
        tets[:, t, 1] = v
        v = v + step[perm[1]]
        tets[:, t, 2] = v
        tets[:, t, 3] = base + step.sum()
    tets = tets.reshape(-1, 4).astype(np.int32)
    o = orientation(tets, x, y, z)
    bad = o <= 0
    tmp = tets[bad, 2].copy()
    tets[bad, 2] = tets[bad, 3]
    tets[bad, 3] = tmp
    return x, y, z, np.ascontiguousarray(tets), int(bad.sum())


def _rotations(st, k):
    q = st.normal(4 * k).reshape(k, 4)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, a, b, c = q.T
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (b * b + c * c)
    R[:, 0, 1] = 2 * (a * b - c * w)
    R[:, 0, 2] = 2 * (a * c + b * w)
    R[:, 1, 0] = 2 * (a * b + c * w)
    R[:, 1, 1] = 1 - 2 * (a * a + c * c)
    R[:, 1, 2] = 2 * (b * c - a * w)
    R[:, 2, 0] = 2 * (a * c - b * w)
    R[:, 2, 1] = 2 * (b * c + a * w)
    R[:, 2, 2] = 1 - 2 * (a * a + b * b)
    return R


REGULAR_TET = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, math.sqrt(3) / 2, 0.0],
                        [0.5, math.sqrt(3) / 6, math.sqrt(2.0 / 3.0)]])


def aspect_tets(n, seed=0, max_log10=4.0, size=0.01):
    """Separate tets with aspect ratio 10^U[0, max_log10]: returns x, y, z of
    4n vertices and tets (n, 4) = [4i, 4i+1, 4i+2, 4i+3], positively oriented."""
    st = Stream(seed, 4)
    a = 10.0 ** st.uniform(n, 0.0, max_log10)
    needle = st.u64(n) & np.uint64(1)
    S = np.ones((n, 3))
    S[:, 2] = np.where(needle == 0, 1.0 / a, a)       # pancake: squash z
    S[:, 1] = np.where(needle == 0, 1.0, a)           # needle: stretch y and z
    P = (REGULAR_TET - REGULAR_TET.mean(0))[None, :, :] * S[:, None, :] * size
    R = _rotations(st, n)
    P = np.einsum("nij,nvj->nvi", R, P)
    P += st.uniform(3 * n, -1.0, 1.0).reshape(n, 1, 3)
    x = np.ascontiguousarray(P[:, :, 0].reshape(-1))
    y = np.ascontiguousarray(P[:, :, 1].reshape(-1))
    z = np.ascontiguousarray(P[:, :, 2].reshape(-1))
    tets = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    o = orientation(tets, x, y, z)
    bad = o < 0
    tmp = tets[bad, 2].copy()
    tets[bad, 2] = tets[bad, 3]
    tets[bad, 3] = tmp
    return x, y, z, np.ascontiguousarray(tets)


# --------------------------------------------------------------- surfaces --
def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return V / np.linalg.norm(V, axis=1, keepdims=True), F


def geodesic_sphere(freq):
    """Unit geodesic sphere: each icosahedron face split into freq^2 triangles;
    10 freq^2 + 2 vertices and 20 freq^2 facets, outward oriented."""
    V0, F0 = _icosahedron()
    ij = [(i, j) for i in range(freq + 1) for j in range(freq + 1 - i)]
    idx = {p: k for k, p in enumerate(ij)}
    I = np.array([p[0] for p in ij], dtype=np.float64)
    J = np.array([p[1] for p in ij], dtype=np.float64)
    tris = []
    for i in range(freq):
        for j in range(freq - i):
            tris.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
            if j < freq - i - 1:
                tris.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
    tris = np.array(tris, dtype=np.int64)
    allV, allF = [], []
    npf = len(ij)
    for f, (a, b, c) in enumerate(F0):
        A, B, C = V0[a], V0[b], V0[c]
        P = (A[None, :] * (freq - I - J)[:, None] + B[None, :] * I[:, None] +
             C[None, :] * J[:, None]) / freq
        allV.append(P)
        allF.append(tris + f * npf)
    V = np.concatenate(allV)
    F = np.concatenate(allF)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    key = np.round(V * 1e9).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    order = np.argsort(first)
    remap = np.empty_like(order)
    remap[order] = np.arange(len(order))
    V = V[first[order]]
    F = remap[inv[F]]
    # outward orientation
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[F].mean(1)) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return np.ascontiguousarray(V), np.ascontiguousarray(F.astype(np.int32))


def _real_sph_harm(l, m, theta, phi):
    """orthonormal real spherical harmonic Y_lm (theta polar, phi azimuth)"""
    from scipy.special import lpmv
    am = abs(m)
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    P = lpmv(am, l, np.cos(theta)) * ((-1) ** am)  # remove Condon-Shortley phase
    if m == 0:
        return norm * P
    if m > 0:
        return math.sqrt(2) * norm * P * np.cos(am * phi)
    return math.sqrt(2) * norm * P * np.sin(am * phi)


def bumpy_sphere(freq=224, n_harm=8, amp=0.05, seed=0):
    V, F = geodesic_sphere(freq)
    st = Stream(seed, 3)
    ls = st.integers(n_harm, 1, 8)
    ms = np.array([int(st.integers(1, -l, l)[0]) for l in ls])
    theta = np.arccos(np.clip(V[:, 2], -1, 1))
    phi = np.arctan2(V[:, 1], V[:, 0])
    r = np.ones(len(V))
    for l, m in zip(ls, ms):
        r += amp * _real_sph_harm(int(l), int(m), theta, phi)
    return np.ascontiguousarray(V * r[:, None]), F


def icosphere_jittered(levels=4, sigma=0.002, seed=0):
    V, F = geodesic_sphere(2 ** levels)
    st = Stream(seed, 5)
    r = 1.0 + sigma * st.normal(len(V))
    return np.ascontiguousarray(V * r[:, None]), F


def facet_normals(V, F):
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    area2 = np.linalg.norm(n, axis=1)
    return n / area2[:, None], 0.5 * area2


def surface_points(V, F, n, seed=0, tag=6, _pre=None):
    """n points uniform by area on the surface: (points, facet ids, normals)"""
    st = Stream(seed, tag)
    if _pre is None:
        nrm, area = facet_normals(V, F)
        cdf = np.cumsum(area)
        cdf /= cdf[-1]
    else:
        nrm, cdf = _pre
    f = np.searchsorted(cdf, st.uniform(n), side="right")
    f = np.minimum(f, len(F) - 1)
    u = st.uniform(n)
    v = st.uniform(n)
    s = np.sqrt(u)
    b0, b1, b2 = 1 - s, s * (1 - v), s * v
    P = (V[F[f, 0]] * b0[:, None] + V[F[f, 1]] * b1[:, None] + V[F[f, 2]] * b2[:, None])
    return P, f, nrm[f]


def query_triangles(V, F, n, eps=1e-3, edge=0.002, seed=0, chunk=1 << 21):
    """(n, 9) near-equilateral triangles: centre uniform by area on the
    surface, in the facet plane, edge * (1 + U[-0.1, 0.1]) long, offset along
    the facet normal by +-U[0, 2 eps]."""
    out = np.empty((n, 9))
    nrm, area = facet_normals(V, F)
    cdf = np.cumsum(area)
    cdf /= cdf[-1]
    for s in range(0, n, chunk):
        k = min(chunk, n - s)
        C, f, N = surface_points(V, F, k, seed, tag=7 + s // chunk, _pre=(nrm, cdf))
        st = Stream(seed, 1000 + s // chunk)
        # tangent frame
        a = np.where(np.abs(N[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
        t1 = np.cross(N, a)
        t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
        t2 = np.cross(N, t1)
        ang = st.uniform(k, 0, 2 * np.pi)
        e = edge * (1 + st.uniform(k, -0.1, 0.1))
        rad = e / math.sqrt(3)
        off = st.uniform(k, -2 * eps, 2 * eps)
        for j in range(3):
            th = ang + j * 2 * np.pi / 3
            P = C + rad[:, None] * (np.cos(th)[:, None] * t1 + np.sin(th)[:, None] * t2) + \
                off[:, None] * N
            out[s:s + k, 3 * j:3 * j + 3] = P
    return out


def envelope_samples(V, F, n, eps=1e-3, band=0.01, seed=0):
    """(n, 3) points at distance ~ eps * (1 + U[-band, band]) from the surface
    along the facet normal, either side."""
    P, f, N = surface_points(V, F, n, seed, tag=11)
    st = Stream(seed, 12)
    d = eps * (1 + st.uniform(n, -band, band))
    sgn = np.where((st.u64(n) & np.uint64(1)) == 0, 1.0, -1.0)
    return np.ascontiguousarray(P + (sgn * d)[:, None] * N)


# ------------------------------------------------------------ one-rings --
_HULLS = {}


def _hull_template(k):
    """triangulated sphere (outward faces) over k seeded points on the unit
    sphere; 2k - 4 faces"""
    if k not in _HULLS:
        from scipy.spatial import ConvexHull
        st = Stream(1000 + k, 13)
        P = st.normal(3 * k).reshape(k, 3)
        P /= np.linalg.norm(P, axis=1, keepdims=True)
        h = ConvexHull(P)
        F = h.simplices.copy()
        n = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
        flip = np.einsum("ij,ij->i", n, P[F].mean(1)) < 0
        F[flip] = F[flip][:, [0, 2, 1]]
        _HULLS[k] = (P, F.astype(np.int64))
    return _HULLS[k]


_EVEN = np.array([[0, 1, 2, 3], [0, 2, 3, 1], [0, 3, 1, 2], [1, 0, 3, 2], [1, 2, 0, 3],
                  [1, 3, 2, 0], [2, 0, 1, 3], [2, 1, 3, 0], [2, 3, 0, 1], [3, 0, 2, 1],
                  [3, 1, 0, 2], [3, 2, 1, 0]])


def ring_sizes(n, seed=0, median=24.0, sigma=0.6, lo=4, hi=200):
    st = Stream(seed, 14)
    s = np.exp(np.log(median) + sigma * st.normal(n))
    s = np.clip(np.round(s), lo, hi).astype(np.int64)
    return s + (s % 2)  # a closed one-ring of a vertex has an even tet count


def smoothing_batch(n, seed=0, surface_frac=0.2, V=None, F=None, radius=0.01, boundary_frac=0.0):
    """n frozen one-ring problems.  Interior rings: cone from the centre over a
    triangulated sphere of k = (size + 4) / 2 jittered points; surface rings
    (a fraction surface_frac, needs the surface V, F): a bipyramid whose
    equator lies on the surface's tangent plane around a surface point, the
    equator faces being the one-ring surface faces.  Centres are displaced
    from the ring's middle so that Newton moves them; every ring tet is
    positively oriented.  Problems of equal ring size are generated together.

    Returns a dict of numpy arrays: x, y, z, center, ring_off, ring_tets (int32
    [n_ring, 4], each tet's vertex order an even permutation, so the centre
    sits at a random slot), kind (uint8), tri_off, tri_ab."""
    st = Stream(seed, 15)
    sizes = ring_sizes(n, seed)
    u = st.uniform(n)
    kind = np.zeros(n, dtype=np.uint8)
    if V is not None and surface_frac > 0:
        kind[u < surface_frac] = 1
        kind[u < surface_frac * boundary_frac] = 2
    sizes[kind > 0] = np.maximum(sizes[kind > 0], 6)
    nverts = np.where(kind == 0, 1 + (sizes + 4) // 2, 1 + sizes // 2 + 2)
    voff = np.zeros(n + 1, dtype=np.int64)
    voff[1:] = np.cumsum(nverts)
    roff = np.zeros(n + 1, dtype=np.int64)
    roff[1:] = np.cumsum(sizes)
    ntri = np.where(kind > 0, sizes // 2, 0)
    toff = np.zeros(n + 1, dtype=np.int64)
    toff[1:] = np.cumsum(ntri)
    pts = np.empty((voff[-1], 3))
    tets = np.empty((roff[-1], 4), dtype=np.int64)
    tri_ab = np.empty((toff[-1], 2), dtype=np.int64)
    if V is not None and (kind > 0).any():
        SP, _, SN = surface_points(V, F, n, seed=seed, tag=16)
    for k in np.unique(sizes[kind == 0]):
        idx = np.nonzero((kind == 0) & (sizes == k))[0]
        m = len(idx)
        kv = (int(k) + 4) // 2
        P, Fh = _hull_template(kv)
        R = _rotations(st, m)
        rad = radius * (1 + 0.1 * st.uniform(m * kv, -1, 1).reshape(m, kv))
        Q = np.einsum("mij,mkj->mki", R, P[None] * rad[:, :, None])
        f0, f1, f2 = Q[:, Fh[:, 0]], Q[:, Fh[:, 1]], Q[:, Fh[:, 2]]
        nrm = np.cross(f1 - f0, f2 - f0)
        nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
        h = np.einsum("mfi,mfi->mf", nrm, f0).min(axis=1)
        dirv = st.normal(3 * m).reshape(m, 3)
        dirv /= np.linalg.norm(dirv, axis=1, keepdims=True)
        ctr = dirv * (h * 0.5 * st.uniform(m))[:, None]
        base = st.uniform(3 * m, -1, 1).reshape(m, 1, 3)
        allp = np.concatenate([ctr[:, None, :], Q], 1) + base
        for j, p in enumerate(idx):
            pts[voff[p]:voff[p + 1]] = allp[j]
        T = np.concatenate([np.zeros((len(Fh), 1), dtype=np.int64), Fh + 1], 1)
        TT = T[None] + voff[idx][:, None, None]
        for j, p in enumerate(idx):
            tets[roff[p]:roff[p + 1]] = TT[j]
    for s in np.unique(sizes[kind > 0]):
        idx = np.nonzero((kind > 0) & (sizes == s))[0]
        m = len(idx)
        ke = int(s) // 2
        c0, nn = SP[idx], SN[idx]
        a = np.where(np.abs(nn[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
        t1 = np.cross(nn, a)
        t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
        t2 = np.cross(nn, t1)
        ang = 2 * np.pi * (np.arange(ke)[None] + 0.3 * st.uniform(m * ke, -1, 1).reshape(m, ke)) / ke
        rr = radius * (1 + 0.1 * st.uniform(m * ke, -1, 1).reshape(m, ke))
        eq = c0[:, None] + rr[:, :, None] * (np.cos(ang)[:, :, None] * t1[:, None] +
                                             np.sin(ang)[:, :, None] * t2[:, None])
        north = c0 + nn * radius
        south = c0 - nn * radius
        off = st.uniform(2 * m, -0.15, 0.15).reshape(m, 2) * radius
        ctr = c0 + off[:, :1] * t1 + off[:, 1:] * t2 + nn * (0.02 * radius * st.uniform(m, -1, 1))[:, None]
        allp = np.concatenate([ctr[:, None], eq, north[:, None], south[:, None]], 1)
        e = np.arange(ke) + 1
        en = np.roll(e, -1)
        faces = np.concatenate([np.stack([np.full(ke, ke + 1), e, en], 1),
                                np.stack([np.full(ke, ke + 2), en, e], 1)])
        fv = allp[:, faces]  # (m, 2ke, 3, 3)
        nrm = np.cross(fv[:, :, 1] - fv[:, :, 0], fv[:, :, 2] - fv[:, :, 0])
        flip = np.einsum("mfi,mfi->mf", nrm, fv.mean(2) - ctr[:, None]) < 0
        Fm = np.broadcast_to(faces, (m,) + faces.shape).copy()
        Fm[flip] = Fm[flip][:, [0, 2, 1]]
        T = np.concatenate([np.zeros((m, len(faces), 1), dtype=np.int64), Fm], 2)
        # one-ring surface faces (ctr, e_i, e_i+1): the centre has the smallest
        # id, so the reference's sorted-and-rotated triple is (ctr, lo, hi);
        # std::sort orders the faces by (lo, hi)
        lo = np.minimum(e, en)
        hi = np.maximum(e, en)
        o = np.lexsort((hi, lo))
        ab = np.stack([lo[o], hi[o]], 1)
        for j, p in enumerate(idx):
            pts[voff[p]:voff[p + 1]] = allp[j]
            tets[roff[p]:roff[p + 1]] = T[j] + voff[p]
            tri_ab[toff[p]:toff[p + 1]] = ab + voff[p]
    perm = _EVEN[st.integers(len(tets), 0, 11)]
    tets = np.take_along_axis(tets, perm, axis=1)
    return {
        "x": np.ascontiguousarray(pts[:, 0]), "y": np.ascontiguousarray(pts[:, 1]),
        "z": np.ascontiguousarray(pts[:, 2]),
        "center": voff[:-1].astype(np.int32), "ring_off": roff.astype(np.int32),
        "ring_tets": np.ascontiguousarray(tets.astype(np.int32)),
        "kind": kind, "tri_off": toff.astype(np.int32),
        "tri_ab": np.ascontiguousarray(tri_ab.astype(np.int32)),
    }


# ------------------------------------------------------- collapse batches --
NOT_SURFACE = 2**31 - 1  # State::NOT_SURFACE


def sheet_mesh(n, jitter=0.15, surf_jitter=0.02, seed=0, level=None):
    """Jittered Kuhn lattice (n cubes per axis) carrying an open surface sheet:
    the horizontal grid plane z = level / n (default level n // 2), inside the
    volume like TetWild's embedded input surface, so every sheet face has a tet
    on each side and the sheet's rim is its boundary in the reference's sense
    (LocalOperations::isEdgeOnBoundary, src/tetwild/LocalOperations.cpp:931-965).
    Sheet vertices move in the plane by N(0, jitter h) and off it by
    N(0, surf_jitter h); the rest by N(0, jitter h) per axis.  Tets are
    re-oriented (2 <-> 3) where the jitter made them non-positive.

    Returns a dict: x, y, z, tets (int32), is_surface_fs (int32 [nt, 4]: 0 for
    the face opposite vertex j when it lies in the sheet, NOT_SURFACE
    otherwise), on_surface / on_boundary (bool per vertex), and the envelope
    inputs: sf_V, sf_F (the sheet triangles, z exactly on the plane) and b_V,
    b_F (the rim segments as degenerate (a, b, b) triangles), the layout TetWild
    builds its boundary tree in (src/tetwild/Preprocess.cpp:191-196)."""
    n1 = n + 1
    h = 1.0 / n
    lv = n // 2 if level is None else level
    x, y, z, tets, _ = kuhn_lattice(n, jitter=jitter, seed=seed)
    vid = np.arange(n1 ** 3)
    gi, gj, gk = vid % n1, (vid // n1) % n1, vid // (n1 * n1)
    on = gk == lv
    st = Stream(seed, 40)
    z[on] = lv * h + st.normal(int(on.sum())) * (surf_jitter * h)
    o = orientation(tets, x, y, z)
    bad = o <= 0
    tmp = tets[bad, 2].copy()
    tets[bad, 2] = tets[bad, 3]
    tets[bad, 3] = tmp
    rim = on & ((gi == 0) | (gi == n) | (gj == 0) | (gj == n))
    fs = np.full(tets.shape, NOT_SURFACE, dtype=np.int32)
    for j in range(4):
        k = [m for m in range(4) if m != j]
        fs[on[tets[:, k[0]]] & on[tets[:, k[1]]] & on[tets[:, k[2]]], j] = 0
    tt, jj = np.nonzero(fs != NOT_SURFACE)
    faces = np.sort(np.stack([tets[tt, (jj + k) % 4] for k in (1, 2, 3)], 1), 1)
    faces = np.unique(faces, axis=0).astype(np.int32)
    V = np.stack([x, y, np.where(on, lv * h, z)], 1)
    segs = []
    for a in range(n):
        for (i0, j0, i1, j1) in ((a, 0, a + 1, 0), (a, n, a + 1, n), (0, a, 0, a + 1), (n, a, n, a + 1)):
            segs.append((i0 + n1 * (j0 + n1 * lv), i1 + n1 * (j1 + n1 * lv)))
    segs = np.array(segs, dtype=np.int32)
    b_F = np.stack([segs[:, 0], segs[:, 1], segs[:, 1]], 1).astype(np.int32)
    return {"x": x, "y": y, "z": z, "tets": np.ascontiguousarray(tets), "is_surface_fs": fs,
            "on_surface": on, "on_boundary": rim, "sf_V": V, "sf_F": faces,
            "b_V": V, "b_F": b_F}


def icosphere_tetmesh(levels=4, sigma=0.002, seed=0, eps_rel=1e-3, edge_rel=1.0 / 20):
    """The end-to-end CPU workload's input surface -- a subdivided icosphere
    (levels=4: 2562 vertices, 5120 triangles, radius 1) jittered radially by
    N(0, sigma) -- embedded in a tet mesh of its bounding region, standing in
    for the stage-2 mesh MeshRefinement::refine works on.

    Tolerances follow State (src/tetwild/State.cpp:19-46) with the default
    Args (include/tetwild/Args.h:21,43,54: eps_rel 1e-3, edge 1/20, stage 1):
    eps_input = eps_rel * bbox diagonal, sampling_dist = eps_input, eps =
    eps_input * (1 - 1/sqrt(3)), ideal edge length L = edge_rel * diagonal.

    The tet mesh is the Delaunay tetrahedralization (scipy) of the surface
    vertices and a jittered cubic lattice of spacing L over [-1.2, 1.2]^3
    without the points within L/2 of the unit sphere; every input triangle is
    a face of it (asserted).  Tets are positively oriented (2 <-> 3 swapped
    where needed); faces on the input surface have is_surface_fs 0; the
    hull's vertices are the bbox vertices.  The surface is closed, so the
    boundary tree is getSimpleMesh's single degenerate triangle at the origin
    (src/tetwild/MeshRefinement.cpp:907-919).

    Returns a dict in sheet_mesh's layout plus on_bbox, eps_input, eps, eps_2,
    sampling_dist, edge_len and diag."""
    from scipy.spatial import Delaunay
    V, F = icosphere_jittered(levels, sigma, seed)
    diag = float(np.linalg.norm(V.max(0) - V.min(0)))
    L = edge_rel * diag
    st = Stream(seed, 60)
    g = np.arange(-1.2, 1.2 + 1e-9, 0.999 * L)
    G = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    G = G + st.normal(G.size).reshape(-1, 3) * (0.1 * L)
    r = np.linalg.norm(G, axis=1)
    G = G[(r < 1 - 0.5 * L) | (r > 1 + 0.5 * L)]
    P = np.ascontiguousarray(np.concatenate([V, G]))
    D = Delaunay(P)
    tets = D.simplices.astype(np.int32)
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    bad = orientation(tets, x, y, z) <= 0
    tets[bad] = tets[bad][:, [0, 1, 3, 2]]
    assert (orientation(tets, x, y, z) > 0).all()
    nv = len(P)
    fkey = lambda f: (f[:, 0] * nv + f[:, 1]) * nv + f[:, 2]  # noqa: E731
    skey = np.sort(fkey(np.sort(F.astype(np.int64), 1)))
    fs = np.full(tets.shape, NOT_SURFACE, dtype=np.int32)
    hit = np.zeros(len(skey), bool)
    for j in range(4):
        f = np.sort(np.stack([tets[:, (j + k) % 4] for k in (1, 2, 3)], 1).astype(np.int64), 1)
        k = fkey(f)
        pos = np.clip(np.searchsorted(skey, k), 0, len(skey) - 1)
        m = skey[pos] == k
        fs[m, j] = 0
        hit[pos[m]] = True
    assert hit.all(), "an input triangle is not a face of the tet mesh"
    on_surface = np.zeros(nv, bool)
    on_surface[:len(V)] = True
    on_bbox = np.zeros(nv, bool)
    on_bbox[np.unique(D.convex_hull)] = True
    eps_input = eps_rel * diag
    eps = eps_input * (1 - 1 / math.sqrt(3))
    return {"x": x, "y": y, "z": z, "tets": np.ascontiguousarray(tets), "is_surface_fs": fs,
            "on_surface": on_surface, "on_boundary": np.zeros(nv, bool), "on_bbox": on_bbox,
            "sf_V": V, "sf_F": F, "b_V": np.zeros((1, 3)), "b_F": np.zeros((1, 3), np.int32),
            "eps_input": eps_input, "eps": eps, "eps_2": eps * eps, "sampling_dist": eps_input,
            "edge_len": L, "diag": diag}


def vertex_tets(tets, nv):
    """conn_tets as CSR (off [nv + 1], tet ids ascending per vertex)"""
    inc_v = tets.reshape(-1).astype(np.int64)
    inc_t = np.repeat(np.arange(len(tets), dtype=np.int64), 4)
    o = np.lexsort((inc_t, inc_v))
    off = np.zeros(nv + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(inc_v, minlength=nv))
    return off, inc_t[o]


def _csr_gather(off, vals, keys):
    """concatenation of vals[off[k]:off[k+1]] for k in keys, and its CSR offsets"""
    cnt = off[keys + 1] - off[keys]
    out_off = np.zeros(len(keys) + 1, dtype=np.int64)
    out_off[1:] = np.cumsum(cnt)
    idx = np.repeat(off[keys] - out_off[:-1], cnt) + np.arange(out_off[-1])
    return vals[idx], out_off, cnt


def collapse_batch(mesh, q, n_cand=None, seed=0, unrounded_frac=0.0):
    """Edge-collapse candidates v1 -> v2 over the directed edges of `mesh`
    (sheet_mesh layout), all of them or n_cand drawn without replacement.
    q: stored slim energy per tet.  Per candidate: v1's one-ring (conn_tets in
    ascending tet id), the ring's stored qualities, the vertex flags, and the
    triangles isCollapsable_epsilon tests (src/tetwild/EdgeCollapser.cpp:742-764:
    v1's one-ring surface faces as sorted vertex triples, in lexicographic order,
    without those containing v2, v1 replaced by v2 in place).

    Returns a dict of numpy arrays: x, y, z, v1, v2, flags (uint8, TWG_CV*),
    ring_off, ring_tets, old_q, tri_off, tri_ids."""
    tets = mesh["tets"]
    nv = len(mesh["x"])
    t64 = tets.astype(np.int64)
    key = np.unique(np.concatenate([t64[:, a] * nv + t64[:, b]
                                    for a in range(4) for b in range(4) if a != b]))
    st = Stream(seed, 41)
    if n_cand is not None and n_cand < len(key):
        key = key[np.argsort(st.u64(len(key)), kind="stable")[:n_cand]]
    v1, v2 = key // nv, key % nv
    off, vt = vertex_tets(tets, nv)
    ring_ids, ring_off, _ = _csr_gather(off, vt, v1)
    # one-ring surface faces per vertex, sorted triples in lexicographic order
    fs = mesh["is_surface_fs"]
    tt, jj = np.nonzero(fs != NOT_SURFACE)
    faces = np.sort(np.stack([tets[tt, (jj + k) % 4] for k in (1, 2, 3)], 1), 1).astype(np.int64)
    assert nv < 2**21, "face keys are packed into int64"
    fkey = np.unique((faces[:, 0] * nv + faces[:, 1]) * nv + faces[:, 2]) if len(faces) else \
        np.zeros(0, np.int64)
    faces = np.stack([fkey // (nv * nv), (fkey // nv) % nv, fkey % nv], 1)
    fv = faces.reshape(-1)
    fi = np.repeat(np.arange(len(faces)), 3)
    o = np.lexsort((fi, fv))  # faces are already lexicographic, so ids ascending = lex order
    foff = np.zeros(nv + 1, dtype=np.int64)
    foff[1:] = np.cumsum(np.bincount(fv, minlength=nv))
    cand_f, f_off, f_cnt = _csr_gather(foff, fi[o], v1)
    owner = np.repeat(np.arange(len(v1)), f_cnt)
    tri = faces[cand_f]
    keep = ~(tri == v2[owner][:, None]).any(1)
    tri, owner = tri[keep], owner[keep]
    tri = np.where(tri == v1[owner][:, None], v2[owner][:, None], tri)
    tri_off = np.zeros(len(v1) + 1, dtype=np.int64)
    tri_off[1:] = np.cumsum(np.bincount(owner, minlength=len(v1)))
    sfv, bdv = mesh["on_surface"], mesh["on_boundary"]
    flags = (sfv[v1] * 1 + bdv[v1] * 2 + sfv[v2] * 4 + bdv[v2] * 8).astype(np.uint8)
    if unrounded_frac > 0:
        flags[st.uniform(len(v1)) < unrounded_frac] |= 16
    return {
        "x": mesh["x"], "y": mesh["y"], "z": mesh["z"],
        "v1": v1.astype(np.int32), "v2": v2.astype(np.int32), "flags": flags,
        "ring_off": ring_off.astype(np.int32),
        "ring_tets": np.ascontiguousarray(tets[ring_ids].astype(np.int32)),
        "old_q": np.ascontiguousarray(np.asarray(q, dtype=np.float64)[ring_ids]),
        "tri_off": tri_off.astype(np.int32),
        "tri_ids": np.ascontiguousarray(tri.astype(np.int32)).reshape(-1, 3),
    }
