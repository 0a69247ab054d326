"""The C++ LocalOperations / VertexSmoother mirrors (tetwild_amd/shim) driven
end to end on a small tet mesh, against the oracle replaying the reference's
per-vertex loops (src/tetwild/VertexSmoother.cpp:114-463) in the same order.

Bar: final vertex positions, stored qualities and counters bit-exact with
the replay, for the reference's vertex order (seq) and for the colour-class
order of the batched variants (batch)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tetwild_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tetwild_amd", "shim", "shim_driver")
NOT_SURFACE = 2**31 - 1


def lattice_mesh(n=5, seed=3):
    x, y, z, tets, _ = W.kuhn_lattice(n, jitter=0.15, seed=seed)
    h = 1.0 / n
    P = np.stack([x, y, z], 1)
    # boundary vertices (on the cube's faces before jitter) are surface vertices;
    # keep them on their face plane so the boundary triangles form the envelope
    g = np.round(P / h).astype(int)
    onb = np.zeros(len(P), bool)
    for a in range(3):
        for val in (0, n):
            m = g[:, a] == val
            P[m, a] = val * h
            onb |= m
    fs = np.full((len(tets), 4), NOT_SURFACE, dtype=np.int32)
    tris = []
    for t in range(len(tets)):
        for j in range(4):
            f = [tets[t][(j + k) % 4] for k in (1, 2, 3)]
            for a in range(3):
                for val in (0, n):
                    if all(g[v, a] == val for v in f):
                        fs[t, j] = 0
                        tris.append(P[f].reshape(9))
    flags = onb.astype(np.uint8)  # surface, not bbox: smoothSurface moves them
    return P, tets.astype(np.int32), fs, flags, np.array(tris)


def write_mesh(path, P, tets, fs, flags, q, sf, eps, d):
    with open(path, "wb") as f:
        f.write(struct.pack("<5q", len(P), len(tets), len(sf), 0, 1))
        f.write(struct.pack("<3d", eps, eps * eps, d))
        f.write(np.ascontiguousarray(P, dtype=np.float64).tobytes())
        f.write(flags.astype(np.uint8).tobytes())
        f.write(np.ascontiguousarray(tets, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(fs, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(q, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(sf, dtype=np.float64).tobytes())


def conn(tets, nv):
    c = [[] for _ in range(nv)]
    for t, tt in enumerate(tets):
        for v in tt:
            c[v].append(t)
    return [sorted(set(a)) for a in c]


def tri_ab_of(v, ring, tets, fs):
    tris = set()
    for t in ring:
        for j in range(4):
            if tets[t][j] != v and fs[t][j] != NOT_SURFACE:
                tris.add(tuple(sorted(tets[t][(j + k) % 4] for k in (1, 2, 3))))
    out = []
    for tri in sorted(tris):
        k = tri.index(v)
        out.append((tri[(k + 1) % 3], tri[(k + 2) % 3]))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def colour_classes(verts, C, tets):
    colour = {}
    classes = []
    for v in verts:
        used = {colour[u] for t in C[v] for u in tets[t] if u != v and u in colour}
        c = 0
        while c in used:
            c += 1
        colour[v] = c
        if c == len(classes):
            classes.append([])
        classes[c].append(v)
    return classes


def replay(P, tets, fs, flags, q, S, eps, d, batched):
    """the reference's smoothSingle + smoothSurface loops on the oracle"""
    P = P.copy()
    q = q.copy()
    nv = len(P)
    C = conn(tets, nv)
    counter = suc = 0

    def one(kind, v):
        x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
        ring = C[v]
        B = {"x": x, "y": y, "z": z, "center": np.array([v], np.int32),
             "ring_off": np.array([0, len(ring)], np.int32),
             "ring_tets": tets[ring], "kind": np.array([kind], np.uint8)}
        ab = tri_ab_of(v, ring, tets, fs) if kind else np.zeros((0, 2), np.int32)
        B["tri_ab"] = ab
        B["tri_off"] = np.array([0, len(ab)], np.int32)
        p, s, nq = O.smooth_batch(B, q[ring], S, S, eps * eps, d)
        return ring, p[0], s[0], nq

    interior = [v for v in range(nv) if not flags[v] & 1]
    groups = colour_classes(interior, C, tets) if batched else [[v] for v in interior]
    for grp in groups:
        res = [one(0, v) for v in grp]  # independent within a group: frozen positions
        for v, (ring, p, s, nq) in zip(grp, res):
            counter += 1
            if s in (0, 3):
                suc += 1
            if s == 0:
                P[v] = p
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    q = O.tet_qualities(tets, x, y, z)
    surf = [v for v in range(nv) if flags[v] & 1]
    groups = colour_classes(surf, C, tets) if batched else [[v] for v in surf]
    for grp in groups:
        res = [one(1, v) for v in grp]
        for v, (ring, p, s, nq) in zip(grp, res):
            counter += 1
            if s == 0:
                suc += 1
                P[v] = p
                q[ring] = nq[:len(ring)]
    return P, q, counter, suc


@pytest.mark.parametrize("mode", ["seq", "batch"])
def test_shim_smooth_matches_replay(ctx, tmp_path, mode):
    assert os.path.exists(DRIVER), "build the shim first (make -C tetwild_amd/shim)"
    P, tets, fs, flags, sf = lattice_mesh()
    eps, d = 2e-3, 2e-3
    q = O.tet_qualities(tets, *(np.ascontiguousarray(P[:, k]) for k in range(3)))
    mesh = tmp_path / "mesh.bin"
    out = tmp_path / "out.bin"
    write_mesh(mesh, P, tets, fs, flags, q, sf, eps, d)
    r = subprocess.run([DRIVER, str(mesh), str(out), mode], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    print(r.stdout.strip())
    raw = np.fromfile(out, dtype=np.float64)
    nv, nt = len(P), len(tets)
    Pg = raw[:3 * nv].reshape(nv, 3)
    qg = raw[3 * nv:3 * nv + nt]
    cnt = raw[3 * nv + nt:3 * nv + nt + 2].view(np.int64)
    # oracle replay, projection tree in the device's facet order
    Vs = sf.reshape(-1, 3)
    Fs = np.arange(len(Vs), dtype=np.int32).reshape(-1, 3)
    surf = ctx.surface(Vs, Fs)
    S = O.Surface(Vs, Fs[surf.order()])
    Pr, qr, counter, suc = replay(P, tets, fs, flags, q, S, eps, d, mode == "batch")
    assert np.array_equal(Pg.view(np.int64), Pr.view(np.int64))
    assert np.array_equal(qg.view(np.int64), qr.view(np.int64))
    assert (cnt[0], cnt[1]) == (counter, suc)
    assert suc > 0.5 * counter
    # NewtonsUpdate / getNewEnergy records and the static energy function
    rec = raw[3 * nv + nt + 2:-1].reshape(-1, 15)
    C = conn(tets, nv)
    interior = [v for v in range(nv) if not flags[v] & 1][:len(rec)]
    x, y, z = (np.ascontiguousarray(Pg[:, k]) for k in range(3))
    for v, row in zip(interior, rec):
        ok, E, J, H, X0 = O.newtons_update(tets[C[v]], v, x, y, z)
        assert row[0] == float(ok) and row[1] == E
        assert np.array_equal(row[2:5], J) and np.array_equal(row[5:14], H.reshape(9))
        assert row[14] == O.new_energy(tets[C[v]], x, y, z)
    T = np.concatenate([Pg[tets[0][j]] for j in range(4)])
    assert raw[-1] == O.amips_ejh_T(T)[0][0]
