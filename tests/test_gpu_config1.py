"""The end-to-end CPU workload -- a subdivided icosphere (2562 vertices, 5120
triangles) jittered radially by N(0, 0.002), eps = 1e-3 * bbox diagonal,
ideal edge length diagonal / 20 -- through the HIP path.

The reference pipeline cannot be built here, so its own energies, decisions
and output tet count are not measurable.  What runs is every in-scope
operator of the refinement loop (src/tetwild/MeshRefinement.cpp:120-183) on a
tet mesh embedding that surface (workloads.icosphere_tetmesh: Delaunay of the
surface vertices and a lattice of spacing L): the C++ VertexSmoother mirror
(smoothSingle + smoothSurface, vertex order and colour-class order), the
EdgeCollapser checks over every directed edge, the EdgeRemover decisions over
every interior edge of 3..5 tets, the energy reductions and the envelope tests
of every surface face.  Each is compared with the oracle replaying the
reference's loops on the same inputs: positions, stored qualities, counters,
return codes and decisions bit-exact."""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from tetwild_amd import workloads as W

from test_gpu_shim import DRIVER, NOT_SURFACE, conn, replay

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mesh():
    M = W.icosphere_tetmesh(levels=4, sigma=0.002, seed=0)
    M["P"] = np.stack([M["x"], M["y"], M["z"]], 1)
    M["q"] = O.tet_qualities(M["tets"], M["x"], M["y"], M["z"])
    M["flags"] = (M["on_surface"] * 1 + M["on_bbox"] * 2 + M["on_boundary"] * 4).astype(np.uint8)
    return M


def write_mesh(path, M, q=None, flags=None):
    P, tets = M["P"], M["tets"]
    sf = M["sf_V"][M["sf_F"]].reshape(-1, 9)
    bf = M["b_V"][M["b_F"]].reshape(-1, 9)
    q = M["q"] if q is None else q
    flags = M["flags"] if flags is None else flags
    with open(path, "wb") as f:
        f.write(struct.pack("<5q", len(P), len(tets), len(sf), len(bf), 1))
        f.write(struct.pack("<3d", M["eps"], M["eps_2"], M["sampling_dist"]))
        f.write(np.ascontiguousarray(P, dtype=np.float64).tobytes())
        f.write(flags.astype(np.uint8).tobytes())
        f.write(np.ascontiguousarray(tets, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(M["is_surface_fs"], dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(q, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(sf).tobytes())
        f.write(np.ascontiguousarray(bf).tobytes())


def run_driver(tmp_path, M, mode, **kw):
    assert os.path.exists(DRIVER), "build the shim first (make -C tetwild_amd/shim)"
    mesh = tmp_path / "mesh.bin"
    out = tmp_path / "out.bin"
    write_mesh(mesh, M, **kw)
    r = subprocess.run([DRIVER, str(mesh), str(out), mode], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    print(mode, r.stdout.strip(), flush=True)
    return r.stdout, out


def test_config1_mesh(mesh):
    M = mesh
    assert len(M["sf_V"]) == 2562 and len(M["sf_F"]) == 5120
    assert abs(M["diag"] - 3.4643) < 1e-3
    assert abs(M["eps"] - M["eps_input"] * (1 - 1 / np.sqrt(3))) <= 1e-18
    assert O.tet_flip(M["tets"], M["x"], M["y"], M["z"]).sum() == 0
    # every input triangle is a tet face (asserted by the generator); count them
    assert (M["is_surface_fs"] != NOT_SURFACE).sum() == 2 * 5120


@pytest.mark.parametrize("mode", ["seq", "serial", "batch"])
def test_config1_smoothing(ctx, mesh, tmp_path, mode):
    """smooth() (smoothSingle then smoothSurface) on the whole mesh"""
    M = mesh
    P, tets, q, flags = M["P"], M["tets"], M["q"], M["flags"]
    stdout, out = run_driver(tmp_path, M, mode)
    info = json.loads(stdout.strip().splitlines()[-1])
    if mode == "seq":
        # the vertex order in far fewer launches than vertices
        assert 0 < info["levels_single"] + info["levels_surface"] < info["counter"] / 4
    raw = np.fromfile(out, dtype=np.float64)
    nv, nt = len(P), len(tets)
    Pg = raw[:3 * nv].reshape(nv, 3)
    qg = raw[3 * nv:3 * nv + nt]
    cnt = raw[3 * nv + nt:3 * nv + nt + 2].view(np.int64)
    Vs = M["sf_V"][M["sf_F"]].reshape(-1, 3)
    Fs = np.arange(len(Vs), dtype=np.int32).reshape(-1, 3)
    surf = ctx.surface(Vs, Fs)
    S = O.Surface(Vs, Fs[surf.order()])
    Pr, qr, counter, suc = replay(P, tets, M["is_surface_fs"], flags, q, S, M["eps"],
                                  M["sampling_dist"], mode == "batch")
    moved = int((Pr != P).any(1).sum())
    print(f"config1 {mode}: counter {counter} suc {suc} moved {moved} "
          f"max E {q.max():.6g} -> {qr.max():.6g}, avg {q.mean():.6g} -> {qr.mean():.6g}")
    assert np.array_equal(Pg.view(np.int64), Pr.view(np.int64))
    assert np.array_equal(qg.view(np.int64), qr.view(np.int64))
    assert (cnt[0], cnt[1]) == (counter, suc)
    assert suc > 0 and qr.mean() < q.mean()


def test_config1_collapse(mesh, tmp_path):
    """EdgeCollapser::checkCollapses over every directed edge"""
    M = mesh
    _, out = run_driver(tmp_path, M, "collapse")
    B = W.collapse_batch(M, M["q"])
    n = len(B["v1"])
    raw = open(out, "rb").read()
    codes = np.frombuffer(raw[:4 * n], dtype=np.int32)
    cnt = np.frombuffer(raw[4 * n:8 * n], dtype=np.int32)
    nq = np.frombuffer(raw[8 * n:8 * n + 8 * int(cnt.sum())], dtype=np.float64)
    so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"]),
                               M["eps_2"], M["sampling_dist"])
    assert np.array_equal(codes, so.astype(np.int32))
    ok = np.isin(so, [0, 4])
    ro = B["ring_off"]
    expect = [nqo[ro[c]:ro[c + 1]] for c in np.nonzero(ok)[0]]
    expect = np.concatenate([e[~np.isnan(e)] for e in expect]) if expect else np.zeros(0)
    assert np.array_equal(nq.view(np.int64), expect.view(np.int64))
    counts = np.bincount(so, minlength=5)
    print("config1 collapse codes (success, flip, quality, envelope, envelope_suc):", counts)
    assert counts[0] > 0 and counts[1] > 0 and counts[3] > 0


def test_config1_remove(mesh, tmp_path):
    """EdgeRemover::checkRemovals over every interior edge of 3..5 tets"""
    M = mesh
    P, tets, q = M["P"], M["tets"], M["q"]
    _, out = run_driver(tmp_path, M, "remove")
    raw = open(out, "rb").read()
    x, y, z = M["x"], M["y"], M["z"]
    C = conn(tets, len(P))
    edges = sorted({(min(a, b), max(a, b)) for t in tets for i, a in enumerate(t) for b in t[i + 1:]})
    eb, pos, kinds = 100, 0, []
    for (a, b) in edges:
        ring = sorted(set(C[a]) & set(C[b]))
        if not 3 <= len(ring) <= 5:
            continue
        others = [v for t in ring for v in tets[t] if v not in (a, b)]
        if any(others.count(v) != 2 for v in set(others)):
            continue
        code, nt, nq, eb = O.remove_edge(a, b, tets[ring], q[ring], x, y, z, equal_buget=eb)
        kind, k = np.frombuffer(raw[pos:pos + 8], np.int32)
        pos += 8
        rec = raw[pos:pos + k * 24]
        pos += k * 24
        assert kind == code, (a, b, kind, code)
        got_t = np.array([np.frombuffer(rec[24 * i:24 * i + 16], np.int32) for i in range(k)]).reshape(-1, 4)
        got_q = np.array([np.frombuffer(rec[24 * i + 16:24 * i + 24], np.float64)[0] for i in range(k)])
        assert np.array_equal(got_t, nt) and np.array_equal(got_q.view(np.int64), nq.view(np.int64))
        kinds.append(code)
    assert np.frombuffer(raw[pos:pos + 4], np.int32)[0] == eb and pos + 4 == len(raw)
    kinds = np.array(kinds)
    print("config1 removal kinds (none, -, -, 3-2, 4-4, 5-6):", np.bincount(kinds, minlength=6))
    assert (kinds == 0).any() and (kinds > 0).any()


def test_config1_energy(mesh, tmp_path):
    """getAvgMaxEnergy / getMaxEnergy / getSecondMaxEnergy / getFilterEnergy"""
    M = mesh
    out, _ = run_driver(tmp_path, M, "energy")
    got = json.loads(out)
    tets, q = M["tets"], M["q"]
    removed = (np.arange(len(tets)) % 7 == 3).astype(np.uint8)
    locked = np.zeros(len(M["P"]), np.uint8)
    s, m, m2, cnt, b = O.energy_stats(q, removed, tets, locked, thres=10.0)
    fe, clean = O.filter_energy(b, thres=10.0)
    assert got["max"] == m and got["second"] == m2
    assert got["filter"] == fe and bool(got["clean"]) == bool(clean)
    assert abs(got["avg"] - s / cnt) <= (len(q) + 64) * 2.0**-53 * np.abs(q).sum() / cnt


def test_config1_envelope_and_quality(ctx, dev, mesh):
    """isFaceOutEnvelop on every tet face with three surface vertices and
    isPointOutEnvelop on every vertex, against the input surface; the
    calTetQualities pass over all tets"""
    import torch
    M = mesh
    tets = M["tets"]
    surf = ctx.surface(M["sf_V"], M["sf_F"])
    S = O.Surface(M["sf_V"], M["sf_F"])
    on = M["on_surface"]
    faces = set()
    for j in range(4):
        f = np.sort(np.stack([tets[:, (j + k) % 4] for k in (1, 2, 3)], 1), 1)
        faces |= set(map(tuple, f[on[f].all(1)]))
    faces = np.array(sorted(faces), dtype=np.int64)
    Q = np.ascontiguousarray(M["P"][faces].reshape(-1, 9))
    out = ctx.envelope_faces(surf, torch.from_numpy(Q).to(dev), M["eps_2"], M["sampling_dist"])
    pts = np.ascontiguousarray(M["P"])
    pout = ctx.envelope_points(surf, torch.from_numpy(pts).to(dev), M["eps_2"])
    tt = torch.from_numpy(tets).to(dev)
    tx, ty, tz = (torch.from_numpy(M[k]).to(dev) for k in "xyz")
    qg, _ = ctx.tet_qualities(tt, tx, ty, tz)
    ctx.sync()
    fo, _ = S.face_out(Q, M["eps_2"], M["sampling_dist"])
    po = S.point_out(pts, M["eps_2"])
    assert np.array_equal(out.cpu().numpy(), fo)
    assert np.array_equal(pout.cpu().numpy(), po)
    assert np.array_equal(qg.cpu().numpy().view(np.int64), M["q"].view(np.int64))
    print(f"config1 envelope: {len(Q)} surface-vertex faces, {int(fo.sum())} out; "
          f"{len(pts)} vertices, {int(po.sum())} out; {len(tets)} tets")
    # the input triangles themselves lie in the envelope; faces spanning the
    # surface from outside it do not
    assert fo.sum() < len(fo) and po[on].sum() == 0 and po[~on].sum() > 0
