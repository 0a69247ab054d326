"""GPU parity of (c): batched one-ring smoothing through the C-ABI against the
CPU oracle, per problem.

Bar: status, final position and (moved surface vertices) the ring's new
qualities bit-exact, for every tile width, queue chunk, FP64 and mixed, eager
launch and hipGraph replay."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tetwild_amd import _capi as C
from tetwild_amd import workloads as W
from tetwild_amd.ops import SmoothBatch

pytestmark = pytest.mark.gpu

EPS2, D = 1e-6, 1e-3


@pytest.fixture(scope="module")
def surf_data():
    V, F = W.bumpy_sphere(freq=40, seed=0)
    Vb, Fb = W.bumpy_sphere(freq=16, seed=0)  # a second ("boundary") tree
    return V, F, Vb, Fb


def bits(a):
    return np.ascontiguousarray(a).view(np.int64)


# row-table entries of a 64-lane tile in the face test (seg_rows<64>): a chunk
# of up to 16 ring faces whose rows fit is sampled segment by segment from the
# table, a chunk whose rows do not by the stored-record sampler
ROW_TABLE = 94


def ring_rows(B, d, chunk=16):
    """Rows of each surface problem's first chunk of ring faces, the centre at
    its input position: per face SampleGen's row count M, the height over the
    longest edge in rows of sqrt(3)/2 d (none for faces of at most one sampling
    distance).  The kernel's faces take the projected centre, so this is an
    estimate of which sampler a ring's face test takes."""
    X = np.stack([B["x"], B["y"], B["z"]], 1)
    to, ab, c = B["tri_off"], B["tri_ab"].reshape(-1, 2), B["center"]
    rows = []
    for p in np.nonzero(B["kind"] > 0)[0]:
        k = np.arange(to[p], min(to[p + 1], to[p] + chunk))
        T = np.stack([np.repeat(X[c[p]][None], len(k), 0), X[ab[k, 0]], X[ab[k, 1]]], 1)
        e = np.stack([T[:, 1] - T[:, 0], T[:, 2] - T[:, 1], T[:, 0] - T[:, 2]], 1)
        l2 = (e ** 2).sum(2)
        i = 2 - np.argmax(l2[:, ::-1], 1)  # the last longest edge
        r = np.arange(len(k))
        v0, v1, v2 = T[r, i], T[r, (i + 1) % 3], T[r, (i + 2) % 3]
        e01 = v1 - v0
        foot = e01 * ((v2 - v0) * e01).sum(1, keepdims=True) / l2[r, i][:, None] + v0
        m = np.floor(np.linalg.norm(foot - v2, axis=1) / (np.sqrt(3) / 2 * d)).astype(int)
        m[np.sqrt(l2[r, i]) / d <= 1] = 0
        rows.append(int(m.sum()))
    return np.array(rows)


def check(batch, ref, B):
    p, s, q = (t.cpu().numpy() for t in (batch.p_out, batch.status, batch.new_q))
    rp, rs, rq = ref
    assert np.array_equal(s, rs), f"status mismatch at {np.nonzero(s != rs)[0][:10]}"
    assert np.array_equal(bits(p), bits(rp))
    moved_surf = (rs == 0) & (B["kind"] > 0)
    for i in np.nonzero(moved_surf)[0]:
        r0, r1 = B["ring_off"][i], B["ring_off"][i + 1]
        assert np.array_equal(bits(q[r0:r1]), bits(rq[r0:r1]))


def test_interior_batch_all_tiles(ctx):
    B = W.smoothing_batch(600, seed=11)
    ref = O.smooth_batch(B)
    batch = SmoothBatch(ctx, B)
    for tile in (16, 32, 64, C.TWG_SMOOTH_TILE_SPLIT):
        for chunk in (1, 4):
            for prec in (0, 1):
                print(f"tile={tile} chunk={chunk} precision={prec}", flush=True)
                batch.p_out.zero_()
                batch.status.fill_(255)
                batch.run(precision=prec, tile=tile, chunk=chunk)
                torch.cuda.synchronize()
                check(batch, ref, B)
    assert (ref[1] == 0).mean() > 0.8


def test_surface_batch(ctx, surf_data):
    V, F, Vb, Fb = surf_data
    S = O.Surface(V, F)
    sf = ctx.surface(V, F)
    bt = ctx.surface(Vb, Fb)
    Sb = O.Surface(Vb, Fb[bt.order()])
    S_sf_order = O.Surface(V, F[sf.order()])
    B = W.smoothing_batch(400, seed=12, surface_frac=0.5, V=V, F=F, boundary_frac=0.3)
    q = O.tet_qualities(B["ring_tets"], B["x"], B["y"], B["z"])
    # the oracle projects with trees in the device's facet order (ties)
    ref = O.smooth_batch(B, q, S_sf_order, Sb, EPS2, D)
    batch = SmoothBatch(ctx, B, sf=sf, bt=bt, eps2=EPS2, sampling_dist=D, old_q=q)
    for tile, prec in ((16, 0), (32, 1), (64, 0), (C.TWG_SMOOTH_TILE_SPLIT, 1),
                       (C.TWG_SMOOTH_TILE_SPLIT, 0)):
        print(f"tile={tile} precision={prec}", flush=True)
        batch.status.fill_(255)
        batch.new_q.zero_()
        batch.run(precision=prec, tile=tile, chunk=2)
        torch.cuda.synchronize()
        check(batch, ref, B)
    st = ref[1]
    print("statuses", np.bincount(st, minlength=8))
    assert ((st == 0) & (B["kind"] == 1)).any() and ((st == 0) & (B["kind"] == 2)).any()
    # the 64-lane tiles sampled these rings from the row table
    assert (ring_rows(B, D) <= ROW_TABLE - 20).mean() > 0.95
    del S


def test_surface_batch_dense_samples(ctx, surf_data):
    """A sampling distance a quarter of the other tests': the rings' rows
    overflow the 64-lane tile's row table, so its face test takes the
    stored-record sampler; eps small enough that some rings leave the
    envelope."""
    V, F, _, _ = surf_data
    sf = ctx.surface(V, F)
    d, eps2 = 2.5e-4, 9e-8
    B = W.smoothing_batch(160, seed=15, surface_frac=1.0, V=V, F=F)
    assert (ring_rows(B, d) > ROW_TABLE + 20).mean() > 0.9
    q = O.tet_qualities(B["ring_tets"], B["x"], B["y"], B["z"])
    S = O.Surface(V, F[sf.order()])
    ref = O.smooth_batch(B, q, S, S, eps2, d)
    st = ref[1]
    print("statuses", np.bincount(st, minlength=8))
    assert (st == 0).sum() > 50 and (st == C.TWG_SMOOTH_STATUS.index("REJECT_ENVELOPE")).sum() > 3
    batch = SmoothBatch(ctx, B, sf=sf, bt=sf, eps2=eps2, sampling_dist=d, old_q=q)
    for tile, prec in ((64, 0), (64, 1), (C.TWG_SMOOTH_TILE_SPLIT, 0), (32, 0)):
        print(f"tile={tile} precision={prec}", flush=True)
        batch.status.fill_(255)
        batch.new_q.zero_()
        batch.run(precision=prec, tile=tile, chunk=2)
        torch.cuda.synchronize()
        check(batch, ref, B)


def test_edge_statuses(ctx, surf_data):
    B = W.smoothing_batch(40, seed=13)
    # problem 0: invert one ring tet -> INVALID
    r0 = B["ring_off"][0]
    B["ring_tets"][r0, [1, 2]] = B["ring_tets"][r0, [2, 1]]
    # problem 1: centre already at the Newton result -> NOT_MOVED (or a tiny move)
    p1 = O.smooth_batch({k: v for k, v in B.items()})[0][1]
    c1 = B["center"][1]
    B["x"][c1], B["y"][c1], B["z"][c1] = p1
    ref = O.smooth_batch(B)  # before the out-of-range edits below
    # problem 2: vertex id out of range -> BAD_INPUT
    B["ring_tets"][B["ring_off"][2] + 1, 3] = len(B["x"]) + 5
    # problem 3: centre id out of range -> BAD_INPUT
    B["center"][3] = -7
    good = np.ones(40, bool)
    good[[2, 3]] = False
    batch = SmoothBatch(ctx, B)
    batch.run(tile=16)
    torch.cuda.synchronize()
    st = batch.status.cpu().numpy()
    assert st[0] == 1 and st[2] == 7 and st[3] == 7
    assert st[1] in (0, 2)
    assert np.array_equal(st[good], ref[1][good])
    assert np.array_equal(bits(batch.p_out.cpu().numpy()[good]), bits(ref[0][good]))


def test_graph_plan_replays(ctx, surf_data):
    V, F, _, _ = surf_data
    sf = ctx.surface(V, F)
    B = W.smoothing_batch(500, seed=14, surface_frac=0.2, V=V, F=F)
    S = O.Surface(V, F[sf.order()])
    q = O.tet_qualities(B["ring_tets"], B["x"], B["y"], B["z"])
    ref = O.smooth_batch(B, q, S, S, EPS2, D)
    batch = SmoothBatch(ctx, B, sf=sf, bt=sf, eps2=EPS2, sampling_dist=D, old_q=q)
    for tile, chunk in ((32, 4), (C.TWG_SMOOTH_TILE_SPLIT, 2)):
        print(f"plan tile={tile} chunk={chunk}", flush=True)
        plan = batch.plan(precision=1, tile=tile, chunk=chunk)
        for rep in range(3):
            batch.status.fill_(255)
            batch.p_out.zero_()
            torch.cuda.synchronize()
            plan.run()
            torch.cuda.synchronize()
            check(batch, ref, B)
            # eager work between replays
            batch.run(precision=0, tile=16, chunk=1)
            torch.cuda.synchronize()
            check(batch, ref, B)
        plan.close()


def test_newtons_method_only(ctx):
    """kind 3: NewtonsMethod alone (is_moved and its returned position)"""
    B = W.smoothing_batch(300, seed=16)
    B["kind"][:] = 3
    ref = O.smooth_batch(B)
    batch = SmoothBatch(ctx, B)
    batch.run(tile=32)
    torch.cuda.synchronize()
    assert np.array_equal(batch.status.cpu().numpy(), ref[1])
    assert np.array_equal(bits(batch.p_out.cpu().numpy()), bits(ref[0]))


def test_autotune_and_empty(ctx):
    B = W.smoothing_batch(300, seed=15)
    ref = O.smooth_batch(B)
    batch = SmoothBatch(ctx, B)
    tile, chunk, ms = batch.autotune()
    assert tile in (16, 32, 64) and chunk in (1, 4, 16) and ms > 0
    torch.cuda.synchronize()
    check(batch, ref, B)
    E = W.smoothing_batch(1, seed=1)
    E0 = {k: (v[:0] if k in ("center", "kind") else v) for k, v in E.items()}
    E0["ring_off"] = E["ring_off"][:1]
    E0["tri_off"] = E["tri_off"][:1]
    SmoothBatch(ctx, E0).run()
    # a NULL descriptor is an argument error, not a crash
    L = C.lib()
    assert L.twg_smooth_batch(ctx.h, None, 0, 0, 0) == C.TWG_EINVAL
    assert L.twg_smooth_autotune(ctx.h, None, 0, None, None, None) == C.TWG_EINVAL


def test_envelope_rejection_batch(ctx):
    """smoothSurface's one-ring face test (src/tetwild/VertexSmoother.cpp:
    412-434) with an envelope tight enough (eps = 6e-5 on one-rings of 0.01)
    that it rejects some moves: statuses (REJECT_ENVELOPE among them),
    positions and qualities equal the oracle's for every tile, FP64 and mixed"""
    V, F = W.bumpy_sphere(freq=24, seed=0)
    sf = ctx.surface(V, F)
    S = O.Surface(V, F[sf.order()])
    B = W.smoothing_batch(120, seed=7, surface_frac=0.4, V=V, F=F, boundary_frac=0.3)
    q = O.tet_qualities(B["ring_tets"], B["x"], B["y"], B["z"])
    eps2, d = 6e-5 ** 2, 1e-3
    ref = O.smooth_batch(B, q, S, S, eps2, d)
    counts = np.bincount(ref[1], minlength=8)
    print("statuses", counts)
    assert counts[6] > 0 and ((ref[1] == 0) & (B["kind"] > 0)).any()
    batch = SmoothBatch(ctx, B, sf=sf, bt=sf, eps2=eps2, sampling_dist=d, old_q=q)
    for tile in (16, 32, 64):
        for prec in (0, 1):
            print(f"tile={tile} precision={prec}", flush=True)
            batch.status.fill_(255)
            batch.new_q.zero_()
            batch.run(precision=prec, tile=tile, chunk=1)
            torch.cuda.synchronize()
            check(batch, ref, B)


def test_newton_step_walks_to_newtons_method(ctx, dev):
    """twg_newton_step: its NewtonsUpdate equals the oracle's, and the host
    halving loop over its candidates (no tet skipped) ends where the batch
    kernel's NewtonsMethod (kind 3) and the oracle end, bit for bit"""
    B = W.smoothing_batch(120, seed=21)
    B3 = dict(B)
    B3["kind"] = np.full(len(B["center"]), 3, np.uint8)
    rp, rs, _ = O.smooth_batch(B3)
    x, y, z = (torch.from_numpy(np.ascontiguousarray(B[k])).to(dev) for k in "xyz")
    for p in range(len(B["center"])):
        r0, r1 = int(B["ring_off"][p]), int(B["ring_off"][p + 1])
        v = int(B["center"][p])
        rt_np = np.ascontiguousarray(B["ring_tets"][r0:r1])
        rt = torch.from_numpy(rt_np).to(dev)
        pos = np.array([B["x"][v], B["y"][v], B["z"][v]])
        moved = False
        for step in range(15):
            o = ctx.newton_step(rt, v, x, y, z, pos)
            assert o.ok >= 0
            if step == 0:
                ok, E, J, H, _ = O.newtons_update(rt_np, v, B["x"], B["y"], B["z"])
                assert bool(o.ok) == bool(ok) and o.energy == E
                assert np.array_equal(np.array(o.J), J) and np.array_equal(np.array(o.H), H.reshape(9))
            if not o.ok:
                break
            taken = tiny = False
            for k in range(20):
                if not o.finite[k] or o.flip[k]:
                    continue
                ne = o.new_energy[k]
                if ne >= o.energy or np.isinf(ne) or np.isnan(ne):
                    continue
                taken, tiny = True, abs(ne - o.energy) < 1e-5
                pos = np.array(o.X[k])
                break
            if taken and tiny:
                taken = False
            if not taken:
                moved = step != 0
                break
            moved = True
        assert moved == (rs[p] == 0), p
        assert np.array_equal(bits(pos), bits(rp[p])), p
