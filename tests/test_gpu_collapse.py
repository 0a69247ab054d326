"""Batched edge-collapse candidate evaluation (twg_collapse_batch) against the
oracle's restatement of EdgeCollapser::collapseAnEdge's checks
(src/tetwild/EdgeCollapser.cpp:284-446, 727-775).

Bar: return codes bit-exact for every candidate; new qualities bit-exact for
every candidate the reference would collapse (SUCCESS / ENVELOP_SUC), NaN for
the tets the collapse removes."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tetwild_amd import _capi as C
from tetwild_amd import workloads as W
from tetwild_amd.ops import CollapseBatch

pytestmark = pytest.mark.gpu


def setup(n=6, seed=0, **kw):
    M = W.sheet_mesh(n, seed=seed)
    q = O.tet_qualities(M["tets"], M["x"], M["y"], M["z"])
    B = W.collapse_batch(M, q, seed=seed, **kw)
    return M, B


def trees(ctx, M):
    sf = ctx.surface(M["sf_V"], M["sf_F"])
    bt = ctx.surface(M["b_V"], M["b_F"])
    return sf, bt


def compare(B, st, nq, so, nqo):
    st = st.cpu().numpy()
    nq = nq.cpu().numpy()
    assert np.array_equal(st, so), (np.nonzero(st != so)[0][:10], st[st != so][:10], so[st != so][:10])
    ok = np.isin(so, [C.TWG_COLLAPSE_SUCCESS, C.TWG_COLLAPSE_ENVELOP_SUC])
    ro = B["ring_off"]
    idx = np.concatenate([np.arange(ro[c], ro[c + 1]) for c in np.nonzero(ok)[0]] or
                         [np.zeros(0, int)])
    assert np.array_equal(nq[idx].view(np.int64), nqo[idx].view(np.int64))
    return ok.sum()


@pytest.mark.parametrize("precision", [C.TWG_FP64, C.TWG_MIXED])
@pytest.mark.parametrize("soft", [None, 40.0])
def test_collapse_matches_oracle(ctx, precision, soft):
    M, B = setup()
    h = 1.0 / 6
    eps = 0.05 * h
    sf, bt = trees(ctx, M)
    so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"]),
                               eps * eps, eps, soft_energy=soft)
    batch = CollapseBatch(ctx, B, sf, bt, eps * eps, eps, soft_energy=soft)
    st, nq = batch.run(precision)
    torch.cuda.synchronize()
    n_ok = compare(B, st, nq, so, nqo)
    counts = np.bincount(so, minlength=8)
    print("codes", counts)
    # the workload exercises every return code
    assert (counts[:5] > 0).all() and n_ok > 0
    sf.close()
    bt.close()


def test_collapse_flags_and_switches(ctx):
    """unrounded v1, quality check off, epsilon check off"""
    M, B = setup(n=5, seed=3, unrounded_frac=0.3)
    eps = 0.05 / 5
    sf, bt = trees(ctx, M)
    S, Bt = O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"])
    for cq, ce in ((True, True), (False, True), (True, False), (False, False)):
        so, nqo = O.collapse_batch(B, S, Bt, eps * eps, eps, check_quality=cq, check_eps=ce)
        batch = CollapseBatch(ctx, B, sf, bt, eps * eps, eps, check_quality=cq, check_eps=ce)
        st, nq = batch.run()
        torch.cuda.synchronize()
        compare(B, st, nq, so, nqo)
        if not ce:
            assert not np.isin(so, [C.TWG_COLLAPSE_ENVELOP_SUC]).any()
        if not cq:
            assert not (so == C.TWG_COLLAPSE_QUALITY).any()


def test_collapse_degenerate_edges(ctx):
    """v2 at v1's position: the point, quality and epsilon checks are skipped"""
    M, B = setup(n=5, seed=1)
    x, y, z = B["x"].copy(), B["y"].copy(), B["z"].copy()
    sel = np.arange(0, len(B["v1"]), 7)
    for c in sel:
        a, b = B["v1"][c], B["v2"][c]
        x[b], y[b], z[b] = x[a], y[a], z[a]
    B = dict(B, x=x, y=y, z=z)
    eps = 0.05 / 5
    sf, bt = trees(ctx, M)
    so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"]),
                               eps * eps, eps)
    st, nq = CollapseBatch(ctx, B, sf, bt, eps * eps, eps).run()
    torch.cuda.synchronize()
    compare(B, st, nq, so, nqo)


def test_collapse_bad_input_and_empty(ctx):
    M, B = setup(n=4, seed=2)
    eps = 0.05 / 4
    sf, bt = trees(ctx, M)
    so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"]),
                               eps * eps, eps)
    bad = dict(B)
    bad["v2"] = B["v2"].copy()
    bad["v2"][3] = len(B["x"])          # out of range
    bad["ring_tets"] = B["ring_tets"].copy()
    r = B["ring_off"][5]
    bad["ring_tets"][r, 0] = -1          # bad ring tet
    bad["v2"][8] = bad["v1"][8]          # v1 == v2
    st, nq = CollapseBatch(ctx, bad, sf, bt, eps * eps, eps).run()
    torch.cuda.synchronize()
    st = st.cpu().numpy()
    assert st[3] == st[5] == st[8] == C.TWG_COLLAPSE_BAD_INPUT
    keep = np.ones(len(st), bool)
    keep[[3, 5, 8]] = False
    assert np.array_equal(st[keep], so[keep])
    # surface candidates without a surface tree are bad input, the rest unaffected
    st2, _ = CollapseBatch(ctx, B, None, None, eps * eps, eps).run()
    torch.cuda.synchronize()
    st2 = st2.cpu().numpy()
    needs = (B["flags"] & (C.TWG_CV1_SURFACE | C.TWG_CV1_BOUNDARY)) != 0
    assert (st2[needs] == C.TWG_COLLAPSE_BAD_INPUT).all()
    assert np.array_equal(st2[~needs], so[~needs])
    # empty batch
    E = {k: (v[:0] if k not in ("x", "y", "z", "ring_off", "tri_off") else v) for k, v in B.items()}
    E["ring_off"] = np.zeros(1, np.int32)
    E["tri_off"] = np.zeros(1, np.int32)
    CollapseBatch(ctx, E, sf, bt, eps * eps, eps).run()
    torch.cuda.synchronize()


def test_collapse_larger_mesh_properties(ctx):
    """a 40^3 lattice (384k tets): all directed edges of 200k candidates;
    oracle parity on a sample, FP64 == mixed on all, and the new qualities of
    every collapsible candidate equal twg_tet_qualities of its new tets"""
    M = W.sheet_mesh(40, seed=5)
    tx, ty, tz = (torch.from_numpy(M[k]).cuda() for k in ("x", "y", "z"))
    tt = torch.from_numpy(M["tets"]).cuda()
    q = ctx.tet_qualities(tt, tx, ty, tz)[0].cpu().numpy()
    B = W.collapse_batch(M, q, n_cand=200000, seed=5)
    eps = 0.05 / 40
    sf, bt = trees(ctx, M)
    batch = CollapseBatch(ctx, B, sf, bt, eps * eps, eps, soft_energy=20.0)
    st, nq = batch.run(C.TWG_FP64)
    st, nq = st.clone(), nq.clone()
    stm, nqm = batch.run(C.TWG_MIXED)
    torch.cuda.synchronize()
    assert torch.equal(st, stm) and torch.equal(nq.view(torch.int64), nqm.view(torch.int64))
    s = st.cpu().numpy()
    counts = np.bincount(s, minlength=8)
    print("codes", counts)
    assert counts[C.TWG_COLLAPSE_BAD_INPUT] == 0
    # oracle on a sample
    sample = np.arange(0, len(s), 97)
    sub = {k: B[k] for k in ("x", "y", "z")}
    ro, to = B["ring_off"], B["tri_off"]
    sub["v1"], sub["v2"], sub["flags"] = B["v1"][sample], B["v2"][sample], B["flags"][sample]
    rs = [np.arange(ro[c], ro[c + 1]) for c in sample]
    ts = [np.arange(to[c], to[c + 1]) for c in sample]
    sub["ring_off"] = np.concatenate([[0], np.cumsum([len(r) for r in rs])]).astype(np.int32)
    sub["tri_off"] = np.concatenate([[0], np.cumsum([len(t) for t in ts])]).astype(np.int32)
    ri, ti = np.concatenate(rs), np.concatenate(ts)
    sub["ring_tets"] = B["ring_tets"][ri]
    sub["old_q"] = B["old_q"][ri]
    sub["tri_ids"] = B["tri_ids"][ti] if len(ti) else np.zeros((0, 3), np.int32)
    so, nqo = O.collapse_batch(sub, O.Surface(M["sf_V"], M["sf_F"]),
                               O.Surface(M["b_V"], M["b_F"]), eps * eps, eps, soft_energy=20.0)
    assert np.array_equal(s[sample], so)
    nqs = nq.cpu().numpy()[ri]
    okr = np.repeat(np.isin(so, [0, 4]), [len(r) for r in rs])
    assert np.array_equal(nqs[okr].view(np.int64), nqo[okr].view(np.int64))
    # new qualities of collapsible candidates = calTetQualities of the new tets
    ok = np.isin(s, [0, 4])
    rows = np.repeat(ok, np.diff(ro))
    rt = B["ring_tets"][rows].copy()
    owner = np.repeat(np.arange(len(s)), np.diff(ro))[rows]
    v1, v2 = B["v1"][owner], B["v2"][owner]
    removed = (rt == v2[:, None]).any(1)
    rt = np.where(rt == v1[:, None], v2[:, None], rt)[~removed]
    qn = ctx.tet_qualities(torch.from_numpy(np.ascontiguousarray(rt)).cuda(), tx, ty, tz)[0]
    qn = qn.cpu().numpy()
    got = nq.cpu().numpy()[rows]
    assert np.isnan(got[removed]).all()
    assert np.array_equal(got[~removed].view(np.int64), qn.view(np.int64))
    sf.close()
    bt.close()


@pytest.mark.parametrize("precision", [C.TWG_FP64, C.TWG_MIXED])
def test_collapse_exact_passes(ctx, precision):
    """an unjittered 8^3 lattice (coordinates k/8, exact): collapses along grid
    lines create exactly flat new tets, and some candidates get an exactly
    collinear epsilon triangle (v2, a, 2a - v2) -- cases the filters cannot
    decide, so the exact ring and face passes must give the oracle's codes"""
    M = W.sheet_mesh(8, jitter=0.0, surf_jitter=0.0, seed=0)
    q = O.tet_qualities(M["tets"], M["x"], M["y"], M["z"])
    B = W.collapse_batch(M, q, seed=0)
    x, y, z = (list(B[k]) for k in ("x", "y", "z"))
    tri = B["tri_ids"].copy()
    to = B["tri_off"]
    n_col = 0
    for c in range(0, len(B["v1"]), 3):
        if to[c + 1] > to[c]:
            k = to[c]
            v2, a = B["v2"][c], tri[k][1] if tri[k][0] == B["v2"][c] else tri[k][0]
            x.append(2 * x[a] - x[v2])
            y.append(2 * y[a] - y[v2])
            z.append(2 * z[a] - z[v2])
            tri[k] = (v2, a, len(x) - 1)
            n_col += 1
    B = dict(B, x=np.array(x), y=np.array(y), z=np.array(z), tri_ids=tri)
    P = np.stack([B["x"], B["y"], B["z"]], 1)
    assert n_col > 0 and sum(O.collinear3(*P[t]) for t in tri) >= n_col
    eps = 0.05 / 8
    sf, bt = trees(ctx, M)
    for soft in (None, 1e40):
        so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]),
                                   O.Surface(M["b_V"], M["b_F"]), eps * eps, eps, soft_energy=soft)
        st, nq = CollapseBatch(ctx, B, sf, bt, eps * eps, eps, soft_energy=soft).run(precision)
        torch.cuda.synchronize()
        compare(B, st, nq, so, nqo)
        counts = np.bincount(so, minlength=8)
        print("codes", counts)
        assert counts[C.TWG_COLLAPSE_FLIP] > 0
        if soft is not None:
            assert counts[C.TWG_COLLAPSE_ENVELOP_SUC] > 0
    sf.close()
    bt.close()


def test_collapse_tiles_chunks_autotune_plan(ctx):
    """codes and new qualities do not depend on tile or chunk; the autotuner
    picks a valid pair; a hipGraph plan replayed several times (with eager work
    and synchronisation between replays) gives the eager result every time"""
    M, B = setup(n=7, seed=6)
    eps = 0.05 / 7
    sf, bt = trees(ctx, M)
    so, nqo = O.collapse_batch(B, O.Surface(M["sf_V"], M["sf_F"]), O.Surface(M["b_V"], M["b_F"]),
                               eps * eps, eps, soft_energy=30.0)
    batch = CollapseBatch(ctx, B, sf, bt, eps * eps, eps, soft_energy=30.0)
    for tile in (16, 32, 64):
        for chunk in (1, 4, 16):
            st, nq = batch.run(C.TWG_FP64, tile, chunk)
            torch.cuda.synchronize()
            compare(B, st, nq, so, nqo)
    for prec in (C.TWG_FP64, C.TWG_MIXED):
        t, c, ms = batch.autotune(prec)
        assert t in (16, 32, 64) and c in (1, 4, 16) and ms > 0
        st, nq = batch.run(prec)
        torch.cuda.synchronize()
        compare(B, st, nq, so, nqo)
        eager_st, eager_nq = st.clone(), nq.clone()
        plan = batch.plan(prec)
        for rep in range(3):
            batch.status.fill_(255)
            batch.new_q.fill_(-1.0)
            torch.cuda.synchronize()
            st, nq = plan.run()
            torch.cuda.synchronize()
            assert torch.equal(st, eager_st), rep
            ok = np.isin(so, [0, 4])
            ro = B["ring_off"]
            idx = np.concatenate([np.arange(ro[c], ro[c + 1]) for c in np.nonzero(ok)[0]])
            assert torch.equal(nq[idx].view(torch.int64), eager_nq[idx].view(torch.int64)), rep
            # eager work on the same context between replays
            batch.run(prec, 16, 4)
            torch.cuda.synchronize()
        plan.close()
    sf.close()
    bt.close()
