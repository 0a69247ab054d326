"""Host-side mirror of the C-ABI over torch device tensors.

Context wraps one twg_ctx and enqueues every call on torch's current stream
of its device, so torch allocations, copies and events order correctly with
the HIP kernels.  Every method checks shapes, dtypes, devices and contiguity
before it reaches the library (the kernels index exactly the extents passed).
"""
import ctypes

import torch

from . import _capi as C


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("tetwild_amd needs a HIP device (torch.cuda.is_available() is False)")
        self.L = C.lib()
        self.device = torch.device("cuda", device)
        h = ctypes.c_void_p()
        C.check(self.L.twg_ctx_create(device, ctypes.byref(h)), "twg_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None:
            self.L.twg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -------------------------------------------------------
    def _bind(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        C.check(self.L.twg_ctx_set_stream(self.h, ctypes.c_void_p(s)), "twg_ctx_set_stream")

    def _dev(self, t, dtype, name, shape=None):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if t.device != self.device:
            raise ValueError(f"{name} must live on {self.device}, got {t.device}")
        if t.dtype != dtype:
            raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        if shape is not None:
            for k, (a, b) in enumerate(zip(t.shape, shape)):
                if b is not None and a != b:
                    raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
            if len(t.shape) != len(shape):
                raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")
        return t

    def _amips_out(self, out, n, mode, gdtype, mixed):
        """caller-supplied outputs of amips_eval / amips_eval_mixed: E (n)
        float64, G (n, 3) and H (n, 9) of gdtype where the mode writes them,
        fallback (n) uint8 for the mixed form; checked like the inputs, since
        the library writes exactly n rows into each"""
        want = 4 if mixed else 3
        if not isinstance(out, (tuple, list)) or len(out) != want:
            raise ValueError(f"out must be a tuple of {want} tensors")
        E, G, H = out[0], out[1], out[2]
        self._dev(E, torch.float64, "out E", (n,))
        if mode >= C.TWG_AMIPS_EG:
            if G is None:
                raise ValueError("out G is required for this mode")
            self._dev(G, gdtype, "out G", (n, 3))
        if mode >= C.TWG_AMIPS_EGH:
            if H is None:
                raise ValueError("out H is required for this mode")
            self._dev(H, gdtype, "out H", (n, 6 if mode == C.TWG_AMIPS_EGH_SYM else 9))
        if mixed and out[3] is not None:
            self._dev(out[3], torch.uint8, "out fallback", (n,))
        return out

    def _verts(self, x, y, z):
        self._dev(x, torch.float64, "x", (None,))
        self._dev(y, torch.float64, "y", (x.shape[0],))
        self._dev(z, torch.float64, "z", (x.shape[0],))
        return x.shape[0]

    def _tets(self, tets):
        self._dev(tets, torch.int32, "tets", (None, 4))
        if tets.data_ptr() % 16:
            raise ValueError("tets must be 16-byte aligned")
        return tets.shape[0]

    def sync(self):
        torch.cuda.synchronize(self.device)

    # -- (a) -----------------------------------------------------------
    def amips_eval(self, tets, x, y, z, mode=C.TWG_AMIPS_EGH, out=None):
        """E (n), G (n, 3), H (n, 9) float64, gradient/Hessian w.r.t. vertex 0
        (mode TWG_AMIPS_EGH_SYM: H (n, 6), the upper triangle row by row)."""
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        dev = self.device
        if out is None:
            hw = 6 if mode == C.TWG_AMIPS_EGH_SYM else 9
            E = torch.empty(n, dtype=torch.float64, device=dev)
            G = torch.empty((n, 3), dtype=torch.float64, device=dev) if mode >= 1 else None
            H = torch.empty((n, hw), dtype=torch.float64, device=dev) if mode >= 2 else None
        else:
            E, G, H = self._amips_out(out, n, mode, torch.float64, False)
        self._bind()
        C.check(self.L.twg_amips_eval(self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv, mode,
                                      _ptr(E), _ptr(G), _ptr(H)), "twg_amips_eval")
        return E, G, H

    def amips_launcher(self, tets, x, y, z, mode, out):
        """A zero-argument callable that enqueues twg_amips_eval with arguments
        checked once (for timing loops: no per-call Python validation)."""
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        E, G, H = self._amips_out(out, n, mode, torch.float64, False)
        self._bind()
        f = self.L.twg_amips_eval
        args = (self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv, mode, _ptr(E), _ptr(G),
                _ptr(H))

        def launch():
            r = f(*args)
            if r:
                C.check(r, "twg_amips_eval")
        return launch

    def amips_eval_mixed(self, tets, x, y, z, mode=C.TWG_AMIPS_EGH, tol=2.0**-16, out=None):
        """E (n) float64, G (n, 3) / H (n, 9) float32 (mode TWG_AMIPS_EGH_SYM:
        H (n, 6)), fallback (n) uint8."""
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        dev = self.device
        if out is None:
            hw = 6 if mode == C.TWG_AMIPS_EGH_SYM else 9
            E = torch.empty(n, dtype=torch.float64, device=dev)
            G = torch.empty((n, 3), dtype=torch.float32, device=dev) if mode >= 1 else None
            H = torch.empty((n, hw), dtype=torch.float32, device=dev) if mode >= 2 else None
            fb = torch.empty(n, dtype=torch.uint8, device=dev)
        else:
            E, G, H, fb = self._amips_out(out, n, mode, torch.float32, True)
        self._bind()
        C.check(self.L.twg_amips_eval_mixed(self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv,
                                            mode, float(tol), _ptr(E), _ptr(G), _ptr(H), _ptr(fb)),
                "twg_amips_eval_mixed")
        return E, G, H, fb

    def tet_qualities(self, tets, x, y, z, precision=C.TWG_FP64, tol=2.0**-16):
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        out = torch.empty(n, dtype=torch.float64, device=self.device)
        fb = torch.empty(n, dtype=torch.uint8, device=self.device)
        self._bind()
        C.check(self.L.twg_tet_qualities(self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv,
                                         precision, float(tol), _ptr(out), _ptr(fb)),
                "twg_tet_qualities")
        return out, fb

    def tet_flip(self, tets, x, y, z):
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        out = torch.empty(n, dtype=torch.uint8, device=self.device)
        self._bind()
        C.check(self.L.twg_tet_flip(self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv,
                                    _ptr(out)), "twg_tet_flip")
        return out

    def tet_groups(self, tets, group_off, x, y, z, want_q=True):
        """isFlip / calTetQualities / getCheckQuality per group of tets:
        returns (flipped[g] uint8, max_q[g], q[t] or None) as device tensors"""
        n_tets = self._tets(tets)
        nv = self._verts(x, y, z)
        group_off = self._dev(group_off, torch.int32, "group_off")
        ng = group_off.shape[0] - 1
        fl = torch.empty(max(ng, 1), dtype=torch.uint8, device=self.device)
        mq = torch.empty(max(ng, 1), dtype=torch.float64, device=self.device)
        q = torch.empty(max(n_tets, 1), dtype=torch.float64, device=self.device) if want_q else None
        self._bind()
        C.check(self.L.twg_tet_groups(self.h, _ptr(tets), _ptr(group_off), ng, n_tets, _ptr(x),
                                      _ptr(y), _ptr(z), nv, _ptr(fl), _ptr(mq),
                                      _ptr(q) if q is not None else None), "twg_tet_groups")
        return fl[:ng], mq[:ng], (q[:n_tets] if q is not None else None)

    def list_stats(self, enable):
        """enable=True: start counting the waves of the envelope list kernels;
        enable=False: stop and return (waves that tested their LDS facet
        list, waves that walked the tree)"""
        out = (ctypes.c_ulonglong * 2)()
        self._bind()
        C.check(self.L.twg_ctx_list_stats(self.h, 1 if enable else 0, out), "twg_ctx_list_stats")
        return None if enable else (int(out[0]), int(out[1]))

    def newton_step(self, ring_tets, v, x, y, z, p, skip=None):
        """One NewtonsMethod step of vertex v at p over ring_tets (device
        tensors): NewtonsUpdate and the twenty line-search candidates with
        their finiteness, flip test (ring tets with skip == 0) and energy.
        Returns a _capi.NewtonStepOut (host copy)."""
        m = self._tets(ring_tets)
        nv = self._verts(x, y, z)
        if not 0 <= int(v) < nv:
            raise ValueError("newton_step: v out of range")
        if skip is not None:
            skip = self._dev(skip, torch.uint8, "skip", (m,))
        out = torch.empty(ctypes.sizeof(C.NewtonStepOut), dtype=torch.uint8, device=self.device)
        self._bind()
        C.check(self.L.twg_newton_step(self.h, _ptr(x), _ptr(y), _ptr(z), nv, _ptr(ring_tets), m,
                                       int(v), float(p[0]), float(p[1]), float(p[2]),
                                       _ptr(skip) if skip is not None else None, _ptr(out)),
                "twg_newton_step")
        host = out.cpu().numpy()
        return C.NewtonStepOut.from_buffer_copy(host.tobytes())

    def energy_stats(self, slim_energy, removed=None, tets=None, locked=None, thres=10.0):
        """(sum, max, second_max, count, buckets[11]) over device arrays"""
        self._dev(slim_energy, torch.float64, "slim_energy", (None,))
        n = slim_energy.shape[0]
        if removed is not None:
            self._dev(removed, torch.uint8, "removed", (n,))
        if (tets is None) != (locked is None):
            raise ValueError("tets and locked go together")
        if tets is not None:
            if self._tets(tets) != n:
                raise ValueError("tets must have one row per energy")
            self._dev(locked, torch.uint8, "locked", (None,))
        nv = locked.shape[0] if locked is not None else 0
        out = C.EnergySummary()
        self._bind()
        C.check(self.L.twg_energy_stats(self.h, _ptr(slim_energy), n,
                                        _ptr(removed) if removed is not None else None,
                                        _ptr(tets) if tets is not None else None,
                                        _ptr(locked) if locked is not None else None, nv,
                                        float(thres), ctypes.byref(out)), "twg_energy_stats")
        return out.sum, out.max, out.second_max, out.count, list(out.buckets)

    # -- (b) -----------------------------------------------------------
    def surface(self, V, F):
        """Device AABB tree over a triangle surface (numpy V (nv,3) float64,
        F (nf,3) int32 on the host)."""
        return Surface(self, V, F)

    def envelope_points(self, surf, p3, eps2, precision=C.TWG_FP64, stats=False):
        self._dev(p3, torch.float64, "p3", (None, 3))
        n = p3.shape[0]
        out = torch.empty(n, dtype=torch.uint8, device=self.device)
        st = torch.zeros(4, dtype=torch.int64, device=self.device) if stats else None
        self._bind()
        C.check(self.L.twg_envelope_points(self.h, surf.h, _ptr(p3), n, float(eps2), precision,
                                           _ptr(out), _ptr(st)), "twg_envelope_points")
        return (out, st) if stats else out

    def envelope_faces(self, surf, tri9, eps2, sampling_dist, precision=C.TWG_FP64, stats=False,
                       out=None):
        self._dev(tri9, torch.float64, "tri9", (None, 9))
        n = tri9.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.uint8, device=self.device)
        else:
            self._dev(out, torch.uint8, "out", (n,))
        st = torch.zeros(4, dtype=torch.int64, device=self.device) if stats else None
        self._bind()
        C.check(self.L.twg_envelope_faces(self.h, surf.h, _ptr(tri9), n, float(eps2),
                                          float(sampling_dist), precision, _ptr(out), _ptr(st)),
                "twg_envelope_faces")
        return (out, st) if stats else out

    def nearest(self, surf, p3):
        """closest point (n,3), squared distance (n), facet tree position (n)"""
        self._dev(p3, torch.float64, "p3", (None, 3))
        n = p3.shape[0]
        cp = torch.empty((n, 3), dtype=torch.float64, device=self.device)
        d2 = torch.empty(n, dtype=torch.float64, device=self.device)
        fid = torch.empty(n, dtype=torch.int32, device=self.device)
        self._bind()
        C.check(self.L.twg_nearest_points(self.h, surf.h, _ptr(p3), n, _ptr(cp), _ptr(d2),
                                          _ptr(fid)), "twg_nearest_points")
        return cp, d2, fid

    def sample_triangles(self, tri9, d):
        """sampleTriangle for every row: (counts (n), offsets (n), pts (total, 3))"""
        self._dev(tri9, torch.float64, "tri9", (None, 9))
        n = tri9.shape[0]
        counts = torch.empty(n, dtype=torch.int64, device=self.device)
        self._bind()
        C.check(self.L.twg_sample_count(self.h, _ptr(tri9), n, float(d), _ptr(counts)),
                "twg_sample_count")
        offsets = torch.cumsum(counts, 0) - counts
        total = int(counts.sum().item()) if n else 0
        pts = torch.empty((max(total, 1), 3), dtype=torch.float64, device=self.device)
        self._bind()
        C.check(self.L.twg_sample_write(self.h, _ptr(tri9), n, float(d), _ptr(offsets), total,
                                        _ptr(pts)), "twg_sample_write")
        return counts, offsets, pts[:total]


class Surface:
    def __init__(self, ctx, V, F):
        import numpy as np
        self.ctx = ctx
        self.V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        self.F = np.ascontiguousarray(F, dtype=np.int32).reshape(-1, 3)
        h = ctypes.c_void_p()
        C.check(ctx.L.twg_bvh_build(ctx.h, self.V.ctypes.data_as(ctypes.c_void_p), self.V.shape[0],
                                    self.F.ctypes.data_as(ctypes.c_void_p), self.F.shape[0],
                                    ctypes.byref(h)), "twg_bvh_build")
        self.h = h
        nf = ctypes.c_int64()
        nn = ctypes.c_int64()
        dep = ctypes.c_int32()
        rmax = ctypes.c_double()
        C.check(ctx.L.twg_bvh_info(h, ctypes.byref(nf), ctypes.byref(nn), ctypes.byref(dep),
                                   ctypes.byref(rmax)), "twg_bvh_info")
        self.nf, self.nnode, self.depth, self.rmax = nf.value, nn.value, dep.value, rmax.value

    def order(self):
        import numpy as np
        o = np.empty(self.nf, dtype=np.int32)
        C.check(self.ctx.L.twg_bvh_facet_order(self.h, o.ctypes.data_as(ctypes.c_void_p)),
                "twg_bvh_facet_order")
        return o

    def close(self):
        if getattr(self, "h", None) is not None:
            self.ctx.L.twg_bvh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SmoothBatch:
    """Device-resident batch of one-ring problems (see twg_smooth_desc).
    arrays: dict of numpy arrays as produced by workloads.smoothing_batch,
    plus optional old_q; sf / bt: Surface objects (or None)."""

    def __init__(self, ctx, arrays, sf=None, bt=None, eps2=0.0, sampling_dist=1.0, old_q=None):
        import numpy as np
        self.ctx = ctx
        dev = ctx.device

        def T(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a).astype(dt, copy=False)).to(dev)
        self.x = T(arrays["x"], np.float64)
        self.y = T(arrays["y"], np.float64)
        self.z = T(arrays["z"], np.float64)
        self.center = T(arrays["center"], np.int32)
        self.ring_off = T(arrays["ring_off"], np.int32)
        self.ring_tets = T(arrays["ring_tets"], np.int32)
        self.kind = T(arrays["kind"], np.uint8)
        self.tri_off = T(arrays["tri_off"], np.int32)
        tri_ab = arrays["tri_ab"]
        self.tri_ab = T(tri_ab if len(tri_ab) else np.zeros((1, 2)), np.int32)
        n = len(arrays["center"])
        nr = len(arrays["ring_tets"])
        self.n, self.n_ring, self.n_tri = n, nr, len(tri_ab)
        if old_q is None:
            old_q = np.zeros(nr)
        self.old_q = T(old_q, np.float64)
        self.p_out = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self.status = torch.empty(n, dtype=torch.uint8, device=dev)
        self.new_q = torch.zeros(max(nr, 1), dtype=torch.float64, device=dev)
        d = C.SmoothDesc()
        d.x, d.y, d.z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
        d.nv = self.x.shape[0]
        d.n = n
        d.center = self.center.data_ptr()
        d.ring_off = self.ring_off.data_ptr()
        d.ring_tets = self.ring_tets.data_ptr()
        d.n_ring = nr
        d.kind = self.kind.data_ptr()
        d.old_q = self.old_q.data_ptr()
        d.tri_off = self.tri_off.data_ptr()
        d.tri_ab = self.tri_ab.data_ptr()
        d.n_tri = self.n_tri
        d.sf_tree = sf.h.value if sf is not None else None
        d.b_tree = bt.h.value if bt is not None else None
        d.eps2 = float(eps2)
        d.sampling_dist = float(sampling_dist)
        d.p_out = self.p_out.data_ptr()
        d.status = self.status.data_ptr()
        d.new_q = self.new_q.data_ptr()
        self.desc = d
        self._keep = (sf, bt)

    def run(self, precision=C.TWG_FP64, tile=0, chunk=0, desc=None):
        self.ctx._bind()
        d = self.desc if desc is None else desc
        C.check(self.ctx.L.twg_smooth_batch(self.ctx.h, ctypes.byref(d), precision, tile, chunk),
                "twg_smooth_batch")
        return self.p_out, self.status, self.new_q

    def problem_desc(self, p):
        """descriptor of problem p alone (n = 1), sharing this batch's arrays"""
        d = C.SmoothDesc()
        ctypes.pointer(d)[0] = self.desc
        d.n = 1
        d.center = self.center.data_ptr() + 4 * p
        d.ring_off = self.ring_off.data_ptr() + 4 * p
        d.kind = self.kind.data_ptr() + p
        d.tri_off = self.tri_off.data_ptr() + 4 * p
        d.p_out = self.p_out.data_ptr() + 24 * p
        d.status = self.status.data_ptr() + p
        return d

    def autotune(self, precision=C.TWG_FP64):
        t = ctypes.c_int()
        c = ctypes.c_int()
        ms = ctypes.c_float()
        self.ctx._bind()
        C.check(self.ctx.L.twg_smooth_autotune(self.ctx.h, ctypes.byref(self.desc), precision,
                                               ctypes.byref(t), ctypes.byref(c), ctypes.byref(ms)),
                "twg_smooth_autotune")
        return t.value, c.value, ms.value

    def plan(self, precision=C.TWG_FP64, tile=0, chunk=0):
        return SmoothPlan(self, precision, tile, chunk)


class SmoothPlan:
    """hipGraph of one persistent smoothing launch over a SmoothBatch."""

    def __init__(self, batch, precision, tile, chunk):
        self.batch = batch
        L = batch.ctx.L
        h = ctypes.c_void_p()
        batch.ctx._bind()
        C.check(L.twg_smooth_plan_create(batch.ctx.h, ctypes.byref(batch.desc), precision, tile,
                                         chunk, ctypes.byref(h)), "twg_smooth_plan_create")
        self.h = h
        t = ctypes.c_int()
        c = ctypes.c_int()
        C.check(L.twg_smooth_plan_info(h, ctypes.byref(t), ctypes.byref(c)), "twg_smooth_plan_info")
        self.tile, self.chunk = t.value, c.value

    def run(self):
        self.batch.ctx._bind()
        C.check(self.batch.ctx.L.twg_smooth_plan_run(self.h), "twg_smooth_plan_run")
        return self.batch.p_out, self.batch.status, self.batch.new_q

    def close(self):
        if getattr(self, "h", None) is not None:
            self.batch.ctx.L.twg_smooth_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CollapseBatch:
    """Device-resident batch of edge-collapse candidates (see twg_collapse_desc).
    arrays: dict of numpy arrays as produced by workloads.collapse_batch (x, y,
    z, v1, v2, flags, ring_off, ring_tets, old_q, tri_off, tri_ids); sf / bt:
    Surface objects (or None)."""

    def __init__(self, ctx, arrays, sf=None, bt=None, eps2=0.0, sampling_dist=1.0,
                 check_quality=True, check_eps=True, soft_energy=None):
        import numpy as np
        self.ctx = ctx
        dev = ctx.device

        def T(a, dt):
            return torch.from_numpy(np.ascontiguousarray(a).astype(dt, copy=False)).to(dev)
        self.x = T(arrays["x"], np.float64)
        self.y = T(arrays["y"], np.float64)
        self.z = T(arrays["z"], np.float64)
        self.v1 = T(arrays["v1"], np.int32)
        self.v2 = T(arrays["v2"], np.int32)
        self.flags = T(arrays["flags"], np.uint8)
        self.ring_off = T(arrays["ring_off"], np.int32)
        rt = arrays["ring_tets"]
        self.ring_tets = T(rt if len(rt) else np.zeros((1, 4)), np.int32)
        self.old_q = T(arrays["old_q"] if len(rt) else np.zeros(1), np.float64)
        self.tri_off = T(arrays["tri_off"], np.int32)
        tri = arrays["tri_ids"]
        self.tri_ids = T(tri if len(tri) else np.zeros((1, 3)), np.int32)
        n = len(arrays["v1"])
        self.n, self.n_ring, self.n_tri = n, len(rt), len(tri)
        self.status = torch.empty(max(n, 1), dtype=torch.uint8, device=dev)
        self.new_q = torch.zeros(max(len(rt), 1), dtype=torch.float64, device=dev)
        d = C.CollapseDesc()
        d.x, d.y, d.z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
        d.nv = self.x.shape[0]
        d.n = n
        d.v1, d.v2, d.flags = self.v1.data_ptr(), self.v2.data_ptr(), self.flags.data_ptr()
        d.ring_off = self.ring_off.data_ptr()
        d.ring_tets = self.ring_tets.data_ptr()
        d.n_ring = self.n_ring
        d.old_q = self.old_q.data_ptr()
        d.tri_off = self.tri_off.data_ptr()
        d.tri_ids = self.tri_ids.data_ptr()
        d.n_tri = self.n_tri
        d.sf_tree = sf.h.value if sf is not None else None
        d.b_tree = bt.h.value if bt is not None else None
        d.eps2 = float(eps2)
        d.sampling_dist = float(sampling_dist)
        d.check_quality = int(bool(check_quality))
        d.check_eps = int(bool(check_eps))
        d.is_soft = int(soft_energy is not None)
        d.soft_energy = float(soft_energy) if soft_energy is not None else 0.0
        d.status = self.status.data_ptr()
        d.new_q = self.new_q.data_ptr()
        self.desc = d
        self._keep = (sf, bt)

    def run(self, precision=C.TWG_FP64, tile=0, chunk=0):
        self.ctx._bind()
        C.check(self.ctx.L.twg_collapse_batch(self.ctx.h, ctypes.byref(self.desc), precision,
                                              tile, chunk), "twg_collapse_batch")
        return self.status[:self.n], self.new_q[:self.n_ring]

    def autotune(self, precision=C.TWG_FP64):
        t = ctypes.c_int()
        c = ctypes.c_int()
        ms = ctypes.c_float()
        self.ctx._bind()
        C.check(self.ctx.L.twg_collapse_autotune(self.ctx.h, ctypes.byref(self.desc), precision,
                                                 ctypes.byref(t), ctypes.byref(c),
                                                 ctypes.byref(ms)), "twg_collapse_autotune")
        return t.value, c.value, ms.value

    def plan(self, precision=C.TWG_FP64, tile=0, chunk=0):
        return CollapsePlan(self, precision, tile, chunk)


class CollapsePlan:
    """hipGraph of every launch of a CollapseBatch."""

    def __init__(self, batch, precision, tile, chunk):
        self.batch = batch
        L = batch.ctx.L
        h = ctypes.c_void_p()
        batch.ctx._bind()
        C.check(L.twg_collapse_plan_create(batch.ctx.h, ctypes.byref(batch.desc), precision, tile,
                                           chunk, ctypes.byref(h)), "twg_collapse_plan_create")
        self.h = h
        t = ctypes.c_int()
        c = ctypes.c_int()
        C.check(L.twg_collapse_plan_info(h, ctypes.byref(t), ctypes.byref(c)),
                "twg_collapse_plan_info")
        self.tile, self.chunk = t.value, c.value

    def run(self):
        self.batch.ctx._bind()
        C.check(self.batch.ctx.L.twg_collapse_plan_run(self.h), "twg_collapse_plan_run")
        return self.batch.status[:self.batch.n], self.batch.new_q[:self.batch.n_ring]

    def close(self):
        if getattr(self, "h", None) is not None:
            self.batch.ctx.L.twg_collapse_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
