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
        """E (n), G (3, n), H (9, n) float64, gradient/Hessian w.r.t. vertex 0."""
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        dev = self.device
        if out is None:
            E = torch.empty(n, dtype=torch.float64, device=dev)
            G = torch.empty((3, n), dtype=torch.float64, device=dev) if mode >= 1 else None
            H = torch.empty((9, n), dtype=torch.float64, device=dev) if mode >= 2 else None
        else:
            E, G, H = out
        self._bind()
        C.check(self.L.twg_amips_eval(self.h, _ptr(tets), n, _ptr(x), _ptr(y), _ptr(z), nv, mode,
                                      _ptr(E), _ptr(G), _ptr(H)), "twg_amips_eval")
        return E, G, H

    def amips_eval_mixed(self, tets, x, y, z, mode=C.TWG_AMIPS_EGH, tol=2.0**-16, out=None):
        """E (n) float64, G (3, n) / H (9, n) float32, fallback (n) uint8."""
        n = self._tets(tets)
        nv = self._verts(x, y, z)
        dev = self.device
        if out is None:
            E = torch.empty(n, dtype=torch.float64, device=dev)
            G = torch.empty((3, n), dtype=torch.float32, device=dev) if mode >= 1 else None
            H = torch.empty((9, n), dtype=torch.float32, device=dev) if mode >= 2 else None
            fb = torch.empty(n, dtype=torch.uint8, device=dev)
        else:
            E, G, H, fb = out
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
