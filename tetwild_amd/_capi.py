"""ctypes binding of libtetwild_amd.so (the C-ABI declared in
include/tetwild_gpu.h).  Loading fails loudly when the library has not been
built: there is no host fallback for any entry point."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtetwild_amd.so")

TWG_OK = 0
TWG_EINVAL = 1
TWG_EDEVICE = 2
TWG_ENOMEM = 3
TWG_FP64 = 0
TWG_MIXED = 1
TWG_AMIPS_E = 0
TWG_AMIPS_EG = 1
TWG_AMIPS_EGH = 2
MAX_ENERGY = 1e50

P = ctypes.c_void_p
i32 = ctypes.c_int
i64 = ctypes.c_int64
f64 = ctypes.c_double
sz = ctypes.c_size_t

# name -> argument types (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    "twg_ctx_create": [i32, ctypes.POINTER(P)],
    "twg_ctx_destroy": [P],
    "twg_ctx_set_stream": [P, P],
    "twg_ctx_stream": [P],
    "twg_ctx_sync": [P],
    "twg_last_error": [],
    "twg_malloc": [P, sz, ctypes.POINTER(P)],
    "twg_free": [P, P],
    "twg_memcpy_h2d": [P, P, P, sz],
    "twg_memcpy_d2h": [P, P, P, sz],
    "twg_amips_eval": [P, P, i64, P, P, P, i64, i32, P, P, P],
    "twg_amips_eval_mixed": [P, P, i64, P, P, P, i64, i32, f64, P, P, P, P],
    "twg_tet_qualities": [P, P, i64, P, P, P, i64, i32, f64, P, P],
    "twg_tet_flip": [P, P, i64, P, P, P, i64, P],
    "twg_bvh_build": [P, P, i64, P, i64, ctypes.POINTER(P)],
    "twg_bvh_destroy": [P],
    "twg_bvh_info": [P, P, P, P, P],
    "twg_bvh_facet_order": [P, P],
    "twg_envelope_points": [P, P, P, i64, f64, i32, P, P],
    "twg_envelope_faces": [P, P, P, i64, f64, f64, i32, P, P],
    "twg_nearest_points": [P, P, P, i64, P, P, P],
    "twg_sample_count": [P, P, i64, f64, P],
    "twg_sample_write": [P, P, i64, f64, P, i64, P],
    "twg_smooth_batch": [P, P, i32, i32, i32],
    "twg_smooth_autotune": [P, P, i32, P, P, P],
    "twg_smooth_plan_create": [P, P, i32, i32, i32, ctypes.POINTER(P)],
    "twg_smooth_plan_run": [P],
    "twg_smooth_plan_info": [P, P, P],
    "twg_smooth_plan_destroy": [P],
}
_RESTYPES = {"twg_last_error": ctypes.c_char_p, "twg_ctx_stream": P}

TWG_SMOOTH_STATUS = ["MOVED", "INVALID", "NOT_MOVED", "FLIP_END", "REJECT_FLIP",
                     "REJECT_QUALITY", "REJECT_ENVELOPE", "BAD_INPUT"]


class SmoothDesc(ctypes.Structure):
    """twg_smooth_desc"""
    _fields_ = [("x", P), ("y", P), ("z", P), ("nv", i64), ("n", i64), ("center", P),
                ("ring_off", P), ("ring_tets", P), ("n_ring", i64), ("kind", P), ("old_q", P),
                ("tri_off", P), ("tri_ab", P), ("n_tri", i64), ("sf_tree", P), ("b_tree", P),
                ("eps2", f64), ("sampling_dist", f64), ("p_out", P), ("status", P),
                ("new_q", P)]

_LIB = None


class TWGError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (or make -C tetwild_amd/csrc); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, i32)
        _LIB = L
    return _LIB


def check(status, what=""):
    if status != TWG_OK:
        msg = lib().twg_last_error()
        raise TWGError(f"{what}: status {status}: {msg.decode() if msg else ''}")
