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
# twg_smooth_batch tile: 64 lanes for each surface problem, 16 for the others
TWG_SMOOTH_TILE_SPLIT = 1
TWG_AMIPS_E = 0
TWG_AMIPS_EG = 1
TWG_AMIPS_EGH = 2
TWG_AMIPS_EGH_SYM = 3  # H as its upper triangle, (n, 6)
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
    "twg_ctx_list_stats": [P, i32, P],
    "twg_sample_count": [P, P, i64, f64, P],
    "twg_sample_write": [P, P, i64, f64, P, i64, P],
    "twg_smooth_batch": [P, P, i32, i32, i32],
    "twg_smooth_batch_noalloc": [P, P, i32, i32, i32],
    "twg_smooth_autotune": [P, P, i32, P, P, P],
    "twg_smooth_plan_create": [P, P, i32, i32, i32, ctypes.POINTER(P)],
    "twg_smooth_plan_run": [P],
    "twg_smooth_plan_info": [P, P, P],
    "twg_smooth_plan_destroy": [P],
    "twg_newton_step": [P, P, P, P, i64, P, i32, i32, f64, f64, f64, P, P],
    "twg_smooth_gather_q": [P, P, P, P, i64],
    "twg_smooth_apply": [P, P, P, P, P, P, P, i64],
    "twg_collapse_batch": [P, P, i32, i32, i32],
    "twg_collapse_autotune": [P, P, i32, P, P, P],
    "twg_collapse_plan_create": [P, P, i32, i32, i32, ctypes.POINTER(P)],
    "twg_collapse_plan_run": [P],
    "twg_collapse_plan_info": [P, P, P],
    "twg_collapse_plan_destroy": [P],
    "twg_tet_groups": [P, P, P, i64, i64, P, P, P, i64, P, P, P],
    "twg_energy_stats": [P, P, i64, P, P, P, i64, f64, P],
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

# return codes of EdgeCollapser::collapseAnEdge, and BAD_INPUT
TWG_COLLAPSE_SUCCESS = 0
TWG_COLLAPSE_FLIP = 1
TWG_COLLAPSE_QUALITY = 2
TWG_COLLAPSE_ENVELOP = 3
TWG_COLLAPSE_ENVELOP_SUC = 4
TWG_COLLAPSE_BAD_INPUT = 7
TWG_CV1_SURFACE = 1
TWG_CV1_BOUNDARY = 2
TWG_CV2_SURFACE = 4
TWG_CV2_BOUNDARY = 8
TWG_CV1_UNROUNDED = 16


class NewtonStepOut(ctypes.Structure):
    """twg_newton_step_out"""
    _fields_ = [("energy", f64), ("J", f64 * 3), ("H", f64 * 9), ("X", (f64 * 3) * 20),
                ("new_energy", f64 * 20), ("ok", ctypes.c_int32), ("finite", ctypes.c_uint8 * 20),
                ("flip", ctypes.c_uint8 * 20)]


class CollapseDesc(ctypes.Structure):
    """twg_collapse_desc"""
    _fields_ = [("x", P), ("y", P), ("z", P), ("nv", i64), ("n", i64), ("v1", P), ("v2", P),
                ("flags", P), ("ring_off", P), ("ring_tets", P), ("n_ring", i64),
                ("old_q", P), ("tri_off", P), ("tri_ids", P), ("n_tri", i64),
                ("sf_tree", P), ("b_tree", P), ("eps2", f64), ("sampling_dist", f64),
                ("check_quality", i32), ("check_eps", i32), ("is_soft", i32),
                ("soft_energy", f64), ("status", P), ("new_q", P)]


TWG_GROUP_BAD_INPUT = 2


class EnergySummary(ctypes.Structure):
    """twg_energy_summary"""
    _fields_ = [("sum", f64), ("max", f64), ("second_max", f64), ("count", i64),
                ("buckets", i64 * 11)]


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
