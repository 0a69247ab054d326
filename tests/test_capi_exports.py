"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

from tetwild_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(twg_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "twg_amips_eval" in names and "twg_ctx_create" in names
    assert len(names) >= 10


def test_library_exports_all_declared():
    L = ctypes.CDLL(_capi.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header():
    missing = sorted(declared_functions() - set(_capi.SIGNATURES))
    assert not missing, f"no ctypes signature for: {missing}"


def test_no_device_no_silent_fallback():
    # without a GPU the context cannot be created and the error says why
    L = _capi.lib()
    h = ctypes.c_void_p()
    r = L.twg_ctx_create(0, ctypes.byref(h))
    if r != 0:
        assert L.twg_last_error()
