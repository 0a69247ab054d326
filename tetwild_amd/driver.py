"""Runs the C++ LocalOperations / VertexSmoother / EdgeCollapser /
EdgeRemover mirrors (tetwild_amd/shim) on a tet mesh through their driver
program, shim_driver, whose mesh and output files are laid out in
tetwild_amd/shim/shim_driver.cpp."""
import json
import os
import struct
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "shim", "shim_driver")


def write_mesh(path, P, tets, is_surface_fs, flags, q, sf_tris, b_tris, eps, eps_2,
               sampling_dist, closed=True, exact=None):
    """flags: uint8 per vertex (1 surface, 2 bbox, 4 boundary, 8 locked, 16
    unrounded); sf_tris / b_tris: float64 [n][9]; exact: the exact positions
    of the unrounded vertices in vertex order, or None"""
    sf_tris = np.ascontiguousarray(sf_tris, dtype=np.float64).reshape(-1, 9)
    b_tris = np.ascontiguousarray(b_tris, dtype=np.float64).reshape(-1, 9)
    with open(path, "wb") as f:
        f.write(struct.pack("<5q", len(P), len(tets), len(sf_tris), len(b_tris), int(closed)))
        f.write(struct.pack("<3d", eps, eps_2, sampling_dist))
        f.write(np.ascontiguousarray(P, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(flags, dtype=np.uint8).tobytes())
        f.write(np.ascontiguousarray(tets, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(is_surface_fs, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(q, dtype=np.float64).tobytes())
        f.write(sf_tris.tobytes())
        f.write(b_tris.tobytes())
        if exact is not None:
            f.write(np.ascontiguousarray(exact, dtype=np.float64).tobytes())


def run(mesh_path, mode, out_path=None, timeout=600):
    """runs shim_driver in `mode`; returns (parsed JSON line or None, out path,
    wall seconds)"""
    if not os.path.exists(DRIVER):
        raise FileNotFoundError(f"{DRIVER} is missing: build it (make -C tetwild_amd/shim)")
    if out_path is None:
        out_path = mesh_path + "." + mode + ".out"
    t0 = time.perf_counter()
    r = subprocess.run([DRIVER, mesh_path, out_path, mode], capture_output=True, text=True,
                       timeout=timeout)
    el = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError(f"shim_driver {mode} failed ({r.returncode}): {r.stderr[-2000:]}")
    line = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
    return (json.loads(line) if line.startswith("{") else None), out_path, el


def tempdir():
    return tempfile.mkdtemp(prefix="twg_driver_")
