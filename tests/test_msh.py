""".msh (Gmsh 2.2 in PyMesh MshSaver's layout, extern/pymesh/MshSaver.cpp) and
Medit .mesh files (saveFinalTetmesh, src/main.cpp:27-70): the C++ writer
(tetwild_amd/shim/msh_io.cpp, driven through msh_tool) and the Python writer
(tetwild_amd/msh.py) against bytes assembled here from the layout the
reference writer produces, and round trips through both readers."""
import os
import struct
import subprocess

import numpy as np
import pytest

from tetwild_amd import msh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOL = os.path.join(ROOT, "tetwild_amd", "shim", "msh_tool")


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(TOOL):
        subprocess.run(["make", "-C", os.path.dirname(TOOL), "msh_tool"], check=True,
                       capture_output=True)
    return TOOL


def tiny():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1.5]], dtype=np.float64)
    T = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int64)
    A = np.array([0.3, 1.25])
    return V, T, A


def golden_binary(V, T, A):
    """PyMesh MshSaver(binary).save_mesh(TET) + save_elem_scalar_field"""
    b = b"$MeshFormat\n2.2 1 8\n" + struct.pack("<i", 1) + b"$EndMeshFormat\n"
    b += b"$Nodes\n%d\n" % len(V)
    for i, v in enumerate(V):
        b += struct.pack("<i3d", i + 1, *v)
    b += b"$EndNodes\n$Elements\n%d\n" % len(T) + struct.pack("<3i", 4, len(T), 0)
    for i, t in enumerate(T):
        b += struct.pack("<5i", i + 1, *(t + 1))
    b += b"$EndElements\n"
    b += b'$ElementData\n1\n"min_dihedral_angle"\n1\n0.0\n3\n0\n1\n%d\n' % len(T)
    for i, a in enumerate(A):
        b += struct.pack("<id", i + 1, a)
    b += b"$EndElementData\n"
    return b


def write_raw(path, V, E, k=4, typ=4, nf=None, ef=None):
    with open(path, "wb") as f:
        f.write(struct.pack("<6q", len(V), len(E), k, typ, nf is not None, ef is not None))
        f.write(np.ascontiguousarray(V, np.float64).tobytes())
        f.write(np.ascontiguousarray(E, np.int32).tobytes())
        if nf is not None:
            f.write(np.ascontiguousarray(nf, np.float64).tobytes())
        if ef is not None:
            f.write(np.ascontiguousarray(ef, np.float64).tobytes())


def read_raw(path):
    d = open(path, "rb").read()
    nv, ne, k, typ, hn, he = struct.unpack("<6q", d[:48])
    p = 48
    V = np.frombuffer(d[p:p + 24 * nv], np.float64).reshape(nv, 3)
    p += 24 * nv
    E = np.frombuffer(d[p:p + 4 * ne * k], np.int32).reshape(ne, k)
    p += 4 * ne * k
    nf = ef = None
    if hn:
        nf = np.frombuffer(d[p:p + 8 * nv], np.float64)
        p += 8 * nv
    if he:
        ef = np.frombuffer(d[p:p + 8 * ne], np.float64)
    return V, E, typ, nf, ef


def test_binary_bytes_match_the_layout(tool, tmp_path):
    V, T, A = tiny()
    gold = golden_binary(V, T, A)
    py = tmp_path / "py.msh"
    msh.write_msh(py, V, T, elem_fields={"min_dihedral_angle": A})
    assert py.read_bytes() == gold
    raw = tmp_path / "in.bin"
    write_raw(raw, V, T, ef=A)
    cc = tmp_path / "cc.msh"
    subprocess.run([tool, "write", str(raw), str(cc)], check=True)
    assert cc.read_bytes() == gold


def test_ascii_text(tool, tmp_path):
    V, T, A = tiny()
    V = V + np.array([0.1234567, 0, 0])
    raw = tmp_path / "in.bin"
    write_raw(raw, V, T, ef=A)
    cc = tmp_path / "cc.msh"
    subprocess.run([tool, "write", str(raw), str(cc), "ascii"], check=True)
    py = tmp_path / "py.msh"
    msh.write_msh(py, V, T, binary=False, elem_fields={"min_dihedral_angle": A})
    text = cc.read_text()
    assert text == py.read_text()
    # C++ stream default formatting: six significant digits, elements with
    # 1-based ids and a space after each
    assert "1 0.123457 0 0\n" in text and "2 4 0 2 3 4 5 \n" in text
    assert text.startswith("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n5\n")


@pytest.mark.parametrize("binary", [True, False])
def test_round_trips(tool, tmp_path, binary):
    rng = np.random.default_rng(4)
    V = rng.normal(size=(300, 3))
    T = rng.integers(0, 300, size=(900, 4))
    nf = rng.normal(size=300)
    ef = rng.uniform(0, 3, size=900)
    # Python writer -> C++ reader
    p = tmp_path / "a.msh"
    msh.write_msh(p, V, T, binary=binary, node_fields={"value": nf},
                  elem_fields={"min_dihedral_angle": ef})
    out = tmp_path / "a.bin"
    subprocess.run([tool, "read", str(p), str(out)], check=True)
    V2, T2, typ, nf2, ef2 = read_raw(out)
    tol = 0 if binary else 1e-5
    assert typ == msh.TET and np.array_equal(T2, T)
    assert np.allclose(V2, V, rtol=tol, atol=tol) and np.allclose(nf2, nf, rtol=tol, atol=tol)
    assert np.allclose(ef2, ef, rtol=tol, atol=tol)
    if binary:
        assert np.array_equal(V2, V) and np.array_equal(ef2, ef)
    # C++ writer -> Python reader
    raw = tmp_path / "b.bin"
    write_raw(raw, V, T, nf=nf, ef=ef)
    q = tmp_path / "b.msh"
    subprocess.run([tool, "write", str(raw), str(q)] + ([] if binary else ["ascii"]), check=True)
    V3, T3, typ3, nfs, efs = msh.read_msh(q)
    assert typ3 == msh.TET and np.array_equal(T3, T)
    assert np.allclose(V3, V, rtol=tol, atol=tol)
    assert np.allclose(nfs["value"], nf, rtol=tol, atol=tol)
    assert np.allclose(efs["min_dihedral_angle"], ef, rtol=tol, atol=tol)


def test_reader_permuted_ids_blocks_and_unknown_sections(tmp_path):
    """node ids out of order, two element blocks (triangles, then tets: tets are
    kept), tags, an unknown section"""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
    b = b"$MeshFormat\n2.2 1 8\n" + struct.pack("<i", 1) + b"\n$EndMeshFormat\n"
    b += b"$PhysicalNames\n1\n3 1 \"vol\"\n$EndPhysicalNames\n"
    b += b"$Nodes\n4\n"
    for i in (3, 1, 0, 2):
        b += struct.pack("<i3d", i + 1, *V[i])
    b += b"$EndNodes\n$Elements\n2\n"
    b += struct.pack("<3i", 2, 1, 2) + struct.pack("<6i", 1, 7, 8, 1, 2, 3)
    b += struct.pack("<3i", 4, 1, 1) + struct.pack("<6i", 2, 9, 1, 2, 3, 4)
    b += b"$EndElements\n"
    p = tmp_path / "g.msh"
    p.write_bytes(b)
    V2, T2, typ, nf, ef = msh.read_msh(p)
    assert np.array_equal(V2, V) and typ == msh.TET and T2.tolist() == [[0, 1, 2, 3]]


def test_medit(tool, tmp_path):
    V, T, _ = tiny()
    V = V + 1e-3 / 3
    p = tmp_path / "py.mesh"
    msh.write_medit(p, V, T)
    raw = tmp_path / "in.bin"
    write_raw(raw, V, T)
    c = tmp_path / "cc.mesh"
    subprocess.run([tool, "mwrite", str(raw), str(c)], check=True)
    assert p.read_text() == c.read_text()
    text = p.read_text()
    assert text.startswith("MeshVersionFormatted 1\nDimension 3\nVertices\n5\n")
    assert "\nTriangles\n0\nTetrahedra\n2\n1 2 3 4 0\n2 3 4 5 0\nEnd" in text
    assert "0.0003333333333333333 " in text  # 16 significant digits
    out = tmp_path / "m.bin"
    subprocess.run([tool, "mread", str(p), str(out)], check=True)
    V2, T2, _, _, _ = read_raw(out)
    assert np.array_equal(T2, T) and np.allclose(V2, V, rtol=1e-15)
    V3, T3 = msh.read_medit(c)
    assert np.array_equal(T3, T) and np.allclose(V3, V, rtol=1e-15)
