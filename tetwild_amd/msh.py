"""Mesh files as TetWild writes them: Gmsh 2.2 .msh in the layout of PyMesh's
MshSaver (extern/pymesh/MshSaver.h:15-54, used by saveFinalTetmesh in
src/main.cpp:27-70) and the Medit .mesh branch of saveFinalTetmesh.

write_msh / read_msh take and return numpy arrays: vertices (n, 3) float64,
elements (m, k) int (0-based), and dicts of node / element fields.  Binary
files are byte-identical to PyMesh's (little-endian int32 ids, float64
values, the endianness word directly followed by $EndMeshFormat); ASCII
values use C++ stream default formatting ("%g")."""
import struct

import numpy as np

TRI, QUAD, TET, HEX = 2, 3, 4, 5
_NODES_PER = {TRI: 3, QUAD: 4, TET: 4, HEX: 8}


def _fmt(v):
    return "%g" % v


def write_msh(path, V, elements, elem_type=TET, binary=True, node_fields=None,
              elem_fields=None):
    """node_fields / elem_fields: {name: array (n,) or (n, 3)}, written in order"""
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
    k = _NODES_PER[elem_type]
    E = np.ascontiguousarray(elements, dtype=np.int64).reshape(-1, k)
    out = bytearray()
    out += b"$MeshFormat\n"
    if binary:
        out += b"2.2 1 8\n" + struct.pack("<i", 1) + b"$EndMeshFormat\n"
    else:
        out += b"2.2 0 8\n$EndMeshFormat\n"
    out += b"$Nodes\n%d\n" % len(V)
    if binary:
        rec = np.zeros(len(V), dtype=[("i", "<i4"), ("v", "<f8", 3)])
        rec["i"] = np.arange(1, len(V) + 1)
        rec["v"] = V
        out += rec.tobytes()
    else:
        for i, v in enumerate(V):
            out += ("%d %s %s %s\n" % (i + 1, _fmt(v[0]), _fmt(v[1]), _fmt(v[2]))).encode()
    out += b"$EndNodes\n$Elements\n%d\n" % len(E)
    if len(E):
        if binary:
            out += struct.pack("<3i", elem_type, len(E), 0)
            rec = np.zeros(len(E), dtype=[("i", "<i4"), ("v", "<i4", k)])
            rec["i"] = np.arange(1, len(E) + 1)
            rec["v"] = E + 1
            out += rec.tobytes()
        else:
            for i, e in enumerate(E):
                out += ("%d %d 0 " % (i + 1, elem_type) + "".join("%d " % (v + 1) for v in e) +
                        "\n").encode()
    out += b"$EndElements\n"
    for section, fields, count in (("NodeData", node_fields, len(V)),
                                   ("ElementData", elem_fields, len(E))):
        for name, f in (fields or {}).items():
            f = np.asarray(f, dtype=np.float64)
            comps = 1 if f.ndim == 1 else 3
            f = f.reshape(count, -1)
            if f.shape[1] < comps:
                f = np.concatenate([f, np.zeros((count, comps - f.shape[1]))], 1)
            out += ("$%s\n1\n\"%s\"\n1\n0.0\n3\n0\n%d\n%d\n" % (section, name, comps, count)).encode()
            if binary:
                rec = np.zeros(count, dtype=[("i", "<i4"), ("v", "<f8", comps)])
                rec["i"] = np.arange(1, count + 1)
                rec["v"] = f[:, :comps]
                out += rec.tobytes()
            else:
                for i in range(count):
                    out += ("%d " % (i + 1) + " ".join(_fmt(v) for v in f[i, :comps]) + "\n").encode()
            out += ("$End%s\n" % section).encode()
    with open(path, "wb") as fh:
        fh.write(bytes(out))


class _Reader:
    def __init__(self, data):
        self.d = data
        self.p = 0

    def token(self):
        d, p = self.d, self.p
        while p < len(d) and d[p] in b" \t\r\n":
            p += 1
        q = p
        while q < len(d) and d[q] not in b" \t\r\n":
            q += 1
        self.p = q
        return d[p:q].decode() if q > p else None

    def end_of_line(self):
        while self.p < len(self.d):
            c = self.d[self.p]
            self.p += 1
            if c == 10:
                return
            if c not in b" \t\r":
                raise ValueError("msh: unexpected data after a count")

    def take(self, n):
        if self.p + n > len(self.d):
            raise ValueError("msh: truncated binary block")
        b = self.d[self.p:self.p + n]
        self.p += n
        return b

    def quoted(self):
        d = self.d
        while d[self.p] in b" \t\r\n":
            self.p += 1
        if d[self.p] == ord('"'):
            q = d.index(b'"', self.p + 1)
            s = d[self.p + 1:q].decode()
            self.p = q + 1
            return s
        return self.token()


def read_msh(path):
    """returns (V (n, 3), elements (m, k) int64, elem_type, node_fields, elem_fields)"""
    with open(path, "rb") as fh:
        r = _Reader(fh.read())
    if r.token() != "$MeshFormat":
        raise ValueError("msh: missing $MeshFormat")
    r.token()
    binary = int(r.token()) == 1
    if int(r.token()) != 8:
        raise ValueError("msh: only 8-byte reals are supported")
    if binary:
        r.end_of_line()
        if struct.unpack("<i", r.take(4))[0] != 1:
            raise ValueError("msh: other-endian binary file")
    if r.token() != "$EndMeshFormat":
        raise ValueError("msh: missing $EndMeshFormat")
    V = np.zeros((0, 3))
    by_type = {}
    nf, ef = {}, {}
    while True:
        tag = r.token()
        if tag is None:
            break
        if tag == "$Nodes":
            n = int(r.token())
            V = np.zeros((n, 3))
            if binary:
                r.end_of_line()
                rec = np.frombuffer(r.take(28 * n), dtype=[("i", "<i4"), ("v", "<f8", 3)])
                V[rec["i"] - 1] = rec["v"]
            else:
                for _ in range(n):
                    i = int(r.token())
                    V[i - 1] = [float(r.token()) for _ in range(3)]
            if not np.isfinite(V).all():
                raise ValueError("msh: NaN or Inf in the nodes")
            assert r.token() == "$EndNodes"
        elif tag == "$Elements":
            m = int(r.token())
            if binary:
                r.end_of_line()
                done = 0
                while done < m:
                    t, cnt, tags = struct.unpack("<3i", r.take(12))
                    k = _NODES_PER[t]
                    rec = np.frombuffer(r.take(4 * (1 + tags + k) * cnt), "<i4").reshape(cnt, -1)
                    by_type.setdefault(t, []).append(rec[:, 1 + tags:] - 1)
                    done += cnt
            else:
                for _ in range(m):
                    _, t, tags = (int(r.token()) for _ in range(3))
                    for _ in range(tags):
                        r.token()
                    by_type.setdefault(t, []).append(
                        np.array([[int(r.token()) - 1 for _ in range(_NODES_PER[t])]]))
            assert r.token() == "$EndElements"
        elif tag in ("$NodeData", "$ElementData"):
            ns = int(r.token())
            names = [r.quoted() for _ in range(ns)]
            for _ in range(int(r.token())):
                r.token()
            ints = [int(r.token()) for _ in range(int(r.token()))]
            comps, count = ints[1], ints[2]
            f = np.zeros((count, comps))
            if binary:
                r.end_of_line()
                rec = np.frombuffer(r.take((4 + 8 * comps) * count),
                                    dtype=[("i", "<i4"), ("v", "<f8", comps)])
                f[rec["i"] - 1] = rec["v"].reshape(count, comps)
            else:
                for _ in range(count):
                    i = int(r.token())
                    f[i - 1] = [float(r.token()) for _ in range(comps)]
            (nf if tag == "$NodeData" else ef)[names[0]] = f[:, 0] if comps == 1 else f
            assert r.token() == "$End" + tag[1:]
        elif tag.startswith("$"):
            end = ("$End" + tag[1:]).encode()
            r.p = r.d.index(end, r.p) + len(end)
        else:
            raise ValueError("msh: unexpected token " + tag)
    for t in (TET, HEX, TRI, QUAD):
        if t in by_type:
            return V, np.concatenate(by_type[t]).astype(np.int64), t, nf, ef
    return V, np.zeros((0, 3), np.int64), TRI, nf, ef


def write_medit(path, V, T):
    """the .mesh branch of saveFinalTetmesh: 16 significant digits, 1-based"""
    V = np.asarray(V, dtype=np.float64).reshape(-1, 3)
    T = np.asarray(T, dtype=np.int64).reshape(-1, 4)
    lines = ["MeshVersionFormatted 1", "Dimension 3", "Vertices", str(len(V))]
    lines += ["%.16g %.16g %.16g 0" % tuple(v) for v in V]
    lines += ["Triangles", "0", "Tetrahedra", str(len(T))]
    lines += ["%d %d %d %d 0" % tuple(t + 1) for t in T]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\nEnd")


def read_medit(path):
    with open(path) as fh:
        tok = fh.read().split()
    V = np.zeros((0, 3))
    T = np.zeros((0, 4), np.int64)
    i = 0
    while i < len(tok):
        w = tok[i]
        if w == "Vertices":
            n = int(tok[i + 1])
            V = np.array(tok[i + 2:i + 2 + 4 * n], dtype=np.float64).reshape(n, 4)[:, :3]
            i += 2 + 4 * n
        elif w == "Triangles":
            i += 2 + 4 * int(tok[i + 1])
        elif w == "Tetrahedra":
            n = int(tok[i + 1])
            T = np.array(tok[i + 2:i + 2 + 5 * n], dtype=np.int64).reshape(n, 5)[:, :4] - 1
            i += 2 + 5 * n
        elif w == "End":
            break
        else:
            i += 1
    return V, T
