"""Mesh assets from disk: Wavefront OBJ and PLY (Import/MeshImporter.swift).

``load_mesh(path)`` picks the importer by extension, as ``Mesh(contentsOf:)``
(MeshImporter.swift:30-40); anything else raises ``MeshLoadingError`` with
the reference's message.

* OBJ (MeshImporter.swift:43-449): ``v``/``vt``/``vn``/``f`` lines, polygon
  faces fanned from their first vertex, one output vertex per distinct
  (position, texcoord, normal) triple in first-use order, missing texcoords
  (0, 0) and normals (0, 0, 0), ``vn`` normalized on read, 1-based and
  negative (relative) indices. Numbers are read the way the reference's byte
  scanner reads them (:292-341): digits accumulated in double, the fraction
  with a running ``scale *= 0.1``, then ``* 10^exponent``, rounded to float
  once; so vertex positions carry the same bits as the reference's.
* PLY (MeshImporter.swift:451-949): ``ascii 1.0`` and
  ``binary_little_endian 1.0``; vertex x/y/z, nx/ny/nz and u|s|texture_u,
  v|t|texture_v; faces from a ``vertex_indices``/``vertex_index`` list,
  fanned; other elements and properties skipped by their declared types.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import os
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

from .scene import Mesh, _normalize

f32 = np.float32


class MeshLoadingError(Exception):
    """MeshLoadingError (MeshImporter.swift:5-23)."""

    def __init__(self, kind: str, reason: str = "", line: int = 0):
        self.kind, self.reason, self.line = kind, reason, line
        if kind == "unsupportedFormat":
            msg = f"Unsupported mesh format '{reason}'."
        elif kind == "invalidOBJ":
            msg = f"Invalid OBJ mesh on line {line}: {reason}."
        elif kind == "invalidPLY":
            msg = f"Invalid PLY mesh on line {line}: {reason}."
        else:
            msg = f"Invalid mesh index '{reason}' on line {line}."
        super().__init__(msg)


# Correctly rounded decimal -> float32, as Swift's Float(String): C strtof
# (parsing to double first and rounding again can differ in the last bit).
_libc = ctypes.CDLL(ctypes.util.find_library("c"))
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p)]


def parse_float32(token: str) -> Optional[float]:
    """Float(token): the whole token must be a decimal (or inf/nan) number;
    returns None otherwise."""
    if not token or token[0].isspace() or token[-1].isspace():
        return None
    raw = token.encode("ascii", "replace")
    end = ctypes.c_char_p()
    buf = ctypes.create_string_buffer(raw)
    v = _libc.strtof(buf, ctypes.byref(end))
    consumed = ctypes.cast(end, ctypes.c_void_p).value - ctypes.addressof(buf)
    if consumed != len(raw) or consumed == 0:
        return None
    return float(f32(v))


def parse_int(token: str) -> Optional[int]:
    """Int(token): optional sign and decimal digits only."""
    s = token[1:] if token[:1] in "+-" else token
    if not s or not s.isdigit() or not s.isascii():
        return None
    return int(token)


def load_mesh(path: str) -> Mesh:
    ext = os.path.splitext(path)[1][1:]
    with open(path, "rb") as f:
        data = f.read()
    if ext.lower() == "obj":
        return parse_obj(data)
    if ext.lower() == "ply":
        return parse_ply(data)
    raise MeshLoadingError("unsupportedFormat", ext)


# --------------------------------------------------------------------- OBJ --
class _ObjScanner:
    """The reference's byte scanner (MeshImporter.swift:205-418)."""

    def __init__(self, data: bytes):
        self.b = data
        self.i = 0
        self.line = 1

    def at_end(self) -> bool:
        return self.i >= len(self.b)

    def at_line_break(self) -> bool:
        return not self.at_end() and self.b[self.i] == 10

    def at_comment(self) -> bool:
        return not self.at_end() and self.b[self.i] == 35

    def skip_ws(self) -> None:
        b, n = self.b, len(self.b)
        while self.i < n and b[self.i] in (32, 9, 13):
            self.i += 1

    def skip_line(self) -> None:
        b, n = self.b, len(self.b)
        while self.i < n:
            c = b[self.i]
            self.i += 1
            if c == 10:
                self.line += 1
                return

    def token(self) -> bytes:
        start = self.i
        b, n = self.b, len(self.b)
        while self.i < n and b[self.i] not in (32, 9, 10, 13, 35):
            self.i += 1
        return b[start:self.i]

    def _consume(self, byte: int) -> bool:
        if not self.at_end() and self.b[self.i] == byte:
            self.i += 1
            return True
        return False

    def _digit(self) -> Optional[int]:
        if self.at_end():
            return None
        c = self.b[self.i]
        if 48 <= c <= 57:
            self.i += 1
            return c - 48
        return None

    def _token_text(self, start: int) -> str:
        end = start
        while end < len(self.b) and self.b[end] not in (32, 9, 10, 13, 35, 47):
            end += 1
        return self.b[start:end].decode("utf-8", "replace")

    def read_float(self) -> float:
        self.skip_ws()
        start = self.i
        sign = 1.0
        if self._consume(45):
            sign = -1.0
        else:
            self._consume(43)
        value, has = 0.0, False
        while (d := self._digit()) is not None:
            has = True
            value = value * 10 + d
        if self._consume(46):
            scale = 0.1
            while (d := self._digit()) is not None:
                has = True
                value += d * scale
                scale *= 0.1
        if not has:
            raise MeshLoadingError("invalidOBJ", f"invalid number '{self._token_text(start)}'", self.line)
        if self._consume(101) or self._consume(69):
            esign = 1
            if self._consume(45):
                esign = -1
            else:
                self._consume(43)
            exp, hexp = 0, False
            while (d := self._digit()) is not None:
                hexp = True
                exp = exp * 10 + d
            if not hexp:
                raise MeshLoadingError("invalidOBJ", f"invalid number '{self._token_text(start)}'", self.line)
            value *= 10.0 ** (exp * esign)
        return float(f32(sign * value))

    def read_int(self) -> int:
        self.skip_ws()
        start = self.i
        sign = 1
        if self._consume(45):
            sign = -1
        else:
            self._consume(43)
        value, has = 0, False
        while (d := self._digit()) is not None:
            has = True
            value = value * 10 + d
        if not has:
            raise MeshLoadingError("invalidIndex", self._token_text(start), self.line)
        return value * sign

    def resolve(self, raw: int, count: int) -> int:
        if raw == 0:
            raise MeshLoadingError("invalidIndex", "0", self.line)
        r = raw - 1 if raw > 0 else count + raw
        if r < 0 or r >= count:
            raise MeshLoadingError("invalidIndex", str(raw), self.line)
        return r

    def face_vertex(self, np_, nt, nn) -> Tuple[int, Optional[int], Optional[int]]:
        p = self.resolve(self.read_int(), np_)
        t = n = None
        if self._consume(47):
            if not self.at_end() and self.b[self.i] != 47 and not self.at_line_break() and not self.at_comment():
                t = self.resolve(self.read_int(), nt)
            if self._consume(47) and not self.at_end() and not self.at_line_break() and not self.at_comment():
                n = self.resolve(self.read_int(), nn)
        return p, t, n


def parse_obj(data: bytes) -> Mesh:
    s = _ObjScanner(data)
    pos: List[Tuple[float, float, float]] = []
    tex: List[Tuple[float, float]] = []
    nrm: List[Tuple[float, float, float]] = []
    out_v: List[Tuple[float, float, float]] = []
    out_t: List[Tuple[float, float]] = []
    out_n: List[Tuple[float, float, float]] = []
    idx: List[int] = []
    packed: Dict[Tuple[int, Optional[int], Optional[int]], int] = {}

    def append(fv):
        k = packed.get(fv)
        if k is not None:
            idx.append(k)
            return
        k = len(out_v)
        packed[fv] = k
        out_v.append(pos[fv[0]])
        out_t.append(tex[fv[1]] if fv[1] is not None else (0.0, 0.0))
        out_n.append(nrm[fv[2]] if fv[2] is not None else (0.0, 0.0, 0.0))
        idx.append(k)

    while not s.at_end():
        s.skip_ws()
        if s.at_line_break():
            s.i += 1
            s.line += 1
            continue
        if s.at_comment():
            s.skip_line()
            continue
        cmd = s.token()
        line = s.line
        if cmd == b"v":
            pos.append((s.read_float(), s.read_float(), s.read_float()))
        elif cmd == b"vt":
            tex.append((s.read_float(), s.read_float()))
        elif cmd == b"vn":
            n = np.array([s.read_float(), s.read_float(), s.read_float()], f32)
            nrm.append(tuple(float(x) for x in _normalize(n)))
        elif cmd == b"f":
            first = prev = None
            count = 0
            while True:
                s.skip_ws()
                if s.at_end() or s.at_line_break() or s.at_comment():
                    break
                fv = s.face_vertex(len(pos), len(tex), len(nrm))
                if count == 0:
                    first = fv
                elif count == 1:
                    prev = fv
                else:
                    append(first)
                    append(prev)
                    append(fv)
                    prev = fv
                count += 1
            if count < 3:
                raise MeshLoadingError("invalidOBJ", "face requires at least three vertices", line)
        s.skip_line()
    if not out_v or not idx:
        raise MeshLoadingError("invalidOBJ", "mesh contains no faces", 0)
    return Mesh(np.array(out_v, f32), np.array(idx, np.uint32), normals=np.array(out_n, f32),
                texcoords=np.array(out_t, f32))


# --------------------------------------------------------------------- PLY --
_PLY_TYPES = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h",
              "ushort": "H", "uint16": "H", "int": "i", "int32": "i", "uint": "I", "uint32": "I",
              "float": "f", "float32": "f", "double": "d", "float64": "d"}
_U_NAMES = ("u", "s", "texture_u", "texture_s")
_V_NAMES = ("v", "t", "texture_v", "texture_t")


def _ply_header(data: bytes):
    end = data.find(b"end_header\n")
    mlen = len(b"end_header\n")
    if end < 0:
        end = data.find(b"end_header\r\n")
        mlen = len(b"end_header\r\n")
    if end < 0:
        raise MeshLoadingError("invalidPLY", "missing end_header", 0)
    try:
        text = data[:end + mlen].decode("utf-8")
    except UnicodeDecodeError:
        raise MeshLoadingError("invalidPLY", "header is not UTF-8", 0)
    lines = [ln for ln in text.replace("\r\n", "\n").replace("\r", "\n").split("\n") if ln]
    if not lines or lines[0] != "ply":
        raise MeshLoadingError("invalidPLY", "missing ply header", 1)
    fmt = None
    elements: List[list] = []  # [name, count, [(name, type|None, count_type|None, value_type|None)]]
    for k, ln in enumerate(lines[1:]):
        no = k + 2
        t = ln.split()
        if not t or t[0] in ("comment", "obj_info", "end_header"):
            continue
        if t[0] == "format":
            if len(t) < 3 or t[2] != "1.0":
                raise MeshLoadingError("invalidPLY", "unsupported format declaration", no)
            if t[1] == "ascii":
                fmt = "ascii"
            elif t[1] == "binary_little_endian":
                fmt = "binary"
            else:
                raise MeshLoadingError("invalidPLY", f"unsupported format '{t[1]}'", no)
        elif t[0] == "element":
            cnt = parse_int(t[2]) if len(t) == 3 else None
            if cnt is None or cnt < 0:
                raise MeshLoadingError("invalidPLY", "invalid element declaration", no)
            elements.append([t[1], cnt, []])
        elif t[0] == "property":
            if not elements:
                raise MeshLoadingError("invalidPLY", "property before element", no)
            if len(t) == 5 and t[1] == "list":
                if t[2] not in _PLY_TYPES or t[3] not in _PLY_TYPES:
                    raise MeshLoadingError("invalidPLY", "unsupported list property type", no)
                elements[-1][2].append((t[4], None, _PLY_TYPES[t[2]], _PLY_TYPES[t[3]]))
            elif len(t) == 3:
                if t[1] not in _PLY_TYPES:
                    raise MeshLoadingError("invalidPLY", f"unsupported property type '{t[1]}'", no)
                elements[-1][2].append((t[2], _PLY_TYPES[t[1]], None, None))
            else:
                raise MeshLoadingError("invalidPLY", "invalid property declaration", no)
    if fmt is None:
        raise MeshLoadingError("invalidPLY", "missing format declaration", 0)
    return fmt, elements, end + mlen


def _append_face(face, nverts: int, out: List[int], line: int) -> None:
    if len(face) < 3:
        return
    for i in face:
        if i < 0 or i >= nverts:
            raise MeshLoadingError("invalidIndex", str(i), line)
    for k in range(1, len(face) - 1):
        out += (face[0], face[k], face[k + 1])


def _vertex_columns(props):
    """Index of each vertex attribute among the scalar properties (the last
    declaration of a name wins, as the reference assigns in order)."""
    cols = {}
    for j, (name, typ, _, _) in enumerate(props):
        if typ is None:
            continue
        for key, names in (("x", ("x",)), ("y", ("y",)), ("z", ("z",)), ("nx", ("nx",)), ("ny", ("ny",)),
                           ("nz", ("nz",)), ("u", _U_NAMES), ("v", _V_NAMES)):
            if name in names:
                cols[key] = j
    return cols


def _finish(positions, normals, texcoords, indices) -> Mesh:
    if len(positions) == 0 or not indices:
        raise MeshLoadingError("invalidPLY", "mesh contains no faces", 0)
    return Mesh(positions, np.array(indices, np.uint32), normals=normals, texcoords=texcoords)


def _vertex_arrays(values: np.ndarray, cols) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """values: (n, scalar props) float64 of one vertex element."""
    n = len(values)

    def col(k):
        return values[:, cols[k]].astype(f32) if k in cols else np.zeros(n, f32)

    p = np.stack([col("x"), col("y"), col("z")], 1)
    nr = np.stack([col("nx"), col("ny"), col("nz")], 1)
    uv = np.stack([col("u"), col("v")], 1)
    return p, nr, uv


def parse_ply(data: bytes) -> Mesh:
    fmt, elements, body = _ply_header(data)
    if fmt == "ascii":
        return _parse_ply_ascii(data[body:], elements)
    return _parse_ply_binary(data, body, elements)


def _parse_ply_ascii(body: bytes, elements) -> Mesh:
    try:
        text = body.decode("utf-8")
    except UnicodeDecodeError:
        raise MeshLoadingError("invalidPLY", "ASCII body is not UTF-8", 0)
    lines = [ln for ln in text.replace("\r\n", "\n").replace("\r", "\n").split("\n") if ln]
    li = 0
    positions = normals = texcoords = np.zeros((0, 3), f32)
    texcoords = np.zeros((0, 2), f32)
    indices: List[int] = []
    nverts = 0
    for name, count, props in elements:
        vrows = []
        for e in range(count):
            if li >= len(lines):
                raise MeshLoadingError("invalidPLY", "unexpected end of file", e + 1)
            toks = lines[li].split()
            li += 1
            ti = 0
            if name == "vertex":
                row = []
                for (pname, typ, _, _) in props:
                    if typ is None:
                        ti = _ascii_list(toks, ti, e + 1)[1]
                    else:
                        if ti >= len(toks):
                            raise MeshLoadingError("invalidPLY", "missing vertex property", e + 1)
                        v = parse_float32(toks[ti])
                        if v is None:
                            raise MeshLoadingError("invalidPLY", f"invalid number '{toks[ti]}'", e + 1)
                        row.append(v)
                        ti += 1
                vrows.append(row)
            elif name == "face":
                for (pname, typ, _, _) in props:
                    if typ is None:
                        vals, ti = _ascii_list(toks, ti, e + 1)
                        if pname in ("vertex_indices", "vertex_index"):
                            _append_face(vals, nverts, indices, e + 1)
                    else:
                        if ti >= len(toks):
                            raise MeshLoadingError("invalidPLY", "missing face property", e + 1)
                        ti += 1
            else:
                for (pname, typ, _, _) in props:
                    if typ is None:
                        ti = _ascii_list(toks, ti, e + 1)[1]
                    else:
                        if ti >= len(toks):
                            raise MeshLoadingError("invalidPLY", "missing property", e + 1)
                        ti += 1
        if name == "vertex":
            scal = [(j, p) for j, p in enumerate(props) if p[1] is not None]
            remap = {j: k for k, (j, _) in enumerate(scal)}
            cols = {key: remap[j] for key, j in _vertex_columns(props).items()}
            if "x" not in cols or "y" not in cols or "z" not in cols:
                if count:
                    raise MeshLoadingError("invalidPLY", "vertex missing x y z", 1)
            vals = np.array(vrows, np.float64).reshape(count, len(scal))
            p, nr, uv = _vertex_arrays(vals, cols)
            positions = np.concatenate([positions, p])
            normals = np.concatenate([normals, nr])
            texcoords = np.concatenate([texcoords, uv])
            nverts = len(positions)
    return _finish(positions, normals, texcoords, indices)


def _ascii_list(toks, ti, line):
    if ti >= len(toks):
        raise MeshLoadingError("invalidPLY", "missing list count", line)
    n = parse_int(toks[ti])
    if n is None:
        raise MeshLoadingError("invalidPLY", f"invalid integer '{toks[ti]}'", line)
    ti += 1
    if n < 0 or ti + n > len(toks):
        raise MeshLoadingError("invalidPLY", "invalid list length", line)
    vals = []
    for k in range(n):
        v = parse_int(toks[ti + k])
        if v is None:
            raise MeshLoadingError("invalidPLY", f"invalid integer '{toks[ti + k]}'", line)
        vals.append(v)
    return vals, ti + n


def _parse_ply_binary(data: bytes, off: int, elements) -> Mesh:
    positions = np.zeros((0, 3), f32)
    normals = np.zeros((0, 3), f32)
    texcoords = np.zeros((0, 2), f32)
    indices: List[int] = []
    n_data = len(data)
    for name, count, props in elements:
        has_list = any(p[1] is None for p in props)
        if not has_list and count:
            # Fixed-size records: one structured read.
            dt = np.dtype([(f"p{j}", "<" + p[1]) for j, p in enumerate(props)])
            need = dt.itemsize * count
            if off + need > n_data:
                raise MeshLoadingError("invalidPLY", "unexpected end of binary body",
                                       (n_data - off) // max(dt.itemsize, 1) + 1)
            rec = np.frombuffer(data, dt, count, off)
            off += need
            if name == "vertex":
                cols = _vertex_columns(props)
                if "x" not in cols or "y" not in cols or "z" not in cols:
                    raise MeshLoadingError("invalidPLY", "vertex missing x y z", 1)
                vals = np.stack([rec[f"p{j}"].astype(np.float64) for j in range(len(props))], 1)
                p, nr, uv = _vertex_arrays(vals, cols)
                positions = np.concatenate([positions, p])
                normals = np.concatenate([normals, nr])
                texcoords = np.concatenate([texcoords, uv])
            continue
        fast = _binary_faces_fast(data, off, count, props) if name == "face" else None
        if fast is not None:
            tri, off = fast
            if len(tri):
                bad = (tri < 0) | (tri >= len(positions))
                if bad.any():
                    r = int(np.argmax(bad.any(1)))
                    raise MeshLoadingError("invalidIndex", str(int(tri[r][bad[r]][0])), r + 1)
            indices += tri.reshape(-1).tolist()
            continue
        vrows = []
        for e in range(count):
            row = []
            for (pname, typ, ct, vt) in props:
                if typ is None:
                    vals, off = _read_list(data, off, ct, vt, e + 1)
                    if name == "face" and pname in ("vertex_indices", "vertex_index"):
                        _append_face(vals, len(positions), indices, e + 1)
                else:
                    sz = struct.calcsize("<" + typ)
                    if off + sz > n_data:
                        raise MeshLoadingError("invalidPLY", "unexpected end of binary body", e + 1)
                    if name == "vertex":
                        row.append(float(struct.unpack_from("<" + typ, data, off)[0]))
                    off += sz
            if name == "vertex":
                vrows.append(row)
        if name == "vertex" and count:
            scal = [(j, p) for j, p in enumerate(props) if p[1] is not None]
            remap = {j: k for k, (j, _) in enumerate(scal)}
            cols = {key: remap[j] for key, j in _vertex_columns(props).items()}
            if "x" not in cols or "y" not in cols or "z" not in cols:
                raise MeshLoadingError("invalidPLY", "vertex missing x y z", 1)
            p, nr, uv = _vertex_arrays(np.array(vrows, np.float64), cols)
            positions = np.concatenate([positions, p])
            normals = np.concatenate([normals, nr])
            texcoords = np.concatenate([texcoords, uv])
    return _finish(positions, normals, texcoords, indices)


def _read_list(data, off, ct, vt, line):
    csz, vsz = struct.calcsize("<" + ct), struct.calcsize("<" + vt)
    if off + csz > len(data):
        raise MeshLoadingError("invalidPLY", "unexpected end of binary body", line)
    n = int(struct.unpack_from("<" + ct, data, off)[0])
    off += csz
    if n < 0:
        raise MeshLoadingError("invalidPLY", "negative list length", line)
    if off + n * vsz > len(data):
        raise MeshLoadingError("invalidPLY", "unexpected end of binary body", line)
    vals = [int(x) for x in struct.unpack_from("<" + vt * n, data, off)]
    return vals, off + n * vsz


def _binary_faces_fast(data, off, count, props):
    """Faces that are exactly one vertex-index list of triangles (the common
    layout): one strided read instead of a loop per face; None otherwise."""
    if len(props) != 1 or props[0][1] is not None or props[0][0] not in ("vertex_indices", "vertex_index"):
        return None
    ct, vt = props[0][2], props[0][3]
    csz, vsz = struct.calcsize("<" + ct), struct.calcsize("<" + vt)
    rec = csz + 3 * vsz
    if count == 0 or off + rec * count > len(data):
        return None
    dt = np.dtype([("n", "<" + ct), ("i", "<" + vt, (3,))])
    arr = np.frombuffer(data, dt, count, off)
    if not (arr["n"] == 3).all():
        return None
    return arr["i"].astype(np.int64), off + rec * count
