"""CPU restatement of the reference's dataset input (TEST INFRASTRUCTURE ONLY: imported by
tests/ as the checker of the C++ loaders in rayspace3d_amd/host/rs3d_io.cpp; the product
never imports it). Pure Python, for small files.

parse_points   query/src/runtime/PointIO.cpp:12-72: fgets into a 1024-byte buffer (a longer
               line arrives in pieces of at most 1023 bytes, each parsed as a line); leading
               blanks skipped; empty, '\\r', '\\n' and '#' pieces ignored; "POINT..." takes
               three strtof after the first '('; a piece starting with a digit or '-' takes
               three strtof from its start and counts when the first one consumed anything.
               strtof is libc's own (ctypes), the function the reference calls.
load_obj       preprocess/src/ObjMeshDatasetLoader.cpp:14-169 over tinyobjloader's LoadObj
               (third-party, not vendored in the reference and absent here; its published
               v1.0.7+/v2.0 behaviour is restated: tryParseDouble / parseReal for numbers,
               fixIndex / parseTriple for face corners, shapes flushed at 'g' and 'o' lines,
               faces with fewer than 3 corners dropped, quads split along the shorter
               diagonal computed in float, quads with a vertex beyond those read so far
               dropped). Vertices of an object are listed in ascending file order (the
               reference iterates an unordered_set: its order is unspecified, the geometry
               is not). Faces with more than 4 corners are rejected (not restated).

Pinning: parse_points and the number parser are checked against hand-computed values and
against strtof on decimals it rounds correctly (tests/test_host_io.py); the OBJ rules
against hand-built files with known triangles. The long-decimal rounding path of
tinyobjloader's parser is restated, not run (the library is absent): parity unpinned there.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

_libc = C.CDLL(None)
_libc.strtof.restype = C.c_float
_libc.strtof.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
LINE_BUF = 1024
_NEG_POW10 = [1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001]


def _strtof(buf: bytes, pos: int) -> tuple[np.float32, int]:
    """libc strtof on buf[pos:] (NUL-terminated); returns (value, end position)."""
    src = C.create_string_buffer(buf[pos:])
    end = C.c_char_p()
    v = _libc.strtof(src, C.byref(end))
    consumed = C.cast(end, C.c_void_p).value - C.addressof(src)
    return np.float32(v), pos + consumed


def _parse_piece(piece: bytes, out: list) -> None:
    piece = piece.split(b"\0", 1)[0]          # C string functions stop at a NUL
    i = 0
    while i < len(piece) and piece[i] in b" \t":
        i += 1
    if i == len(piece) or piece[i] in b"\r\n#":
        return
    if piece.startswith(b"POINT", i):
        j = piece.find(b"(", i)
        if j < 0:
            return
        x, e = _strtof(piece, j + 1)
        y, e = _strtof(piece, e)
        z, e = _strtof(piece, e)
        out.append((x, y, z))
    elif piece[i] in b"0123456789-":
        x, e = _strtof(piece, i)
        y, e = _strtof(piece, e)
        z, e = _strtof(piece, e)
        if e != i:
            out.append((x, y, z))


def parse_points(data: bytes) -> np.ndarray:
    """Points of a text point file's bytes, as the reference's loader reads them (float32 [n,3])."""
    out: list = []
    pos = 0
    while pos < len(data):
        nl = data.find(b"\n", pos)
        line_end = len(data) if nl < 0 else nl + 1
        while pos < line_end:
            n = min(line_end - pos, LINE_BUF - 1)
            _parse_piece(data[pos:pos + n], out)
            pos += n
    return np.array(out, np.float32).reshape(-1, 3)


# ---- tinyobjloader number and index rules --------------------------------------------

def _digit(c: str) -> bool:
    return "0" <= c <= "9"


def tobj_parse_double(s: str):
    """tryParseDouble over the token s; None when it fails."""
    if not s:
        return None
    i, n = 0, len(s)
    neg = exp_neg = lead_dot = False
    m, ex = 0.0, 0
    if s[0] in "+-":
        neg = s[0] == "-"
        i = 1
        if i < n and s[i] == ".":
            lead_dot = True
    elif s[0] == ".":
        lead_dot = True
    elif not _digit(s[0]):
        return None
    if not lead_dot:
        d = 0
        while i < n and _digit(s[i]):
            m = m * 10.0
            m = m + int(s[i])
            i += 1
            d += 1
        if d == 0:
            return None
    if i < n:
        try_exp = True
        if s[i] == ".":
            i += 1
            k = 1
            while i < n and _digit(s[i]):
                m = m + int(s[i]) * (_NEG_POW10[k] if k < 8 else math.pow(10.0, -k))
                k += 1
                i += 1
            try_exp = i < n
        elif s[i] not in "eE":
            try_exp = False
        if try_exp and s[i] in "eE":
            i += 1
            if i < n and s[i] in "+-":
                exp_neg = s[i] == "-"
                i += 1
            elif not (i < n and _digit(s[i])):
                return None
            d = 0
            while i < n and _digit(s[i]):
                if ex > 2147483647 // 10:
                    return None
                ex = ex * 10 + int(s[i])
                i += 1
                d += 1
            if exp_neg:
                ex = -ex
            if d == 0:
                return None
    v = math.ldexp(m * math.pow(5.0, ex), ex) if ex else m
    return -v if neg else v


def _parse_real(line: str, pos: int):
    while pos < len(line) and line[pos] in " \t":
        pos += 1
    end = pos
    while end < len(line) and line[end] not in " \t\r":
        end += 1
    v = tobj_parse_double(line[pos:end])
    return np.float32(0.0 if v is None else v), end


def _atoi(line: str, pos: int) -> int:
    """C atoi: leading whitespace, optional sign, digits."""
    while pos < len(line) and line[pos] in " \t\n\v\f\r":
        pos += 1
    sign = 1
    if pos < len(line) and line[pos] in "+-":
        sign = -1 if line[pos] == "-" else 1
        pos += 1
    v = 0
    while pos < len(line) and _digit(line[pos]):
        v = v * 10 + int(line[pos])
        pos += 1
    return sign * v


def _fix_index(idx: int, n: int, allow_zero: bool):
    if idx > 0:
        return idx - 1
    if idx == 0:
        return -1 if allow_zero else None
    r = n + idx
    return r if r >= 0 else None


def _skip_to(line: str, pos: int, stops: str) -> int:
    while pos < len(line) and line[pos] not in stops:
        pos += 1
    return pos


def _corner(line: str, pos: int, nv: int, nvt: int, nvn: int):
    v = _fix_index(_atoi(line, pos), nv, False)
    if v is None:
        return None, pos
    pos = _skip_to(line, pos, "/ \t\r")
    if pos >= len(line) or line[pos] != "/":
        return v, pos
    pos += 1
    if pos < len(line) and line[pos] == "/":
        pos += 1
        if _fix_index(_atoi(line, pos), nvn, True) is None:
            return None, pos
        return v, _skip_to(line, pos, "/ \t\r")
    if _fix_index(_atoi(line, pos), nvt, True) is None:
        return None, pos
    pos = _skip_to(line, pos, "/ \t\r")
    if pos >= len(line) or line[pos] != "/":
        return v, pos
    pos += 1
    if _fix_index(_atoi(line, pos), nvn, True) is None:
        return None, pos
    return v, _skip_to(line, pos, "/ \t\r")


def _lines(text: str):
    """Lines split at '\\n', '\\r\\n' or '\\r' (tinyobjloader's safeGetline)."""
    i, n = 0, len(text)
    while i < n:
        j = i
        while j < n and text[j] not in "\r\n":
            j += 1
        yield text[i:j]
        if j < n and text[j] == "\r" and j + 1 < n and text[j + 1] == "\n":
            i = j + 2
        else:
            i = j + 1


def _export(faces, v, tris):
    for f in faces:
        if len(f) < 3:
            continue
        if len(f) == 3:
            tris.append(tuple(f))
            continue
        if len(f) > 4:
            raise ValueError("faces with more than 4 vertices are not supported")
        if any(k * 3 + 2 >= len(v) for k in f):
            continue
        p = [np.array(v[k * 3:k * 3 + 3], np.float32) for k in f]
        e02 = p[2] - p[0]
        e13 = p[3] - p[1]
        d02 = e02[0] * e02[0] + e02[1] * e02[1] + e02[2] * e02[2]
        d13 = e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2]
        i0, i1, i2, i3 = f
        if d02 < d13:
            tris += [(i0, i1, i2), (i0, i2, i3)]
        else:
            tris += [(i0, i1, i3), (i1, i2, i3)]


def obj_shapes(text: str):
    """(vertex floats, [triangle list per shape]) or None when LoadObj would fail."""
    v: list = []
    nvt = nvn = 0
    shapes, faces, tris = [], [], []

    def flush():
        nonlocal faces, tris
        _export(faces, v, tris)
        faces = []
        if tris:
            shapes.append(tris)
        tris = []

    for line in _lines(text):
        line = line.split("\0", 1)[0]
        t = 0
        while t < len(line) and line[t] in " \t":
            t += 1
        s = line[t:]
        if not s or s[0] == "#":
            continue
        c1 = s[1] if len(s) > 1 else "\0"
        c2 = s[2] if len(s) > 2 else "\0"
        if s[0] == "v" and c1 in " \t":
            pos = 2
            for _ in range(3):
                val, pos = _parse_real(s, pos)
                v.append(val)
        elif s[0] == "v" and c1 == "t" and c2 in " \t":
            nvt += 1
        elif s[0] == "v" and c1 == "n" and c2 in " \t":
            nvn += 1
        elif s[0] == "f" and c1 in " \t":
            pos = 2
            while pos < len(s) and s[pos] in " \t":
                pos += 1
            face = []
            while pos < len(s) and s[pos] not in "\r\n":
                vi, pos = _corner(s, pos, len(v) // 3, nvt, nvn)
                if vi is None:
                    return None
                face.append(vi)
                while pos < len(s) and s[pos] in " \t\r":
                    pos += 1
            faces.append(face)
        elif s[0] in "go" and c1 in " \t":
            flush()
    flush()
    return v, shapes


def load_obj(text: str):
    """GeometryData of an OBJ file as the reference's loader builds it: (vertices float32
    [n,3], indices uint32 [m,3], tri_to_obj int32 [m]); None when loading fails."""
    r = obj_shapes(text)
    if r is None or not r[1]:
        return None
    v, shapes = r
    nv = len(v) // 3
    verts, idx, obj = [], [], []
    o = 0
    for tris in shapes:
        used = sorted({k for t in tris for k in t})
        remap = {}
        for k in used:
            if k < nv:
                remap[k] = len(verts)
                verts.append(v[k * 3:k * 3 + 3])
        for t in tris:
            if all(k in remap for k in t):
                idx.append([remap[k] for k in t])
                obj.append(o)
        o += 1
    return (np.array(verts, np.float32).reshape(-1, 3), np.array(idx, np.uint32).reshape(-1, 3),
            np.array(obj, np.int32))
