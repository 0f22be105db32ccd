"""The .pre binary geometry format (common/include/BinaryIO.h:11-180).

v2 (current): 64-byte header {u32 magic 'R3DB', u32 version, u64 numVertices,
numIndices, numMappings, numEdges, numEdgeSourceObjects, totalTriangles, u8 hasGrid,
u8 hasEdges, 6 pad}, then float3 vertices, uint3 indices, int32 triangle->object,
[edge payload: float3 starts, float3 ends, int32 offsets, int32 counts, int32 sources],
[grid: 40-byte GridParams {float min[3], max[3], u32 resolution[3], u32 pad} and
GridCell[res^3] {u32 CenterCount, u32 TouchCount, f32 AvgSizeMean, f32 VolRatio}].
v1 (older files): 48-byte header {magic, version, numVertices, numIndices, numMappings,
totalTriangles, u8 hasGrid, 7 pad}, no edge payload, same grid block.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from .gen import EdgeSet, Mesh

MAGIC = 0x52334442
VERSION = 2
GRID_CELL = np.dtype([("center", "<u4"), ("touch", "<u4"), ("avg", "<f4"), ("ratio", "<f4")])


@dataclass
class Grid:
    min_bound: np.ndarray
    max_bound: np.ndarray
    resolution: tuple
    cells: np.ndarray


@dataclass
class PreFile:
    mesh: Mesh
    version: int
    total_triangles: int
    grid: Grid | None


def read_pre(path: str) -> PreFile:
    with open(path, "rb") as f:
        buf = f.read()
    magic, ver = struct.unpack_from("<II", buf, 0)
    if magic != MAGIC:
        raise ValueError(f"{path}: invalid magic {magic:#x}")
    if ver == 2:
        nv, ni, nm, ne, nes, tt, hg, he = struct.unpack_from("<QQQQQQBB", buf, 8)
        off = 64
    elif ver == 1:
        nv, ni, nm, tt, hg = struct.unpack_from("<QQQQB", buf, 8)
        ne = nes = he = 0
        off = 48
    else:
        raise ValueError(f"{path}: unsupported version {ver}")

    def take(dtype, count, shape=None):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.reshape(shape) if shape else a

    V = take("<f4", 3 * nv, (nv, 3)).copy()
    I = take("<u4", 3 * ni, (ni, 3)).copy()
    O = take("<i4", nm).copy()
    edges = None
    if he:
        S = take("<f4", 3 * ne, (ne, 3)).copy()
        E = take("<f4", 3 * ne, (ne, 3)).copy()
        offs = take("<i4", ne).copy()
        cnts = take("<i4", ne).copy()
        src = take("<i4", nes).copy()
        edges = EdgeSet(S, E, offs, cnts, src)
    grid = None
    if hg:
        gp = struct.unpack_from("<6f4I", buf, off)
        off += 40
        res = gp[6:9]
        cells = take(GRID_CELL, res[0] * res[1] * res[2]).copy()
        grid = Grid(np.array(gp[0:3], np.float32), np.array(gp[3:6], np.float32), tuple(res), cells)
    nobj = int(O.max()) + 1 if len(O) else 0
    mesh = Mesh(V, I, O, num_objects=nobj, edges=edges)
    return PreFile(mesh, ver, int(tt), grid)


def write_pre(path: str, mesh: Mesh, grid: Grid | None = None) -> None:
    """Writes format v2 (padding bytes are written as zeros)."""
    e = mesh.edges
    has_edges = e is not None and e.num_edges > 0
    hdr = struct.pack("<IIQQQQQQBB6x", MAGIC, VERSION, len(mesh.vertices), len(mesh.indices),
                      len(mesh.tri_to_obj), e.num_edges if has_edges else 0,
                      len(e.src_objects) if has_edges else 0, len(mesh.indices),
                      1 if grid is not None else 0, 1 if has_edges else 0)
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(mesh.vertices, "<f4").tobytes())
        f.write(np.ascontiguousarray(mesh.indices, "<u4").tobytes())
        f.write(np.ascontiguousarray(mesh.tri_to_obj, "<i4").tobytes())
        if has_edges:
            for a, dt in ((e.starts, "<f4"), (e.ends, "<f4"), (e.src_offsets, "<i4"),
                          (e.src_counts, "<i4"), (e.src_objects, "<i4")):
                f.write(np.ascontiguousarray(a, dt).tobytes())
        if grid is not None:
            f.write(struct.pack("<6f4I", *grid.min_bound.tolist(), *grid.max_bound.tolist(),
                                *grid.resolution, 0))
            f.write(np.ascontiguousarray(grid.cells, GRID_CELL).tobytes())
