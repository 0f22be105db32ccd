"""Preprocessing of generated meshes into .pre files (the role of
preprocess/src/preprocess_dataset.cpp:183-385 for this build's synthetic inputs):
unique-edge payload and, optionally, the grid histogram computed on the GPU."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, gen, preio
from .device import Context, to_dev, to_host


def extract_edges_gpu(ctx: Context, mesh: gen.Mesh) -> gen.EdgeSet:
    """Unique edges of a mesh on the GPU (rs3d_extract_edges): the same records, in the
    same order, as gen.extract_edges."""
    dV, dI, dO = to_dev(mesh.vertices), to_dev(mesh.indices), to_dev(mesh.tri_to_obj)
    ne, ns = C.c_uint64(), C.c_uint64()
    torch.cuda.synchronize()
    args = (ctx.handle, C.c_void_p(dV.data_ptr()), len(mesh.vertices), C.c_void_p(dI.data_ptr()),
            len(mesh.indices), C.c_void_p(dO.data_ptr()))
    # a mesh of t triangles has at most 3t edges and 3t edge sources: one pass with outputs
    # of that size (HBM is plentiful), the sizing pass only if that were ever exceeded
    n_e = n_s = 3 * len(mesh.indices)
    p = lambda t: C.c_void_p(t.data_ptr())
    while True:
        st = torch.empty((max(n_e, 1), 3), dtype=torch.float32, device="cuda")
        en = torch.empty_like(st)
        off = torch.empty(max(n_e, 1), dtype=torch.int32, device="cuda")
        cnt = torch.empty_like(off)
        src = torch.empty(max(n_s, 1), dtype=torch.int32, device="cuda")
        rc = ctx.lib.rs3d_extract_edges(*args, p(st), p(en), p(off), p(cnt), p(src), n_e, n_s,
                                        C.byref(ne), C.byref(ns))
        if rc == _lib.RS3D_E_CAPACITY and (ne.value > n_e or ns.value > n_s):
            n_e, n_s = ne.value, ns.value
            continue
        _lib.check(rc, "rs3d_extract_edges")
        break
    n_e, n_s = ne.value, ns.value
    # to_host returns fresh host arrays (no second copy needed)
    return gen.EdgeSet(to_host(st[:n_e]), to_host(en[:n_e]), to_host(off[:n_e]), to_host(cnt[:n_e]),
                       to_host(src[:n_s]))


def grid_stats_gpu(ctx: Context, mesh: gen.Mesh, res: int, world_size: float = -1.0) -> preio.Grid:
    n_obj = int(mesh.tri_to_obj.max()) + 1 if len(mesh.tri_to_obj) else 0
    cells = torch.zeros(res ** 3 * 16, dtype=torch.uint8, device="cuda")
    dV, dI, dO = to_dev(mesh.vertices), to_dev(mesh.indices), to_dev(mesh.tri_to_obj)
    b = (C.c_float * 6)()
    torch.cuda.synchronize()
    _lib.check(ctx.lib.rs3d_grid_stats(ctx.handle, C.c_void_p(dV.data_ptr()), len(mesh.vertices),
                                       C.c_void_p(dI.data_ptr()), len(mesh.indices),
                                       C.c_void_p(dO.data_ptr()), n_obj, res, world_size,
                                       C.c_void_p(cells.data_ptr()), b), "rs3d_grid_stats")
    c = to_host(cells).view(preio.GRID_CELL).copy()
    return preio.Grid(np.array(b[:3], np.float32), np.array(b[3:], np.float32), (res, res, res), c)


def write_dataset(path: str, mesh: gen.Mesh, ctx: Context | None = None, grid_res: int = 0,
                  world_size: float = -1.0) -> None:
    """Writes mesh (+ unique edges, + grid histogram when grid_res > 0) as .pre v2. Edges
    and grid come from the GPU when a context is given."""
    if mesh.edges is None:
        mesh.edges = extract_edges_gpu(ctx, mesh) if ctx is not None else gen.extract_edges(mesh)
    grid = grid_stats_gpu(ctx, mesh, grid_res, world_size) if grid_res > 0 else None
    preio.write_pre(path, mesh, grid)


def write_points_wkt(path: str, pts: np.ndarray) -> None:
    with open(path, "w") as f:
        for p in pts.astype(np.float64):
            f.write(f"POINT Z ({float(p[0])!r} {float(p[1])!r} {float(p[2])!r})\n")
