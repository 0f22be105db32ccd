"""Python mirror of the join's host interfaces over librs3d (device arrays are torch
tensors on a ROCm device; torch only provides HBM allocations and host<->device copies).

Names follow the reference's host classes:
  Context                <- OptixContext          (query/src/optix/OptixContext.cpp:4-8)
  Bvh                    <- OptixAccelerationStructure (query/src/optix/OptixAccelerationStructure.h:9-34)
  RayLauncher            <- RayLauncher           (query/src/raytracing/RayLauncher.h:9-40)
  FilterRefinePipeline   <- FilterRefinePipeline  (query/src/raytracing/FilterRefine.h:27-48)
  overlap_join           <- executeTwoPassQueryEdgesOptimized / the hash query
                            (query/src/applications/rayspace_mesh_overlap_main.cpp:42-146)
Every call goes to the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check

_TORCH_OF = {
    np.dtype(np.float32): torch.float32,
    np.dtype(np.int32): torch.int32,
    np.dtype(np.uint32): torch.int32,
    np.dtype(np.int64): torch.int64,
    np.dtype(np.uint64): torch.int64,
    np.dtype(np.uint8): torch.uint8,
}


def to_dev(a: np.ndarray, device="cuda") -> torch.Tensor:
    """Copies a numpy array to HBM (unsigned types travel as their signed bit patterns)."""
    a = np.ascontiguousarray(a)
    view = a
    if a.dtype == np.uint32:
        view = a.view(np.int32)
    elif a.dtype == np.uint64:
        view = a.view(np.int64)
    return torch.from_numpy(view).to(device)


def to_host(t: torch.Tensor, dtype=None) -> np.ndarray:
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


# ---- argument checks ------------------------------------------------------------------
# The C ABI takes raw device pointers and cannot see dtypes, shapes or lengths, so the
# wrappers check them first and raise ValueError: a wrong dtype or a short buffer would
# otherwise be read or written out of bounds on the device. Checks run in the order dtype,
# shape, layout, index values, device.

_INT4 = tuple(t for t in (torch.int32, getattr(torch, "uint32", None)) if t is not None)


def _need(cond, msg):
    if not cond:
        raise ValueError(msg)


def _on_device(t, device, name):
    _need(t.is_cuda and (t.device.index if t.device.index is not None else torch.cuda.current_device()) == device,
          f"{name} must be on cuda:{device} (got {t.device})")


def check_points(points, device: int, name: str = "points"):
    # the common case in one test, without building the messages (a few us per call)
    if (isinstance(points, torch.Tensor) and points.dtype == torch.float32 and points.dim() == 2
            and points.shape[1] == 3 and points.is_contiguous() and points.is_cuda
            and points.get_device() == device):
        return
    _need(isinstance(points, torch.Tensor), f"{name} must be a torch tensor")
    _need(points.dtype == torch.float32, f"{name} must be float32 (got {points.dtype})")
    _need(points.dim() == 2 and points.shape[1] == 3, f"{name} must have shape (n, 3) (got {tuple(points.shape)})")
    _need(points.is_contiguous(), f"{name} must be contiguous")
    _on_device(points, device, name)


def check_index_map(index_map, n_points: int, device: int):
    """index_map: 4-byte integers, every value in [0, n_points)."""
    _need(isinstance(index_map, torch.Tensor), "index_map must be a torch tensor")
    _need(index_map.dtype in _INT4, f"index_map must hold 4-byte integers (int32), got {index_map.dtype}")
    _need(index_map.dim() == 1 and index_map.is_contiguous(), "index_map must be a contiguous 1-D tensor")
    if index_map.numel():
        lo, hi = int(index_map.min().item()), int(index_map.max().item())
        _need(0 <= lo and hi < n_points, f"index_map values must lie in [0, {n_points}) (got [{lo}, {hi}])")
    _on_device(index_map, device, "index_map")


_OUT_KINDS = {_lib.OUT_FLAG_U8: (torch.uint8, 1), _lib.OUT_OBJ_I32: (torch.int32, 1),
              _lib.OUT_RAY_RESULT: (torch.int32, 2)}


def check_out(out, out_kind: int, slots: int, device: int, name: str = "out"):
    """out: the dtype of out_kind with at least `slots` result slots (ray results: (m, 2))."""
    _need(out_kind in _OUT_KINDS, f"unknown out_kind {out_kind}")
    dt, width = _OUT_KINDS[out_kind]
    if (isinstance(out, torch.Tensor) and out.dtype == dt and out.dim() == (1 if width == 1 else 2)
            and (width == 1 or out.shape[1] == 2) and out.shape[0] >= slots and out.is_contiguous()
            and out.is_cuda and out.get_device() == device):
        return
    _need(isinstance(out, torch.Tensor), f"{name} must be a torch tensor")
    _need(out.dtype == dt, f"{name} must be {dt} for out_kind {out_kind} (got {out.dtype})")
    if width == 1:
        _need(out.dim() == 1, f"{name} must be 1-D (got shape {tuple(out.shape)})")
    else:
        _need(out.dim() == 2 and out.shape[1] == 2, f"{name} must have shape (m, 2) (got {tuple(out.shape)})")
    _need(out.shape[0] >= slots, f"{name} has {out.shape[0]} slots, {slots} needed")
    _need(out.is_contiguous(), f"{name} must be contiguous")
    _on_device(out, device, name)


def check_mesh_arrays(vertices, indices, tri_to_obj, device: int):
    _need(vertices.dtype == torch.float32 and vertices.dim() == 2 and vertices.shape[1] == 3,
          f"vertices must be float32 (nV, 3) (got {vertices.dtype} {tuple(vertices.shape)})")
    _need(indices.dtype in _INT4 and indices.dim() == 2 and indices.shape[1] == 3,
          f"indices must be 4-byte integers (nT, 3) (got {indices.dtype} {tuple(indices.shape)})")
    _need(vertices.is_contiguous() and indices.is_contiguous(), "vertices and indices must be contiguous")
    if tri_to_obj is not None:
        _need(tri_to_obj.dtype in _INT4 and tri_to_obj.dim() == 1 and tri_to_obj.shape[0] == indices.shape[0]
              and tri_to_obj.is_contiguous(), "tri_to_obj must be a contiguous int32 array, one per triangle")
    if indices.numel():
        lo, hi = int(indices.min().item()), int(indices.max().item())
        _need(0 <= lo and hi < vertices.shape[0],
              f"indices must lie in [0, {vertices.shape[0]}) (got [{lo}, {hi}])")
    for t, nm in ((vertices, "vertices"), (indices, "indices"), (tri_to_obj, "tri_to_obj")):
        if t is not None:
            _on_device(t, device, nm)


class Context:
    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.handle = C.c_void_p()
        torch.cuda.init()
        check(self.lib.rs3d_ctx_create(device, C.byref(self.handle)), "rs3d_ctx_create")
        self.device = device

    def sync(self):
        check(self.lib.rs3d_ctx_sync(self.handle), "rs3d_ctx_sync")

    def set_option(self, option: int, value: int):
        check(self.lib.rs3d_ctx_set_option(self.handle, option, value), "rs3d_ctx_set_option")

    def get_option(self, option: int) -> int:
        v = C.c_int64(0)
        check(self.lib.rs3d_ctx_get_option(self.handle, option, C.byref(v)), "rs3d_ctx_get_option")
        return v.value

    def profile(self, enable: bool):
        check(self.lib.rs3d_ctx_profile(self.handle, int(enable)), "rs3d_ctx_profile")

    def profile_read(self):
        ms = C.c_double()
        n = C.c_uint64()
        check(self.lib.rs3d_ctx_profile_read(self.handle, C.byref(ms), C.byref(n)),
              "rs3d_ctx_profile_read")
        return ms.value, n.value

    def close(self):
        if self.handle:
            self.lib.rs3d_ctx_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Bvh:
    """Linear BVH over a device-resident indexed mesh."""

    def __init__(self, ctx: Context, vertices: torch.Tensor, indices: torch.Tensor,
                 tri_to_obj: torch.Tensor | None, leaf_size: int = 1):
        self.ctx = ctx
        self.handle = C.c_void_p()
        check_mesh_arrays(vertices, indices, tri_to_obj, ctx.device)
        opts = _lib.BvhOpts()
        opts.leaf_size = leaf_size
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_bvh_build(ctx.handle, _ptr(vertices), vertices.shape[0], _ptr(indices),
                                     indices.shape[0], _ptr(tri_to_obj), C.byref(opts),
                                     C.byref(self.handle)), "rs3d_bvh_build")
        # keep the source arrays alive with the structure
        self._src = (vertices, indices, tri_to_obj)

    @classmethod
    def from_mesh(cls, ctx: Context, mesh, leaf_size: int = 1) -> "Bvh":
        b = cls(ctx, to_dev(mesh.vertices), to_dev(mesh.indices), to_dev(mesh.tri_to_obj),
                leaf_size)
        b.num_objects = int(mesh.tri_to_obj.max()) + 1 if len(mesh.tri_to_obj) else 0
        return b

    def view(self) -> _lib.MeshView:
        """The source mesh as an rs3d_mesh_view (rays that start on this mesh)."""
        v, i, o = self._src
        n_obj = getattr(self, "num_objects", None)
        if n_obj is None:
            n_obj = int(o.max().item()) + 1 if o is not None and o.numel() else 0
        return _lib.MeshView(v.data_ptr(), v.shape[0], i.data_ptr(), i.shape[0],
                             o.data_ptr() if o is not None else None, n_obj)

    def export(self):
        """Host copy of the tree: (nodes float32 [m, 12] child boxes as (lo.x, hi.x, lo.y,
        hi.y, lo.z, hi.z) of child 0 then child 1, links int32 [m, 4] = (link 0, link 1,
        count 0, count 1), triangles float32 [n, 3, 4] whose w lanes carry the original
        index and the object id as int32 bits)."""
        inf = self.info()
        nodes = torch.empty((max(inf.num_nodes, 1), 16), dtype=torch.int32, device="cuda")
        tris = torch.empty((max(inf.num_triangles, 1), 3, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        check(self.ctx.lib.rs3d_bvh_export(self.handle, _ptr(nodes), _ptr(tris)), "rs3d_bvh_export")
        nh = to_host(nodes[:inf.num_nodes]).copy()
        return nh[:, :12].view(np.float32), nh[:, 12:], to_host(tris[:inf.num_triangles]).copy()

    def info(self) -> _lib.BvhInfo:
        inf = _lib.BvhInfo()
        check(self.ctx.lib.rs3d_bvh_info_get(self.handle, C.byref(inf)), "rs3d_bvh_info_get")
        return inf

    def free(self):
        if self.handle:
            self.ctx.lib.rs3d_bvh_free(self.handle)
            self.handle = C.c_void_p()


def trace_points(ctx: Context, bvh: Bvh, points: torch.Tensor, axis: int = 2,
                 out_kind: int = _lib.OUT_OBJ_I32, index_map: torch.Tensor | None = None,
                 out: torch.Tensor | None = None, sync: bool = True) -> torch.Tensor:
    check_points(points, ctx.device)
    n = points.shape[0]
    _need(out_kind in _OUT_KINDS, f"unknown out_kind {out_kind}")
    if index_map is not None:
        check_index_map(index_map, n, ctx.device)
    if out is None:
        if out_kind == _lib.OUT_FLAG_U8:
            out = torch.empty(n, dtype=torch.uint8, device=points.device)
        elif out_kind == _lib.OUT_OBJ_I32:
            out = torch.empty(n, dtype=torch.int32, device=points.device)
        else:
            out = torch.empty((n, 2), dtype=torch.int32, device=points.device)
    # results land at out[index_map[i]] (index_map values < n) or out[i]
    check_out(out, out_kind, n, ctx.device)
    nrays = index_map.shape[0] if index_map is not None else n
    if sync:
        torch.cuda.synchronize()
    check(ctx.lib.rs3d_trace_points(ctx.handle, bvh.handle, _ptr(points), nrays, _ptr(index_map),
                                    axis, out_kind, _ptr(out)), "rs3d_trace_points")
    if sync:
        ctx.sync()
    return out


def trace_stats(ctx: Context, bvh: Bvh, points: torch.Tensor, axis: int = 2) -> dict:
    check_points(points, ctx.device)
    _need(axis in (0, 1, 2), f"axis must be 0, 1 or 2 (got {axis})")
    st = _lib.TraceStats()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_trace_points_stats(ctx.handle, bvh.handle, _ptr(points), points.shape[0],
                                          axis, C.byref(st)), "rs3d_trace_points_stats")
    return {k: int(getattr(st, k)) for k, _ in st._fields_}


def axis_grid_prepare(ctx: Context, bvh: Bvh, axis: int = 2) -> dict:
    """Builds (once) the axis grid of `bvh` for rays along +axis (rs3d_axis_grid_prepare);
    state 1 = accepted (point launches use it), 2 = rejected (they stay on the BVH)."""
    info = _lib.AxisGridInfo()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_axis_grid_prepare(ctx.handle, bvh.handle, axis, C.byref(info)), "rs3d_axis_grid_prepare")
    return {k: getattr(info, k) for k, _ in info._fields_}


def seg_occupancy_prepare(ctx: Context, bvh: Bvh) -> dict:
    """Builds (once) the segment occupancy bits of `bvh` (rs3d_seg_occupancy_prepare);
    state 1 = built (edge rays test them first), 2 = rejected (edge rays walk the tree)."""
    info = _lib.SegOccupancyInfo()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_seg_occupancy_prepare(ctx.handle, bvh.handle, C.byref(info)), "rs3d_seg_occupancy_prepare")
    return {k: getattr(info, k) for k, _ in info._fields_}


def axis_grid_stats(ctx: Context, bvh: Bvh, points: torch.Tensor, axis: int = 2) -> dict:
    """Work counters of a point launch on the axis grid: nodes_visited = cell lookups,
    tris_tested = 64-byte records read."""
    check_points(points, ctx.device)
    st = _lib.TraceStats()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_axis_grid_stats(ctx.handle, bvh.handle, _ptr(points), points.shape[0], axis, C.byref(st)),
          "rs3d_axis_grid_stats")
    return {k: int(getattr(st, k)) for k, _ in st._fields_}


@dataclass
class FilterRefineResult:
    final_results: torch.Tensor   # int32 per point, -1 outside
    candidate_count: int
    inside_count: int


class FilterRefinePipeline:
    """Dataset-box filter then exact containment (query/src/raytracing/FilterRefine.cpp:81-122)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def execute(self, bvh: Bvh, points: torch.Tensor, axis: int = 2,
                out: torch.Tensor | None = None, sync: bool = True) -> FilterRefineResult:
        check_points(points, self.ctx.device)
        n = points.shape[0]
        if out is None:
            out = torch.empty(n, dtype=torch.int32, device=points.device)
        check_out(out, _lib.OUT_OBJ_I32, n, self.ctx.device)
        cand = C.c_uint64()
        inside = C.c_uint64()
        if sync:
            torch.cuda.synchronize()
        check(self.ctx.lib.rs3d_filter_refine(self.ctx.handle, bvh.handle, _ptr(points), n, axis,
                                              _ptr(out), C.byref(cand), C.byref(inside)),
              "rs3d_filter_refine")
        return FilterRefineResult(out, cand.value, inside.value)


class DevEdges:
    """Device copy of an EdgeSet plus the C struct that points into it."""

    def __init__(self, edges):
        self.t = [to_dev(edges.starts), to_dev(edges.ends), to_dev(edges.src_offsets),
                  to_dev(edges.src_counts), to_dev(edges.src_objects if len(edges.src_objects)
                                                    else np.zeros(1, np.int32))]
        self.s = _lib.Edges(self.t[0].data_ptr(), self.t[1].data_ptr(), self.t[2].data_ptr(),
                            self.t[3].data_ptr(), self.t[4].data_ptr(), edges.num_edges)

    def spatial(self, ctx) -> "DevEdges":
        """The same edges in Morton order of their midpoints (rs3d_edges_spatial_order):
        faster edge-ray joins, identical join results. Shares the source list."""
        out = DevEdges.__new__(DevEdges)
        out.t = [torch.empty_like(self.t[0]), torch.empty_like(self.t[1]), torch.empty_like(self.t[2]),
                 torch.empty_like(self.t[3]), self.t[4]]
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_edges_spatial_order(ctx.handle, C.byref(self.s), _ptr(out.t[0]), _ptr(out.t[1]),
                                               _ptr(out.t[2]), _ptr(out.t[3])), "rs3d_edges_spatial_order")
        out.s = _lib.Edges(out.t[0].data_ptr(), out.t[1].data_ptr(), out.t[2].data_ptr(), out.t[3].data_ptr(),
                           out.t[4].data_ptr(), self.s.num_edges)
        return out


def edge_count(ctx, bvh, dedges: DevEdges) -> torch.Tensor:
    counts = torch.empty(max(dedges.s.num_edges, 1), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_edge_count(ctx.handle, bvh.handle, C.byref(dedges.s), _ptr(counts)),
          "rs3d_edge_count")
    ctx.sync()
    return counts[:dedges.s.num_edges]


def edge_stats(ctx, bvh, dedges: DevEdges) -> dict:
    """Work counters of the edge rays of `dedges` against `bvh` (rs3d_edge_stats)."""
    st = _lib.EdgeStats()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_edge_stats(ctx.handle, bvh.handle, C.byref(dedges.s), C.byref(st)),
          "rs3d_edge_stats")
    return {k: int(getattr(st, k)) for k, _ in st._fields_}


def overlap_join(ctx, bvh1, dedges1, bvh2, dedges2, mode: int = 0, max_out: int = 1 << 22) -> np.ndarray:
    """Sorted unique (mesh1, mesh2) keys of both edge directions (uint64 a << 32 | b)."""
    while True:
        out = torch.empty(max(max_out, 1), dtype=torch.int64, device="cuda")
        cnt = C.c_uint64()
        torch.cuda.synchronize()
        rc = ctx.lib.rs3d_overlap_join(ctx.handle, bvh1.handle, C.byref(dedges1.s), bvh2.handle,
                                       C.byref(dedges2.s), mode, _ptr(out), max_out, C.byref(cnt))
        if rc == _lib.RS3D_E_CAPACITY and cnt.value > max_out:
            max_out = int(cnt.value)
            continue
        check(rc, "rs3d_overlap_join")
        return to_host(out[:cnt.value], np.uint64)


def tri_intersect(ctx, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Exact triangle-triangle predicate on pairs (rs3d_tri_intersect): a, b float32
    device tensors [n, 3, 3] -> uint8 [n]."""
    _need(a.dtype == torch.float32 and b.dtype == torch.float32, "tri_intersect: a and b must be float32")
    _need(a.numel() % 9 == 0 and b.numel() % 9 == 0, "tri_intersect: a and b must hold 9 floats per triangle")
    a = a.reshape(-1, 9).contiguous()
    b = b.reshape(-1, 9).contiguous()
    _on_device(a, ctx.device, "a")
    _on_device(b, ctx.device, "b")
    if a.shape != b.shape:
        raise ValueError("tri_intersect: a and b must hold the same number of triangles")
    out = torch.empty(max(a.shape[0], 1), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_tri_intersect(ctx.handle, _ptr(a), _ptr(b), a.shape[0], _ptr(out)), "rs3d_tri_intersect")
    ctx.sync()
    return out[:a.shape[0]]


def tri_join(ctx, bvh_a, bvh_b, max_out: int = 1 << 22) -> np.ndarray:
    """Exact intersection join (rs3d_tri_join): sorted unique (a << 32 | b) keys."""
    while True:
        out = torch.empty(max(max_out, 1), dtype=torch.int64, device="cuda")
        cnt = C.c_uint64()
        torch.cuda.synchronize()
        rc = ctx.lib.rs3d_tri_join(ctx.handle, bvh_a.handle, bvh_b.handle, _ptr(out), max_out, C.byref(cnt))
        if rc == _lib.RS3D_E_CAPACITY and cnt.value > max_out:
            max_out = int(cnt.value)
            continue
        check(rc, "rs3d_tri_join")
        return to_host(out[:cnt.value], np.uint64)


def tri_join_stats(ctx, bvh_a, bvh_b) -> dict:
    st = _lib.TriJoinStats()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_tri_join_stats(ctx.handle, bvh_a.handle, bvh_b.handle, C.byref(st)), "rs3d_tri_join_stats")
    return {k: int(getattr(st, k)) for k, _ in st._fields_}


def sort_unique_u64(ctx, keys: torch.Tensor) -> torch.Tensor:
    _need(keys.dtype == torch.int64 and keys.dim() == 1 and keys.is_contiguous(),
          "sort_unique_u64: keys must be a contiguous 1-D int64 tensor (u64 bit patterns)")
    _on_device(keys, ctx.device, "keys")
    cnt = C.c_uint64()
    torch.cuda.synchronize()
    check(ctx.lib.rs3d_sort_unique_u64(ctx.handle, _ptr(keys), keys.shape[0], C.byref(cnt)),
          "rs3d_sort_unique_u64")
    return keys[:cnt.value]


def _join_out(fn, max_out):
    while True:
        out = torch.empty(max(max_out, 1), dtype=torch.int64, device="cuda")
        cnt = C.c_uint64()
        rc, extra = fn(out, max_out, cnt)
        if rc == _lib.RS3D_E_CAPACITY and cnt.value > max_out:
            max_out = int(cnt.value)
            continue
        check(rc, "join")
        return to_host(out[:cnt.value], np.uint64), extra


def intersection_join(ctx, bvh1, dedges1, bvh2, dedges2, overlap_max_iterations=100,
                      max_out=1 << 20):
    """Sorted unique keys of the intersection join (overlap + containment fallback)."""
    v1, v2 = bvh1.view(), bvh2.view()
    torch.cuda.synchronize()

    def call(out, cap, cnt):
        over = C.c_uint64()
        rc = ctx.lib.rs3d_intersection_join(ctx.handle, C.byref(v1), bvh1.handle, C.byref(dedges1.s),
                                            C.byref(v2), bvh2.handle, C.byref(dedges2.s),
                                            overlap_max_iterations, 0, _ptr(out), cap, C.byref(cnt),
                                            C.byref(over))
        return rc, over.value
    return _join_out(call, max_out)


def containment_join(ctx, bvha, dedgesa, bvhb, dedgesb, max_out=1 << 20):
    """(containment keys, overlap keys, overflow) of the containment join, (a << 32 | b)."""
    va, vb = bvha.view(), bvhb.view()
    torch.cuda.synchronize()
    while True:
        cont = torch.empty(max_out, dtype=torch.int64, device="cuda")
        ov = torch.empty(max_out, dtype=torch.int64, device="cuda")
        nc, no, over = C.c_uint64(), C.c_uint64(), C.c_uint64()
        rc = ctx.lib.rs3d_containment_join(ctx.handle, C.byref(va), bvha.handle, C.byref(dedgesa.s),
                                           C.byref(vb), bvhb.handle, C.byref(dedgesb.s), 0,
                                           _ptr(cont), max_out, C.byref(nc), _ptr(ov), max_out,
                                           C.byref(no), C.byref(over))
        if rc == _lib.RS3D_E_CAPACITY:
            max_out = int(max(nc.value, no.value))
            continue
        check(rc, "rs3d_containment_join")
        return (to_host(cont[:nc.value], np.uint64), to_host(ov[:no.value], np.uint64), over.value)


def _batch_array(queries, device: int):
    n = len(queries)
    arr = (_lib.BatchQuery * max(n, 1))()
    for i, q in enumerate(queries):
        b = arr[i]
        if q["kind"] == "contain":
            check_points(q["points"], device, f"queries[{i}]['points']")
            check_out(q["out"], _lib.OUT_OBJ_I32, q["points"].shape[0], device, f"queries[{i}]['out']")
            b.kind = 0
            b.axis = q.get("axis", 2)
            b.bvh1 = q["bvh"].handle.value
            b.points = q["points"].data_ptr()
            b.num_points = q["points"].shape[0]
            b.out_obj = q["out"].data_ptr()
        else:
            b.kind = 1
            b.bvh1 = q["bvh1"].handle.value
            b.bvh2 = q["bvh2"].handle.value
            b.edges1 = C.cast(C.pointer(q["edges1"].s), C.c_void_p).value
            b.edges2 = C.cast(C.pointer(q["edges2"].s), C.c_void_p).value
    return arr


def batch_execute(ctx, queries, max_pairs=1 << 22):
    """Runs many queries in one persistent launch. Each query is a dict:
    {"kind": "contain", "bvh": Bvh, "points": tensor (n,3), "out": int32 tensor (n,), "axis": 2}
    {"kind": "overlap", "bvh1": Bvh, "edges1": DevEdges, "bvh2": Bvh, "edges2": DevEdges}
    Returns a list with the sorted pair keys (uint64 numpy) of each query (empty for
    containment queries, whose ids land in their "out" tensors)."""
    n = len(queries)
    arr = _batch_array(queries, ctx.device)
    offs = (C.c_uint64 * (n + 1))()
    torch.cuda.synchronize()
    while True:
        out = torch.empty(max(max_pairs, 1), dtype=torch.int64, device="cuda")
        total = C.c_uint64()
        rc = ctx.lib.rs3d_batch_execute(ctx.handle, arr, n, _ptr(out), max_pairs, offs, C.byref(total))
        if rc == _lib.RS3D_E_CAPACITY and total.value > max_pairs:
            max_pairs = int(total.value)
            continue
        check(rc, "rs3d_batch_execute")
        keys = to_host(out[:total.value], np.uint64)
        return [keys[offs[i]:offs[i + 1]].copy() for i in range(n)]


class CapturedBatch:
    """A batch planned once and captured as a hipGraph (rs3d_batch_*): run() replays it on
    the current contents of the queries' inputs and returns what batch_execute returns."""

    def __init__(self, ctx, queries):
        self.ctx, self.n = ctx, len(queries)
        self._arr = _batch_array(queries, ctx.device)
        self._keep = queries
        self.handle = C.c_void_p()
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_batch_create(ctx.handle, self._arr, self.n, C.byref(self.handle)),
              "rs3d_batch_create")

    def run(self, max_pairs=1 << 22):
        offs = (C.c_uint64 * (self.n + 1))()
        torch.cuda.synchronize()
        while True:
            out = torch.empty(max(max_pairs, 1), dtype=torch.int64, device="cuda")
            total = C.c_uint64()
            rc = self.ctx.lib.rs3d_batch_run(self.handle, _ptr(out), max_pairs, offs, C.byref(total))
            if rc == _lib.RS3D_E_CAPACITY and total.value > max_pairs:
                max_pairs = int(total.value)
                continue
            check(rc, "rs3d_batch_run")
            keys = to_host(out[:total.value], np.uint64)
            return [keys[offs[i]:offs[i + 1]].copy() for i in range(self.n)]

    def close(self):
        if self.handle:
            check(self.ctx.lib.rs3d_batch_destroy(self.handle), "rs3d_batch_destroy")
            self.handle = C.c_void_p()


class Server:
    """Request server with continuous batching over resident datasets (rs3d_server_*)."""

    def __init__(self, ctx, max_slots: int = 64, memory_budget: int = 8 << 30):
        self.ctx = ctx
        self.handle = C.c_void_p()
        check(ctx.lib.rs3d_server_create(ctx.handle, max_slots, memory_budget, C.byref(self.handle)),
              "rs3d_server_create")
        self._keep = []

    def add_dataset(self, bvh: Bvh, dedges: "DevEdges | None" = None, grid_res: int = 0) -> int:
        view = bvh.view()
        i = C.c_int()
        torch.cuda.synchronize()
        check(self.ctx.lib.rs3d_server_add_dataset(self.handle, C.byref(view), bvh.handle,
                                                   C.byref(dedges.s) if dedges is not None else None,
                                                   grid_res, C.byref(i)), "rs3d_server_add_dataset")
        self._keep.append((bvh, dedges))
        return i.value

    def submit(self, kind: int, dataset: int, query_dataset: int = -1, points=None, out=None,
               axis: int = 2) -> int:
        if kind == _lib.REQ_CONTAIN:
            check_points(points, self.ctx.device)
            check_out(out, _lib.OUT_OBJ_I32, points.shape[0], self.ctx.device)
        r = _lib.Request(kind, dataset, query_dataset, axis,
                         points.data_ptr() if points is not None else None,
                         points.shape[0] if points is not None else 0,
                         out.data_ptr() if out is not None else None)
        t = C.c_int64()
        check(self.ctx.lib.rs3d_server_submit(self.handle, C.byref(r), C.byref(t)), "rs3d_server_submit")
        return t.value

    def step(self):
        a, c = C.c_int(), C.c_int()
        check(self.ctx.lib.rs3d_server_step(self.handle, C.byref(a), C.byref(c)), "rs3d_server_step")
        return a.value, c.value

    def result(self, ticket: int):
        done = C.c_int()
        n = C.c_uint64()
        check(self.ctx.lib.rs3d_server_result(self.handle, ticket, C.byref(done), None, 0, C.byref(n)),
              "rs3d_server_result")
        keys = np.empty(max(n.value, 1), np.uint64)
        if done.value == 1 and n.value:
            check(self.ctx.lib.rs3d_server_result(self.handle, ticket, C.byref(done),
                                                  keys.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
                  "rs3d_server_result")
        return done.value, keys[:n.value]

    def estimate(self, ticket: int) -> int:
        """The pair estimate that sized an overlap/intersection request (-1 before a step)."""
        e = C.c_int64()
        check(self.ctx.lib.rs3d_server_request_estimate(self.handle, ticket, C.byref(e)),
              "rs3d_server_request_estimate")
        return e.value

    def stats(self):
        q, m, s = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(self.ctx.lib.rs3d_server_stats(self.handle, C.byref(q), C.byref(m), C.byref(s)),
              "rs3d_server_stats")
        return {"queued": q.value, "peak_memory": m.value, "steps": s.value}

    def close(self):
        if self.handle:
            self.ctx.lib.rs3d_server_destroy(self.handle)
            self.handle = C.c_void_p()


def object_rays(ctx, src: "Bvh | None", target: "Bvh", swap: bool = False, max_targets: int = 256,
                chain_mode: int = _lib.OBJ_ALL_HITS, max_iter: int = 0, origins: torch.Tensor | None = None,
                track: bool = False, max_pair: int = 256, with_stats: bool = False, two_pass: bool = False,
                form: int = _lib.OBJ_FALLBACK, overlap_table=None):
    """rs3d_object_rays: per-object +z containment rays with the reference's launch options.
    Returns a dict: keys (sorted unique (a << 32 | b) as uint64), overflow, and when asked the
    per-source tracking arrays and the counters (rs3d_intersection_stats as a dict)."""
    n_src = origins.shape[0] if origins is not None else src.view().num_objects
    if origins is not None:
        check_points(origins, ctx.device, "origins")
    o = _lib.ObjectRayOpts()
    o.chain_mode, o.max_iterations = chain_mode, max_iter
    keep = []

    def dev(shape, dtype, fill=0):
        t = torch.full(shape, fill, dtype=dtype, device="cuda")
        keep.append(t)
        return t
    if origins is not None:
        o.origins, o.num_sources = origins.data_ptr(), n_src
    tr = None
    if track:
        tr = {"pair_target_ids": dev((max(n_src, 1), max_pair), torch.int32, -7),
              "pair_target_hits": dev((max(n_src, 1), max_pair), torch.int32, 7),
              "iterations": dev((max(n_src, 1),), torch.int32, 7), "candidate_count": dev((max(n_src, 1),), torch.int32, 7),
              "candidate_overflow": dev((max(n_src, 1),), torch.int32, 7)}
        o.pair_target_ids, o.pair_target_hits = tr["pair_target_ids"].data_ptr(), tr["pair_target_hits"].data_ptr()
        o.max_pair_targets = max_pair
        o.iterations, o.candidate_count = tr["iterations"].data_ptr(), tr["candidate_count"].data_ptr()
        o.candidate_overflow = tr["candidate_overflow"].data_ptr()
    st = None
    if with_stats:
        st = dev((13,), torch.int64, 0)
        o.stats = st.data_ptr()
    view = src.view() if src is not None else None
    over = C.c_uint64()
    ht = None
    if two_pass:
        counts = dev((max(n_src, 1),), torch.int32, -1)
        o.counts = counts.data_ptr()
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_object_rays(ctx.handle, C.byref(view) if view is not None else None, target.handle, form,
                                       int(swap), max_targets, C.byref(o), None, None, C.byref(over)), "rs3d_object_rays")
        c = to_host(counts)[:n_src].astype(np.int64)
        offs = np.concatenate([[0], np.cumsum(c)[:-1]]) if n_src else np.zeros(0, np.int64)
        doffs = to_dev(offs.astype(np.int64)) if n_src else dev((1,), torch.int64)
        res = dev((max(int(c.sum()), 1), 2), torch.int32)
        o.counts, o.offsets, o.results = None, doffs.data_ptr(), res.data_ptr()
        if st is not None:
            st.zero_()
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_object_rays(ctx.handle, C.byref(view) if view is not None else None, target.handle, form,
                                       int(swap), max_targets, C.byref(o), None, None, C.byref(over)), "rs3d_object_rays")
        pr = to_host(res)[:int(c.sum())]
        keys = (pr[:, 0].astype(np.uint32).astype(np.uint64) << np.uint64(32)) | pr[:, 1].astype(np.uint32).astype(np.uint64)
        keys = np.unique(keys)
    else:
        slots = 1 << 12
        while slots < 4 * max(n_src, 1) * 8:
            slots <<= 1
        tab = dev((slots + 1,), torch.int64, -1)
        tab[slots] = 0
        ht = _lib.HashTable(tab.data_ptr(), slots, 1, 0, tab.data_ptr() + 8 * slots)
        torch.cuda.synchronize()
        check(ctx.lib.rs3d_object_rays(ctx.handle, C.byref(view) if view is not None else None, target.handle, form,
                                       int(swap), max_targets, C.byref(o), C.byref(ht),
                                       C.byref(overlap_table) if overlap_table is not None else None, C.byref(over)),
              "rs3d_object_rays")
        t = to_host(tab)
        assert t[slots] == 0, "pair table overflow"
        keys = np.unique(t[:slots][t[:slots] != -1].view(np.uint64))
    out = {"keys": keys, "overflow": over.value}
    if tr is not None:
        out["tracking"] = {k: to_host(v)[:n_src].view(np.uint32 if k != "pair_target_ids" else np.int32)
                           for k, v in tr.items()}
    if st is not None:
        raw = to_host(st).view(np.uint8).tobytes()
        sv = _lib.IntersectionStats.from_buffer_copy(raw[:C.sizeof(_lib.IntersectionStats)])
        out["stats"] = {k: int(getattr(sv, k)) for k, _ in sv._fields_}
    return out
