"""rs3d_filter_refine_noalloc (the entry point whose scratch is a fixed device array inside
the library) against rs3d_filter_refine and the oracle: two calls in a row on each path,
the second with a smaller input; no device memory is allocated by its calls."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as orc
from rayspace3d_amd import _lib, gen
from rayspace3d_amd.device import Bvh, Context, to_dev, to_host

pytestmark = pytest.mark.gpu


def _call(ctx, fn, bvh, pts, axis):
    n = pts.shape[0]
    out = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    cand, inside = C.c_uint64(), C.c_uint64()
    torch.cuda.synchronize()
    rc = fn(ctx.handle, bvh.handle, C.c_void_p(pts.data_ptr()), n, axis, C.c_void_p(out.data_ptr()),
            C.byref(cand), C.byref(inside))
    _lib.check(rc, fn.__name__)
    return to_host(out), cand.value, inside.value


def _workload(seed, n_points):
    # many spheres in a cube, points over a slightly larger cube so the filter rejects some
    mesh, _ = gen.many_spheres_workload(seed=seed, n_objects=1500, n_points=1, world=150.0)
    rng = np.random.Generator(np.random.PCG64(seed + 100))
    pts = gen.uniform_points(rng, n_points, -10.0, 160.0)
    return mesh, pts


@pytest.mark.parametrize("axis", [2, 0])
def test_noalloc_bvh_path_equals_filter_refine(axis):
    print(f"noalloc BVH path, axis {axis}", flush=True)
    with Context(0) as ctx:  # a fresh context: its own scratch buffers are still empty
        mesh, pts = _workload(31 + axis, 400_000)
        bvh = Bvh.from_mesh(ctx, mesh)
        ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=axis, threads=8)
        # one warm-up call: the runtime loads a kernel's code object on its first launch
        _call(ctx, ctx.lib.rs3d_filter_refine_noalloc, bvh, to_dev(pts[:1000]), axis)
        for n in (400_000, 123_457):  # second call smaller: stale scratch must not leak through
            dp = to_dev(pts[:n])
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            got, cand, inside = _call(ctx, ctx.lib.rs3d_filter_refine_noalloc, bvh, dp, axis)
            torch.cuda.synchronize()
            assert torch.cuda.mem_get_info()[0] == free0, "the no-alloc call allocated device memory"
            assert np.array_equal(got, ref[:n]), f"{np.count_nonzero(got != ref[:n])} differ from the oracle"
            ctx.set_option(_lib.OPT_AXIS_GRID, 0)
            exp, cand_e, inside_e = _call(ctx, ctx.lib.rs3d_filter_refine, bvh, dp, axis)
            ctx.set_option(_lib.OPT_AXIS_GRID, 1)
            assert np.array_equal(got, exp)
            assert (cand, inside) == (cand_e, inside_e)
            assert inside == int((ref[:n] >= 0).sum()) > 0 and 0 < cand < n
        bvh.free()


def test_noalloc_grid_path_equals_filter_refine():
    print("noalloc grid path, axis 2", flush=True)
    with Context(0) as ctx:
        mesh, pts = _workload(41, 600_000)
        bvh = Bvh.from_mesh(ctx, mesh)
        info = _lib.AxisGridInfo()
        _lib.check(ctx.lib.rs3d_axis_grid_prepare(ctx.handle, bvh.handle, 2, C.byref(info)), "prepare")
        assert info.state == 1
        ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2, threads=8)
        _call(ctx, ctx.lib.rs3d_filter_refine_noalloc, bvh, to_dev(pts[:1000]), 2)  # warm-up
        for n in (600_000, 77_777):
            dp = to_dev(pts[:n])
            exp, cand_e, inside_e = _call(ctx, ctx.lib.rs3d_filter_refine, bvh, dp, 2)
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            got, cand, inside = _call(ctx, ctx.lib.rs3d_filter_refine_noalloc, bvh, dp, 2)
            torch.cuda.synchronize()
            assert torch.cuda.mem_get_info()[0] == free0, "the no-alloc call allocated device memory"
            assert np.array_equal(got, exp) and np.array_equal(got, ref[:n])
            assert (cand, inside) == (cand_e, inside_e)
        bvh.free()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    print("noalloc size limit", flush=True)
    with Context(0) as ctx:
        mesh, _ = _workload(51, 1)
        bvh = Bvh.from_mesh(ctx, mesh)
        n = _lib.NOALLOC_MAX_POINTS + 1
        dp = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
        out = torch.empty(n, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        rc = ctx.lib.rs3d_filter_refine_noalloc(ctx.handle, bvh.handle, C.c_void_p(dp.data_ptr()), n, 2,
                                                C.c_void_p(out.data_ptr()), None, None)
        assert rc == _lib.RS3D_E_CAPACITY
        assert b"scratch" in ctx.lib.rs3d_last_error()
        bvh.free()
