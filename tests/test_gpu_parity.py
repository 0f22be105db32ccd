"""GPU parity: the HIP library (through the C ABI) against the CPU oracle, bit-exact for
all integer outputs (flags, object ids, pair sets, counts)."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as orc
from rayspace3d_amd import _lib, gen
from rayspace3d_amd.device import axis_grid_prepare, axis_grid_stats
from rayspace3d_amd.device import (Bvh, DevEdges, FilterRefinePipeline, batch_execute, containment_join,
                                   edge_count, intersection_join, overlap_join, sort_unique_u64,
                                   to_dev, to_host, trace_points, trace_stats)

pytestmark = pytest.mark.gpu


def _contain(ctx, mesh, pts, axis, leaf=1, kind=_lib.OUT_OBJ_I32):
    bvh = Bvh.from_mesh(ctx, mesh, leaf_size=leaf)
    out = to_host(trace_points(ctx, bvh, to_dev(pts), axis=axis, out_kind=kind))
    info = bvh.info()
    bvh.free()
    return out, info


@pytest.mark.parametrize("axis", [2, 0, 1])
def test_single_sphere_points(ctx, axis):
    mesh, pts = gen.single_sphere_workload(seed=11, n_points=200_000)
    print(f"single sphere L5, 200k points, axis {axis}", flush=True)
    out, info = _contain(ctx, mesh, pts, axis)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=axis)
    assert info.num_triangles == 20480 and info.num_nodes == 20479
    assert np.array_equal(out, ref), f"{np.count_nonzero(out != ref)} mismatches"
    frac = (out >= 0).mean()
    assert 0.12 < frac < 0.19


def test_output_kinds_agree(ctx):
    mesh, pts = gen.single_sphere_workload(seed=12, n_points=50_000, level=4)
    obj, _ = _contain(ctx, mesh, pts, 2, kind=_lib.OUT_OBJ_I32)
    flag, _ = _contain(ctx, mesh, pts, 2, kind=_lib.OUT_FLAG_U8)
    rr, _ = _contain(ctx, mesh, pts, 2, kind=_lib.OUT_RAY_RESULT)
    assert np.array_equal(flag.astype(bool), obj >= 0)
    assert np.array_equal(rr[:, 0], np.arange(len(pts)))
    assert np.array_equal(rr[:, 1], obj)


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_many_spheres_points_leaf_invariant(ctx, leaf):
    mesh, pts = gen.many_spheres_workload(seed=21, n_objects=2000, n_points=300_000, world=200.0)
    print(f"2000 spheres in 200^3, 300k points, leaf {leaf}", flush=True)
    out, _ = _contain(ctx, mesh, pts, 2, leaf=leaf)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    assert np.array_equal(out, ref), f"{np.count_nonzero(out != ref)} mismatches"
    assert (out >= 0).sum() > 100


def test_hit_buffer_restart(ctx):
    # 40 concentric spheres: +z rays cross up to 80 surfaces (> the 8-entry hit buffer)
    n = 40
    mesh = gen.spheres(2, np.zeros((n, 3)), np.linspace(1.0, 3.0, n))
    rng = np.random.Generator(np.random.PCG64(5))
    pts = gen.uniform_points(rng, 20000, -3.2, 3.2)
    out, _ = _contain(ctx, mesh, pts, 2)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    assert np.array_equal(out, ref)
    st = trace_stats(ctx, Bvh.from_mesh(ctx, mesh), to_dev(pts), 2)
    assert st["restarts"] > 0


def test_degenerate_sizes(ctx):
    # one triangle, zero points, empty mesh
    V = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], np.float32)
    m = gen.Mesh(V, np.array([[0, 1, 2]], np.uint32), np.array([7], np.int32), 1)
    pts = np.array([[0.2, 0.2, 0.0], [0.9, 0.9, 0.0], [0.2, 0.2, 2.0]], np.float32)
    out, info = _contain(ctx, m, pts, 2)
    assert list(out) == [7, -1, -1] and info.num_triangles == 1
    out0, _ = _contain(ctx, m, np.zeros((0, 3), np.float32), 2)
    assert out0.shape == (0,)
    empty = gen.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), np.zeros(0, np.int32))
    oute, _ = _contain(ctx, empty, pts, 2)
    assert list(oute) == [-1, -1, -1]


@pytest.mark.parametrize("grid", [1, 0])
@pytest.mark.parametrize("axis", [2, 0])
def test_filter_refine(ctx, grid, axis):
    # grid 1: the one-pass filter + axis-grid kernel; grid 0: filter, compaction, BVH walk
    mesh, pts = gen.many_spheres_workload(seed=31, n_objects=1500, n_points=200_000, world=150.0)
    # points beyond the dataset box on every side, and on its faces
    rng = np.random.Generator(np.random.PCG64(3))
    vmin, vmax = mesh.vertices.min(0), mesh.vertices.max(0)
    face = gen.uniform_points(rng, 6000, -10.0, 160.0)
    for k in range(3):
        face[2000 * k: 2000 * k + 1000, k] = vmin[k]
        face[2000 * k + 1000: 2000 * k + 2000, k] = vmax[k]
    pts = np.concatenate([pts, gen.uniform_points(rng, 20000, -30.0, 180.0), face]).astype(np.float32)
    print(f"filter-refine grid {grid} axis {axis}: {len(pts)} points", flush=True)
    saved = ctx.get_option(_lib.OPT_AXIS_GRID)
    ctx.set_option(_lib.OPT_AXIS_GRID, grid)
    try:
        bvh = Bvh.from_mesh(ctx, mesh)
        res = FilterRefinePipeline(ctx).execute(bvh, to_dev(pts), axis=axis)
        out = to_host(res.final_results)
        bvh.free()
    finally:
        ctx.set_option(_lib.OPT_AXIS_GRID, saved)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=axis)
    assert np.array_equal(out, ref), np.count_nonzero(out != ref)
    assert res.inside_count == int((ref >= 0).sum())
    import boxcases as bc
    V, T = bc.box_mesh(tuple(vmin) + tuple(vmax))
    passes = orc.points_contain(V, T, np.zeros(12, np.int32), pts, axis=axis, use_grid=False) >= 0
    assert res.candidate_count == int(passes.sum())
    # off the face planes the filter is the open box
    on_face = np.any((pts == vmin) | (pts == vmax), axis=1)
    in_box = np.all((pts > vmin) & (pts < vmax), axis=1)
    assert np.array_equal(passes[~on_face], in_box[~on_face])


def test_filter_refine_counts_repeat(ctx):
    # the grid launch leaves its counters zeroed for the next call: calls in a row on
    # different point sets (and sizes) each report their own candidate and inside counts
    mesh, pts = gen.many_spheres_workload(seed=33, n_objects=400, n_points=300_000, world=80.0)
    bvh = Bvh.from_mesh(ctx, mesh)
    pipe = FilterRefinePipeline(ctx)
    vmin, vmax = mesh.vertices.min(0), mesh.vertices.max(0)
    for k, m in enumerate((300_000, 1_000, 120_000, 300_000)):
        p = np.ascontiguousarray(pts[:m] if k != 2 else pts[-m:])
        print(f"filter-refine repeat {k}: {m} points", flush=True)
        res = pipe.execute(bvh, to_dev(p), axis=2)
        ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, p, axis=2)
        assert np.array_equal(to_host(res.final_results), ref)
        assert res.inside_count == int((ref >= 0).sum())
        assert res.candidate_count == int(np.all((p > vmin) & (p < vmax), axis=1).sum())
    bvh.free()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_filter_box_margins(ctx, axis):
    # points at and around the thresholds of the filter's fast classification (face
    # distance m = 2^-16 mag + 1e-5, the faces' diagonal), checked against the oracle's
    # parity over the whole box mesh
    import boxcases as bc
    rng = np.random.Generator(np.random.PCG64(17 + axis))
    for box in bc.MARGIN_BOXES:
        pts = bc.margin_points(box, rng)
        print(f"filter box margins {box}, axis {axis}, {len(pts)} points", flush=True)
        V, T = bc.box_mesh(box)
        ref = orc.points_contain(V, T, np.zeros(12, np.int32), pts, axis=axis, use_grid=False) >= 0
        dp = to_dev(pts)
        flags = torch.zeros(len(pts), dtype=torch.uint8, device="cuda")
        hb = (C.c_float * 6)(*[float(np.float32(v)) for v in box])
        torch.cuda.synchronize()
        _lib.check(ctx.lib.rs3d_filter_box(ctx.handle, hb, C.c_void_p(dp.data_ptr()), len(pts), axis,
                                           C.c_void_p(flags.data_ptr())), "filter_box")
        ctx.sync()
        got = to_host(flags).astype(bool)
        assert np.array_equal(got, ref), (box, np.count_nonzero(got != ref))
        assert 0 < int(ref.sum()) < len(pts)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_filter_box_flags(ctx, axis):
    # rs3d_filter_box tests the 4 triangles perpendicular to the axis; the oracle parity runs
    # over the whole 12-triangle box mesh, on points on and one ulp off every face plane
    import boxcases as bc
    for box in bc.BOXES:
        pts = bc.box_points(box)
        print(f"filter box {box}, axis {axis}, {len(pts)} points", flush=True)
        V, T = bc.box_mesh(box)
        ref = orc.points_contain(V, T, np.zeros(12, np.int32), pts, axis=axis, use_grid=False) >= 0
        dp = to_dev(pts)
        flags = torch.zeros(len(pts), dtype=torch.uint8, device="cuda")
        hb = (C.c_float * 6)(*[float(np.float32(v)) for v in box])
        torch.cuda.synchronize()
        _lib.check(ctx.lib.rs3d_filter_box(ctx.handle, hb, C.c_void_p(dp.data_ptr()), len(pts), axis,
                                           C.c_void_p(flags.data_ptr())), "filter_box")
        ctx.sync()
        got = to_host(flags).astype(bool)
        assert np.array_equal(got, ref), (box, np.count_nonzero(got != ref))


def test_scan_compact_sort(ctx):
    L = ctx.lib
    rng = np.random.Generator(np.random.PCG64(9))
    for n in [0, 1, 5, 2047, 2048, 2049, 3_000_001]:
        x = rng.integers(0, 5, size=n).astype(np.int32)
        dx = to_dev(x) if n else torch.zeros(1, dtype=torch.int32, device="cuda")
        out = torch.zeros(max(n, 1), dtype=torch.int64, device="cuda")
        tot = C.c_int64()
        torch.cuda.synchronize()
        _lib.check(L.rs3d_scan_i32_i64(ctx.handle, C.c_void_p(dx.data_ptr()),
                                       C.c_void_p(out.data_ptr()), n, C.byref(tot)), "scan")
        exp = np.concatenate([[0], np.cumsum(x, dtype=np.int64)])
        assert tot.value == exp[-1]
        assert np.array_equal(to_host(out)[:n], exp[:n])
        f = (x > 2).astype(np.uint8)
        df = to_dev(f) if n else torch.zeros(1, dtype=torch.uint8, device="cuda")
        idx = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        cnt = C.c_uint64()
        torch.cuda.synchronize()
        _lib.check(L.rs3d_compact_flags(ctx.handle, C.c_void_p(df.data_ptr()), n,
                                        C.c_void_p(idx.data_ptr()), C.byref(cnt)), "compact")
        assert np.array_equal(to_host(idx)[:cnt.value], np.nonzero(f)[0].astype(np.int32))
    for n in [1, 100, 70_000, 2_500_000]:
        keys = rng.integers(0, 1 << 40, size=n, dtype=np.uint64)
        keys[: n // 3] = keys[n // 3: 2 * (n // 3)]  # duplicates
        got = to_host(sort_unique_u64(ctx, to_dev(keys)), np.uint64)
        assert np.array_equal(got, np.unique(keys))
        k32 = rng.integers(0, 1 << 30, size=n, dtype=np.uint32)
        v32 = np.arange(n, dtype=np.uint32)
        dk, dv = to_dev(k32), to_dev(v32)
        torch.cuda.synchronize()
        _lib.check(L.rs3d_sort_pairs_u32(ctx.handle, C.c_void_p(dk.data_ptr()),
                                         C.c_void_p(dv.data_ptr()), n, 30), "sort_pairs")
        ctx.sync()
        perm = np.argsort(k32, kind="stable")
        assert np.array_equal(to_host(dk, np.uint32), k32[perm])
        assert np.array_equal(to_host(dv, np.uint32), v32[perm])


def _overlap_case(seed, n, world):
    a = gen.overlap_dataset(seed, n_objects=n, world=world, levels=(2, 3))
    b = gen.overlap_dataset(seed + 1, n_objects=n, world=world, levels=(2, 3))
    a.edges = gen.extract_edges(a)
    b.edges = gen.extract_edges(b)
    return a, b


@pytest.mark.parametrize("leaf", [1, 4])
def test_edge_counts_and_overlap_join(ctx, leaf):
    a, b = _overlap_case(41, 300, 120.0)
    print(f"overlap 300 x 300 objects, {a.edges.num_edges} + {b.edges.num_edges} edges, "
          f"leaf {leaf}", flush=True)
    ba, bb = Bvh.from_mesh(ctx, a, leaf_size=leaf), Bvh.from_mesh(ctx, b, leaf_size=leaf)
    ea, eb = DevEdges(a.edges), DevEdges(b.edges)
    keys, counts = orc.edge_pairs(a.edges, b.vertices, b.indices, b.tri_to_obj, with_counts=True)
    got_counts = to_host(edge_count(ctx, bb, ea))
    assert np.array_equal(got_counts, counts)
    ref = orc.overlap_pairs(a, a.edges, b, b.edges)
    assert len(ref) > 50
    for mode in (0, 1):
        got = overlap_join(ctx, ba, ea, bb, eb, mode=mode)
        assert np.array_equal(got, ref), f"mode {mode}: {len(got)} vs {len(ref)} pairs"
    # the same edges in spatial order: a permutation of the records, the same join
    sa, sb = ea.spatial(ctx), eb.spatial(ctx)
    rec = lambda d: np.concatenate([to_host(d.t[0]).view(np.uint32), to_host(d.t[1]).view(np.uint32),
                                    to_host(d.t[2])[:, None].view(np.uint32), to_host(d.t[3])[:, None].view(np.uint32)], 1)
    for d, s_ in ((ea, sa), (eb, sb)):
        r0, r1 = rec(d), rec(s_)
        assert not np.array_equal(r0, r1)
        assert np.array_equal(r0[np.lexsort(r0.T[::-1])], r1[np.lexsort(r1.T[::-1])])
    s_counts = to_host(edge_count(ctx, bb, sa))
    assert np.array_equal(np.sort(s_counts), np.sort(counts))
    for mode in (0, 1):
        assert np.array_equal(overlap_join(ctx, ba, sa, bb, sb, mode=mode), ref)
    ba.free()
    bb.free()


def test_reference_named_dedup(ctx):
    L = ctx.lib
    rng = np.random.Generator(np.random.PCG64(4))
    p1 = rng.integers(0, 50, size=(1000, 2)).astype(np.int32)
    p2 = rng.integers(0, 50, size=(700, 2)).astype(np.int32)
    d1, d2 = to_dev(p1), to_dev(p2)
    out = torch.zeros((1700, 2), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    n = L.merge_and_deduplicate_pairs_gpu(C.c_void_p(d1.data_ptr()), 1000, C.c_void_p(d2.data_ptr()),
                                          700, C.c_void_p(out.data_ptr()))
    got = to_host(out)[:n]
    allp = np.concatenate([p1, p2])
    # reference order: u64 of the little-endian struct = (b << 32 | a)
    key = (allp[:, 1].astype(np.uint64) << np.uint64(32)) | allp[:, 0].astype(np.uint64)
    uk = np.unique(key)
    exp = np.stack([(uk & np.uint64(0xffffffff)).astype(np.int32), (uk >> np.uint64(32)).astype(np.int32)], 1)
    assert np.array_equal(got, exp)
    x = rng.integers(0, 9, size=5000).astype(np.int32)
    dx = to_dev(x)
    dout = torch.zeros(5000, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    tot = L.exclusive_scan_gpu(C.c_void_p(dx.data_ptr()), C.c_void_p(dout.data_ptr()), 5000)
    assert tot == int(x.sum())
    assert np.array_equal(to_host(dout)[1:], np.cumsum(x)[:-1])


def test_hash_table_compact(ctx):
    L = ctx.lib
    size = 1 << 12
    slots = torch.full((size,), -1, dtype=torch.int64, device="cuda")
    keys = np.unique(np.random.Generator(np.random.PCG64(8)).integers(0, 1 << 40, 900, dtype=np.uint64))
    table = np.full(size, np.uint64(0xFFFFFFFFFFFFFFFF))
    for k in keys:
        assert orc.ht_insert(table, True, 5000, int(k))
    slots.copy_(to_dev(table))
    out = torch.zeros((size, 2), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    n = L.compact_hash_table_pairs(C.c_void_p(slots.data_ptr()), size, C.c_void_p(out.data_ptr()), size)
    got = to_host(out)[:n]
    gk = (got[:, 0].astype(np.uint32).astype(np.uint64) << np.uint64(32)) | got[:, 1].astype(np.uint32).astype(np.uint64)
    # slot order: exactly the occupied slots of the oracle table, in slot order
    assert np.array_equal(gk, table[table != np.uint64(0xFFFFFFFFFFFFFFFF)])
    m = L.compact_hash_table_pairs(C.c_void_p(slots.data_ptr()), size, C.c_void_p(out.data_ptr()), 10)
    assert m == 10


def test_estimate_matches_oracle(ctx):
    a = gen.overlap_dataset(3, n_objects=400, world=100.0, levels=(2,))
    b = gen.overlap_dataset(4, n_objects=400, world=100.0, levels=(2,))
    ca, mn, mx = orc.grid_stats(a.vertices, a.indices, a.tri_to_obj, 16, world_size=100.0)
    cb, _, _ = orc.grid_stats(b.vertices, b.indices, b.tri_to_obj, 16, world_size=100.0)
    cv = float(np.prod((mx - mn) / 16))
    ref = orc.estimate_raw(ca, cb, cv)
    dA, dB = to_dev(ca.view(np.uint8)), to_dev(cb.view(np.uint8))
    r = C.c_double()
    torch.cuda.synchronize()
    _lib.check(ctx.lib.rs3d_estimate_raw(ctx.handle, C.c_void_p(dA.data_ptr()), C.c_void_p(dB.data_ptr()),
                                         len(ca), cv, 0.001, 0.8, C.byref(r)), "estimate")
    assert ref > 0 and abs(r.value - ref) <= 1e-5 * ref


def _nested_case(seed, n=200, world=200.0):
    """A: radius-5 spheres; B: half nested inside A objects (radius 2 at A centres), half
    random radius-3 spheres, so the joins see overlap, containment and disjoint pairs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ca = rng.uniform(10, world - 10, size=(n, 3))
    a = gen.spheres(rng.choice([2, 3], size=n), ca, np.full(n, 5.0),
                    rotations=gen.random_rotations(rng, n))
    cb = np.concatenate([ca[: n // 2] + rng.uniform(-1, 1, size=(n // 2, 3)),
                         rng.uniform(10, world - 10, size=(n - n // 2, 3))])
    rb = np.concatenate([np.full(n // 2, 2.0), np.full(n - n // 2, 3.0)])
    b = gen.spheres(2, cb, rb, rotations=gen.random_rotations(rng, n))
    a.edges, b.edges = gen.extract_edges(a), gen.extract_edges(b)
    return a, b


def test_intersection_join(ctx):
    a, b = _nested_case(51)
    ref = orc.intersection_pairs(a, b)
    ov = orc.overlap_pairs(a, a.edges, b, b.edges)
    assert len(ref) > len(ov) >= 5  # the fallback adds the nested pairs
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    got, over = intersection_join(ctx, ba, DevEdges(a.edges), bb, DevEdges(b.edges))
    got_s, over_s = intersection_join(ctx, ba, DevEdges(a.edges).spatial(ctx), bb, DevEdges(b.edges).spatial(ctx))
    assert np.array_equal(got_s, got) and over_s == over
    assert over == 0
    assert np.array_equal(got, ref), f"{len(got)} vs {len(ref)}"
    ba.free()
    bb.free()


def test_containment_join(ctx):
    a, b = _nested_case(61)
    ov_ref, ct_ref = orc.containment_pairs(a, b)
    assert len(ct_ref) >= 50 and len(ov_ref) > 5
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    ct, ov, over = containment_join(ctx, ba, DevEdges(a.edges), bb, DevEdges(b.edges))
    ct_s, ov_s, over_s = containment_join(ctx, ba, DevEdges(a.edges).spatial(ctx), bb, DevEdges(b.edges).spatial(ctx))
    assert np.array_equal(ct_s, ct) and np.array_equal(ov_s, ov) and over_s == over
    assert over == 0
    assert np.array_equal(ov, ov_ref)
    assert np.array_equal(ct, ct_ref)
    # every nested B object (first half: radius 2 within sqrt(3) of the centre of a
    # radius-5 A object, so wholly inside it) is contained in the A object it was placed in
    pairs = {(int(k >> np.uint64(32)), int(k & np.uint64(0xffffffff))) for k in ct}
    assert all((i, i) in pairs for i in range(100))
    ba.free()
    bb.free()


def test_grid_stats_and_estimate(ctx):
    from rayspace3d_amd import preio
    import os
    golden = os.path.join(os.path.dirname(__file__), "golden", "check_n_geometry.pre")
    for mesh, res, world in [(preio.read_pre(golden).mesh, 5, -1.0),
                             (gen.overlap_dataset(3, n_objects=3000, world=100.0, levels=(2,)), 16, 100.0)]:
        ref, mn, mx = orc.grid_stats(mesh.vertices, mesh.indices, mesh.tri_to_obj, res, world)
        cells = torch.zeros(res ** 3 * 16, dtype=torch.uint8, device="cuda")
        dV, dI, dO = to_dev(mesh.vertices), to_dev(mesh.indices), to_dev(mesh.tri_to_obj)
        b = (C.c_float * 6)()
        torch.cuda.synchronize()
        _lib.check(ctx.lib.rs3d_grid_stats(ctx.handle, C.c_void_p(dV.data_ptr()), len(mesh.vertices),
                                           C.c_void_p(dI.data_ptr()), len(mesh.indices),
                                           C.c_void_p(dO.data_ptr()), int(mesh.tri_to_obj.max()) + 1,
                                           res, world, C.c_void_p(cells.data_ptr()), b), "grid_stats")
        got = to_host(cells).view(orc.GRID_CELL)
        assert np.allclose(list(b)[:3], mn) and np.allclose(list(b)[3:], mx)
        assert np.array_equal(got["center"], ref["center"]) and np.array_equal(got["touch"], ref["touch"])
        assert np.allclose(got["avg"], ref["avg"], rtol=1e-5)
        assert np.allclose(got["ratio"], ref["ratio"], rtol=1e-4)
    # estimate with the alpha correction against the oracle (sizes tables only: tolerance)
    a = gen.overlap_dataset(3, n_objects=3000, world=100.0, levels=(2,))
    bm = gen.overlap_dataset(4, n_objects=3000, world=100.0, levels=(2,))
    ca, mn, mx = orc.grid_stats(a.vertices, a.indices, a.tri_to_obj, 16, 100.0)
    cb, _, _ = orc.grid_stats(bm.vertices, bm.indices, bm.tri_to_obj, 16, 100.0)
    ref = orc.estimate_pairs(ca, cb, mn, mx, [16, 16, 16])
    dA, dB = to_dev(ca.view(np.uint8)), to_dev(cb.view(np.uint8))
    est = C.c_int64()
    torch.cuda.synchronize()
    _lib.check(ctx.lib.rs3d_estimate_pairs(ctx.handle, C.c_void_p(dA.data_ptr()), C.c_void_p(dB.data_ptr()),
                                           len(ca), (C.c_float * 3)(*mn), (C.c_float * 3)(*mx),
                                           (C.c_uint32 * 3)(16, 16, 16), 0.001, 0.8, 1, C.byref(est)),
               "estimate_pairs")
    assert ref > 0 and abs(est.value - ref) <= max(1, 1e-4 * ref)


def test_ray_reordering_is_invisible(ctx):
    mesh, pts = gen.many_spheres_workload(seed=23, n_objects=1000, n_points=200_000, world=120.0)
    bvh = Bvh.from_mesh(ctx, mesh)
    dp = to_dev(pts)
    outs = []
    for thresh in (0, 1):  # 0: launch order; 1: every launch reordered
        ctx.set_option(_lib.OPT_RAY_SORT_MIN, thresh)
        outs.append(to_host(trace_points(ctx, bvh, dp, axis=2, out_kind=_lib.OUT_RAY_RESULT)))
    ctx.set_option(_lib.OPT_RAY_SORT_MIN, 1 << 16)
    bvh.free()
    assert np.array_equal(outs[0], outs[1])
    assert np.array_equal(outs[0][:, 0], np.arange(len(pts)))


@pytest.mark.parametrize("per_wave", [0, 64, 100, 512, 5000])
def test_lane_refill_matches_oracle(ctx, per_wave):
    # lane refill on and off, run lengths that do and do not divide the ray count; the
    # nested-sphere case forces hit-buffer restarts inside refilled lanes
    print(f"lane refill, rays per wave {per_wave}", flush=True)
    ctx.set_option(_lib.OPT_RAYS_PER_WAVE, per_wave)
    try:
        mesh, pts = gen.many_spheres_workload(seed=29, n_objects=800, n_points=150_001, world=100.0)
        out, _ = _contain(ctx, mesh, pts, 2, kind=_lib.OUT_RAY_RESULT)
        ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
        assert np.array_equal(out[:, 1], ref) and np.array_equal(out[:, 0], np.arange(len(pts)))
        n = 30
        nested = gen.spheres(2, np.zeros((n, 3)), np.linspace(1.0, 3.0, n))
        p2 = gen.uniform_points(np.random.Generator(np.random.PCG64(6)), 70_000, -3.2, 3.2)
        out2, _ = _contain(ctx, nested, p2, 0)
        ref2 = orc.points_contain(nested.vertices, nested.indices, nested.tri_to_obj, p2, axis=0)
        assert np.array_equal(out2, ref2)
    finally:
        ctx.set_option(_lib.OPT_RAYS_PER_WAVE, 64)


def test_autotune_restores_options(ctx):
    from rayspace3d_amd import autotune
    mesh, pts = gen.many_spheres_workload(seed=41, n_objects=200, n_points=40_000, world=60.0)
    dv, di, do = to_dev(mesh.vertices), to_dev(mesh.indices), to_dev(mesh.tri_to_obj)
    opts = (_lib.OPT_RAY_SORT_MIN, _lib.OPT_RAYS_PER_WAVE, _lib.OPT_AXIS_GRID, _lib.OPT_GRID_RAYS_PER_WAVE)
    before = tuple(ctx.get_option(o) for o in opts)
    best, trials = autotune.tune_traversal(ctx, dv, di, do, to_dev(pts), leaf_sizes=(2, 4),
                                           rays_per_wave=(0, 128), reps=1, grid_rays_per_wave=(256, 1024))
    walk = [t for t in trials if not t["axis_grid"]]
    assert len(walk) == 8 and len(trials) - len(walk) in (0, 2)
    assert best in trials and all(t["ms"] > 0 for t in trials)
    assert tuple(ctx.get_option(o) for o in opts) == before
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    for trial in (best, min(walk, key=lambda t: t["ms"])):
        autotune.apply(ctx, trial)
        try:
            out, _ = _contain(ctx, mesh, pts, 2, leaf=trial["leaf_size"])
        finally:
            for o, v in zip(opts, before):
                ctx.set_option(o, v)
        assert np.array_equal(out, ref), trial


def test_edge_stats_counters(ctx):
    from rayspace3d_amd.device import edge_stats
    a, b = _overlap_case(71, 60, 40.0)
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    ea = DevEdges(a.edges)
    st = edge_stats(ctx, bb, ea)
    counts = to_host(edge_count(ctx, bb, ea))
    assert st["rays"] == a.edges.num_edges
    assert st["pair_emits"] == int(counts.sum()) > 0
    assert st["source_refs"] == int(a.edges.src_counts.sum())
    assert 0 < st["rays_walked"] <= st["rays"] and st["occ_words"] > 0
    assert st["nodes_visited"] >= st["rays_walked"] and st["tris_tested"] >= st["hits"] > 0
    ba.free()
    bb.free()


def _crafted_segments(mesh, rng, n_tris=3000):
    """Segments against the triangles of `mesh` that sit on the boundaries of the hit rules:
    ending on a triangle (t = len, inside len + 1e-6), ending just short of it, starting on
    it (t = 0, below 1e-6), along an edge (coplanar), through a vertex, grazing the plane,
    plus short random segments near the surface and long ones across the whole field."""
    T = mesh.vertices[mesh.indices[rng.integers(0, len(mesh.indices), n_tris)]].astype(np.float32)
    v0, v1, v2 = T[:, 0], T[:, 1], T[:, 2]
    c = ((v0 + v1 + v2) / np.float32(3)).astype(np.float32)
    nrm = np.cross(v1 - v0, v2 - v0).astype(np.float32)
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30).astype(np.float32)
    off = np.float32(0.3)
    S, E = [], []
    S.append(c + off * nrm); E.append(c)                                  # ends on the triangle
    S.append(c + off * nrm); E.append(c + np.float32(1e-7) * nrm)         # ends 1e-7 short
    S.append(c); E.append(c + off * nrm)                                  # starts on it
    S.append(v0); E.append(v1)                                            # along an edge
    S.append(v0 + off * nrm); E.append(v0 - off * nrm)                    # through a vertex
    S.append(c + np.float32(1e-6) * nrm); E.append(c + np.float32(1e-6) * nrm + (v1 - v0))  # grazing
    S.append(c + rng.normal(0, 0.5, (n_tris, 3)).astype(np.float32))
    E.append(S[-1] + rng.normal(0, 0.5, (n_tris, 3)).astype(np.float32))  # short, near the surface
    lo, hi = mesh.vertices.min(0), mesh.vertices.max(0)
    a = rng.uniform(lo, hi, (200, 3)).astype(np.float32)
    S.append(a); E.append(rng.uniform(lo, hi, (200, 3)).astype(np.float32))  # long: no occupancy test
    S = np.ascontiguousarray(np.concatenate(S), np.float32)
    E = np.ascontiguousarray(np.concatenate(E), np.float32)
    n = len(S)
    return gen.EdgeSet(S, E, np.arange(n, dtype=np.int32), np.ones(n, np.int32), np.zeros(n, np.int32))


@pytest.mark.parametrize("pct", [200, 25, 1000])
def test_seg_occupancy_same_hits(ctx, pct):
    # edge rays with the occupancy test (fine, default and coarse cells) against the walk
    # alone and the oracle, on field edges and on segments at the hit rules' boundaries
    from rayspace3d_amd.device import edge_stats, seg_occupancy_prepare
    a, b = _overlap_case(73, 300, 120.0)
    rng = np.random.Generator(np.random.PCG64(pct))
    crafted = _crafted_segments(b, rng)
    saved = (ctx.get_option(_lib.OPT_SEG_OCCUPANCY), ctx.get_option(_lib.OPT_SEG_CELL_PCT))
    try:
        ctx.set_option(_lib.OPT_SEG_CELL_PCT, pct)
        bb = Bvh.from_mesh(ctx, b)
        info = seg_occupancy_prepare(ctx, bb)
        print(f"occupancy bits, cell {pct} %: {info}", flush=True)
        assert info["state"] == 1 and 0 < info["occupied_cells"] <= info["cells_x"] * info["cells_y"] * info["cells_z"]
        for es in (a.edges, crafted):
            de = DevEdges(es)
            ctx.set_option(_lib.OPT_SEG_OCCUPANCY, 1)
            on = to_host(edge_count(ctx, bb, de)).copy()
            st = edge_stats(ctx, bb, de)
            ctx.set_option(_lib.OPT_SEG_OCCUPANCY, 0)
            off = to_host(edge_count(ctx, bb, de)).copy()
            _, ref = orc.edge_pairs(es, b.vertices, b.indices, b.tri_to_obj, with_counts=True)
            assert np.array_equal(on, off) and np.array_equal(on, ref), np.count_nonzero(on != ref)
            assert st["rays_walked"] <= st["rays"]
            if pct <= 200:
                assert st["rays_walked"] < st["rays"]
        assert int(ref.sum()) > 100  # the crafted segments do hit
        bb.free()
    finally:
        ctx.set_option(_lib.OPT_SEG_OCCUPANCY, saved[0])
        ctx.set_option(_lib.OPT_SEG_CELL_PCT, saved[1])


def test_batched_equals_per_query(ctx):
    from rayspace3d_amd.device import batch_execute
    rng = np.random.Generator(np.random.PCG64(91))
    queries, expect, keep = [], [], []
    for k in range(10):
        if k % 3 == 2:
            a, b = _overlap_case(100 + k, int(rng.integers(20, 120)), 60.0)
            ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
            ea, eb = DevEdges(a.edges), DevEdges(b.edges)
            queries.append({"kind": "overlap", "bvh1": ba, "edges1": ea, "bvh2": bb, "edges2": eb})
            expect.append(("pairs", overlap_join(ctx, ba, ea, bb, eb, mode=1)))
            keep += [ba, bb, ea, eb]
        else:
            n_obj = int(rng.integers(5, 400))
            npts = int(rng.integers(1000, 150_000))
            mesh, pts = gen.many_spheres_workload(seed=200 + k, n_objects=n_obj, n_points=npts, world=80.0)
            bvh = Bvh.from_mesh(ctx, mesh)
            dp = to_dev(pts)
            out = torch.full((npts,), -7, dtype=torch.int32, device="cuda")
            queries.append({"kind": "contain", "bvh": bvh, "points": dp, "out": out, "axis": k % 3})
            expect.append(("ids", to_host(trace_points(ctx, bvh, dp, axis=k % 3))))
            keep += [bvh]
    got = batch_execute(ctx, queries)
    for q, (kind, ref), g in zip(queries, expect, got):
        if kind == "pairs":
            assert np.array_equal(g, ref)
        else:
            assert len(g) == 0
            assert np.array_equal(to_host(q["out"]), ref)
    assert sum(len(r) for k, r in expect if k == "pairs") > 20
    for b in keep:
        if isinstance(b, Bvh):
            b.free()


@pytest.mark.parametrize("sizes", [(98300, 65537, 131000), (40000, 3000, 70000)])
def test_batched_containment_reorder_layouts(ctx, sizes):
    # sizes whose whole-chunk padding is small (per-query segmented reorder sort) and
    # large (one global sort on (query, Morton)); both must equal the single-query path
    from rayspace3d_amd.device import batch_execute
    queries, expect, keep = [], [], []
    for k, npts in enumerate(sizes):
        mesh, pts = gen.many_spheres_workload(seed=500 + k, n_objects=250, n_points=npts, world=70.0)
        bvh = Bvh.from_mesh(ctx, mesh)
        dp = to_dev(pts)
        queries.append({"kind": "contain", "bvh": bvh, "points": dp, "axis": (2, 0, 2)[k],
                        "out": torch.full((npts,), -7, dtype=torch.int32, device="cuda")})
        expect.append(to_host(trace_points(ctx, bvh, dp, axis=(2, 0, 2)[k])))
        keep.append(bvh)
    print(f"batched containment, sizes {sizes}", flush=True)
    batch_execute(ctx, queries)
    for q, ref in zip(queries, expect):
        assert np.array_equal(to_host(q["out"]), ref)
    for b in keep:
        b.free()


def test_captured_batch_replays(ctx):
    # graph replays must equal eager batch execution, replay after replay, with eager work
    # and synchronisations in between, and must read inputs rewritten in place
    from rayspace3d_amd.device import CapturedBatch, batch_execute
    queries, keep = [], []
    for k in range(6):
        if k % 3 == 1:
            a, b = _overlap_case(300 + k, 80, 50.0)
            ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
            ea, eb = DevEdges(a.edges), DevEdges(b.edges)
            queries.append({"kind": "overlap", "bvh1": ba, "edges1": ea, "bvh2": bb, "edges2": eb})
            keep += [ba, bb, ea, eb]
        else:
            mesh, pts = gen.many_spheres_workload(seed=400 + k, n_objects=300, n_points=90_000 + k,
                                                  world=70.0)
            bvh = Bvh.from_mesh(ctx, mesh)
            queries.append({"kind": "contain", "bvh": bvh, "points": to_dev(pts), "axis": k % 3,
                            "out": torch.full((len(pts),), -7, dtype=torch.int32, device="cuda")})
            keep += [bvh]
    cb = CapturedBatch(ctx, queries)
    try:
        for rep in range(4):
            print(f"captured batch replay {rep}", flush=True)
            if rep == 2:  # rewrite every point set in place (same size and address)
                for q in queries:
                    if q["kind"] == "contain":
                        q["points"].mul_(0.9).add_(3.0)
            got = cb.run()
            outs = [to_host(q["out"]).copy() for q in queries if q["kind"] == "contain"]
            # eager reference of the same inputs
            for q in queries:
                if q["kind"] == "contain":
                    q["out"].fill_(-7)
            ref = batch_execute(ctx, queries)
            refs = [to_host(q["out"]) for q in queries if q["kind"] == "contain"]
            torch.cuda.synchronize()
            assert all(np.array_equal(g, r) for g, r in zip(got, ref))
            assert all(np.array_equal(o, r) for o, r in zip(outs, refs))
            assert sum(len(r) for r in ref) > 0
            # the eager path per query as well
            ci = 0
            for q in queries:
                if q["kind"] == "contain":
                    alone = to_host(trace_points(ctx, q["bvh"], q["points"], axis=q["axis"]))
                    assert np.array_equal(alone, outs[ci])
                    ci += 1
    finally:
        cb.close()
    for b in keep:
        if isinstance(b, Bvh):
            b.free()


def test_server_alone_equals_mixed(ctx):
    from rayspace3d_amd.device import Server
    rng = np.random.Generator(np.random.PCG64(17))
    resident, _ = gen.many_spheres_workload(seed=401, n_objects=1500, n_points=10, world=80.0)
    resident.edges = gen.extract_edges(resident)
    qb = gen.overlap_dataset(402, n_objects=60, world=80.0, levels=(2, 3))
    qc = gen.overlap_dataset(403, n_objects=40, world=80.0, levels=(2,))
    qb.edges, qc.edges = gen.extract_edges(qb), gen.extract_edges(qc)
    bA, bB, bC = Bvh.from_mesh(ctx, resident), Bvh.from_mesh(ctx, qb), Bvh.from_mesh(ctx, qc)
    eA, eB, eC = DevEdges(resident.edges), DevEdges(qb.edges), DevEdges(qc.edges)
    srv = Server(ctx, max_slots=3)
    A = srv.add_dataset(bA, eA, grid_res=32)
    B = srv.add_dataset(bB, eB, grid_res=32)
    Cq = srv.add_dataset(bC, eC, grid_res=32)
    pts = [to_dev(gen.uniform_points(rng, n, 0.0, 80.0)) for n in (5000, 120_000, 700)]
    outs = [torch.full((p.shape[0],), -9, dtype=torch.int32, device="cuda") for p in pts]
    for p, o in zip(pts, outs):
        srv.submit(_lib.REQ_CONTAIN, A, points=p, out=o)
    t_ov = srv.submit(_lib.REQ_OVERLAP, A, B)
    t_in = srv.submit(_lib.REQ_INTERSECT, A, Cq)
    t_ov2 = srv.submit(_lib.REQ_OVERLAP, A, Cq)
    steps = 0
    while srv.stats()["queued"]:
        srv.step()
        steps += 1
    assert steps >= 2  # three slots: the six requests needed several iterations
    for p, o in zip(pts, outs):
        assert np.array_equal(to_host(o), to_host(trace_points(ctx, bA, p)))
    for t, ref in ((t_ov, overlap_join(ctx, bB, eB, bA, eA, mode=1)),
                   (t_ov2, overlap_join(ctx, bC, eC, bA, eA, mode=1)),
                   (t_in, intersection_join(ctx, bC, eC, bA, eA)[0])):
        done, keys = srv.result(t)
        assert done == 1 and np.array_equal(keys, ref), (t, len(keys), len(ref))
    srv.close()
    for b in (bA, bB, bC):
        b.free()


def _edges_equal(e, ref):
    S, E, off, cnt, src = ref
    return (np.array_equal(e.starts, S) and np.array_equal(e.ends, E)
            and np.array_equal(e.src_offsets, off) and np.array_equal(e.src_counts, cnt)
            and np.array_equal(e.src_objects, src))


def test_gpu_edge_extraction_matches_oracle(ctx):
    import os
    from rayspace3d_amd import preio
    from rayspace3d_amd.prep import extract_edges_gpu
    # the reference's committed fixture (540 closed objects at large coordinates)
    golden = os.path.join(os.path.dirname(__file__), "golden", "check_n_geometry.pre")
    m = preio.read_pre(golden).mesh
    print(f"edge extraction: fixture {len(m.indices)} triangles", flush=True)
    assert _edges_equal(extract_edges_gpu(ctx, m), orc.extract_edges(m.vertices, m.indices, m.tri_to_obj))
    # mixed icosphere levels, rotated, anisotropic
    m2, lv = gen.overlap_dataset(23, n_objects=120, world=90.0, levels=(2, 3, 4), return_levels=True)
    print(f"edge extraction: {len(m2.indices)} triangles, mixed levels", flush=True)
    e2 = extract_edges_gpu(ctx, m2)
    assert _edges_equal(e2, orc.extract_edges(m2.vertices, m2.indices, m2.tri_to_obj))
    ref_np = gen.extract_edges(m2)
    assert np.array_equal(e2.starts, ref_np.starts) and np.array_equal(e2.src_objects, ref_np.src_objects)


def test_gpu_edge_extraction_edge_cases(ctx):
    from rayspace3d_amd.prep import extract_edges_gpu
    rng = np.random.Generator(np.random.PCG64(77))
    # a coarse lattice of positions (many exact duplicates under llround(x * 1e5)), negative
    # coordinates, positions duplicated under different vertex indices, degenerate
    # triangles (repeated index, zero-length edges), one edge shared by several objects in
    # interleaved triangle order, and triangles with an out-of-range index (skipped)
    base = rng.integers(-40, 40, size=(300, 3)).astype(np.float32) * np.float32(0.37)
    V = np.concatenate([base, base[:150], base + np.float32(1e-7)]).astype(np.float32)
    nV = len(V)
    I = rng.integers(0, nV, size=(4000, 3)).astype(np.uint32)
    I[::37, 1] = I[::37, 0]                      # degenerate
    I[5::101, 2] = nV + 3                        # out of range
    O = rng.integers(0, 9, size=len(I)).astype(np.int32)
    m = gen.Mesh(V, I, O, 9)
    print(f"edge extraction edge cases: {len(I)} triangles, {nV} vertices", flush=True)
    ref = orc.extract_edges(V, I, O)
    e = extract_edges_gpu(ctx, m)
    assert _edges_equal(e, ref)
    assert e.num_edges > 100 and (e.src_counts > 1).any()
    # empty mesh
    em = gen.Mesh(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), np.zeros(0, np.int32), 0)
    assert extract_edges_gpu(ctx, em).num_edges == 0


def _check_bvh(bvh, mesh):
    """Local invariants that make a BVH correct, over the nodes reachable from the root
    (records inside collapsed leaves are never visited): leaves hold every triangle once,
    each leaf box holds its triangles, each inner child box holds both boxes of that child."""
    boxes, links, tris = bvh.export()
    n = len(mesh.indices)
    orig = tris[:, 0, 3].view(np.int32)
    assert np.array_equal(np.sort(orig), np.arange(n))
    V = mesh.vertices[mesh.indices[orig]]                                # [n, 3, 3]
    assert np.array_equal(tris[:, :, :3], V)
    assert np.array_equal(tris[:, 1, 3].view(np.int32), mesh.tri_to_obj[orig])
    reach = np.zeros(len(links), bool)
    frontier = np.array([0])
    depth = 0
    while len(frontier):
        assert not reach[frontier].any()                                 # a tree: no node twice
        reach[frontier] = True
        nxt = [links[frontier, c][(links[frontier, 2 + c] == 0) & (links[frontier, c] >= 0)] for c in (0, 1)]
        frontier = np.concatenate(nxt)
        depth += 1
    assert depth <= bvh.info().max_depth
    covered = np.zeros(n, np.int64)
    for c in (0, 1):
        lo, hi = boxes[:, 6 * c:6 * c + 6:2], boxes[:, 6 * c + 1:6 * c + 6:2]
        link, cnt = links[:, c], links[:, 2 + c]
        leaf = reach & (cnt > 0)
        for k in range(int(cnt[leaf].max()) if leaf.any() else 0):
            sel = leaf & (cnt > k)
            t = link[sel] + k
            np.add.at(covered, t, 1)
            tv = tris[t, :, :3]
            assert np.all(tv.min(axis=1) >= lo[sel]) and np.all(tv.max(axis=1) <= hi[sel])
        inner = reach & (cnt == 0) & (link >= 0)
        kids = link[inner]
        for cc in (0, 1):
            klo, khi = boxes[kids, 6 * cc:6 * cc + 6:2], boxes[kids, 6 * cc + 1:6 * cc + 6:2]
            nonempty = klo[:, 0] <= khi[:, 0]
            assert np.all(klo[nonempty] >= lo[inner][nonempty]) and np.all(khi[nonempty] <= hi[inner][nonempty])
    assert np.all(covered == 1)


@pytest.mark.parametrize("n_obj", [300, 5000])
def test_bvh_structure(ctx, n_obj):
    # 5000 mixed-level objects: ~4 M triangles, so the refit runs all three phases with
    # 16 phase-2 spans of 262,144 leaves
    mesh = gen.overlap_dataset(71, n_objects=n_obj, world=400.0, levels=(2, 3))
    print(f"bvh structure: {len(mesh.indices)} triangles", flush=True)
    for leaf in (1, 4):
        b = Bvh.from_mesh(ctx, mesh, leaf_size=leaf)
        _check_bvh(b, mesh)
        b.free()


def _grid_and_bvh(ctx, mesh, pts, axis, leaf=1):
    """Point containment with the axis grid and with the BVH walk on the same tree."""
    b = Bvh.from_mesh(ctx, mesh, leaf_size=leaf)
    info = axis_grid_prepare(ctx, b, axis)
    dp = to_dev(pts)
    try:
        ctx.set_option(_lib.OPT_AXIS_GRID, 1)
        g = to_host(trace_points(ctx, b, dp, axis=axis, out_kind=_lib.OUT_OBJ_I32)).copy()
        ctx.set_option(_lib.OPT_AXIS_GRID, 0)
        v = to_host(trace_points(ctx, b, dp, axis=axis, out_kind=_lib.OUT_OBJ_I32)).copy()
    finally:
        ctx.set_option(_lib.OPT_AXIS_GRID, 1)
    return b, info, g, v


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_grid_matches_bvh_and_oracle(ctx, axis):
    mesh, pts = gen.many_spheres_workload(seed=31, n_objects=3000, n_points=400_000, world=300.0)
    print(f"axis grid: {len(mesh.indices)} triangles, {len(pts)} points, axis {axis}", flush=True)
    b, info, g, v = _grid_and_bvh(ctx, mesh, pts, axis)
    assert info["state"] == 1 and info["entries"] >= len(mesh.indices)
    assert np.array_equal(g, v)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=axis, threads=8)
    assert np.array_equal(g, ref)
    assert (g >= 0).sum() > 100
    st = axis_grid_stats(ctx, b, to_dev(pts), axis)
    assert st["rays"] == len(pts) and 0 < st["nodes_visited"] <= len(pts) and st["tris_tested"] > 0
    b.free()


def test_axis_grid_restarts_and_nesting(ctx):
    # 6 nested spheres (12 surfaces per column, more than either hit buffer) plus
    # points on lattice coordinates shared with vertices, at leaf sizes 1 and 4
    mesh = gen.spheres(3, np.zeros((6, 3)), np.linspace(1.0, 2.5, 6))
    rng = np.random.Generator(np.random.PCG64(8))
    pts = np.concatenate([gen.uniform_points(rng, 30000, -2.7, 2.7),
                          mesh.vertices[rng.integers(0, len(mesh.vertices), 3000)] * np.float32(0.5)])
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    for leaf in (1, 4):
        b, info, g, v = _grid_and_bvh(ctx, mesh, pts, 2, leaf)
        assert info["state"] == 1, info
        assert np.array_equal(g, v) and np.array_equal(g, ref)
        assert axis_grid_stats(ctx, b, to_dev(pts), 2)["restarts"] > 0
        b.free()


@pytest.mark.parametrize("pct", [10, 50, 200])
def test_axis_grid_cell_size_invariant(ctx, pct):
    # 10 % of the mean extent would list ~100 records per triangle: the edge is widened
    # until the 16-per-triangle budget holds; every cell size gives the same answers
    mesh, pts = gen.many_spheres_workload(seed=33, n_objects=1000, n_points=100_000, world=100.0)
    n = len(mesh.indices)
    saved = ctx.get_option(_lib.OPT_GRID_CELL_PCT)
    ctx.set_option(_lib.OPT_GRID_CELL_PCT, pct)
    try:
        b, info, g, v = _grid_and_bvh(ctx, mesh, pts, 2)
    finally:
        ctx.set_option(_lib.OPT_GRID_CELL_PCT, saved)
    print(f"cell pct {pct}: {info}", flush=True)
    assert info["state"] == 1
    assert n <= info["entries"] <= (16 if pct < 100 else 6) * n + (1 << 20)
    assert np.array_equal(g, v)
    assert np.array_equal(g, orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2))
    b.free()


@pytest.mark.parametrize("axis,pct,centre", [(2, 35, 0.0), (0, 10, 400.0), (1, 100, 400.0), (2, 20, -3000.0)])
def test_axis_grid_footprint_edges_and_slivers(ctx, axis, pct, centre):
    # Cells list a triangle only where its projected footprint, widened by the fp32 edge
    # test's error bound, meets them. Triangles spanning many cells (slivers and ordinary
    # ones, near the origin and far from it), and rays on, a few ulps beside and 1e-6
    # beside their projected edges and at their vertices, must give the walk's and the
    # oracle's answers.
    rng = np.random.Generator(np.random.PCG64(40 + axis + pct))
    tris = []
    for k in range(2400):
        c = centre + rng.uniform(-20.0, 20.0, 3)
        size = 10.0 ** rng.uniform(-0.5, 1.0)
        a, b = rng.normal(size=3), rng.normal(size=3)
        if k % 3 == 0:  # sliver: the third vertex close to the first edge
            b = a * rng.uniform(0.0, 1.0) + rng.normal(size=3) * 10.0 ** rng.uniform(-6.0, -2.0)
        tris.append([c, c + a * size, c + b * size])
    V = np.asarray(tris, np.float64).reshape(-1, 3).astype(np.float32)
    I = np.arange(len(V), dtype=np.uint32).reshape(-1, 3)
    O = (np.arange(len(I)) % 97).astype(np.int32)
    mesh = gen.Mesh(V, I, O, num_objects=97)
    U, W = (axis + 1) % 3, (axis + 2) % 3
    pts = []
    for t in range(len(I)):
        P = V[I[t]]
        below = np.float32(P[:, axis].min() - 1.0)
        for e in range(3):
            A, B = P[e], P[(e + 1) % 3]
            for s in np.float32([0.0, 0.25, 0.5, 1.0 / 3.0, 0.999]):
                q = np.empty(3, np.float32)
                q[axis] = below
                for ax in (U, W):
                    q[ax] = np.float32(A[ax] + s * (B[ax] - A[ax]))
                pts.append(q.copy())
                for d in (np.float32(1e-6), np.float32(-1e-6)):  # beside the edge, both ways
                    r = q.copy()
                    r[U] = np.float32(r[U] + d * max(1.0, abs(float(r[U]))))
                    pts.append(r)
                r = q.copy()
                r[W] = np.nextafter(np.nextafter(r[W], np.float32(np.inf)), np.float32(np.inf))
                pts.append(r)
    pts = np.asarray(pts, np.float32)
    saved = ctx.get_option(_lib.OPT_GRID_CELL_PCT)
    ctx.set_option(_lib.OPT_GRID_CELL_PCT, pct)
    try:
        b, info, g, v = _grid_and_bvh(ctx, mesh, pts, axis)
    finally:
        ctx.set_option(_lib.OPT_GRID_CELL_PCT, saved)
    print(f"footprint: {len(I)} triangles, {len(pts)} points, axis {axis}, pct {pct}: {info}", flush=True)
    assert info["state"] == 1 and info["entries"] > (2 if pct >= 100 else 4) * len(I)  # over several cells
    ref = orc.points_contain(V, I, O, pts, axis=axis, threads=8)
    assert np.array_equal(g, v), np.count_nonzero(g != v)
    assert np.array_equal(g, ref), np.count_nonzero(g != ref)
    assert (g >= 0).sum() > 100
    b.free()


def test_axis_grid_rejected_falls_back(ctx):
    # 150 stacked copies of one large square: every ray's cell would list 300 records, so
    # the grid is rejected and launches stay on the BVH with the same results
    n = 150
    base = np.array([[0, 0, 0], [10, 0, 0], [10, 10, 0], [0, 10, 0]], np.float32)
    V = np.concatenate([base + np.float32([0, 0, 0.05 * k]) for k in range(n)]).astype(np.float32)
    I = np.concatenate([np.array([[0, 1, 2], [0, 2, 3]], np.uint32) + 4 * k for k in range(n)])
    O = np.repeat(np.arange(n, dtype=np.int32), 2)
    mesh = gen.Mesh(V, I, O, num_objects=n)
    rng = np.random.Generator(np.random.PCG64(3))
    pts = gen.uniform_points(rng, 5000, -1.0, 11.0)
    b, info, g, v = _grid_and_bvh(ctx, mesh, pts, 2)
    assert info["state"] == 2 and info["entries"] == 0
    assert np.array_equal(g, v)
    assert np.array_equal(g, orc.points_contain(V, I, O, pts, axis=2))
    b.free()


def test_occupancy_selection_path_large(ctx):
    # 1.3 M triangles and 2.0 M edges per side: edge launches and the tri join take the
    # selection + compaction path (2^20 items or more); counts, pairs and the tri join must
    # equal the oracle and the walk without bits
    from rayspace3d_amd.device import tri_join
    a = gen.overlap_dataset(91, n_objects=400, world=150.0, levels=(3, 4))
    b = gen.overlap_dataset(92, n_objects=400, world=150.0, levels=(3, 4))
    a.edges = gen.extract_edges(a)
    print(f"selection path: {a.num_triangles} x {b.num_triangles} triangles, {a.edges.num_edges} edges",
          flush=True)
    assert a.edges.num_edges >= (1 << 20) and a.num_triangles >= (1 << 20)
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    ea = DevEdges(a.edges)
    keys, ref_counts = orc.edge_pairs(a.edges, b.vertices, b.indices, b.tri_to_obj, with_counts=True)
    got = to_host(edge_count(ctx, bb, ea)).copy()
    assert np.array_equal(got, ref_counts), np.count_nonzero(got != ref_counts)
    ref_tri = orc.tri_join(a, b, threads=8)
    assert len(ref_tri) > 100
    assert np.array_equal(tri_join(ctx, ba, bb), ref_tri)
    # both join forms (count + scan + write, and hash) through the selection path
    b.edges = gen.extract_edges(b)
    eb = DevEdges(b.edges)
    ref_pairs = orc.overlap_pairs(a, a.edges, b, b.edges)
    for mode in (0, 1):
        assert np.array_equal(overlap_join(ctx, ba, ea, bb, eb, mode=mode), ref_pairs), mode
    saved = ctx.get_option(_lib.OPT_SEG_OCCUPANCY)
    try:
        ctx.set_option(_lib.OPT_SEG_OCCUPANCY, 0)
        assert np.array_equal(to_host(edge_count(ctx, bb, ea)), ref_counts)
        assert np.array_equal(tri_join(ctx, ba, bb), ref_tri)
    finally:
        ctx.set_option(_lib.OPT_SEG_OCCUPANCY, saved)
    ba.free()
    bb.free()


def test_seg_occupancy_flat_and_rejected(ctx):
    # a flat target (zero extent along z) gets bits over one cell layer; a target with one
    # triangle spanning more than 4096 cells has its bits rejected (edge rays walk); the
    # counts equal the oracle either way
    from rayspace3d_amd.device import seg_occupancy_prepare
    rng = np.random.Generator(np.random.PCG64(17))
    n = 400
    base = rng.uniform(0, 50, (n, 2)).astype(np.float32)
    quads = []
    for k in range(n):
        x, y = base[k]
        quads.append([[x, y, 0], [x + 1, y, 0], [x + 1, y + 1, 0], [x, y + 1, 0]])
    V = np.asarray(quads, np.float32).reshape(-1, 3)
    I = np.concatenate([np.array([[0, 1, 2], [0, 2, 3]], np.uint32) + 4 * k for k in range(n)])
    O = np.repeat(np.arange(n, dtype=np.int32), 2)
    flat = gen.Mesh(V, I, O, num_objects=n)
    big_v = np.array([[-1000, -1000, 5], [1000, -1000, 5], [0, 1000, 5]], np.float32)
    spanning = gen.Mesh(np.concatenate([V, big_v]), np.concatenate([I, np.array([[4 * n, 4 * n + 1, 4 * n + 2]], np.uint32)]),
                        np.concatenate([O, [n]]).astype(np.int32), num_objects=n + 1)
    s = np.concatenate([rng.uniform(-5, 55, (20000, 2)), rng.uniform(-3, 8, (20000, 1))], 1).astype(np.float32)
    e = (s + rng.normal(0, 2.0, s.shape)).astype(np.float32)
    m = len(s)
    segs = gen.EdgeSet(s, e, np.arange(m, dtype=np.int32), np.ones(m, np.int32), np.zeros(m, np.int32))
    for mesh, state in ((flat, 1), (spanning, 2)):
        b = Bvh.from_mesh(ctx, mesh)
        info = seg_occupancy_prepare(ctx, b)
        print(f"occupancy on {mesh.num_triangles} triangles: {info}", flush=True)
        assert info["state"] == state
        got = to_host(edge_count(ctx, b, DevEdges(segs)))
        _, ref = orc.edge_pairs(segs, mesh.vertices, mesh.indices, mesh.tri_to_obj, with_counts=True)
        assert np.array_equal(got, ref) and ref.sum() > 100
        b.free()


@pytest.mark.parametrize("grid", [1, 0])
def test_index_map_launch(ctx, grid):
    # rays given as an index list (a shuffled subset): every listed point gets its answer in
    # each output form, the others keep what the caller wrote there
    mesh, pts = gen.many_spheres_workload(seed=41, n_objects=800, n_points=120_000, world=120.0)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    rng = np.random.Generator(np.random.PCG64(41))
    sel = rng.permutation(len(pts))[: len(pts) // 3].astype(np.int32)
    saved = ctx.get_option(_lib.OPT_AXIS_GRID)
    ctx.set_option(_lib.OPT_AXIS_GRID, grid)
    try:
        b = Bvh.from_mesh(ctx, mesh)
        dp, dm = to_dev(pts), to_dev(sel)
        for kind in (_lib.OUT_OBJ_I32, _lib.OUT_FLAG_U8, _lib.OUT_RAY_RESULT):
            if kind == _lib.OUT_RAY_RESULT:
                out = torch.full((len(pts), 2), -7, dtype=torch.int32, device="cuda")
            elif kind == _lib.OUT_FLAG_U8:
                out = torch.full((len(pts),), 9, dtype=torch.uint8, device="cuda")
            else:
                out = torch.full((len(pts),), -7, dtype=torch.int32, device="cuda")
            got = to_host(trace_points(ctx, b, dp, axis=2, out_kind=kind, index_map=dm, out=out))
            rest = np.setdiff1d(np.arange(len(pts)), sel)
            if kind == _lib.OUT_RAY_RESULT:
                assert np.array_equal(got[sel, 1], ref[sel]) and np.array_equal(got[sel, 0], sel)
                assert (got[rest] == -7).all()
            elif kind == _lib.OUT_FLAG_U8:
                assert np.array_equal(got[sel], (ref[sel] >= 0).astype(np.uint8)) and (got[rest] == 9).all()
            else:
                assert np.array_equal(got[sel], ref[sel]) and (got[rest] == -7).all()
        b.free()
    finally:
        ctx.set_option(_lib.OPT_AXIS_GRID, saved)


def test_grid_sessions_with_deferred_passes(ctx):
    # nearly every point lies over the sphere and below its top, so a 256-point session
    # queues far more than the 128 entries of one walk pass: the rest wait for later passes
    # (contiguous launches, index-map launches, filter-refine and the batch engine)
    mesh, _ = gen.single_sphere_workload(seed=13, n_points=0, level=5)
    rng = np.random.Generator(np.random.PCG64(13))
    n = 150_000
    pts = np.empty((n, 3), np.float32)
    pts[:, 0] = rng.uniform(-0.6, 0.6, n)
    pts[:, 1] = rng.uniform(-0.6, 0.6, n)
    pts[:, 2] = rng.uniform(-1.5, 0.9, n)
    ref = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=2)
    assert 0.3 < (ref >= 0).mean() < 0.9
    print("dense sessions: single sphere L5, 150k points over the sphere", flush=True)
    b = Bvh.from_mesh(ctx, mesh)
    assert axis_grid_prepare(ctx, b, 2)["state"] == 1
    st = axis_grid_stats(ctx, b, to_dev(pts), axis=2)
    assert st["nodes_visited"] > 0.9 * n  # cell-table reads: nearly every point walks
    dp = to_dev(pts)
    got = to_host(trace_points(ctx, b, dp, axis=2))
    assert np.array_equal(got, ref), f"contiguous: {np.count_nonzero(got != ref)} differ"
    perm = rng.permutation(n).astype(np.int32)
    out = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    got = to_host(trace_points(ctx, b, dp, axis=2, index_map=to_dev(perm), out=out))
    assert np.array_equal(got, ref), f"index map: {np.count_nonzero(got != ref)} differ"
    fr = FilterRefinePipeline(ctx).execute(b, dp, axis=2)
    assert np.array_equal(to_host(fr.final_results), ref)
    vmin, vmax = mesh.vertices.min(0), mesh.vertices.max(0)
    assert fr.candidate_count == int(np.all((pts > vmin) & (pts < vmax), axis=1).sum())
    assert fr.inside_count == int((ref >= 0).sum())
    bo = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    batch_execute(ctx, [{"kind": "contain", "bvh": b, "points": dp, "out": bo, "axis": 2}])
    assert np.array_equal(to_host(bo), ref)
    b.free()


def test_server_estimates_match_oracle(ctx):
    # the server computes the pending requests' pair estimates in one launch at the step's
    # start, over the non-empty cells of the sparser histogram: rs3d_estimate_pairs' formula
    # (alpha-corrected, the target's bounds set the cell volume), up to the float rounding of
    # the device-built histograms (tolerance: sizing only)
    from rayspace3d_amd.device import Server
    a = gen.overlap_dataset(3, n_objects=3000, world=100.0, levels=(2,))
    b = gen.overlap_dataset(4, n_objects=3000, world=100.0, levels=(2,))
    c = gen.overlap_dataset(5, n_objects=20, world=100.0, levels=(2, 3))
    for m in (a, b, c):
        m.edges = gen.extract_edges(m)
    bvhs = [Bvh.from_mesh(ctx, m) for m in (a, b, c)]
    srv = Server(ctx, max_slots=8)
    ids = [srv.add_dataset(bv, DevEdges(m.edges), grid_res=16) for bv, m in zip(bvhs, (a, b, c))]
    reqs = [(_lib.REQ_OVERLAP, 0, 1), (_lib.REQ_INTERSECT, 1, 0), (_lib.REQ_OVERLAP, 0, 2), (_lib.REQ_OVERLAP, 2, 2)]
    tickets = [srv.submit(k, ids[d], ids[q]) for k, d, q in reqs]
    assert all(srv.estimate(t) == -1 for t in tickets)
    srv.step()
    grids = [orc.grid_stats(m.vertices, m.indices, m.tri_to_obj, 16) for m in (a, b, c)]
    for t, (_, d, q) in zip(tickets, reqs):
        cd, mn, mx = grids[d]
        ref = max(orc.estimate_pairs(grids[q][0], cd, mn, mx, [16, 16, 16]), 256)
        est = srv.estimate(t)
        assert abs(est - ref) <= max(2, 1e-3 * ref), (d, q, est, ref)
    assert orc.estimate_pairs(grids[1][0], grids[0][0], grids[0][1], grids[0][2], [16, 16, 16]) > 1000
    srv.close()
    for bv in bvhs:
        bv.free()
