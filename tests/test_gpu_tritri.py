"""Exact triangle-triangle intersection on the GPU (rs3d_tri_intersect, rs3d_tri_join*)
against the oracle's predicate and join, bit for bit."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as orc
from rayspace3d_amd import _lib, gen
from rayspace3d_amd.device import Bvh, DevEdges, overlap_join, to_dev, to_host, tri_intersect, tri_join, tri_join_stats
from test_tritri_oracle import _pair_sets

pytestmark = pytest.mark.gpu


def test_tri_intersect_matches_oracle(ctx):
    rng = np.random.default_rng(9)
    for k, (A, B) in enumerate(_pair_sets(rng)):
        print(f"tri pairs set {k}: {len(A)} pairs", flush=True)
        got = to_host(tri_intersect(ctx, to_dev(A), to_dev(B))).astype(bool)
        np.testing.assert_array_equal(got, orc.tri_tri(A, B))
    # large random batch (many blocks), every branch of the predicate mixed in one launch
    A = rng.integers(-3, 4, (300000, 3, 3)).astype(np.float32) * np.float32(0.25)
    B = rng.integers(-3, 4, (300000, 3, 3)).astype(np.float32) * np.float32(0.25)
    got = to_host(tri_intersect(ctx, to_dev(A), to_dev(B))).astype(bool)
    np.testing.assert_array_equal(got, orc.tri_tri(A, B))


def _cubes(origins, size=1.0):
    """Axis-aligned cubes (12 triangles each), one object per cube."""
    corners = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32) * size
    faces = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1],
                      [2, 3, 7], [2, 7, 6], [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    V = (origins[:, None, :].astype(np.float32) + corners[None]).reshape(-1, 3).astype(np.float32)
    I = (faces[None] + 8 * np.arange(len(origins), dtype=np.uint32)[:, None, None]).reshape(-1, 3)
    O = np.repeat(np.arange(len(origins), dtype=np.int32), 12)
    return gen.Mesh(V, I, O, num_objects=len(origins))


@pytest.mark.parametrize("leaf", [1, 4])
def test_tri_join_matches_oracle(ctx, leaf):
    a = gen.overlap_dataset(51, n_objects=400, world=120.0, levels=(2, 3))
    b = gen.overlap_dataset(52, n_objects=400, world=120.0, levels=(2, 3))
    print(f"tri join {len(a.indices)} x {len(b.indices)} triangles, leaf {leaf}", flush=True)
    ba, bb = Bvh.from_mesh(ctx, a, leaf_size=leaf), Bvh.from_mesh(ctx, b, leaf_size=leaf)
    ref = orc.tri_join(a, b, threads=8)
    assert len(ref) > 50
    got = tri_join(ctx, ba, bb)
    assert np.array_equal(got, ref), f"{len(got)} vs {len(ref)} pairs"
    # B -> A gives the same pairs, swapped
    back = tri_join(ctx, bb, ba)
    sw = np.sort((back << np.uint64(32)) | (back >> np.uint64(32)))
    assert np.array_equal(sw, ref)
    st = tri_join_stats(ctx, ba, bb)
    assert st["queries"] == len(a.indices) and st["predicate_tests"] >= st["hits"] > 0
    assert st["tris_loaded"] >= st["predicate_tests"]
    ba.free()
    bb.free()


def test_tri_join_touching_cubes(ctx):
    # Cubes of B sit exactly on cubes of A (shared faces, edges or corners, or a gap of one
    # float ulp): the contacts are coplanar or touching, which edge rays need not report.
    rng = np.random.default_rng(5)
    oa = np.stack(np.meshgrid(np.arange(0, 40, 4), np.arange(0, 40, 4), np.arange(0, 8, 4)), -1).reshape(-1, 3)
    oa = oa.astype(np.float32)
    shift = rng.integers(0, 4, len(oa))
    step = np.array([[1, 0, 0], [1, 1, 0], [1, 1, 1], [0, 0, 1]], np.float32)[shift]
    ob = oa + step
    gap = rng.random(len(oa)) < 0.25
    ob[gap, 0] = np.nextafter(ob[gap, 0], np.float32(np.inf))   # a one-ulp gap in x
    a, b = _cubes(oa), _cubes(ob.astype(np.float32))
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    ref = orc.tri_join(a, b, threads=8)
    got = tri_join(ctx, ba, bb)
    assert np.array_equal(got, ref)
    pairs = {(int(k >> np.uint64(32)), int(k & np.uint64(0xffffffff))) for k in got}
    for i in range(len(oa)):
        touching = not gap[i] or step[i][0] == 0            # the gap separates only x contacts
        assert ((i, i) in pairs) == touching, i
    # the edge-ray join on the same input, for the record
    a.edges, b.edges = gen.extract_edges(a), gen.extract_edges(b)
    ea, eb = DevEdges(a.edges), DevEdges(b.edges)
    edge_pairs = overlap_join(ctx, ba, ea, bb, eb, mode=1)
    print(f"touching cubes: tri join {len(got)} pairs, edge-ray join {len(edge_pairs)}", flush=True)
    ba.free()
    bb.free()


def test_tri_join_hash_swap(ctx):
    a = gen.overlap_dataset(61, n_objects=100, world=60.0, levels=(2,))
    b = gen.overlap_dataset(62, n_objects=100, world=60.0, levels=(2,))
    ba, bb = Bvh.from_mesh(ctx, a), Bvh.from_mesh(ctx, b)
    ref = orc.tri_join(a, b, threads=8)
    size = 1 << 16
    slots = torch.full((size + 1,), -1, dtype=torch.int64, device="cuda")
    slots[size] = 0
    ht = _lib.HashTable(C.c_void_p(slots.data_ptr()), size, 1, 0, C.c_void_p(slots.data_ptr() + 8 * size))
    torch.cuda.synchronize()
    _lib.check(ctx.lib.rs3d_tri_join_hash(ctx.handle, ba.handle, bb.handle, 1, C.byref(ht)), "tri_join_hash")
    ctx.sync()
    s = to_host(slots, np.uint64)
    assert s[size] == 0
    keys = np.sort(s[:size][s[:size] != np.uint64(0xffffffffffffffff)])
    sw = np.sort((ref << np.uint64(32)) | (ref >> np.uint64(32)))
    assert np.array_equal(keys, sw)
    ba.free()
    bb.free()
