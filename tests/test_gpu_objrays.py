"""Object rays with the reference's launch options (rs3d_object_rays: all-hits and both
closest-hit loops, swap, explicit origins, two-pass output, per-source tracking and the
profiling counters) and the overlap counters (rs3d_edge_profile), against the oracle."""
import numpy as np
import pytest
import torch

import oracle as orc
from rayspace3d_amd import _lib, gen
from rayspace3d_amd.device import Bvh, DevEdges, object_rays, to_dev

pytestmark = pytest.mark.gpu

UNIT_TRI = np.array([[-2, -1], [0, -1], [1, 1]], np.float32)


def _planes(zs, objs=None):
    zs = np.asarray(zs, np.float32)
    n = len(zs)
    V = np.zeros((3 * n, 3), np.float32)
    V[:, :2] = np.tile(UNIT_TRI, (n, 1))
    V[:, 2] = np.repeat(zs, 3)
    O = np.arange(n, dtype=np.int32) if objs is None else np.asarray(objs, np.int32)
    return gen.Mesh(V, np.arange(3 * n, dtype=np.uint32).reshape(n, 3), O, int(O.max()) + 1)


def _origin_source():
    V = np.array([[0, 0, 0], [1e-3, 0, -1], [0, 1e-3, -1]], np.float32)
    return gen.Mesh(V, np.array([[0, 1, 2]], np.uint32), np.zeros(1, np.int32), 1)


def _sphere_case(seed, n_t=40, n_s=300):
    rng = np.random.Generator(np.random.PCG64(seed))
    tgt = gen.spheres(2, rng.uniform(0, 12, size=(n_t, 3)), rng.uniform(1.0, 3.0, size=n_t), noise=0.05, rng=rng)
    src = gen.spheres(1, rng.uniform(0, 12, size=(n_s, 3)), rng.uniform(0.2, 0.8, size=n_s),
                      rotations=gen.random_rotations(rng, n_s))
    return src, tgt


def _compare(ctx, src, tgt, swap, max_targets, mode, max_iter, max_pair=16):
    bs, bt = Bvh.from_mesh(ctx, src), Bvh.from_mesh(ctx, tgt)
    got = object_rays(ctx, bs, bt, swap=swap, max_targets=max_targets, chain_mode=mode, max_iter=max_iter,
                      track=True, max_pair=max_pair, with_stats=True)
    keys, over, tr, st = orc.object_parity_ex(src, tgt, max_targets, swap, mode, max_iter, track=True,
                                              max_pair=max_pair, max_out=256)
    assert np.array_equal(got["keys"], np.unique(keys)), (mode, len(got["keys"]), len(keys))
    assert got["overflow"] == over
    for k in ("iterations", "candidate_count", "candidate_overflow", "pair_target_ids", "pair_target_hits"):
        assert np.array_equal(got["tracking"][k].astype(np.int64), tr[k].astype(np.int64)), k
    g = got["stats"]
    assert (g["containment_rays_total"], g["containment_iterations_total"], g["containment_hits_total"],
            g["containment_max_iterations_per_ray"], g["containment_same_hit_suppressed"],
            g["containment_candidate_additions"], g["containment_candidate_toggles"],
            g["containment_candidate_overflow"], g["containment_targets_total"]) == \
        tuple(st[k] for k in orc.STAT_NAMES), (g, st)
    bs.free()
    bt.free()
    return got


@pytest.mark.parametrize("mode", [_lib.OBJ_ALL_HITS, _lib.OBJ_CHAIN_NEXTAFTER, _lib.OBJ_CHAIN_STEP])
@pytest.mark.parametrize("swap", [False, True])
def test_object_ray_modes_on_spheres(ctx, mode, swap):
    print(f"object rays on spheres, mode {mode}, swap {swap}", flush=True)
    src, tgt = _sphere_case(31)
    got = _compare(ctx, src, tgt, swap, 256, mode, 0)
    assert len(got["keys"]) > 20


@pytest.mark.parametrize("mode", [_lib.OBJ_ALL_HITS, _lib.OBJ_CHAIN_NEXTAFTER, _lib.OBJ_CHAIN_STEP])
def test_object_ray_modes_on_plane_stacks(ctx, mode):
    # ties, folds, caps and buffer refills: 300 planes (many more than the 8-hit buffer),
    # pairs of planes 4e-6 apart on one object, equal heights on two objects, shuffled ids
    rng = np.random.Generator(np.random.PCG64(7))
    zs = list(0.5 + 0.25 * np.arange(300))
    objs = list(rng.permutation(300))
    zs += [80.0, 80.000004, 90.0, 90.0, 100.0, 100.00005, 3000.0]
    objs += [300, 300, 301, 302, 303, 304, 305]
    m = _planes(zs, objs)
    print(f"object rays on plane stacks, mode {mode}", flush=True)
    for max_targets, max_iter in ((512, 0), (256, 0), (512, 100), (64, 1999)):
        _compare(ctx, _origin_source(), m, False, max_targets, mode, max_iter, max_pair=8)


def test_object_ray_origins_two_pass_and_point_in_mesh(ctx):
    src, tgt = _sphere_case(41, n_t=25, n_s=200)
    bs, bt = Bvh.from_mesh(ctx, src), Bvh.from_mesh(ctx, tgt)
    first = np.stack([src.vertices[src.indices[np.nonzero(src.tri_to_obj == o)[0][0], 0]] for o in range(200)])
    for mode in (_lib.OBJ_ALL_HITS, _lib.OBJ_CHAIN_STEP):
        a = object_rays(ctx, bs, bt, chain_mode=mode)
        b = object_rays(ctx, None, bt, chain_mode=mode, origins=to_dev(first.astype(np.float32)))
        c = object_rays(ctx, bs, bt, chain_mode=mode, two_pass=True)
        ref = np.unique(orc.object_parity_ex(src, tgt, 256, False, mode)[0])
        assert np.array_equal(a["keys"], ref) and np.array_equal(b["keys"], ref) and np.array_equal(c["keys"], ref)
        assert len(ref) > 10
    # point-in-mesh form: (target, source) pairs not in the overlap table
    ref = np.unique(orc.object_parity_ex(src, tgt, 512, True, _lib.OBJ_CHAIN_STEP)[0])
    drop = ref[::3]
    size = 1 << 12
    tab = torch.full((size + 1,), -1, dtype=torch.int64, device="cuda")
    tab[size] = 0
    table = np.full(size, np.uint64(0xFFFFFFFFFFFFFFFF))
    for k in drop:
        assert orc.ht_insert(table, True, 5000, int(k))
    tab[:size].copy_(to_dev(table))
    ov = _lib.HashTable(tab.data_ptr(), size, 1, 0, tab.data_ptr() + 8 * size)
    got = object_rays(ctx, bs, bt, max_targets=512, chain_mode=_lib.OBJ_CHAIN_STEP, form=_lib.OBJ_POINT_IN_MESH,
                      overlap_table=ov)
    assert np.array_equal(got["keys"], np.setdiff1d(ref, drop)) and len(drop) > 3
    bs.free()
    bt.free()


@pytest.mark.parametrize("chain_rule,cap", [(0, 100), (0, 3), (1, 256)])
def test_edge_profile_counters(ctx, chain_rule, cap):
    import ctypes as C
    a = gen.overlap_dataset(61, n_objects=60, world=40.0, levels=(2, 3))
    b = gen.overlap_dataset(62, n_objects=60, world=40.0, levels=(2,))
    ea = gen.extract_edges(a)
    bb = Bvh.from_mesh(ctx, b)
    de = DevEdges(ea)
    st = torch.zeros(13, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(ctx.lib.rs3d_edge_profile(ctx.handle, bb.handle, C.byref(de.s), chain_rule, cap,
                                         C.c_void_p(st.data_ptr())), "rs3d_edge_profile")
    sv = _lib.IntersectionStats.from_buffer_copy(st.cpu().numpy().view(np.uint8).tobytes())
    _, counts = orc.edge_pairs(ea, b.vertices, b.indices, b.tri_to_obj, chain_rule=chain_rule, max_iter=cap,
                               with_counts=True)
    d = (ea.ends - ea.starts).astype(np.float32)
    ln = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)
    live = (ln >= np.float32(1e-6)) & (ea.src_counts > 0)
    sc = ea.src_counts.astype(np.int64)
    h = np.where(live, counts // np.maximum(sc, 1), 0)
    it = np.minimum(h + 1, cap)
    assert sv.overlap_trace_calls == int(sc[live].sum())
    assert sv.overlap_iterations_total == int((it * sc)[live].sum())
    assert sv.overlap_hits_total == int(counts[live].sum()) > 0
    assert sv.overlap_max_iterations_per_trace == int(it[live].max())
    bb.free()
