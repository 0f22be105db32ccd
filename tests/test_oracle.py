"""Pins the CPU oracle against closed forms, independent methods, invariants and the
reference's committed .pre fixture (tests/golden/check_n_geometry.pre)."""
import os

import numpy as np
import pytest

import oracle as orc
from rayspace3d_amd import gen, preio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "check_n_geometry.pre")


# ---- triangle test ---------------------------------------------------------------------

def test_tri_closed_forms():
    tri = ([0, 0, 5], [1, 0, 5], [0, 1, 5])
    assert orc.tri_hit([0.25, 0.25, 0], [0, 0, 1], *tri) == (True, 5.0)
    # back side (culling off) and reversed winding
    assert orc.tri_hit([0.25, 0.25, 10], [0, 0, -1], *tri) == (True, 5.0)
    assert orc.tri_hit([0.25, 0.25, 0], [0, 0, 1], tri[0], tri[2], tri[1]) == (True, 5.0)
    assert not orc.tri_hit([0.75, 0.75, 0], [0, 0, 1], *tri)[0]
    assert not orc.tri_hit([-0.01, 0.25, 0], [0, 0, 1], *tri)[0]
    # behind the origin: negative t
    hit, t = orc.tri_hit([0.25, 0.25, 7], [0, 0, 1], *tri)
    assert hit and t == -2.0
    # ray parallel to the plane
    assert not orc.tri_hit([0.25, -1, 5], [0, 1, 0], *tri)[0]


def test_tri_watertight_shared_edge_and_vertex():
    # quad split along the diagonal (0,0)-(1,1); rays exactly on the diagonal hit both halves
    a, b, c, d = [0, 0, 2], [1, 0, 2], [1, 1, 2], [0, 1, 2]
    for x in np.linspace(0.01, 0.99, 37):
        o = [x, x, 0]
        h1 = orc.tri_hit(o, [0, 0, 1], a, b, c)
        h2 = orc.tri_hit(o, [0, 0, 1], a, c, d)
        assert h1[0] and h2[0] and h1[1] == 2.0 == h2[1]
    # a ray through the shared vertex of a fan hits every triangle of the fan
    ring = [[np.cos(k * np.pi / 3), np.sin(k * np.pi / 3), 1.0] for k in range(6)]
    hits = [orc.tri_hit([0, 0, 0], [0, 0, 1], [0, 0, 1], ring[k], ring[(k + 1) % 6])[0]
            for k in range(6)]
    assert all(hits)


def _moller_trumbore(o, d, v0, v1, v2):
    o, d, v0, v1, v2 = (np.asarray(x, np.float64) for x in (o, d, v0, v1, v2))
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-12:
        return None
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return u, v, t


def test_tri_against_double_precision_moller_trumbore():
    rng = np.random.Generator(np.random.PCG64(1))
    agree = 0
    for _ in range(4000):
        v = rng.uniform(-2, 2, size=(3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, size=3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        mt = _moller_trumbore(o, d, *v)
        if mt is None:
            continue
        u, w, t = mt
        margin = min(u, w, 1 - u - w)
        if abs(margin) < 1e-4:
            continue  # too close to an edge for a double/float comparison
        hit, tf = orc.tri_hit(o, d, *v)
        assert hit == (margin > 0)
        if hit:
            assert abs(tf - t) <= 1e-5 * max(1.0, abs(t))
        agree += 1
    assert agree > 3500


# ---- point containment (A.1) -----------------------------------------------------------

def test_points_analytic_sphere():
    mesh = gen.spheres(4, np.zeros((1, 3)), np.ones(1))
    rng = np.random.Generator(np.random.PCG64(2))
    pts = gen.uniform_points(rng, 60000, -1.3, 1.3)
    for axis in (0, 1, 2):
        out = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis=axis)
        r = np.linalg.norm(pts.astype(np.float64), axis=1)
        assert np.all(out[r < 0.99] == 0)
        assert np.all(out[r > 1.0] == -1)
        frac_in = (out == 0).mean()
        assert abs(frac_in - (4 / 3 * np.pi) / 2.6 ** 3) < 0.01


def test_points_grid_equals_brute_force():
    mesh, pts = gen.single_sphere_workload(seed=3, n_points=3000, level=3)
    for axis in (0, 1, 2):
        a = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis, True)
        b = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, axis, False)
        assert np.array_equal(a, b)


def test_points_nested_and_overlapping_objects():
    # total-crossing parity: a point inside two overlapping objects is "outside"
    mesh = gen.spheres(3, np.array([[0, 0, 0], [0.5, 0, 0], [5, 5, 5]], float), np.array([1.0, 1.0, 0.5]))
    pts = np.array([[0.2, 0.01, 0.02], [-0.8, 0.01, 0.02], [1.3, 0.01, 0.02], [5, 5, 5.1],
                    [9, 9, 9]], np.float32)
    out = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, 2, False)
    assert list(out) == [-1, 0, 1, 2, -1]


def test_points_first_hit_object_and_outside_box():
    mesh = gen.spheres(2, np.array([[0, 0, 0.0], [0, 0, 10.0]]), np.array([1.0, 1.0]))
    pts = np.array([[0.05, 0.03, 0.1], [0.05, 0.03, 10.2], [0.05, 0.03, 5.0], [3, 3, 0]], np.float32)
    out = orc.points_contain(mesh.vertices, mesh.indices, mesh.tri_to_obj, pts, 2, True)
    assert list(out) == [0, 1, -1, -1]


# ---- edges (A.2) ------------------------------------------------------------------------

def _two_spheres(d, r2=1.0, level=3):
    a = gen.spheres(level, np.zeros((1, 3)), np.ones(1))
    b = gen.spheres(level, np.array([[d, 0.013, 0.021]]), np.array([r2]))
    a.edges, b.edges = gen.extract_edges(a), gen.extract_edges(b)
    return a, b


def test_edges_analytic_sphere_pairs():
    a, b = _two_spheres(1.5)
    assert list(orc.overlap_pairs(a, a.edges, b, b.edges)) == [0]
    a, b = _two_spheres(2.2)
    assert len(orc.overlap_pairs(a, a.edges, b, b.edges)) == 0
    a, b = _two_spheres(0.0, r2=0.5)  # nested, boundaries disjoint
    assert len(orc.overlap_pairs(a, a.edges, b, b.edges)) == 0


def test_edges_grid_equals_brute_force():
    A = gen.overlap_dataset(5, n_objects=40, world=25.0, levels=(2,))
    B = gen.overlap_dataset(6, n_objects=40, world=25.0, levels=(2,))
    ea = gen.extract_edges(A)
    k1, c1 = orc.edge_pairs(ea, B.vertices, B.indices, B.tri_to_obj, use_grid=True, with_counts=True)
    k2, c2 = orc.edge_pairs(ea, B.vertices, B.indices, B.tri_to_obj, use_grid=False, with_counts=True)
    assert np.array_equal(k1, k2) and np.array_equal(c1, c2) and len(k1) > 10
    ks = orc.edge_pairs(ea, B.vertices, B.indices, B.tri_to_obj, swap=True)
    sw = ((k1 & np.uint64(0xffffffff)) << np.uint64(32)) | (k1 >> np.uint64(32))
    assert np.array_equal(ks, sw)


def test_edge_chain_cap_and_self_join():
    # the committed fixture's self join contains every (i, i)
    pre = preio.read_pre(GOLDEN)
    m = pre.mesh
    sub = m.tri_to_obj < 20
    V, I, O = m.vertices, m.indices[sub], m.tri_to_obj[sub]
    sm = gen.Mesh(V, I, O, 20)
    e = gen.extract_edges(sm)
    keys = orc.sort_unique(orc.edge_pairs(e, V, I, O))
    diag = {int(k >> np.uint64(32)) for k in keys if (k >> np.uint64(32)) == (k & np.uint64(0xffffffff))}
    assert diag == set(range(20))


# ---- hashing, sorting, estimation --------------------------------------------------------

M64 = (1 << 64) - 1


def _unfmix64(h):
    c1inv = pow(0xFF51AFD7ED558CCD, -1, 1 << 64)
    c2inv = pow(0xC4CEB9FE1A85EC53, -1, 1 << 64)
    h ^= h >> 33
    h = (h * c2inv) & M64
    h ^= h >> 33
    h = (h * c1inv) & M64
    h ^= h >> 33
    return h


def test_fmix64_is_the_murmur3_finaliser():
    rng = np.random.Generator(np.random.PCG64(3))
    assert orc.fmix64(0) == 0
    for k in rng.integers(0, 1 << 63, size=200, dtype=np.uint64):
        assert _unfmix64(orc.fmix64(int(k))) == int(k)
    assert orc.fmix64(1) != 1


def test_hash_insert_probe_rules():
    table = np.full(64, np.uint64(M64))
    keys = [(i << 32) | (i * 7) for i in range(40)]
    for k in keys:
        assert orc.ht_insert(table, True, 5000, k)
        assert orc.ht_insert(table, True, 5000, k)  # duplicate: found, not reinserted
    assert sorted(int(x) for x in table if x != np.uint64(M64)) == sorted(keys)
    full = np.full(8, np.uint64(M64))
    for i in range(8):
        assert orc.ht_insert(full, False, 8, (i << 32) | 1)
    assert not orc.ht_insert(full, False, 8, (99 << 32) | 1)  # saturation is reported


def test_sort_unique():
    rng = np.random.Generator(np.random.PCG64(5))
    k = rng.integers(0, 1000, size=5000, dtype=np.uint64)
    assert np.array_equal(orc.sort_unique(k), np.unique(k))
    assert len(orc.sort_unique(np.zeros(0, np.uint64))) == 0


def test_estimate_hand_computed():
    a = np.zeros(2, orc.GRID_CELL)
    b = np.zeros(2, orc.GRID_CELL)
    a[0] = (1, 4, 1.0, 0.5)
    b[0] = (1, 2, 2.0, 0.5)
    a[1] = (0, 3, 1.0, 1.0)   # b[1].touch == 0: contributes nothing
    # cell 0: (1 + 2 + 0.001)^3 / 1000 * (sqrt(0.25))^0.8 = 0.027027... * 0.574349...
    p = (3.001 ** 3) / 1000.0 * 0.5 ** 0.8
    exp = 4 * 2 * p
    assert abs(orc.estimate_raw(a, b, 1000.0) - exp) < 1e-5 * exp
    # probability clamps at 1
    assert abs(orc.estimate_raw(a, b, 1e-3) - 8.0) < 1e-6


# ---- the reference's committed .pre fixture ---------------------------------------------

def test_golden_pre_v1_header_and_closedness():
    pre = preio.read_pre(GOLDEN)
    m = pre.mesh
    assert pre.version == 1 and pre.total_triangles == 159840
    assert m.vertices.shape == (81000, 3) and m.indices.shape == (159840, 3)
    assert m.num_objects == 540
    assert np.all(np.bincount(m.tri_to_obj) == 296)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    assert np.allclose(lo, [2.97, 0.72, 16.28], atol=0.01)
    assert np.allclose(hi, [713.59, 581.61, 3013.81], atol=0.01)
    # closed: every undirected edge used by exactly two triangles of its object
    I = m.indices.astype(np.int64)
    e = np.concatenate([I[:, [0, 1]], I[:, [1, 2]], I[:, [2, 0]]])
    e.sort(axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert np.all(cnt == 2)


def test_golden_grid_stats_recomputed():
    pre = preio.read_pre(GOLDEN)
    g = pre.grid
    assert g.resolution == (5, 5, 5)
    assert int(g.cells["center"].sum()) == 540 and int(g.cells["touch"].sum()) == 850
    m = pre.mesh
    cells, mn, mx = orc.grid_stats(m.vertices, m.indices, m.tri_to_obj, 5)
    assert np.allclose(mn, g.min_bound) and np.allclose(mx, g.max_bound)
    assert np.array_equal(cells["center"], g.cells["center"])
    assert np.array_equal(cells["touch"], g.cells["touch"])
    assert np.allclose(cells["avg"], g.cells["avg"], rtol=1e-5)
    assert np.allclose(cells["ratio"], g.cells["ratio"], rtol=1e-4)


def test_golden_edge_extraction():
    m = preio.read_pre(GOLDEN).mesh
    S, E, off, cnt, src = orc.extract_edges(m.vertices, m.indices, m.tri_to_obj)
    # closed manifold objects: 3F/2 edges per object, each with exactly one source
    assert len(S) == 540 * 296 * 3 // 2
    assert np.all(cnt == 1)
    assert np.array_equal(np.bincount(src), np.full(540, 444))
    # the product's extractor produces the identical payload
    e = gen.extract_edges(m)
    assert np.array_equal(e.starts, S) and np.array_equal(e.ends, E)
    assert np.array_equal(e.src_offsets, off) and np.array_equal(e.src_counts, cnt)
    assert np.array_equal(e.src_objects, src)
    # key order: lexicographic on the llround(x * 1e5) quantised, min-first endpoints
    qa = (np.floor(np.abs(S.astype(np.float64) * 1e5) + 0.5) * np.sign(S)).astype(np.int64)
    qb = (np.floor(np.abs(E.astype(np.float64) * 1e5) + 0.5) * np.sign(E)).astype(np.int64)
    keys = [min(tuple(a), tuple(b)) + max(tuple(a), tuple(b)) for a, b in zip(qa[:3000], qb[:3000])]
    assert keys == sorted(keys) and len(set(keys)) == len(keys)


def test_pre_v2_roundtrip(tmp_path):
    m = gen.overlap_dataset(9, n_objects=30, world=50.0, levels=(2,))
    m.edges = gen.extract_edges(m)
    cells, mn, mx = orc.grid_stats(m.vertices, m.indices, m.tri_to_obj, 8)
    g = preio.Grid(mn, mx, (8, 8, 8), cells)
    p = tmp_path / "x.pre"
    preio.write_pre(str(p), m, g)
    r = preio.read_pre(str(p))
    assert r.version == 2
    assert np.array_equal(r.mesh.vertices, m.vertices) and np.array_equal(r.mesh.indices, m.indices)
    assert np.array_equal(r.mesh.edges.src_objects, m.edges.src_objects)
    assert np.array_equal(r.grid.cells, cells)
    with open(p, "r+b") as f:
        f.write(b"XXXX")
    with pytest.raises(ValueError):
        preio.read_pre(str(p))


def test_sphere_field_edges_match_extractor():
    m, _ = gen.many_spheres_workload(seed=5, n_objects=60, n_points=1)
    e1, e2 = gen.extract_edges(m), gen.sphere_field_edges(m, 2)
    rec = lambda e: set(zip(map(tuple, e.starts.tolist()), map(tuple, e.ends.tolist()),
                            np.repeat(e.src_objects, 1).tolist()))
    assert e1.num_edges == e2.num_edges == 60 * 480
    assert rec(e1) == rec(e2)
    # mixed levels (the overlap generator): same records as the extractor
    m3, lv = gen.overlap_dataset(17, n_objects=40, world=60.0, levels=(2, 3, 4), return_levels=True)
    f1, f2 = gen.extract_edges(m3), gen.sphere_field_edges(m3, lv)
    assert f1.num_edges == f2.num_edges == int(sum({2: 480, 3: 1920, 4: 7680}[int(l)] for l in lv))
    assert rec(f1) == rec(f2)


def test_threaded_oracle_matches_single_thread():
    m, p = gen.many_spheres_workload(seed=8, n_objects=300, n_points=60_000, world=80.0)
    a = orc.points_contain(m.vertices, m.indices, m.tri_to_obj, p, 2)
    b = orc.points_contain(m.vertices, m.indices, m.tri_to_obj, p, 2, threads=4)
    assert np.array_equal(a, b) and (a >= 0).any()


# ---- box filter: side faces never count -----------------------------------------------

@pytest.mark.parametrize("axis", [0, 1, 2])
def test_box_side_faces_never_count(axis):
    """The containment filter tests only the 4 triangles perpendicular to the ray axis. On
    the oracle, the parity over those 4 equals the parity over the full 12-triangle box mesh
    for points on, one ulp off and between every face plane, signed-zero and subnormal boxes
    included -- a side face that could count a hit fails this."""
    import boxcases as bc
    for box in bc.BOXES:
        pts = bc.box_points(box)
        V, T = bc.box_mesh(box)
        full = orc.points_contain(V, T, np.zeros(12, np.int32), pts, axis=axis, use_grid=False)
        Vp, Tp = bc.perp_mesh(box, axis)
        perp = orc.points_contain(Vp, Tp, np.zeros(4, np.int32), pts, axis=axis, use_grid=False)
        assert np.array_equal(full >= 0, perp >= 0), box
        # an interior point is inside when the far face lies inside the [1e-6, 1e16) t window
        if all(box[3 + k] > box[k] for k in range(3)) and 1e-3 < box[3 + axis] - box[axis] < 1e15:
            mid = np.array([[0.5 * box[k] + 0.5 * box[3 + k] for k in range(3)]], np.float32)
            if np.all((mid > np.float32(box[:3])) & (mid < np.float32(box[3:]))):
                assert orc.points_contain(V, T, np.zeros(12, np.int32), mid, axis=axis,
                                          use_grid=False)[0] == 0, box
