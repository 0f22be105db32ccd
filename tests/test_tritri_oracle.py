"""The exact triangle-triangle predicate of the oracle (oracle/tritri.c) pinned three ways:
hand-computed cases, an exact separating-axis test in integer arithmetic on lattice
triangles (where every fp32 operation of the predicate is exact), and the reference's own
TriInt compiled from its source (oracle/_ref/libtriint.so) on real-valued and degenerate
pairs. Then the oracle's join against the all-pairs predicate."""
import itertools

import numpy as np
import pytest

import oracle as orc
from rayspace3d_amd import gen


def T(*pts):
    return np.array(pts, np.float32).reshape(3, 3)


def test_known_answers():
    base = T((0, 0, 0), (4, 0, 0), (0, 4, 0))
    cases = [
        (T((1, 1, -1), (1, 1, 1), (2, 3, 1)), True),        # pierces the interior
        (T((0, 0, 1), (4, 0, 1), (0, 4, 1)), False),        # parallel plane
        (T((4, 0, 0), (5, 0, 1), (5, 1, 1)), True),         # shares one vertex
        (T((4, 0, 0), (0, 4, 0), (3, 3, 2)), True),         # shares an edge
        (T((1, 1, 0), (5, 1, 0), (1, 5, 0)), True),         # coplanar, overlapping
        (T((5, 5, 0), (7, 5, 0), (5, 7, 0)), False),        # coplanar, apart
        (T((2, 2, 0), (6, 2, 0), (2, 6, 0)), True),         # coplanar, touching at (2, 2)
        (T((1, 1, 0), (1, 1, 3), (2, 2, 3)), True),         # one vertex on the interior
        (T((5, 5, 0), (5, 5, 3), (6, 6, 3)), False),        # one vertex on the plane, outside
        (T((-1, 2, -1), (-1, 2, 1), (1, 2, 0)), True),      # crosses the edge x = 0
        (T((3, 3, -1), (3, 3, 1), (5, 5, 0)), False),       # passes beside the hypotenuse
    ]
    A = np.stack([base] * len(cases))
    B = np.stack([c for c, _ in cases])
    got = orc.tri_tri(A, B)
    assert got.tolist() == [e for _, e in cases]
    assert orc.tri_tri(B, A).tolist() == [e for _, e in cases]


def _sat_intersect(t1, t2):
    """Exact closed-triangle intersection by separating axes (integer arithmetic)."""
    t1 = [tuple(int(c) for c in p) for p in t1]
    t2 = [tuple(int(c) for c in p) for p in t2]
    sub = lambda a, b: (a[0] - b[0], a[1] - b[1], a[2] - b[2])
    cross = lambda a, b: (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    e1 = [sub(t1[(k + 1) % 3], t1[k]) for k in range(3)]
    e2 = [sub(t2[(k + 1) % 3], t2[k]) for k in range(3)]
    n1, n2 = cross(e1[0], e1[1]), cross(e2[0], e2[1])
    axes = [n1, n2] + [cross(a, b) for a in e1 for b in e2] + [cross(n1, a) for a in e1] + \
           [cross(n2, b) for b in e2]
    for ax in axes:
        if ax == (0, 0, 0):
            continue
        p1 = [dot(ax, p) for p in t1]
        p2 = [dot(ax, p) for p in t2]
        if max(p1) < min(p2) or max(p2) < min(p1):
            return False
    return True


def test_lattice_pairs_equal_exact_separating_axes():
    # Small integer coordinates: every difference, cross product, dot product and product
    # of the predicate is an integer below 2^24, so fp32 evaluates it exactly, and the
    # predicate must agree with the exact test on every non-degenerate pair -- touching,
    # shared vertices and edges, and coplanar pairs included (a 3^3 lattice makes them common).
    rng = np.random.default_rng(3)
    A = rng.integers(-1, 2, (6000, 3, 3)).astype(np.float32)
    B = rng.integers(-1, 2, (6000, 3, 3)).astype(np.float32)
    area = lambda t: np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1)
    keep = (area(A) > 0) & (area(B) > 0)
    A, B = A[keep], B[keep]
    got = orc.tri_tri(A, B)
    exp = np.array([_sat_intersect(a, b) for a, b in zip(A, B)])
    assert len(A) > 3000 and 0.2 < exp.mean() < 0.9
    coplanar = np.array([abs(np.dot(np.cross(a[1] - a[0], a[2] - a[0]), b[k] - a[0])) == 0
                         for a, b in zip(A, B) for k in (0,)])
    assert coplanar.sum() > 100
    np.testing.assert_array_equal(got, exp)


def _pair_sets(rng):
    """Real-valued, lattice, and 1-ulp-perturbed near-degenerate triangle pairs."""
    sets = []
    sets.append((rng.normal(size=(4000, 3, 3)), rng.normal(size=(4000, 3, 3))))
    sets.append((rng.integers(-2, 3, (4000, 3, 3)), rng.integers(-2, 3, (4000, 3, 3))))
    L = rng.integers(-2, 3, (4000, 3, 3)).astype(np.float32) * np.float32(0.1)
    M = rng.integers(-2, 3, (4000, 3, 3)).astype(np.float32) * np.float32(0.1)
    bump = rng.integers(-1, 2, M.shape).astype(np.float32)
    M = np.where(bump > 0, np.nextafter(M, np.float32(np.inf)), np.where(bump < 0, np.nextafter(M, np.float32(-np.inf)), M))
    sets.append((L, M))
    # coplanar pairs in a tilted plane (rounding makes the distances tiny, not always zero)
    u, v = np.array([1, 2, 0.5]), np.array([-0.3, 0.7, 1.1])
    c = rng.uniform(-1, 1, (4000, 6, 2))
    P = c[..., :1] * u + c[..., 1:] * v
    sets.append((P[:, :3], P[:, 3:]))
    return [(np.asarray(a, np.float32), np.asarray(b, np.float32)) for a, b in sets]


def test_oracle_equals_reference_triint():
    ref = orc.ref_tri_int()
    if ref is None:
        pytest.skip("oracle/_ref/libtriint.so not built (the reference tree is absent)")
    rng = np.random.default_rng(9)
    for A, B in _pair_sets(rng):
        got = orc.tri_tri(A, B)
        exp = np.array([ref(a, b) for a, b in zip(A, B)])
        np.testing.assert_array_equal(got, exp)
        assert 0.0 < exp.mean() < 1.0


def test_join_equals_all_pairs_predicate():
    # The join's box filter removes no pair that the predicate accepts: compare with the
    # predicate on every triangle pair (the reference's MeshInt loops over all pairs).
    a = gen.overlap_dataset(5, n_objects=12, world=12.0, levels=(1,))
    b = gen.overlap_dataset(6, n_objects=12, world=12.0, levels=(1,))
    keys = orc.tri_join(a, b, threads=4)
    ta = a.vertices[a.indices]
    tb = b.vertices[b.indices]
    ia, ib = np.meshgrid(np.arange(len(ta)), np.arange(len(tb)), indexing="ij")
    hit = orc.tri_tri(ta[ia.ravel()], tb[ib.ravel()])
    pairs = {(int(a.tri_to_obj[i]), int(b.tri_to_obj[j])) for i, j in zip(ia.ravel()[hit], ib.ravel()[hit])}
    assert len(pairs) > 5
    assert {(int(k >> np.uint64(32)), int(k & np.uint64(0xffffffff))) for k in keys} == pairs
    assert np.all(np.diff(keys.astype(np.int64)) > 0)
