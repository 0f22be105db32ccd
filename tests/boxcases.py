"""Box-mesh cases for the containment filter: the 12-triangle mesh of a bounding box
(query/src/geometry/BoundingBox.cpp:26-49 vertex order and triangulation) and points placed
on, next to and between its face planes, including signed-zero and subnormal boxes."""
import numpy as np

from rayspace3d_amd import gen

BOX_TRIS = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                     [3, 6, 2], [3, 7, 6], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
# triangles of the two faces perpendicular to each axis
PERP = {0: [8, 9, 10, 11], 1: [4, 5, 6, 7], 2: [0, 1, 2, 3]}

BOXES = [
    (0.0, 0.0, 0.0, 1.0, 1.0, 1.0),
    (-0.0, -0.0, -0.0, 1.0, 2.0, 3.0),
    (-1.0, -2.0, -3.0, 0.0, -0.0, 0.0),
    (1e-44, 2e-44, 1e-45, 4e-44, 8e-44, 6e-45),
    (-3e-45, -3e-45, -3e-45, 3e-45, 3e-45, 3e-45),
    (1e6, -1e6, 123.25, 1e6 + 1.0, -1e6 + 0.5, 123.5),
    (1e30, 2e30, -3e30, 3e30, 5e30, -1e30),
    (0.5, 0.5, 0.5, 0.5, 1.5, 1.5),  # flat in x
]


def box_mesh(box, sides=True):
    x0, y0, z0, x1, y1, z1 = (np.float32(v) for v in box)
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    return V, BOX_TRIS if sides else None


def perp_mesh(box, axis):
    V, _ = box_mesh(box)
    return V, BOX_TRIS[PERP[axis]]


def _around(lo, hi):
    lo, hi = np.float32(lo), np.float32(hi)
    inf = np.float32(np.inf)
    mid = np.float32(0.5) * lo + np.float32(0.5) * hi
    span = max(abs(float(hi) - float(lo)), 1e-30)
    vals = [lo, np.nextafter(lo, -inf), np.nextafter(lo, inf), mid, hi, np.nextafter(hi, -inf),
            np.nextafter(hi, inf), np.float32(float(lo) - 3 * span), np.float32(float(hi) + 3 * span),
            np.float32(0.0), np.float32(-0.0)]
    return np.array(vals, np.float32)


def box_points(box):
    """Every combination of per-axis coordinates on, one ulp either side of, between and
    beyond the two face planes (plus both zeros)."""
    ax = [_around(box[k], box[3 + k]) for k in range(3)]
    g = np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3)
    return np.ascontiguousarray(g, np.float32)


def mesh_of(V, T):
    return gen.Mesh(V, np.ascontiguousarray(T, np.uint32), np.zeros(len(T), np.int32), 1)


MARGIN_BOXES = [
    (0.0, 0.0, 0.0, 1.0, 1.0, 1.0),
    (-3.0, 10.0, 2.5, 997.0, 1010.0, 1002.5),
    (1e6, -1e6, 123.25, 1e6 + 8.0, -1e6 + 0.5, 131.5),
    (-2e11, 1e11, 3e11, 1e11, 4e11, 5e11),
    (0.0, 0.0, 0.0, 1e-4, 2e-4, 3e-5),
]


def margin_points(box, rng, n=20000):
    """Points around the fast classification's thresholds of the dataset-box filter: per
    coordinate on, beside (multiples of half of m = 2^-16 of the largest magnitude + 1e-5,
    and a few ulps) and between the face planes, and points on and near the diagonal of
    the faces perpendicular to each axis."""
    lo = np.array(box[:3], np.float32)
    hi = np.array(box[3:], np.float32)
    mag = float(np.abs(np.concatenate([lo, hi])).max())
    m = np.float32(mag * 2.0 ** -16 + 1e-5)
    out = np.empty((n, 3), np.float32)
    for k in range(3):
        base = np.where(rng.random(n) < 0.5, lo[k], hi[k]).astype(np.float32)
        step = (rng.integers(-6, 7, n) * np.float32(0.5) * m).astype(np.float32)
        near = (base + step).astype(np.float32)
        ulp = base.copy()
        for _ in range(3):
            sel = rng.random(n) < 0.5
            ulp = np.where(sel, np.nextafter(ulp, np.float32(np.inf)), np.nextafter(ulp, np.float32(-np.inf)))
        inside = (lo[k] + rng.random(n).astype(np.float32) * (hi[k] - lo[k])).astype(np.float32)
        pick = rng.integers(0, 3, n)
        out[:, k] = np.where(pick == 0, near, np.where(pick == 1, ulp, inside))
    # a quarter on or near a face diagonal (lo-lo to hi-hi in the plane of each axis)
    q = n // 4
    for k in range(3):
        u, w = (k + 1) % 3, (k + 2) % 3
        sl = slice(k * q // 3, (k + 1) * q // 3)
        s = rng.uniform(-0.1, 1.1, sl.stop - sl.start).astype(np.float32)
        out[sl, u] = lo[u] + s * (hi[u] - lo[u])
        out[sl, w] = lo[w] + s * (hi[w] - lo[w])
        bump = rng.integers(-3, 4, sl.stop - sl.start).astype(np.float32) * np.float32(0.5) * m
        out[sl, w] = (out[sl, w] + np.where(rng.random(sl.stop - sl.start) < 0.5, bump, 0)).astype(np.float32)
    return out
