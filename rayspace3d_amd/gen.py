"""Seeded synthetic inputs for the ray-cast join workloads.

Meshes are closed icospheres (icosahedron with L midpoint subdivisions projected to the
unit sphere: F = 20*4^L triangles, V = 10*4^L + 2 vertices), scaled, rotated and placed
per object; points are uniform in a box. All randomness comes from numpy's PCG64 with
an explicit seed, so every input is reproducible from (workload, seed).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

_T = (1.0 + 5.0 ** 0.5) / 2.0
_ICO_V = np.array(
    [(-1, _T, 0), (1, _T, 0), (-1, -_T, 0), (1, -_T, 0), (0, -1, _T), (0, 1, _T),
     (0, -1, -_T), (0, 1, -_T), (_T, 0, -1), (_T, 0, 1), (-_T, 0, -1), (-_T, 0, 1)],
    dtype=np.float64)
_ICO_F = np.array(
    [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
     (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
     (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)


@lru_cache(maxsize=None)
def icosphere(level: int):
    """Unit icosphere (float64 vertices, int64 faces, outward counter-clockwise)."""
    v = _ICO_V / np.linalg.norm(_ICO_V, axis=1, keepdims=True)
    f = _ICO_F.copy()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        m = len(f)
        ab, bc, ca = inv[:m] + base, inv[m:2 * m] + base, inv[2 * m:] + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1),
                            np.stack([c, ca, bc], 1), np.stack([ab, bc, ca], 1)])
    v.setflags(write=False)
    f.setflags(write=False)
    return v, f


@dataclass
class Mesh:
    """Indexed triangle soup of many objects, in the layout of the .pre payload."""
    vertices: np.ndarray          # float32 (nV, 3)
    indices: np.ndarray           # uint32 (nT, 3)
    tri_to_obj: np.ndarray        # int32 (nT,)
    num_objects: int = 0
    edges: "EdgeSet | None" = None

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class EdgeSet:
    """Unique edges with their source objects (EdgeData of the .pre format)."""
    starts: np.ndarray            # float32 (nE, 3)
    ends: np.ndarray              # float32 (nE, 3)
    src_offsets: np.ndarray       # int32 (nE,)
    src_counts: np.ndarray        # int32 (nE,)
    src_objects: np.ndarray       # int32 (sum counts,)

    @property
    def num_edges(self) -> int:
        return int(self.starts.shape[0])


def random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices (n, 3, 3) from unit quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)


def spheres(levels, centers, radii, scales=None, rotations=None, noise=0.0, rng=None) -> Mesh:
    """Many icospheres. levels: int or (n,) ints; radii (n,); scales (n,3); rotations (n,3,3)."""
    n = len(centers)
    levels = np.broadcast_to(np.asarray(levels), (n,))
    vparts, fparts, oparts = [], [], []
    voff = 0
    for lv in np.unique(levels):
        sel = np.nonzero(levels == lv)[0]
        uv, uf = icosphere(int(lv))
        k = len(sel)
        p = np.broadcast_to(uv, (k,) + uv.shape).copy()
        if noise:
            p *= (1.0 + rng.uniform(-noise, noise, size=(k, len(uv), 1)))
        if scales is not None:
            p *= scales[sel][:, None, :]
        if rotations is not None:
            p = np.einsum("kij,kvj->kvi", rotations[sel], p)
        p = p * radii[sel][:, None, None] + centers[sel][:, None, :]
        vparts.append((sel, p.astype(np.float32)))
        fparts.append((sel, uf, len(uv)))
    # assemble in object order so triangle blocks are contiguous per object
    order_v = {}
    for (sel, p), (_, uf, nv) in zip(vparts, fparts):
        for j, obj in enumerate(sel):
            order_v[int(obj)] = (p[j], uf, nv)
    V, F, O = [], [], []
    for obj in range(n):
        p, uf, nv = order_v[obj]
        V.append(p)
        F.append(uf + voff)
        O.append(np.full(len(uf), obj, dtype=np.int32))
        voff += nv
    return Mesh(np.ascontiguousarray(np.concatenate(V), dtype=np.float32),
                np.ascontiguousarray(np.concatenate(F), dtype=np.uint32),
                np.concatenate(O), num_objects=n)


def uniform_points(rng: np.random.Generator, n: int, lo, hi) -> np.ndarray:
    return rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)


# ---- workloads ----------------------------------------------------------------------

def single_sphere_workload(seed: int = 1, n_points: int = 1_000_000, level: int = 5):
    """One closed icosphere (level 5: 20,480 triangles), r = 1 at the origin, vertices
    radially perturbed by U(-0.05, 0.05); points uniform in [-1.5, 1.5]^3."""
    rng = np.random.Generator(np.random.PCG64(seed))
    m = spheres(level, np.zeros((1, 3)), np.ones(1), noise=0.05, rng=rng)
    pts = uniform_points(rng, n_points, -1.5, 1.5)
    return m, pts


def many_spheres_workload(seed: int = 2, n_objects: int = 10_000, n_points: int = 10_000_000,
                          world: float = 1000.0, level: int = 2):
    """n_objects closed icospheres (level 2: 320 triangles each), radii log-uniform in
    [0.5, 5], centres uniform in [0, world]^3; points uniform in the same cube."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centers = rng.uniform(0.0, world, size=(n_objects, 3))
    radii = np.exp(rng.uniform(np.log(0.5), np.log(5.0), size=n_objects))
    m = spheres(level, centers, radii)
    pts = uniform_points(rng, n_points, 0.0, world)
    return m, pts


def overlap_dataset(seed: int, n_objects: int = 50_000, world: float = 1000.0,
                    levels=(2, 3, 4), r0: float = 5.0, return_levels: bool = False):
    """Anisotropically scaled (0.5-2x per axis), randomly rotated icospheres of levels
    drawn uniformly from `levels`, centres uniform in [0, world]^3. With return_levels,
    also returns the per-object level array (for sphere_field_edges)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lv = rng.choice(np.asarray(levels), size=n_objects)
    centers = rng.uniform(0.0, world, size=(n_objects, 3))
    scales = rng.uniform(0.5, 2.0, size=(n_objects, 3))
    rots = random_rotations(rng, n_objects)
    mesh = spheres(lv, centers, np.full(n_objects, r0), scales=scales, rotations=rots)
    return (mesh, lv) if return_levels else mesh


# ---- unique-edge extraction (preprocess/src/EdgePreprocessor.cpp:41-119) --------------

def _llround(x: np.ndarray) -> np.ndarray:
    """C llround: round half away from zero."""
    a = np.floor(np.abs(x) + 0.5)
    return (np.sign(x) * a).astype(np.int64)


def extract_edges(mesh: Mesh) -> EdgeSet:
    """Unique edges keyed by endpoints quantised with llround(x * 1e5), in lexicographic
    key order; start/end are the endpoints of the edge's first occurrence (triangle
    order, edges (v0,v1),(v1,v2),(v2,v0)); sources are the distinct objects in order of
    first appearance."""
    V = mesh.vertices
    I = mesh.indices.astype(np.int64)
    nT = len(I)
    if nT == 0:
        z3 = np.zeros((0, 3), np.float32)
        zi = np.zeros(0, np.int32)
        return EdgeSet(z3, z3.copy(), zi, zi.copy(), zi.copy())
    a = I[:, [0, 1, 2]].reshape(-1)
    b = I[:, [1, 2, 0]].reshape(-1)
    order = np.arange(3 * nT, dtype=np.int64)
    qa = _llround(V[a].astype(np.float64) * 1e5)
    qb = _llround(V[b].astype(np.float64) * 1e5)
    # lexicographic min/max of the two quantised endpoints
    lt = (qb[:, 0] < qa[:, 0]) | ((qb[:, 0] == qa[:, 0]) & (
        (qb[:, 1] < qa[:, 1]) | ((qb[:, 1] == qa[:, 1]) & (qb[:, 2] < qa[:, 2]))))
    k0 = np.where(lt[:, None], qb, qa)
    k1 = np.where(lt[:, None], qa, qb)
    perm = np.lexsort((order, k1[:, 2], k1[:, 1], k1[:, 0], k0[:, 2], k0[:, 1], k0[:, 0]))
    ks = np.concatenate([k0, k1], axis=1)[perm]
    new = np.ones(len(perm), dtype=bool)
    new[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    gid = np.cumsum(new) - 1
    firsts = perm[new]
    starts = V[a[firsts]]
    ends = V[b[firsts]]
    src = mesh.tri_to_obj[perm // 3].astype(np.int64)
    # distinct (group, source) in first-appearance order within each group
    key = gid.astype(np.int64) * (int(src.max()) + 2 if len(src) else 1) + (src + 1)
    _, first_idx = np.unique(key, return_index=True)
    first_idx.sort()
    sg = gid[first_idx]
    sobj = src[first_idx].astype(np.int32)
    ne = int(gid[-1]) + 1
    counts = np.bincount(sg, minlength=ne).astype(np.int32)
    offsets = np.zeros(ne, dtype=np.int32)
    offsets[1:] = np.cumsum(counts)[:-1]
    return EdgeSet(np.ascontiguousarray(starts, np.float32), np.ascontiguousarray(ends, np.float32),
                   offsets, counts, np.ascontiguousarray(sobj))


# ---- batched and serving workloads -----------------------------------------------------

def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size=size)).astype(np.int64)


def sphere_field(rng, n_objects: int, density_world=1000.0, density_objects=10_000, level=2):
    """n_objects level-`level` icospheres (radii log-uniform [0.5, 5]) in a cube sized so
    the object density equals that of density_objects spheres in density_world^3."""
    side = density_world * (n_objects / density_objects) ** (1.0 / 3.0)
    centers = rng.uniform(0.0, side, size=(n_objects, 3))
    radii = np.exp(rng.uniform(np.log(0.5), np.log(5.0), size=n_objects))
    return spheres(level, centers, radii), side


def batched_workload(seed: int = 4, n_queries: int = 256, tri_range=(1e3, 5e5),
                     point_range=(1e3, 1e7), overlap_every: int = 4):
    """Independent queries with mesh sizes log-uniform in tri_range triangles (unions of
    level-2 icospheres at the density of the many-spheres workload) and point counts
    log-uniform in point_range. Every `overlap_every`-th query is an overlap query between
    two such meshes (anisotropic, rotated icospheres of levels 2-4); the rest are
    containment queries. Returns a list of dicts with numpy inputs."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for q in range(n_queries):
        tris = int(log_uniform_int(rng, *tri_range))
        if overlap_every and q % overlap_every == overlap_every - 1:
            n_obj = max(2, tris // 2240)
            side = 1000.0 * (n_obj / 50_000) ** (1.0 / 3.0)
            a = overlap_dataset(int(rng.integers(1 << 31)), n_objects=n_obj, world=side)
            b = overlap_dataset(int(rng.integers(1 << 31)), n_objects=n_obj, world=side)
            a.edges, b.edges = extract_edges(a), extract_edges(b)
            out.append({"kind": "overlap", "a": a, "b": b})
        else:
            n_obj = max(1, tris // 320)
            mesh, side = sphere_field(rng, n_obj)
            npts = int(log_uniform_int(rng, *point_range))
            pts = (rng.random(size=(npts, 3), dtype=np.float32) * np.float32(side)).astype(np.float32)
            out.append({"kind": "contain", "mesh": mesh, "points": pts})
    return out


@lru_cache(maxsize=None)
def _template_edges(level: int):
    """Unique edges of the unit icosphere as vertex-index pairs in first-occurrence
    direction (triangle order, (v0,v1),(v1,v2),(v2,v0))."""
    _, f = icosphere(level)
    a = f[:, [0, 1, 2]].reshape(-1)
    b = f[:, [1, 2, 0]].reshape(-1)
    key = np.minimum(a, b) * (len(f) * 3 + 7) + np.maximum(a, b)
    _, first = np.unique(key, return_index=True)
    first.sort()
    return a[first], b[first]


def sphere_field_edges(mesh: Mesh, level) -> EdgeSet:
    """Edge payload of a mesh built by spheres(): the same (start, end, source) records as
    extract_edges (objects share no vertices), grouped by icosphere level and object
    instead of sorted by quantised key; join results do not depend on edge order. `level`
    is one level for every object or the per-object level array."""
    n_obj = mesh.num_objects
    lv = np.broadcast_to(np.asarray(level), (n_obj,))
    nv_of = {int(l): len(icosphere(int(l))[0]) for l in np.unique(lv)}
    nv = np.array([nv_of[int(l)] for l in lv], dtype=np.int64) if len(nv_of) > 1 else \
        np.full(n_obj, next(iter(nv_of.values())) if nv_of else 0, dtype=np.int64)
    voff = np.concatenate([[0], np.cumsum(nv)[:-1]]) if n_obj else np.zeros(0, np.int64)
    s_parts, e_parts, src_parts = [], [], []
    for l in sorted(nv_of):
        sel = np.nonzero(lv == l)[0]
        e0, e1 = _template_edges(l)
        base = voff[sel][:, None]
        s_parts.append(mesh.vertices[(base + e0[None, :]).reshape(-1)])
        e_parts.append(mesh.vertices[(base + e1[None, :]).reshape(-1)])
        src_parts.append(np.repeat(sel.astype(np.int32), len(e0)))
    if not s_parts:
        z3 = np.zeros((0, 3), np.float32)
        zi = np.zeros(0, np.int32)
        return EdgeSet(z3, z3.copy(), zi, zi.copy(), zi.copy())
    starts = np.ascontiguousarray(np.concatenate(s_parts))
    ends = np.ascontiguousarray(np.concatenate(e_parts))
    src = np.concatenate(src_parts)
    ne = len(src)
    return EdgeSet(starts, ends, np.arange(ne, dtype=np.int32), np.ones(ne, dtype=np.int32), src)
