"""ctypes wrapper of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker. The product never imports it."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_P = C.c_void_p
_I64 = C.c_int64
_sig = {
    "orc_ray_pre": (None, [_P, _P]),
    "orc_tri_hit": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "orc_points_contain": (_I64, [_P, _I64, _P, _I64, _P, _P, _I64, C.c_int, C.c_int, _P]),
    "orc_points_contain_mt": (_I64, [_P, _I64, _P, _I64, _P, _P, _I64, C.c_int, C.c_int, _P, C.c_int]),
    "orc_edge_pairs": (_I64, [_P, _P, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, C.c_int, C.c_int,
                              _P, _I64, _P]),
    "orc_edge_pairs_ex": (_I64, [_P, _P, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, C.c_int, C.c_int,
                                 C.c_int, C.c_int, _P, _I64, _P]),
    "orc_object_parity": (_I64, [_P, _P, _P, _I64, _I64, _P, _I64, _P, _I64, _P, C.c_int, C.c_int,
                                 _P, _I64, _P]),
    "orc_object_parity_ex": (_I64, [_P, _P, _P, _I64, _I64, _P, _I64, _P, _I64, _P, C.c_int, C.c_int,
                                    _P, _P, _I64, _P]),
    "orc_sort_unique_u64": (_I64, [_P, _I64]),
    "orc_fmix64": (C.c_uint64, [C.c_uint64]),
    "orc_ht_insert": (C.c_int, [_P, C.c_uint64, C.c_int, C.c_int, C.c_uint64]),
    "orc_estimate_raw": (C.c_double, [_P, _P, _I64, C.c_float, C.c_float, C.c_float]),
    "orc_estimate_pairs": (_I64, [_P, _P, _I64, _P, _P, _P, C.c_float, C.c_float, C.c_int]),
    "orc_grid_stats": (None, [_P, _I64, _P, _I64, _P, C.c_int, C.c_float, _P, _P]),
    "orc_extract_edges": (_I64, [_P, _I64, _P, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "orc_tri_tri_many": (None, [_P, _P, _I64, _P]),
    "orc_tri_join": (_I64, [_P, _I64, _P, _P, _I64, _P, _I64, _P, _P, _I64, _P, _I64, C.c_int]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        L = C.CDLL(LIB)
        for k, (r, a) in _sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


GRID_CELL = np.dtype([("center", "<u4"), ("touch", "<u4"), ("avg", "<f4"), ("ratio", "<f4")])


def tri_hit(o, d, v0, v1, v2):
    """(hit, t) of the canonical fp32 watertight test."""
    rp = (C.c_byte * 32)()
    dd = _c(d, np.float32)
    lib().orc_ray_pre(_p(dd), C.cast(rp, C.c_void_p))
    t = np.zeros(1, np.float32)
    h = lib().orc_tri_hit(_p(_c(o, np.float32)), C.cast(rp, C.c_void_p), _p(_c(v0, np.float32)),
                          _p(_c(v1, np.float32)), _p(_c(v2, np.float32)), _p(t))
    return bool(h), float(t[0])


def points_contain(V, I, tri_obj, P, axis=2, use_grid=True, threads=1):
    V = _c(V, np.float32); I = _c(I, np.uint32); O = _c(tri_obj, np.int32); P = _c(P, np.float32)
    out = np.empty(len(P), np.int32)
    lib().orc_points_contain_mt(_p(V), len(V), _p(I), len(I), _p(O), _p(P), len(P), axis,
                                1 if use_grid else 0, _p(out), int(threads))
    return out


def edge_pairs(edges, V, I, tri_obj, swap=False, use_grid=True, with_counts=False,
               chain_rule=0, max_iter=100):
    """Raw pair keys (a << 32 | b) in edge order, and optionally per-edge counts."""
    S = _c(edges.starts, np.float32); E = _c(edges.ends, np.float32)
    off = _c(edges.src_offsets, np.int32); cnt = _c(edges.src_counts, np.int32)
    so = _c(edges.src_objects, np.int32)
    V = _c(V, np.float32); I = _c(I, np.uint32); O = _c(tri_obj, np.int32)
    nE = len(S)
    counts = np.empty(nE, np.int32)
    args = (_p(S), _p(E), _p(off), _p(cnt), _p(so), nE, _p(V), len(V), _p(I), len(I), _p(O),
            int(swap), int(use_grid), int(chain_rule), int(max_iter))
    total = lib().orc_edge_pairs_ex(*args, None, 0, _p(counts))
    keys = np.empty(max(total, 1), np.uint64)
    lib().orc_edge_pairs_ex(*args, _p(keys), total, None)
    keys = keys[:total]
    return (keys, counts) if with_counts else keys


def object_parity(src_mesh, mesh, max_targets=256, swap=False):
    """A.3/A.4 per-object +z containment rays: odd (src, tgt) keys (swapped when swap)."""
    sV = _c(src_mesh.vertices, np.float32); sI = _c(src_mesh.indices, np.uint32)
    sO = _c(src_mesh.tri_to_obj, np.int32)
    nobj = int(sO.max()) + 1 if len(sO) else 0
    V = _c(mesh.vertices, np.float32); I = _c(mesh.indices, np.uint32); O = _c(mesh.tri_to_obj, np.int32)
    over = C.c_int64(0)
    args = (_p(sV), _p(sI), _p(sO), len(sI), nobj, _p(V), len(V), _p(I), len(I), _p(O),
            int(max_targets), int(swap))
    total = lib().orc_object_parity(*args, None, 0, C.cast(C.pointer(over), C.c_void_p))
    keys = np.empty(max(total, 1), np.uint64)
    lib().orc_object_parity(*args, _p(keys), total, C.cast(C.pointer(over), C.c_void_p))
    return keys[:total], over.value


class ObjRayOpts(C.Structure):
    _fields_ = [("chain_mode", C.c_int), ("max_iter", C.c_int), ("origins", _P),
                ("iterations", _P), ("cand_count", _P), ("cand_overflow", _P),
                ("pair_ids", _P), ("pair_hits", _P), ("max_pair", C.c_int), ("max_out", C.c_int),
                ("stats", C.c_uint64 * 9)]


STAT_NAMES = ("rays", "iterations", "hits", "max_iterations", "suppressed", "additions", "toggles",
              "overflow", "targets")


def object_parity_ex(src_mesh, mesh, max_targets=256, swap=False, chain_mode=0, max_iter=0,
                     origins=None, n_src=None, track=False, max_pair=256, max_out=0):
    """A.3/A.4 object rays with the reference's launch options (chain_mode 0 all hits, 1
    nextafter closest-hit loop, 2 t + 1e-4 closest-hit loop). Returns (keys, overflow,
    tracking dict or None, stats dict)."""
    sV = _c(src_mesh.vertices, np.float32) if src_mesh is not None else np.zeros((1, 3), np.float32)
    sI = _c(src_mesh.indices, np.uint32) if src_mesh is not None else np.zeros((1, 3), np.uint32)
    sO = _c(src_mesh.tri_to_obj, np.int32) if src_mesh is not None else np.zeros(0, np.int32)
    if n_src is None:
        n_src = len(origins) if origins is not None else (int(sO.max()) + 1 if len(sO) else 0)
    V = _c(mesh.vertices, np.float32); I = _c(mesh.indices, np.uint32); O = _c(mesh.tri_to_obj, np.int32)
    org = _c(origins, np.float32) if origins is not None else None
    tr = None
    if track:
        tr = {"iterations": np.zeros(n_src, np.uint32), "candidate_count": np.zeros(n_src, np.uint32),
              "candidate_overflow": np.zeros(n_src, np.uint32),
              "pair_target_ids": np.zeros((n_src, max_pair), np.int32),
              "pair_target_hits": np.zeros((n_src, max_pair), np.uint32)}

    def opts():
        o = ObjRayOpts()
        o.chain_mode, o.max_iter, o.origins, o.max_out = chain_mode, max_iter, _p(org), max_out
        if tr is not None:
            o.iterations, o.cand_count = _p(tr["iterations"]), _p(tr["candidate_count"])
            o.cand_overflow = _p(tr["candidate_overflow"])
            o.pair_ids, o.pair_hits, o.max_pair = _p(tr["pair_target_ids"]), _p(tr["pair_target_hits"]), max_pair
        return o
    over = C.c_int64(0)
    args = (_p(sV), _p(sI), _p(sO), len(sO), n_src, _p(V), len(V), _p(I), len(I), _p(O),
            int(max_targets), int(swap))
    o1 = opts()
    total = lib().orc_object_parity_ex(*args, C.cast(C.pointer(o1), C.c_void_p), None, 0,
                                       C.cast(C.pointer(over), C.c_void_p))
    keys = np.empty(max(total, 1), np.uint64)
    o2 = opts()
    lib().orc_object_parity_ex(*args, C.cast(C.pointer(o2), C.c_void_p), _p(keys), total,
                               C.cast(C.pointer(over), C.c_void_p))
    stats = {k: int(v) for k, v in zip(STAT_NAMES, o2.stats)}
    return keys[:total], over.value, tr, stats


def intersection_pairs(m1, m2, max_iter=100):
    """Intersection join: overlap edge rays both ways plus the containment fallback both ways."""
    k1 = edge_pairs(m1.edges, m2.vertices, m2.indices, m2.tri_to_obj, False, max_iter=max_iter)
    k2 = edge_pairs(m2.edges, m1.vertices, m1.indices, m1.tri_to_obj, True, max_iter=max_iter)
    c1, _ = object_parity(m1, m2, 256, False)
    c2, _ = object_parity(m2, m1, 256, True)
    return sort_unique(np.concatenate([k1, k2, c1, c2]))


def containment_pairs(ma, mb):
    """Containment join: (overlap keys, strict containment keys), both (a << 32 | b)."""
    o1 = edge_pairs(mb.edges, ma.vertices, ma.indices, ma.tri_to_obj, True, chain_rule=1, max_iter=256)
    o2 = edge_pairs(ma.edges, mb.vertices, mb.indices, mb.tri_to_obj, False, chain_rule=1, max_iter=256)
    ov = sort_unique(np.concatenate([o1, o2]))
    cand, _ = object_parity(mb, ma, 512, True)  # (a << 32 | b)
    cand = sort_unique(cand)
    return ov, cand[~np.isin(cand, ov)]


def sort_unique(keys):
    k = np.array(keys, dtype=np.uint64, copy=True)
    n = lib().orc_sort_unique_u64(_p(k), len(k))
    return k[:n]


def overlap_pairs(mesh1, edges1, mesh2, edges2, use_grid=True):
    """A.2 join: sorted unique keys of edges1 -> mesh2 and edges2 -> mesh1 (swapped)."""
    k1 = edge_pairs(edges1, mesh2.vertices, mesh2.indices, mesh2.tri_to_obj, False, use_grid)
    k2 = edge_pairs(edges2, mesh1.vertices, mesh1.indices, mesh1.tri_to_obj, True, use_grid)
    return sort_unique(np.concatenate([k1, k2]))


def fmix64(k):
    return int(lib().orc_fmix64(C.c_uint64(k)))


def ht_insert(table, bitwise, max_probes, key):
    return bool(lib().orc_ht_insert(_p(table), len(table), int(bitwise), int(max_probes), key))


def estimate_raw(cells_a, cells_b, cell_volume, eps=0.001, gamma=0.8):
    a = np.ascontiguousarray(cells_a, GRID_CELL); b = np.ascontiguousarray(cells_b, GRID_CELL)
    return float(lib().orc_estimate_raw(_p(a), _p(b), len(a), cell_volume, eps, gamma))


def estimate_pairs(cells_a, cells_b, minb, maxb, res, eps=0.001, gamma=0.8, alpha=True):
    a = np.ascontiguousarray(cells_a, GRID_CELL); b = np.ascontiguousarray(cells_b, GRID_CELL)
    return int(lib().orc_estimate_pairs(_p(a), _p(b), len(a), _p(_c(minb, np.float32)),
                                        _p(_c(maxb, np.float32)), _p(_c(res, np.uint32)), eps,
                                        gamma, int(alpha)))


def grid_stats(V, I, tri_obj, res, world_size=-1.0):
    V = _c(V, np.float32); I = _c(I, np.uint32); O = _c(tri_obj, np.int32)
    cells = np.zeros(res ** 3, GRID_CELL)
    bounds = np.zeros(6, np.float32)
    lib().orc_grid_stats(_p(V), len(V), _p(I), len(I), _p(O), res, world_size, _p(cells), _p(bounds))
    return cells, bounds[:3].copy(), bounds[3:].copy()


def extract_edges(V, I, tri_obj):
    """Returns (starts, ends, offsets, counts, sources)."""
    V = _c(V, np.float32); I = _c(I, np.uint32); O = _c(tri_obj, np.int32)
    ns = C.c_int64(0)
    ne = lib().orc_extract_edges(_p(V), len(V), _p(I), len(I), _p(O), None, None, None, None, None,
                                 C.cast(C.pointer(ns), C.c_void_p))
    S = np.empty((ne, 3), np.float32); E = np.empty((ne, 3), np.float32)
    off = np.empty(ne, np.int32); cnt = np.empty(ne, np.int32); src = np.empty(max(ns.value, 1), np.int32)
    lib().orc_extract_edges(_p(V), len(V), _p(I), len(I), _p(O), _p(S), _p(E), _p(off), _p(cnt),
                            _p(src), C.cast(C.pointer(ns), C.c_void_p))
    return S, E, off, cnt, src[:ns.value]


def tri_tri(A, B):
    """Exact triangle-triangle predicate on pairs: A, B float32 [n, 3, 3] -> bool [n]."""
    A = _c(np.asarray(A, np.float32).reshape(-1, 9), np.float32)
    B = _c(np.asarray(B, np.float32).reshape(-1, 9), np.float32)
    out = np.zeros(len(A), np.uint8)
    lib().orc_tri_tri_many(_p(A), _p(B), len(A), _p(out))
    return out.astype(bool)


def tri_join(ma, mb, threads=1):
    """Exact intersection join of two meshes: sorted unique keys (a << 32 | b)."""
    args = []
    for m in (ma, mb):
        V = _c(m.vertices, np.float32); I = _c(m.indices, np.uint32); O = _c(m.tri_to_obj, np.int32)
        args.append((V, I, O))
    (Va, Ia, Oa), (Vb, Ib, Ob) = args
    def call(out, cap):
        return lib().orc_tri_join(_p(Va), len(Va), _p(Ia), _p(Oa), len(Ia), _p(Vb), len(Vb), _p(Ib), _p(Ob),
                                  len(Ib), _p(out), cap, threads)
    n = call(None, 0)
    out = np.empty(max(n, 1), np.uint64)
    call(out, n)
    return out[:n]


def ref_tri_int():
    """The reference's own TriInt compiled from its source (oracle/_ref/libtriint.so), or
    None when it was not built. Returns f(t1, t2) -> bool on 9-float triangles."""
    path = os.path.join(HERE, "_ref", "libtriint.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    f = getattr(L, "_ZN6tdbase6TriIntEPKfS1_")
    f.restype = C.c_bool
    f.argtypes = [_P, _P]

    def call(t1, t2):
        a = _c(np.asarray(t1, np.float32).reshape(9), np.float32)
        b = _c(np.asarray(t2, np.float32).reshape(9), np.float32)
        return bool(f(_p(a), _p(b)))
    return call
