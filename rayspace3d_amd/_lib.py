"""ctypes binding of librs3d.so (the C ABI declared in include/rs3d.h).

The library is built in-tree by ``__graft_entry__.build()`` (``make`` in
``rayspace3d_amd/csrc``) into ``rayspace3d_amd/lib/librs3d.so``. There is no fallback:
if the library is missing or fails to load, ``load()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "librs3d.so")

RS3D_OK = 0
RS3D_E_ARG = -1
RS3D_E_HIP = -2
RS3D_E_NOMEM = -3
RS3D_E_CAPACITY = -4
RS3D_E_DEPTH = -5
NOALLOC_MAX_POINTS = 1 << 23

OUT_RAY_RESULT = 0
OUT_FLAG_U8 = 1
OUT_OBJ_I32 = 2
OPT_RAY_SORT_MIN = 1
OPT_RAYS_PER_WAVE = 2
OPT_AXIS_GRID = 3
OPT_GRID_RAYS_PER_WAVE = 4
OPT_GRID_CELL_PCT = 5
OPT_SEG_OCCUPANCY = 6
OPT_SEG_CELL_PCT = 7
REQ_CONTAIN = 0
REQ_OVERLAP = 1
REQ_INTERSECT = 2


class BvhOpts(C.Structure):
    _fields_ = [("leaf_size", C.c_int32), ("reserved", C.c_int32 * 7)]


class BvhInfo(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_uint64),
        ("num_nodes", C.c_uint64),
        ("max_depth", C.c_int32),
        ("leaf_size", C.c_int32),
        ("bounds_min", C.c_float * 3),
        ("bounds_max", C.c_float * 3),
        ("device_bytes", C.c_uint64),
    ]


class AxisGridInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("cells_u", C.c_int32), ("cells_w", C.c_int32), ("entries", C.c_uint64),
                ("expected_list", C.c_double), ("device_bytes", C.c_uint64), ("build_ms", C.c_double)]


class TraceStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("nodes_visited", C.c_uint64),
        ("tris_tested", C.c_uint64),
        ("hits", C.c_uint64),
        ("restarts", C.c_uint64),
    ]


class EdgeStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("rays", "nodes_visited", "tris_tested", "hits",
                                           "restarts", "pair_emits", "source_refs", "rays_walked",
                                           "occ_words")]


class SegOccupancyInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("cells_x", C.c_int32), ("cells_y", C.c_int32), ("cells_z", C.c_int32),
                ("cell_edge", C.c_double), ("occupied_cells", C.c_uint64), ("device_bytes", C.c_uint64),
                ("build_ms", C.c_double)]


class TriJoinStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("queries", "nodes_visited", "tris_loaded", "predicate_tests",
                                           "hits")]


class Edges(C.Structure):
    _fields_ = [
        ("starts", C.c_void_p),
        ("ends", C.c_void_p),
        ("src_offsets", C.c_void_p),
        ("src_counts", C.c_void_p),
        ("src_objects", C.c_void_p),
        ("num_edges", C.c_uint64),
    ]


class MeshView(C.Structure):
    _fields_ = [
        ("vertices", C.c_void_p),
        ("num_vertices", C.c_uint64),
        ("indices", C.c_void_p),
        ("num_triangles", C.c_uint64),
        ("tri_to_obj", C.c_void_p),
        ("num_objects", C.c_uint64),
    ]


class HashTable(C.Structure):
    _fields_ = [
        ("slots", C.c_void_p),
        ("size", C.c_uint64),
        ("bitwise", C.c_int32),
        ("max_probes", C.c_int32),
        ("failed", C.c_void_p),
    ]


class BatchQuery(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("axis", C.c_int32),
        ("bvh1", C.c_void_p),
        ("points", C.c_void_p),
        ("num_points", C.c_uint64),
        ("out_obj", C.c_void_p),
        ("edges1", C.c_void_p),
        ("bvh2", C.c_void_p),
        ("edges2", C.c_void_p),
    ]


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("dataset", C.c_int32),
        ("query_dataset", C.c_int32),
        ("axis", C.c_int32),
        ("points", C.c_void_p),
        ("num_points", C.c_uint64),
        ("out_obj", C.c_void_p),
    ]


P = C.c_void_p
U64 = C.c_uint64
I32 = C.c_int
PU64 = C.POINTER(C.c_uint64)
PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol declared in include/rs3d.h
class IntersectionStats(C.Structure):
    _fields_ = [("overlap_trace_calls", C.c_uint64), ("overlap_iterations_total", C.c_uint64),
                ("overlap_hits_total", C.c_uint64), ("overlap_max_iterations_per_trace", C.c_uint32),
                ("containment_rays_total", C.c_uint64), ("containment_iterations_total", C.c_uint64),
                ("containment_hits_total", C.c_uint64), ("containment_max_iterations_per_ray", C.c_uint32),
                ("containment_same_hit_suppressed", C.c_uint64),
                ("containment_candidate_additions", C.c_uint64),
                ("containment_candidate_toggles", C.c_uint64),
                ("containment_candidate_overflow", C.c_uint64), ("containment_targets_total", C.c_uint64)]


class ObjectRayOpts(C.Structure):
    _fields_ = [("chain_mode", C.c_int32), ("max_iterations", C.c_int32), ("origins", C.c_void_p),
                ("num_sources", C.c_uint64), ("pair_target_ids", C.c_void_p), ("pair_target_hits", C.c_void_p),
                ("max_pair_targets", C.c_int32), ("iterations", C.c_void_p), ("candidate_count", C.c_void_p),
                ("candidate_overflow", C.c_void_p), ("stats", C.c_void_p), ("counts", C.c_void_p),
                ("offsets", C.c_void_p), ("results", C.c_void_p)]


OBJ_ALL_HITS = 0
OBJ_CHAIN_NEXTAFTER = 1
OBJ_CHAIN_STEP = 2
OBJ_FALLBACK = 0
OBJ_POINT_IN_MESH = 1


SIGNATURES = {
    "rs3d_last_error": (C.c_char_p, []),
    "rs3d_ctx_create": (I32, [I32, C.POINTER(P)]),
    "rs3d_ctx_destroy": (I32, [P]),
    "rs3d_ctx_sync": (I32, [P]),
    "rs3d_ctx_stream": (P, [P]),
    "rs3d_ctx_set_option": (I32, [P, I32, C.c_int64]),
    "rs3d_ctx_get_option": (I32, [P, I32, C.POINTER(C.c_int64)]),
    "rs3d_ctx_profile": (I32, [P, I32]),
    "rs3d_ctx_profile_read": (I32, [P, C.POINTER(C.c_double), PU64]),
    "rs3d_bvh_build": (I32, [P, P, U64, P, U64, P, C.POINTER(BvhOpts), C.POINTER(P)]),
    "rs3d_bvh_free": (I32, [P]),
    "rs3d_bvh_info_get": (I32, [P, C.POINTER(BvhInfo)]),
    "rs3d_trace_points": (I32, [P, P, P, U64, P, I32, I32, P]),
    "rs3d_trace_points_stats": (I32, [P, P, P, U64, I32, C.POINTER(TraceStats)]),
    "rs3d_axis_grid_prepare": (I32, [P, P, I32, C.POINTER(AxisGridInfo)]),
    "rs3d_axis_grid_query": (I32, [P, I32, C.POINTER(AxisGridInfo)]),
    "rs3d_seg_occupancy_prepare": (I32, [P, P, C.POINTER(SegOccupancyInfo)]),
    "rs3d_axis_grid_stats": (I32, [P, P, P, U64, I32, C.POINTER(TraceStats)]),
    "rs3d_edge_stats": (I32, [P, P, C.POINTER(Edges), C.POINTER(EdgeStats)]),
    "rs3d_extract_edges": (I32, [P, P, U64, P, U64, P, P, P, P, P, P, U64, U64, PU64, PU64]),
    "rs3d_tri_intersect": (I32, [P, P, P, U64, P]),
    "rs3d_bvh_export": (I32, [P, P, P]),
    "rs3d_edges_spatial_order": (I32, [P, C.POINTER(Edges), P, P, P, P]),
    "rs3d_tri_join_hash": (I32, [P, P, P, I32, C.POINTER(HashTable)]),
    "rs3d_tri_join": (I32, [P, P, P, P, U64, PU64]),
    "rs3d_tri_join_stats": (I32, [P, P, P, C.POINTER(TriJoinStats)]),
    "rs3d_filter_refine": (I32, [P, P, P, U64, I32, P, PU64, PU64]),
    "rs3d_filter_refine_noalloc": (I32, [P, P, P, U64, I32, P, PU64, PU64]),
    "rs3d_filter_box": (I32, [P, C.POINTER(C.c_float), P, U64, I32, P]),
    "rs3d_scan_i32_i64": (I32, [P, P, P, U64, PI64]),
    "rs3d_compact_hits": (I32, [P, P, U64, P, PU64]),
    "rs3d_compact_flags": (I32, [P, P, U64, P, PU64]),
    "rs3d_sort_unique_u64": (I32, [P, P, U64, PU64]),
    "rs3d_sort_pairs_u32": (I32, [P, P, P, U64, I32]),
    "rs3d_edge_count": (I32, [P, P, C.POINTER(Edges), P]),
    "rs3d_edge_write": (I32, [P, P, C.POINTER(Edges), P, I32, P]),
    "rs3d_ht_clear": (I32, [P, C.POINTER(HashTable)]),
    "rs3d_edge_hash": (I32, [P, P, C.POINTER(Edges), I32, C.POINTER(HashTable)]),
    "rs3d_ht_compact": (I32, [P, C.POINTER(HashTable), P, U64, PU64]),
    "rs3d_overlap_join": (I32, [P, P, C.POINTER(Edges), P, C.POINTER(Edges), I32, P, U64, PU64]),
    "rs3d_edge_hash_ex": (I32, [P, P, C.POINTER(Edges), I32, I32, I32, C.POINTER(HashTable)]),
    "rs3d_object_rays": (I32, [P, C.POINTER(MeshView), P, I32, I32, I32, C.POINTER(ObjectRayOpts),
                               C.POINTER(HashTable), C.POINTER(HashTable), PU64]),
    "rs3d_edge_profile": (I32, [P, P, C.POINTER(Edges), I32, I32, P]),
    "rs3d_edge_count_ex": (I32, [P, P, C.POINTER(Edges), I32, I32, P]),
    "rs3d_edge_write_ex": (I32, [P, P, C.POINTER(Edges), I32, I32, P, I32, P]),
    "rs3d_containment_fallback": (I32, [P, C.POINTER(MeshView), P, I32, I32, C.POINTER(HashTable), PU64]),
    "rs3d_point_in_mesh_join": (I32, [P, C.POINTER(MeshView), P, I32, C.POINTER(HashTable),
                                      C.POINTER(HashTable), PU64]),
    "rs3d_intersection_join": (I32, [P, C.POINTER(MeshView), P, C.POINTER(Edges), C.POINTER(MeshView),
                                     P, C.POINTER(Edges), I32, U64, P, U64, PU64, PU64]),
    "rs3d_containment_join": (I32, [P, C.POINTER(MeshView), P, C.POINTER(Edges), C.POINTER(MeshView),
                                    P, C.POINTER(Edges), U64, P, U64, PU64, P, U64, PU64, PU64]),
    "rs3d_batch_execute": (I32, [P, C.POINTER(BatchQuery), I32, P, U64, PU64, PU64]),
    "rs3d_batch_create": (I32, [P, C.POINTER(BatchQuery), I32, C.POINTER(P)]),
    "rs3d_batch_run": (I32, [P, P, U64, PU64, PU64]),
    "rs3d_batch_destroy": (I32, [P]),
    "rs3d_server_create": (I32, [P, I32, U64, C.POINTER(P)]),
    "rs3d_server_destroy": (I32, [P]),
    "rs3d_server_add_dataset": (I32, [P, C.POINTER(MeshView), P, C.POINTER(Edges), I32, C.POINTER(C.c_int)]),
    "rs3d_server_submit": (I32, [P, C.POINTER(Request), PI64]),
    "rs3d_server_step": (I32, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "rs3d_server_result": (I32, [P, C.c_int64, C.POINTER(C.c_int), P, U64, PU64]),
    "rs3d_server_stats": (I32, [P, PU64, PU64, PU64]),
    "rs3d_server_request_estimate": (I32, [P, C.c_int64, C.POINTER(C.c_int64)]),
    "rs3d_estimate_raw": (I32, [P, P, P, U64, C.c_float, C.c_float, C.c_float, C.POINTER(C.c_double)]),
    "rs3d_grid_stats": (I32, [P, P, U64, P, U64, P, U64, I32, C.c_float, P, C.POINTER(C.c_float)]),
    "rs3d_estimate_pairs": (I32, [P, P, P, U64, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(C.c_uint32), C.c_float, C.c_float, I32, PI64]),
    "count_hits_gpu": (I32, [P, I32]),
    "compact_hits_gpu": (None, [P, P, I32, C.POINTER(C.c_int)]),
    "exclusive_scan_gpu": (C.c_longlong, [P, P, I32]),
    "merge_and_deduplicate_pairs_gpu": (C.c_longlong, [P, C.c_longlong, P, C.c_longlong, P]),
    "compact_hash_table_pairs": (I32, [P, C.c_ulonglong, P, I32]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads librs3d.so and binds every declared entry point. Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"librs3d.so not found at {path}: build it with __graft_entry__.build() "
            "(the HIP library has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class Rs3dError(RuntimeError):
    def __init__(self, code: int, where: str):
        msg = load().rs3d_last_error()
        super().__init__(f"{where} failed with code {code}: {msg.decode() if msg else ''}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != RS3D_OK:
        raise Rs3dError(code, where)
