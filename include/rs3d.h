/*
 * rs3d.h -- C ABI of the MI355X-native ray-cast spatial join (librs3d.so).
 *
 * This is the drop-in boundary under RaySpace3D's host classes. Every entry point takes
 * plain pointers and sizes. Unless a comment says "host", pointer arguments are DEVICE
 * pointers (HBM) and the call is asynchronous on the context's stream; calls that
 * return counts to the host synchronise that stream. Every call returns 0 on success or
 * a negative RS3D_E* code; rs3d_last_error() describes the last failure of the calling
 * thread. Reference interfaces each entry replaces are cited per declaration
 * (paths relative to the reference repository root).
 */
#ifndef RS3D_H
#define RS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RS3D_OK = 0,
    RS3D_E_ARG = -1,      /* invalid argument */
    RS3D_E_HIP = -2,      /* HIP runtime error */
    RS3D_E_NOMEM = -3,    /* device allocation failed */
    RS3D_E_CAPACITY = -4, /* an output or table capacity was exceeded (retry larger) */
    RS3D_E_DEPTH = -5,    /* BVH deeper than the traversal stack */
};

/* Wire types, byte-identical to common/include/common.h:14-37. */
typedef struct {
    int32_t ray_id;
    int32_t polygon_index;
} rs3d_ray_result; /* RayResult */

typedef struct {
    int32_t object_id_mesh1;
    int32_t object_id_mesh2;
} rs3d_pair; /* MeshQueryResult */

typedef struct rs3d_ctx rs3d_ctx;
typedef struct rs3d_bvh rs3d_bvh;

const char* rs3d_last_error(void);

/* ---- context -------------------------------------------------------------------------
 * Replaces OptixContext (query/src/optix/OptixContext.cpp:4-8): one device, one stream. */
int rs3d_ctx_create(int device, rs3d_ctx** out);
int rs3d_ctx_destroy(rs3d_ctx* ctx);
int rs3d_ctx_sync(rs3d_ctx* ctx);
/* The hipStream_t the context launches on (as void*), for event timing by callers. */
void* rs3d_ctx_stream(rs3d_ctx* ctx);
/* Context options. RS3D_OPT_RAY_SORT_MIN: point launches with at least this many rays
 * are first reordered along a 2-D Morton curve of the ray origins (perpendicular to the ray
 * axis) so that the lanes of a wave walk the same BVH nodes; results are unchanged (each
 * ray still writes its own slot). 0 disables; default 65536.
 * RS3D_OPT_RAYS_PER_WAVE: point launches give each 64-lane wave a run of this many
 * consecutive rays and refill a lane from the run as soon as its ray finishes; 0 uses one
 * ray per lane. Results are unchanged. Default 64; 0 or 64..2^20.
 * RS3D_OPT_AXIS_GRID: 1 (default) lets point launches use an axis grid of the BVH's
 * triangles (see rs3d_axis_grid_prepare) where one is accepted for the ray axis; 0 keeps
 * them on the BVH. Results are unchanged either way. Grid launches take the rays in the
 * caller's order (no reordering) with RS3D_OPT_GRID_RAYS_PER_WAVE rays per wave run
 * (default 256; 64..2^20), halved down to 64 for launches that would otherwise have fewer
 * than 16,384 runs. RS3D_OPT_GRID_CELL_PCT: cell edge of grids built afterwards,
 * in percent of the triangles' mean extent (default 35; 10..1000; widened while the grid
 * would hold more than 16 records per triangle; without a caller's value, grids under 2^19
 * records take finer cells, down to 10 %, while they stay that small). Results are unchanged.
 * RS3D_OPT_SEG_OCCUPANCY: 1 (default) lets edge rays first test the target's segment
 * occupancy bits (see rs3d_seg_occupancy_prepare); 0 always walks the BVH. Results are
 * unchanged. RS3D_OPT_SEG_CELL_PCT: their cell edge for bits built afterwards, in percent
 * of the triangles' mean extent (default 200; 25..10000). */
enum { RS3D_OPT_RAY_SORT_MIN = 1, RS3D_OPT_RAYS_PER_WAVE = 2, RS3D_OPT_AXIS_GRID = 3,
       RS3D_OPT_GRID_RAYS_PER_WAVE = 4, RS3D_OPT_GRID_CELL_PCT = 5, RS3D_OPT_SEG_OCCUPANCY = 6,
       RS3D_OPT_SEG_CELL_PCT = 7 };
int rs3d_ctx_set_option(rs3d_ctx* ctx, int option, int64_t value);
int rs3d_ctx_get_option(const rs3d_ctx* ctx, int option, int64_t* value);

/* Kernel timing: when enabled, every point-traversal launch on this context is bracketed
 * by HIP events on the context's stream; profile_read returns the summed device time (ms)
 * and the launch count since the last read (and resets both). Replaces the per-launch
 * addMeasurement timings of query/src/timer.h:9-49 for the traversal kernel. */
int rs3d_ctx_profile(rs3d_ctx* ctx, int enable);
int rs3d_ctx_profile_read(rs3d_ctx* ctx, double* total_ms, uint64_t* launches);

/* ---- BVH (replaces OptixAccelerationStructure::buildInternal,
 *      query/src/optix/OptixAccelerationStructure.cpp:22-68) ------------------------------
 * LBVH over an indexed fp32 triangle mesh: 30-bit Morton codes, LDS-staged radix sort,
 * Karras hierarchy, bottom-up refit. vertices: float3[nV]; indices: uint3[nT];
 * tri_to_obj: int32[nT] (may be NULL: object id = triangle index). */
typedef struct {
    int32_t leaf_size; /* 1..8, 0 = default (1: one triangle per leaf) */
    int32_t reserved[7];
} rs3d_bvh_opts;

typedef struct {
    uint64_t num_triangles;
    uint64_t num_nodes;
    int32_t max_depth;  /* deepest primitive in the binary hierarchy */
    int32_t leaf_size;
    float bounds_min[3];
    float bounds_max[3];
    uint64_t device_bytes;
} rs3d_bvh_info;

int rs3d_bvh_build(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                   const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                   const rs3d_bvh_opts* opts, rs3d_bvh** out);
int rs3d_bvh_free(rs3d_bvh* bvh);
int rs3d_bvh_info_get(const rs3d_bvh* bvh, rs3d_bvh_info* info); /* info: host */
/* Copy of the tree's device arrays (no counterpart in the reference, whose OptiX handle is
 * opaque; for inspection and structural checks): nodes = num_nodes records of 64 bytes
 * (child 0 box x lo/hi, y lo/hi, z lo/hi, child 1 box likewise, as 12 floats, then int32
 * link 0, link 1, count 0, count 1: count > 0 = leaf of that many triangles starting at the
 * link, count 0 with link >= 0 = internal node, link < 0 = empty child); tris =
 * num_triangles records of 48 bytes (3 vertices as float4; v0.w = original triangle index
 * and v1.w = object id, as int32 bits). Either output may be null. Device pointers. */
int rs3d_bvh_export(const rs3d_bvh* bvh, void* nodes, void* tris);

/* ---- point containment (replaces __raygen__rg, query/src/cuda/raytracing.cu:9-78, and
 *      RayLauncher::launch, query/src/raytracing/RayLauncher.cpp:44-58) ------------------
 * One ray per point along +axis (0=+x, 1=+y, 2=+z; the reference uses +z). Odd chain
 * length => inside; the object is that of the first chain hit. Output kinds: */
enum {
    RS3D_OUT_RAY_RESULT = 0, /* rs3d_ray_result[n]: {i, obj or -1} */
    RS3D_OUT_FLAG_U8 = 1,    /* uint8[n]: 1 inside, 0 outside */
    RS3D_OUT_OBJ_I32 = 2,    /* int32[n]: obj or -1 (as uint32: 0xFFFFFFFF = none) */
};
/* points: float3[n]. index_map: optional uint32[n]; when given, ray i uses point
 * points[index_map[i]] and writes its result at out[index_map[i]] (ray_id = index_map[i]). */
int rs3d_trace_points(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                      const uint32_t* index_map, int axis, int out_kind, void* out);

/* Traversal counters of the last call with counters enabled (canonical work figures). */
typedef struct {
    uint64_t rays;
    uint64_t nodes_visited; /* 64-byte node fetches */
    uint64_t tris_tested;   /* triangle record fetches */
    uint64_t hits;          /* candidate hits collected */
    uint64_t restarts;      /* traversals restarted after hit-buffer overflow */
} rs3d_trace_stats;
int rs3d_trace_points_stats(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                            int axis, rs3d_trace_stats* stats /* host */);

/* Axis grid (no counterpart in the reference, whose rays run on RT cores): for rays along
 * +axis, a uniform 2-D grid over the plane perpendicular to the axis lists, per cell, the
 * triangles whose padded box (the BVH's leaf box) overlaps the cell and whose projected
 * footprint, widened by the fp32 triangle test's error bound, meets it, ordered by the
 * triangle's top along the axis. A ray reads its cell and tests the listed triangles that
 * reach its origin -- every triangle the test can report for the ray, so the same hits and
 * results as the BVH walk -- in one lookup instead of a walk down the tree. The cell edge is
 * RS3D_OPT_GRID_CELL_PCT percent of the triangles' mean extent in the plane (default 35),
 * widened while the records would exceed 16 per triangle + 2^20. A grid is built once per
 * (BVH, axis), lazily by the first point launch along that axis or here, and rejected (the
 * BVH is used) when its records would exceed 6 per triangle + 2^20 even at the mean extent,
 * or the lists average more than 64 records per cell, or more than 256 per record (dense
 * spots). Freed with the BVH. */
typedef struct {
    int32_t state;            /* 0 not built yet, 1 built, 2 rejected, 3 refused for lack of
                                 device memory (launches use the BVH; prepare tries again) */
    int32_t cells_u, cells_w; /* cells along the two perpendicular axes */
    uint64_t entries;         /* (triangle, cell) records */
    double expected_list;     /* records per cell: the mean list a uniformly placed ray reads */
    uint64_t device_bytes;
    double build_ms;          /* device time of the build (0 when not built here) */
} rs3d_axis_grid_info;
int rs3d_axis_grid_prepare(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, rs3d_axis_grid_info* info /* host, may be null */);
/* The grid's current state and sizes without building it (info: host). */
int rs3d_axis_grid_query(const rs3d_bvh* bvh, int axis, rs3d_axis_grid_info* info);
/* Work counters of a point launch on the axis grid (every ray reads one occupancy bit;
 * nodes_visited = 8-byte cell ranges read for occupied cells, tris_tested = 64-byte
 * records read); RS3D_E_ARG when no grid is accepted for the axis. */
int rs3d_axis_grid_stats(rs3d_ctx* ctx, rs3d_bvh* bvh, const float* points, uint64_t n, int axis,
                         rs3d_trace_stats* stats);
/* ---- filter-refine (replaces FilterRefinePipeline::execute,
 *      query/src/raytracing/FilterRefine.cpp:81-122) --------------------------------------
 * Filter: +axis ray parity against the 12-triangle box mesh of the dataset AABB
 * (query/src/geometry/BoundingBox.cpp:4-50). Refine: point containment of the stably
 * compacted candidates. out_obj: int32[n] (-1 outside). candidate_count: host. */
int rs3d_filter_refine(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                       int axis, int32_t* out_obj, uint64_t* candidate_count,
                       uint64_t* inside_count);
/* Filter-refine that allocates nothing. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, answers and counts as
 * rs3d_filter_refine, which it shares kernels with. Its scratch is a fixed device array of
 * RS3D_NOALLOC_SCRATCH_BYTES inside librs3d, held by every GPU the library runs on until
 * the process ends; no call allocates or frees device or host memory. An axis grid is used
 * only when rs3d_axis_grid_prepare (or an earlier point launch) has already built it for
 * this BVH and axis; then n is limited only as in rs3d_filter_refine. Otherwise the
 * three-phase BVH form runs with its flags, positions and candidates in the array, and
 * n > RS3D_NOALLOC_MAX_POINTS returns RS3D_E_CAPACITY. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU race and must be serialised by the caller. */
#define RS3D_NOALLOC_MAX_POINTS (1ull << 23)
#define RS3D_NOALLOC_SCRATCH_BYTES 75538688ull
int rs3d_filter_refine_noalloc(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                               int axis, int32_t* out_obj, uint64_t* candidate_count,
                               uint64_t* inside_count);
/* Box-mesh filter flags only: flags[i] = 1 when point i passes. box: host float[6]. */
int rs3d_filter_box(rs3d_ctx* ctx, const float* box_min_max, const float* points, uint64_t n,
                    int axis, uint8_t* flags);

/* ---- scan / compaction (replace query/src/cuda/scan_utils.cu:11-27 and
 *      query/src/cuda/result_compaction.cu:5-61) -------------------------------------- */
/* out[i] = sum(in[0..i)); *total (host) = sum(in). */
int rs3d_scan_i32_i64(rs3d_ctx* ctx, const int32_t* in, int64_t* out, uint64_t n, int64_t* total);
/* Stable compaction of rs3d_ray_result entries with polygon_index != -1. count: host. */
int rs3d_compact_hits(rs3d_ctx* ctx, const rs3d_ray_result* in, uint64_t n, rs3d_ray_result* out,
                      uint64_t* count);
/* Stable compaction of indices i with flags[i] != 0 into idx_out. count: host. */
int rs3d_compact_flags(rs3d_ctx* ctx, const uint8_t* flags, uint64_t n, uint32_t* idx_out,
                       uint64_t* count);

/* ---- sort / unique (replace deduplicate_inplace,
 *      query/src/cuda/mesh_query_deduplication.cu:56-90) ---------------------------------
 * LSD radix sort of u64 keys in place, then unique. unique_count: host. */
int rs3d_sort_unique_u64(rs3d_ctx* ctx, uint64_t* keys, uint64_t n, uint64_t* unique_count);
/* Stable LSD radix sort of u32 keys with u32 values, bits [0, end_bit). In place. */
int rs3d_sort_pairs_u32(rs3d_ctx* ctx, uint32_t* keys, uint32_t* values, uint64_t n, int end_bit);

/* ---- edge rays (replace __raygen__mesh_overlap_edges,
 *      query/src/cuda/mesh_overlap_edges.cu:106-216) ------------------------------------ */
typedef struct {
    const float* starts;         /* float3[num_edges] */
    const float* ends;           /* float3[num_edges] */
    const int32_t* src_offsets;  /* int32[num_edges] */
    const int32_t* src_counts;   /* int32[num_edges] */
    const int32_t* src_objects;  /* int32[sum(src_counts)] */
    uint64_t num_edges;
} rs3d_edges;

/* Pass 1: counts[e] = chain hits * source objects of edge e. */
int rs3d_edge_count(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int32_t* counts);
/* Work counters of the edge rays of `edges` against `target` (one counting pass of the
 * A.2 chain, as rs3d_edge_count): pair_emits = sum over edges of chain hits x sources
 * (the hash inserts of a hash-mode join), source_refs = sum of the edges' source counts.
 * No reference counterpart (measurement of the algorithmic bytes). */
typedef struct {
    uint64_t rays, nodes_visited, tris_tested, hits, restarts, pair_emits, source_refs;
    uint64_t rays_walked;    /* rays that passed the occupancy test (all, without bits) */
    uint64_t occ_words;      /* occupancy words read */
} rs3d_edge_stats_t;
int rs3d_edge_stats(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges,
                    rs3d_edge_stats_t* stats /* host */);
/* Pass 2: writes raw pairs at offsets[e] (exclusive scan of counts). swap: emit (tgt, src). */
int rs3d_edge_write(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges,
                    const int64_t* offsets, int swap, rs3d_pair* out);

/* Segment occupancy bits (no counterpart in the reference, whose rays run on RT cores): one
 * bit per cube cell of a uniform 3-D grid over the target BVH's padded bounds, set where a
 * triangle, widened by its box padding, meets the cell. An edge ray whose segment box
 * (widened by 2^-14 of its coordinates and 2e-6) covers no set bit cannot hit the target
 * and skips the walk; results are unchanged. Built once per BVH by the first edge launch
 * against it, or here; freed with the BVH. state: 1 built, 2 rejected (a triangle box over
 * more than 4096 cells, or non-finite bounds: edge rays walk the tree), 3 refused for lack
 * of device memory (edge rays walk the tree; this call tries again). */
typedef struct {
    int state;
    int cells_x, cells_y, cells_z;
    double cell_edge;
    uint64_t occupied_cells;
    uint64_t device_bytes;
    double build_ms;
} rs3d_seg_occupancy_info;
int rs3d_seg_occupancy_prepare(rs3d_ctx* ctx, rs3d_bvh* bvh, rs3d_seg_occupancy_info* info);

/* Spatial order of an edge payload (no counterpart in the reference; for callers that keep
 * edges resident): writes the edge records (start, end, source offset, source count)
 * permuted along a 30-bit Morton curve of the segment midpoints, stable for equal codes;
 * src_objects is shared unchanged, so the reordered view = the outputs + edges->src_objects.
 * Edge rays traced in this order walk neighbouring nodes in neighbouring lanes (the
 * lexicographic .pre order spreads a wave over a thin slab); every join result is a sorted
 * set and does not change. Outputs: device arrays of num_edges records, not aliasing the
 * inputs. Synchronous. */
int rs3d_edges_spatial_order(rs3d_ctx* ctx, const rs3d_edges* edges, float* starts_out, float* ends_out,
                             int32_t* src_offsets_out, int32_t* src_counts_out);

/* Open-addressing pair hash table of u64 keys (a << 32 | b), empty = ~0. */
typedef struct {
    uint64_t* slots;
    uint64_t size;          /* slot count */
    int32_t bitwise;        /* 1: size is a power of two, index = hash & (size-1) */
    int32_t max_probes;     /* 0 = unlimited (bounded by size) */
    uint64_t* failed;       /* device u64 counter of inserts that exhausted max_probes */
} rs3d_hash_table;
int rs3d_ht_clear(rs3d_ctx* ctx, rs3d_hash_table* ht);
/* Hash mode: inserts every emitted pair into ht (on-the-fly dedup). */
int rs3d_edge_hash(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int swap,
                   rs3d_hash_table* ht);
/* Stable scan-based compaction of the table into keys_out (slot order). count: host. */
int rs3d_ht_compact(rs3d_ctx* ctx, const rs3d_hash_table* ht, uint64_t* keys_out, uint64_t max_out,
                    uint64_t* count);

/* Full overlap join A->B and B->A (edge rays of each mesh against the other's BVH),
 * result = sorted unique keys (a << 32 | b) of (mesh1 object, mesh2 object).
 * mode 0: two-pass count/scan/write + radix sort/unique; mode 1: hash + compact + sort.
 * keys_out capacity max_out; count: host. Returns RS3D_E_CAPACITY if it does not fit
 * (count then holds the required size). */
int rs3d_overlap_join(rs3d_ctx* ctx, const rs3d_bvh* bvh1, const rs3d_edges* edges1,
                      const rs3d_bvh* bvh2, const rs3d_edges* edges2, int mode,
                      uint64_t* keys_out, uint64_t max_out, uint64_t* count);

/* Edge-ray hash insert with an explicit chain rule and iteration cap:
 * RS3D_CHAIN_EPS (A->B overlap, query/src/cuda/mesh_overlap_edges.cu:128-178 and the
 * intersection module's trace_edge_multi_hits, query/src/cuda/mesh_intersection.cu:16-91);
 * RS3D_CHAIN_NEXTAFTER (containment join edge check, query/src/cuda/mesh_containment.cu:76-172). */
enum { RS3D_CHAIN_EPS = 0, RS3D_CHAIN_NEXTAFTER = 1 };
int rs3d_edge_hash_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int swap,
                      int chain_rule, int max_iterations, rs3d_hash_table* ht);

/* A mesh as the device arrays of the .pre payload (needed where rays start on the
 * source mesh itself). num_objects: dense object ids 0..num_objects-1. */
typedef struct {
    const float* vertices;     /* float3[num_vertices] */
    uint64_t num_vertices;
    const uint32_t* indices;   /* uint3[num_triangles] */
    uint64_t num_triangles;
    const int32_t* tri_to_obj; /* int32[num_triangles] */
    uint64_t num_objects;
} rs3d_mesh_view;

/* Containment fallback of the intersection join (replaces __raygen__mesh_containment_anyhit,
 * query/src/cuda/mesh_intersection_anyhit.cu:46-186): one +z ray per source object from
 * the first vertex of its first triangle; every target object crossed an odd number of
 * times yields (src, tgt), or (tgt, src) when swap, inserted into ht. max_targets: distinct
 * targets tracked per ray (reference: 256); *overflow (host) counts targets beyond it. */
int rs3d_containment_fallback(rs3d_ctx* ctx, const rs3d_mesh_view* src, const rs3d_bvh* target,
                              int swap, int max_targets, rs3d_hash_table* ht, uint64_t* overflow);

/* Point-in-mesh phase of the containment join (replaces __raygen__point_in_mesh,
 * query/src/cuda/mesh_containment.cu:258-377): for each B object, A objects containing its
 * first vertex give (a, b), kept when (a, b) is not in overlap_ht. Reference max_targets: 512. */
int rs3d_point_in_mesh_join(rs3d_ctx* ctx, const rs3d_mesh_view* src_b, const rs3d_bvh* target_a,
                            int max_targets, const rs3d_hash_table* overlap_ht,
                            rs3d_hash_table* out_ht, uint64_t* overflow);

/* Work counters of the intersection join, field for field the reference's
 * MeshIntersectionProfilingStats (query/src/cuda/mesh_intersection.h:7-22), accumulated on
 * the device (clear it before the first launch). */
typedef struct {
    uint64_t overlap_trace_calls;
    uint64_t overlap_iterations_total;
    uint64_t overlap_hits_total;
    uint32_t overlap_max_iterations_per_trace;
    uint64_t containment_rays_total;
    uint64_t containment_iterations_total;
    uint64_t containment_hits_total;
    uint32_t containment_max_iterations_per_ray;
    uint64_t containment_same_hit_suppressed;
    uint64_t containment_candidate_additions;
    uint64_t containment_candidate_toggles;
    uint64_t containment_candidate_overflow;
    uint64_t containment_targets_total;
} rs3d_intersection_stats;

/* Object rays with the reference's launch options (the containment fields of
 * MeshIntersectionLaunchParams, query/src/cuda/mesh_intersection.h:54-88, and of
 * MeshContainmentLaunchParams, query/src/cuda/mesh_containment.h:31-47).
 * chain_mode RS3D_OBJ_ALL_HITS: every hit in (t, triangle) order over [1e-4, 1e10], a hit of
 *   the same object within 1e-5 of the previous hit merged (the reference's default any-hit
 *   forms, mesh_intersection_anyhit.cu:46-240 and mesh_containment.cu:196-256);
 * RS3D_OBJ_CHAIN_NEXTAFTER: one closest hit per iteration, next tmin = nextafterf(t, 1e10),
 *   at most max_iterations (0: 2048) (use_anyhit_containment = 0, mesh_intersection.cu:134-230);
 * RS3D_OBJ_CHAIN_STEP: one closest hit per iteration over [tmin, 1e10), next tmin =
 *   fl(t + 1e-4), at most max_iterations (0: 2000) (use_anyhit_point_in_mesh = 0,
 *   mesh_containment.cu:336-366).
 * origins: optional device float3 per source object (num_sources of them) instead of the
 * first vertex of each object's first triangle (MeshContainmentLaunchParams::b_first_vertices).
 * Optional per-source device outputs: pair_target_ids / pair_target_hits
 * [num_sources * max_pair_targets] (odd targets in first-hit order and their hit counts, -1 /
 * 0 beyond; PairHitTracking.h:28-110), iterations / candidate_count / candidate_overflow
 * [num_sources] (ContainmentTracking.h:20-95); stats accumulates the containment counters.
 * Two-pass output of the fallback form (instead of a hash table): counts (pass 1, odd
 * targets per source, at most 256), offsets + results (pass 2). */
enum { RS3D_OBJ_ALL_HITS = 0, RS3D_OBJ_CHAIN_NEXTAFTER = 1, RS3D_OBJ_CHAIN_STEP = 2 };
enum { RS3D_OBJ_FALLBACK = 0, RS3D_OBJ_POINT_IN_MESH = 1 };
typedef struct {
    int32_t chain_mode;
    int32_t max_iterations;
    const float* origins;
    uint64_t num_sources;
    int32_t* pair_target_ids;
    uint32_t* pair_target_hits;
    int32_t max_pair_targets;
    uint32_t* iterations;
    uint32_t* candidate_count;
    uint32_t* candidate_overflow;
    rs3d_intersection_stats* stats;
    int32_t* counts;
    const int64_t* offsets;
    rs3d_pair* results;
} rs3d_object_ray_opts;
/* form RS3D_OBJ_FALLBACK: (src, tgt), or (tgt, src) when swap, for every odd target (the
 * first 256), into ht or the two-pass outputs; RS3D_OBJ_POINT_IN_MESH: (tgt, src) for every
 * odd target whose pair is not in overlap_ht, into ht. src may be null when opts->origins is
 * given. *overflow (host) counts targets beyond max_targets. */
int rs3d_object_rays(rs3d_ctx* ctx, const rs3d_mesh_view* src, const rs3d_bvh* target, int form,
                     int swap, int max_targets, const rs3d_object_ray_opts* opts, rs3d_hash_table* ht,
                     const rs3d_hash_table* overlap_ht, uint64_t* overflow);

/* Overlap counters of edge rays (the overlap fields of rs3d_intersection_stats, as
 * trace_edge_multi_hits counts them per source object, query/src/cuda/mesh_intersection.cu:
 * 16-91): one trace call per source object of every edge of length >= 1e-6, min(hits + 1,
 * max_iterations) iterations each. Accumulates into stats (device). */
int rs3d_edge_profile(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                      int max_iterations, rs3d_intersection_stats* stats);
/* Counting and writing passes of edge rays with an explicit chain rule and cap (the two-pass
 * form of the intersection module's overlap launches). */
int rs3d_edge_count_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                       int max_iterations, int32_t* counts);
int rs3d_edge_write_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                       int max_iterations, const int64_t* offsets, int swap, rs3d_pair* out);

/* Whole intersection join (raytracer_intersection_estimated / raytracer_mesh_intersection,
 * query/src/applications/raytracer_intersection_estimated_main.cpp:77-176): overlap edge
 * rays 1->2 and 2->1 (cap overlap_max_iterations, reference default 100) plus the
 * containment fallback both ways, deduplicated in one table (table_slots 0 = automatic;
 * grown and rerun whenever an insert fails), sorted unique keys (a << 32 | b). */
int rs3d_intersection_join(rs3d_ctx* ctx, const rs3d_mesh_view* m1, const rs3d_bvh* b1,
                           const rs3d_edges* e1, const rs3d_mesh_view* m2, const rs3d_bvh* b2,
                           const rs3d_edges* e2, int overlap_max_iterations, uint64_t table_slots,
                           uint64_t* keys_out, uint64_t max_out, uint64_t* count,
                           uint64_t* overflow);

/* Whole containment join (raytracer_containment, query/src/applications/
 * raytracer_containment_main.cpp:247-357): overlap set from B->A and A->B edge rays
 * (nextafter chain, cap 256), then strict containment pairs (a, b) from B's first vertices.
 * contain_out / overlap_out: sorted unique keys (a << 32 | b); overlap_out may be NULL. */
int rs3d_containment_join(rs3d_ctx* ctx, const rs3d_mesh_view* ma, const rs3d_bvh* ba,
                          const rs3d_edges* ea, const rs3d_mesh_view* mb, const rs3d_bvh* bb,
                          const rs3d_edges* eb, uint64_t table_slots, uint64_t* contain_out,
                          uint64_t max_contain, uint64_t* n_contain, uint64_t* overlap_out,
                          uint64_t max_overlap, uint64_t* n_overlap, uint64_t* overflow);

/* ---- batched form: many independent queries in one persistent launch --------------
 * No reference counterpart (the reference runs one query per launch); used for the
 * batched workload. Containment queries write out_obj like rs3d_trace_points; overlap
 * queries produce the same sorted unique pair keys as rs3d_overlap_join (both edge
 * directions). Object ids of overlap queries must be < 2^24; at most 1023 containment
 * and 32767 queries in total per batch. */
enum { RS3D_BATCH_CONTAIN = 0, RS3D_BATCH_OVERLAP = 1 };
typedef struct {
    int32_t kind;
    int32_t axis;                /* containment ray axis (0,1,2) */
    const rs3d_bvh* bvh1;        /* containment: target; overlap: mesh1 */
    const float* points;         /* containment: float3[num_points] */
    uint64_t num_points;
    int32_t* out_obj;            /* containment: int32[num_points] */
    const rs3d_edges* edges1;    /* overlap: mesh1 edges, traced against bvh2 */
    const rs3d_bvh* bvh2;        /* overlap: mesh2 */
    const rs3d_edges* edges2;    /* overlap: mesh2 edges, traced against bvh1 (swapped) */
} rs3d_batch_query;
/* pair_keys_out: all overlap pairs, query by query; pair_offsets (host, num_queries + 1):
 * query q owns [pair_offsets[q], pair_offsets[q+1]). */
int rs3d_batch_execute(rs3d_ctx* ctx, const rs3d_batch_query* queries, int num_queries,
                       uint64_t* pair_keys_out, uint64_t max_pairs, uint64_t* pair_offsets,
                       uint64_t* total_pairs);

/* Captured batch (no reference counterpart; the batched form of the hot path).
 * rs3d_batch_create plans the queries once, executes them once (which sizes every
 * buffer; that execution's results are discarded) and captures the device work of an
 * execution -- ray reorder keys and sort, pair-table clear, queue reset and the persistent
 * traversal launch -- into a hipGraph on the context's stream. rs3d_batch_run replays the
 * graph and then compacts, sorts and splits the pairs exactly as rs3d_batch_execute does
 * (same results). Each run reads the current contents of the queries' inputs (points,
 * meshes, edges, BVHs), so callers may rewrite them in place between runs, keeping their
 * sizes and addresses; containment ids land in the queries' out_obj arrays. If a run finds
 * more pairs than the captured table holds it reruns uncaptured and captures again. The
 * batch keeps its own scratch memory; it must be destroyed before the context. */
typedef struct rs3d_batch rs3d_batch;
int rs3d_batch_create(rs3d_ctx* ctx, const rs3d_batch_query* queries, int num_queries,
                      rs3d_batch** out);
int rs3d_batch_run(rs3d_batch* batch, uint64_t* pair_keys_out, uint64_t max_pairs,
                   uint64_t* pair_offsets, uint64_t* total_pairs);
int rs3d_batch_destroy(rs3d_batch* batch);

/* ---- serving: resident datasets, request queue, continuous batching ----------------
 * No reference counterpart (the reference answers one query per process run). Datasets
 * (device mesh arrays kept alive by the caller, their BVH, optional edges, optional grid
 * histogram of the given resolution for the estimator) stay resident. Requests queue FIFO;
 * each rs3d_server_step admits queued requests while slots are free and their estimated
 * memory fits the budget, runs them all in one persistent launch, and completes them.
 * memory_budget (0: 8 GiB) bounds the device memory the server keeps resident (per dataset
 * its first-triangle list and grid histogram, and the axis grids and segment occupancy
 * bits of the datasets' BVHs) plus the per-step memory of admitted requests; the caller's
 * BVHs and mesh arrays are not counted. A grid or occupancy bits are built, before a step,
 * only when the upper bound of their size fits the budget; otherwise the request runs on
 * the BVH walk with identical results. peak_memory (rs3d_server_stats) includes the
 * resident bytes. Overlap / intersection pairs are (query-dataset object, resident-dataset
 * object), identical to rs3d_overlap_join / rs3d_intersection_join of the two datasets.
 * Calls on one server are serial (single host thread). */
typedef struct rs3d_server rs3d_server;
enum { RS3D_REQ_CONTAIN = 0, RS3D_REQ_OVERLAP = 1, RS3D_REQ_INTERSECT = 2 };
typedef struct {
    int32_t kind;
    int32_t dataset;        /* resident dataset id */
    int32_t query_dataset;  /* overlap / intersection: the request's own dataset id */
    int32_t axis;           /* containment ray axis */
    const float* points;    /* containment: float3[num_points] (device) */
    uint64_t num_points;
    int32_t* out_obj;       /* containment: int32[num_points] (device) */
} rs3d_request;
int rs3d_server_create(rs3d_ctx* ctx, int max_slots, uint64_t memory_budget, rs3d_server** out);
int rs3d_server_destroy(rs3d_server* server);
int rs3d_server_add_dataset(rs3d_server* server, const rs3d_mesh_view* mesh, const rs3d_bvh* bvh,
                            const rs3d_edges* edges, int grid_resolution, int* id);
int rs3d_server_submit(rs3d_server* server, const rs3d_request* request, int64_t* ticket);
int rs3d_server_step(rs3d_server* server, int* admitted, int* completed);
/* done: 0 queued, 1 complete, 2 failed; keys (host): the request's sorted pair keys. */
int rs3d_server_result(rs3d_server* server, int64_t ticket, int* done, uint64_t* keys,
                       uint64_t max_keys, uint64_t* n_keys);
int rs3d_server_stats(rs3d_server* server, uint64_t* queued, uint64_t* peak_memory,
                      uint64_t* steps);
/* The grid-histogram pair estimate that sized an overlap or intersection request's table
 * (rs3d_estimate_pairs with alpha correction over the two datasets' histograms, the target's
 * bounds setting the cell volume; at least 256): -1 until a step has computed it, 0 for a
 * containment request. No reference counterpart (for monitoring and tests). */
int rs3d_server_request_estimate(rs3d_server* server, int64_t ticket, int64_t* estimate);

/* ---- selectivity estimation (replaces estimateOverlapSelectivity,
 *      query/src/cuda/estimated_overlap.cu:43-77) --------------------------------------
 * cells: GridCell[num_cells] (device). Deterministic reduction. result: host. */
int rs3d_estimate_raw(rs3d_ctx* ctx, const void* cells_a, const void* cells_b, uint64_t num_cells,
                      float cell_volume, float epsilon, float gamma, double* result);

/* Grid statistics of a mesh on the GPU (replaces generateGridStats,
 * preprocess/src/preprocess_dataset.cpp:37-179): cells = GridCell[resolution^3] (device),
 * bounds_out (host float[6]) = grid min/max: the vertex bounds, or [0, world_size]^3 when
 * world_size > 0. */
int rs3d_grid_stats(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                    const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                    uint64_t num_objects, int resolution, float world_size, void* cells,
                    float* bounds_out);

/* Unique-edge extraction on the GPU (replaces EdgePreprocessor::extractEdges,
 * preprocess/src/EdgePreprocessor.cpp:52-130, the producer of the .pre edge payload):
 * edges keyed by endpoints quantised with llround(x * 1e5), in lexicographic order of
 * (smaller, larger) quantised endpoint; start/end = endpoints of the edge's first
 * occurrence in triangle order; sources = distinct objects (tri_to_obj; the triangle index
 * when tri_to_obj is null) in order of first appearance. Triangles with an index out of
 * range are skipped. Outputs (device): starts/ends float3[max_edges], src_offsets /
 * src_counts int32[max_edges], src_objects int32[max_sources]. num_edges/num_sources
 * (host) are always set; RS3D_E_CAPACITY when they exceed the capacities (call with zero
 * capacities to size the buffers). Synchronous. */
int rs3d_extract_edges(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                       const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                       float* starts, float* ends, int32_t* src_offsets, int32_t* src_counts,
                       int32_t* src_objects, uint64_t max_edges, uint64_t max_sources,
                       uint64_t* num_edges, uint64_t* num_sources);

/* ---- exact triangle-triangle intersection --------------------------------------------
 * Predicate: Guigue-Devillers ("Fast and Robust Triangle-Triangle Overlap Test Using
 * Orientation Predicates", JGT 8(1), 2003) with the decisions of the reference's TriInt
 * (preprocess/src/tdbase_lib/geometry/TriInt.cpp:400-492), bit-identical in fp32; closed
 * triangles (touching and coplanar overlap count).
 * rs3d_tri_intersect replaces TriInt / TriInt_batch_gpu (tdbase_lib/include/geometry.h:164-166):
 * a, b = device float[9 n] (three vertices per triangle), out = device uint8[n]. Async. */
int rs3d_tri_intersect(rs3d_ctx* ctx, const float* a, const float* b, uint64_t n, uint8_t* out);
/* Exact intersection join: pairs (object of bvh_a, object of bvh_b) having a triangle pair
 * that intersects, the exact counterpart of the edge-ray overlap join (edge rays miss
 * touching and coplanar contacts). Every triangle of bvh_a queries bvh_b with its bounding
 * box. _hash: inserts into ht ((b, a) when swap), async. rs3d_tri_join: sorted unique keys
 * (a << 32 | b); RS3D_E_CAPACITY when they exceed max_out (count = required size). */
int rs3d_tri_join_hash(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b, int swap,
                       rs3d_hash_table* ht);
int rs3d_tri_join(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b, uint64_t* keys_out,
                  uint64_t max_out, uint64_t* count);
/* Work counters of one tri join (for the bytes-per-query roofline). Synchronous. */
typedef struct {
    uint64_t queries, nodes_visited, tris_loaded, predicate_tests, hits;
} rs3d_tri_join_stats_t;
int rs3d_tri_join_stats(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b,
                        rs3d_tri_join_stats_t* stats /* host */);

/* Estimated overlap pairs from two grids (host sizing of
 * query/src/applications/raytracer_overlap_estimated_main.cpp:133-195): raw estimate / alpha
 * (alpha_correction = 0: raw). grid_min/max/resolution: host arrays of grid A. */
int rs3d_estimate_pairs(rs3d_ctx* ctx, const void* cells_a, const void* cells_b, uint64_t num_cells,
                        const float* grid_min, const float* grid_max, const uint32_t* resolution,
                        float epsilon, float gamma, int alpha_correction, int64_t* estimate);

/* ---- reference-named entry points (same names and signatures as the reference's
 *      C ABI; default stream semantics: synchronous) --------------------------------- */
/* query/src/cuda/result_compaction.h:7-16 */
int count_hits_gpu(const rs3d_ray_result* d_results, int num_results);
void compact_hits_gpu(const rs3d_ray_result* d_input, rs3d_ray_result* d_output, int num_results,
                      int* h_num_hits);
/* query/src/cuda/scan_utils.h:7 */
long long exclusive_scan_gpu(const int* d_input, long long* d_output, int num_elements);
/* query/src/cuda/mesh_query_deduplication.h:6-15 */
long long merge_and_deduplicate_pairs_gpu(const rs3d_pair* d_results1, long long num_results1,
                                          const rs3d_pair* d_results2, long long num_results2,
                                          rs3d_pair* d_merged_output);
int compact_hash_table_pairs(const unsigned long long* d_hash_table, unsigned long long table_size,
                             rs3d_pair* d_output, int max_output_size);

#ifdef __cplusplus
}
#endif
#endif
