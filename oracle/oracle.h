/*
 * oracle.h -- CPU restatement of RaySpace3D's ray-cast spatial-join semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product path
 * (librs3d.so) is compared against. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it. Nothing in the product links or calls it.
 *
 * Every function follows a reference file (cited in oracle.c). The reference itself
 * (OptiX + CUDA) cannot be built here, so the fp32 triangle test, the closest-hit
 * chain rules and the tie rules are the canonical ones stated in oracle.c and
 * implemented identically, operation for operation, in the HIP kernels.
 *
 * Pinning (tests/test_oracle.py): closed-form triangle cases, an independent
 * double-precision Moller-Trumbore test, analytic sphere containment and sphere-sphere
 * overlap, brute force vs binned equality, the committed check_n_geometry.pre fixture
 * (v1 header, closedness, grid statistics), fmix64 inverse, and hand-computed
 * estimator cells. Parity unpinned: the edge/vertex tie behaviour of OptiX itself
 * (only an OptiX run could pin it).
 */
#ifndef RS3D_ORACLE_H
#define RS3D_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int kx, ky, kz;
    float Sx, Sy, Sz;
} orc_raypre;

void orc_ray_pre(const float d[3], orc_raypre* r);
/* Returns 1 and the fp32 hit distance in *t when the (two-sided) triangle is hit. */
int orc_tri_hit(const float o[3], const orc_raypre* r, const float v0[3], const float v1[3],
                const float v2[3], float* t);

/* A.1 point containment along +axis (0=x, 1=y, 2=z). out_obj[i] = object id of the
 * first chain hit if the chain length is odd, else -1. use_grid=0: brute force over all
 * triangles; 1: 2-D binned candidates (same result). Returns the number inside. */
int64_t orc_points_contain(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                           const int32_t* tri_obj, const float* P, int64_t n, int axis,
                           int use_grid, int32_t* out_obj);
/* Same, over `threads` OpenMP threads (results identical for any thread count). */
int64_t orc_points_contain_mt(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                              const int32_t* tri_obj, const float* P, int64_t n, int axis,
                              int use_grid, int32_t* out_obj, int threads);

/* A.2 edge rays against a target mesh. Writes up to cap raw pair keys
 * ((u64)a << 32 | b, a = mesh1 object) into out_keys in edge order, and the per-edge
 * raw result count (hits * sources) into per_edge_counts when non-null. Returns the
 * total number of raw results (may exceed cap; only cap are written). */
int64_t orc_edge_pairs(const float* starts, const float* ends, const int32_t* src_off,
                       const int32_t* src_cnt, const int32_t* src_obj, int64_t nE,
                       const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                       const int32_t* tri_obj, int swap, int use_grid, uint64_t* out_keys,
                       int64_t cap, int32_t* per_edge_counts);

/* As orc_edge_pairs with an explicit chain rule (0: t + 1e-6 / 2e-4, 1: nextafterf) and
 * iteration cap. */
int64_t orc_edge_pairs_ex(const float* starts, const float* ends, const int32_t* src_off,
                          const int32_t* src_cnt, const int32_t* src_obj, int64_t nE,
                          const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                          const int32_t* tri_obj, int swap, int use_grid, int chain_rule,
                          int max_iter, uint64_t* out_keys, int64_t cap, int32_t* per_edge_counts);

/* A.3/A.4 per-object +z rays over [1e-4, 1e10] from the first vertex of each source
 * object's first triangle; hits walked in (t, tri) order, same-object hits within 1e-5 of
 * the previous hit skipped, per-target parity over at most max_targets targets. Emits
 * (src << 32 | tgt), or (tgt << 32 | src) when swap, for odd targets. */
int64_t orc_object_parity(const float* srcV, const uint32_t* srcI, const int32_t* src_tri_obj,
                          int64_t src_nT, int64_t n_src_obj, const float* V, int64_t nV,
                          const uint32_t* I, int64_t nT, const int32_t* tri_obj, int max_targets,
                          int swap, uint64_t* out_keys, int64_t cap, int64_t* overflow);

/* Options of orc_object_parity_ex. chain_mode 0: all hits in (t, tri) order with the
 * same-object 1e-5 merge (the any-hit forms); 1: closest-hit loop with a nextafterf step
 * (query/src/cuda/mesh_intersection.cu:134-230, cap max_iter, default 2048); 2: closest-hit
 * loop with a t + 1e-4 step and t < 1e10 (query/src/cuda/mesh_containment.cu:336-366, cap
 * max_iter, default 2000). origins: optional float3 per source object (instead of the first
 * vertex of its first triangle). Optional per-source tracking: iterations, cand_count,
 * cand_overflow [n_src]; pair_ids / pair_hits [n_src * max_pair] (odd targets in candidate
 * order, -1 / 0 beyond). stats accumulates: rays, iterations, hits, max iterations per ray,
 * suppressed, additions, toggles, overflow, odd targets (at most 256 per ray). The fallback
 * form reports at most 256 odd targets per source (max_out = 256,
 * query/src/cuda/mesh_intersection_anyhit.cu:175, mesh_intersection.cu:268). */
typedef struct {
    int chain_mode, max_iter;
    const float* origins;
    uint32_t* iterations;
    uint32_t* cand_count;
    uint32_t* cand_overflow;
    int32_t* pair_ids;
    uint32_t* pair_hits;
    int max_pair;
    int max_out;  /* odd targets reported per source, first ones in candidate order (0: all) */
    uint64_t stats[9];
} orc_objray_opts;
int64_t orc_object_parity_ex(const float* srcV, const uint32_t* srcI, const int32_t* src_tri_obj,
                             int64_t src_nT, int64_t n_src_obj, const float* V, int64_t nV,
                             const uint32_t* I, int64_t nT, const int32_t* tri_obj, int max_targets,
                             int swap, orc_objray_opts* opt, uint64_t* out_keys, int64_t cap,
                             int64_t* overflow);

/* Sorts ascending and removes duplicates in place; returns the unique count. */
int64_t orc_sort_unique_u64(uint64_t* keys, int64_t n);

/* MurmurHash3 fmix64 finaliser and the reference's open-addressing insert. */
uint64_t orc_fmix64(uint64_t k);
/* Returns 1 if inserted or already present, 0 if the probe cap was exhausted. */
int orc_ht_insert(uint64_t* table, uint64_t size, int bitwise, int max_probes, uint64_t key);

/* A.5 selectivity estimate of one grid pair. cells are GridCell records
 * {u32 CenterCount, u32 TouchCount, f32 AvgSizeMean, f32 VolRatio}. */
double orc_estimate_raw(const void* cellsA, const void* cellsB, int64_t ncells, float cell_volume,
                        float epsilon, float gamma);
/* Full estimate including the alpha correction, as the estimated mains compute it. */
int64_t orc_estimate_pairs(const void* cellsA, const void* cellsB, int64_t ncells,
                           const float minb[3], const float maxb[3], const uint32_t res[3],
                           float epsilon, float gamma, int alpha_correction);

/* Grid statistics of a mesh (preprocess_dataset.cpp generateGridStats). cells_out has
 * res^3 GridCell records; bounds_out = minBound[3], maxBound[3]. */
void orc_grid_stats(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                    const int32_t* tri_obj, int res, float world_size, void* cells_out,
                    float* bounds_out);

/* Unique-edge extraction (EdgePreprocessor::extractEdges). Pass null outputs to get
 * the counts: returns number of edges, *n_sources = flattened source count. */
int64_t orc_extract_edges(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                          const int32_t* tri_obj, float* starts, float* ends, int32_t* offsets,
                          int32_t* counts, int32_t* sources, int64_t* n_sources);

/* Exact triangle-triangle predicate (Guigue-Devillers, decision form of the reference's
 * TriInt): t1, t2 = 9 floats (three vertices). 1 when the closed triangles intersect. */
int orc_tri_tri(const float* t1, const float* t2);
/* orc_tri_tri on n pairs (A[i], B[i]), 9 floats each. */
void orc_tri_tri_many(const float* A, const float* B, int64_t n, uint8_t* out);

/* Exact intersection join: sorted unique keys (a << 32 | b) of (object of A, object of B)
 * with a triangle pair whose closed boxes overlap and orc_tri_tri reports intersecting.
 * Returns the pair count (keys written up to cap). */
int64_t orc_tri_join(const float* Va, int64_t nVa, const uint32_t* Ia, const int32_t* Oa, int64_t nTa,
                     const float* Vb, int64_t nVb, const uint32_t* Ib, const int32_t* Ob, int64_t nTb,
                     uint64_t* out_keys, int64_t cap, int threads);

#ifdef __cplusplus
}
#endif
#endif
