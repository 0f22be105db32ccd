// contain.hip -- per-object containment rays of the intersection and containment joins.
//
// Replaces __raygen__mesh_containment_anyhit + __anyhit__ah_containment
// (query/src/cuda/mesh_intersection_anyhit.cu:46-240) and __raygen__point_in_mesh +
// __anyhit__ah (query/src/cuda/mesh_containment.cu:196-377). One lane per source object:
// the ray starts at the first vertex of the object's first triangle and runs along +z
// over [1e-4, 1e10]. All hits are walked in (t, triangle) order; a hit is skipped when
// it belongs to the same target object as the previous hit and lies within 1e-5 of it
// (the any-hit suppression of mesh_containment.cu:222-231); every other hit toggles its
// target object's parity. Targets with odd parity contain the source point. The walk
// uses the hit buffer of the traversal kernels and restarts strictly after the last
// processed (t, triangle) when more hits exist, so no hit is lost for any count.
#include <cstring>

#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

__global__ __launch_bounds__(256) void k_first_tri_init(int* __restrict__ first, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) first[i] = 0x7fffffff;
}

__global__ __launch_bounds__(256) void k_first_tri(const int32_t* __restrict__ tri_obj, uint64_t nT,
                                                   uint64_t n_obj, int* __restrict__ first) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int o = tri_obj[t];
    if (o >= 0 && (uint64_t)o < n_obj) atomicMin(&first[o], (int)t);
}

enum { OBJ_FALLBACK = 0, OBJ_POINT_IN_MESH = 1 };
static_assert(OBJ_FALLBACK == RS3D_OBJ_FALLBACK && OBJ_POINT_IN_MESH == RS3D_OBJ_POINT_IN_MESH, "forms");
static_assert(OBJ_ALL_HITS == RS3D_OBJ_ALL_HITS && OBJ_CHAIN_NEXTAFTER == RS3D_OBJ_CHAIN_NEXTAFTER &&
                  OBJ_CHAIN_STEP == RS3D_OBJ_CHAIN_STEP,
              "chain modes");

struct ObjRayArgs {
    const float* src_v;          // source mesh: first vertex of each object's first triangle
    const uint32_t* src_i;
    const int* first_tri;
    const float* origins;        // or one origin per source object (float3)
    uint64_t n_obj;
    int max_targets, form, swap, chain_mode, max_iter;
    int* cand_ids;               // n_obj * max_targets scratch
    uint8_t* cand_par;
    uint32_t* cand_hits;         // per-candidate hit counts (pair tracking only)
    HashArgs out_ht, overlap_ht;
    int32_t* counts;             // fallback two-pass, pass 1: odd targets per source
    const int64_t* offsets;      // pass 2: where each source's pairs go
    rs3d_pair* results;
    int32_t* pair_ids;           // optional per-source tracking
    uint32_t* pair_hits;
    int max_pair;
    uint32_t* iterations;
    uint32_t* cand_count;
    uint32_t* cand_overflow;
    rs3d_intersection_stats* stats;
    unsigned long long* overflow;
};

// The fallback form reports at most this many odd targets per source object, as both
// reference forms do (mesh_intersection_anyhit.cu:175, mesh_intersection.cu:268).
constexpr int FALLBACK_MAX_OUT = 256;

template <int DEPTH>
__global__ __launch_bounds__(TR_BLOCK) void k_object_rays(const BNode* __restrict__ nodes,
                                                          const TriRec* __restrict__ tris, ObjRayArgs A) {
    __shared__ int stack_mem[DEPTH * TR_BLOCK];
    const uint64_t s = (uint64_t)blockIdx.x * TR_BLOCK + threadIdx.x;
    if (s >= A.n_obj) return;
    float3 o;
    if (A.origins) {
        o = ld3f(A.origins, s);
    } else {
        const int ft = A.first_tri[s];
        if (ft == 0x7fffffff) {  // no triangle: no ray
            if (A.iterations) {
                A.iterations[s] = 0;
                A.cand_count[s] = 0;
                A.cand_overflow[s] = 0;
            }
            if (A.pair_ids)
                for (int j = 0; j < A.max_pair; ++j) {
                    A.pair_ids[s * (uint64_t)A.max_pair + j] = -1;
                    A.pair_hits[s * (uint64_t)A.max_pair + j] = 0;
                }
            if (A.counts) A.counts[s] = 0;
            return;
        }
        o = ld3f(A.src_v, A.src_i[3 * (uint64_t)ft]);
    }
    int* ids = A.cand_ids + s * (uint64_t)A.max_targets;
    uint8_t* par = A.cand_par + s * (uint64_t)A.max_targets;
    uint32_t* hc = A.cand_hits ? A.cand_hits + s * (uint64_t)A.max_targets : nullptr;
    ObjRayCtr c;
    int count = 0;
    if (nodes)
        count = object_ray_walk<DEPTH>(nodes, tris, o, stack_mem + threadIdx.x, ids, par, hc, A.max_targets,
                                       A.chain_mode, A.max_iter, c);
    else
        c.iters = 1;  // an empty target: the one trace misses
    if (c.overflow) atomicAdd(A.overflow, (unsigned long long)c.overflow);
    const int max_out = A.form == OBJ_FALLBACK ? FALLBACK_MAX_OUT : A.max_targets;
    int odd = 0;
    int64_t w = A.results ? A.offsets[s] : 0;
    for (int j = 0; j < count; ++j) {
        if (!(par[j] & 1)) continue;
        if (odd >= max_out) break;
        const uint32_t tgt = (uint32_t)ids[j];
        const uint32_t src = (uint32_t)s;
        if (A.pair_ids && odd < A.max_pair) {
            A.pair_ids[s * (uint64_t)A.max_pair + odd] = (int32_t)tgt;
            A.pair_hits[s * (uint64_t)A.max_pair + odd] = hc ? hc[j] : 0u;
        }
        ++odd;
        if (A.form == OBJ_FALLBACK) {
            const uint32_t a = A.swap ? tgt : src, b = A.swap ? src : tgt;
            if (A.results) {
                rs3d_pair p;
                p.object_id_mesh1 = (int32_t)a;
                p.object_id_mesh2 = (int32_t)b;
                A.results[w++] = p;
            } else if (!A.counts) {
                ht_insert(A.out_ht, ((uint64_t)a << 32) | b);
            }
        } else {
            // source = B object, targets = A objects; pair (a, b) unless it overlaps
            const uint64_t key = ((uint64_t)tgt << 32) | src;
            if (!ht_contains(A.overlap_ht, key)) ht_insert(A.out_ht, key);
        }
    }
    if (A.counts) A.counts[s] = odd;
    if (A.pair_ids)
        for (int j = odd; j < A.max_pair; ++j) {
            A.pair_ids[s * (uint64_t)A.max_pair + j] = -1;
            A.pair_hits[s * (uint64_t)A.max_pair + j] = 0;
        }
    if (A.iterations) {
        A.iterations[s] = c.iters;
        A.cand_count[s] = (uint32_t)count;
        A.cand_overflow[s] = c.overflow;
    }
    if (A.stats) {
        rs3d_intersection_stats* st = A.stats;
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_rays_total), 1ull);
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_iterations_total), (unsigned long long)c.iters);
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_hits_total), (unsigned long long)c.hits);
        atomicMax(&st->containment_max_iterations_per_ray, c.iters);
        if (c.suppressed)
            atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_same_hit_suppressed),
                      (unsigned long long)c.suppressed);
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_candidate_additions), (unsigned long long)c.additions);
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_candidate_toggles), (unsigned long long)c.toggles);
        if (c.overflow)
            atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_candidate_overflow), (unsigned long long)c.overflow);
        atomicAdd(reinterpret_cast<unsigned long long*>(&st->containment_targets_total),
                  (unsigned long long)(odd < FALLBACK_MAX_OUT ? odd : FALLBACK_MAX_OUT));
    }
}

static int first_triangles(rs3d_ctx* ctx, const int32_t* tri_obj, uint64_t nT, uint64_t n_obj,
                           int* first) {
    unsigned nb;
    RS3D_TRY(grid_blocks(n_obj, 256, &nb));
    k_first_tri_init<<<nb, 256, 0, ctx->stream>>>(first, n_obj);
    if (nT) {
        RS3D_TRY(grid_blocks(nT, 256, &nb));
        k_first_tri<<<nb, 256, 0, ctx->stream>>>(tri_obj, nT, n_obj, first);
    }
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int first_triangles_dev(rs3d_ctx* ctx, const int32_t* tri_obj, uint64_t nT, uint64_t n_obj, int* first) {
    return first_triangles(ctx, tri_obj, nT, n_obj, first);
}

static HashArgs hash_args(const rs3d_hash_table* ht) {
    HashArgs h{};
    if (ht) {
        h.slots = reinterpret_cast<unsigned long long*>(ht->slots);
        h.size = ht->size;
        h.bitwise = ht->bitwise;
        h.max_probes = ht->max_probes;
        h.failed = reinterpret_cast<unsigned long long*>(ht->failed);
    }
    return h;
}

// Object rays of every source object (src's first vertices, or opts->origins) against
// target in the given form; opts may be null (all hits, no tracking, hash output).
int object_rays_internal(rs3d_ctx* ctx, const rs3d_mesh_view* src, const rs3d_bvh* target,
                         int form, int swap, int max_targets, const rs3d_object_ray_opts* opts,
                         const rs3d_hash_table* out_ht, const rs3d_hash_table* overlap_ht,
                         uint64_t* overflow_host) {
    if (overflow_host) *overflow_host = 0;
    const bool by_origin = opts && opts->origins;
    const uint64_t n_obj = by_origin ? opts->num_sources : (src ? src->num_objects : 0);
    if (n_obj == 0) return RS3D_OK;
    const bool track = opts && (opts->pair_target_ids || opts->iterations);
    if (target->num_tris == 0 && !track && !(opts && opts->counts)) return RS3D_OK;
    int depth = 24;
    if (target->num_tris) RS3D_TRY(pick_depth(target, &depth));
    static thread_local DevBuf scratch;
    const bool hits = opts && opts->pair_target_ids;
    const size_t ids_bytes = n_obj * (size_t)max_targets * sizeof(int);
    const size_t par_bytes = (n_obj * (size_t)max_targets + 255) & ~size_t(255);
    const size_t hit_bytes = hits ? n_obj * (size_t)max_targets * sizeof(uint32_t) : 0;
    const size_t first_bytes = by_origin ? 0 : n_obj * sizeof(int);
    RS3D_TRY(scratch.ensure(ids_bytes + par_bytes + hit_bytes + first_bytes + 1024));
    uint8_t* base = scratch.as<uint8_t>();
    ObjRayArgs A{};
    A.cand_ids = reinterpret_cast<int*>(base);
    A.cand_par = base + ids_bytes;
    A.cand_hits = hits ? reinterpret_cast<uint32_t*>(base + ids_bytes + par_bytes) : nullptr;
    int* first = reinterpret_cast<int*>(base + ((ids_bytes + par_bytes + hit_bytes + 255) & ~size_t(255)));
    unsigned long long* over = reinterpret_cast<unsigned long long*>(
        base + ((ids_bytes + par_bytes + hit_bytes + first_bytes + 511) & ~size_t(255)));
    if (by_origin) {
        A.origins = opts->origins;
    } else {
        RS3D_TRY(first_triangles(ctx, src->tri_to_obj, src->num_triangles, n_obj, first));
        A.src_v = src->vertices;
        A.src_i = src->indices;
        A.first_tri = first;
    }
    RS3D_TRY(fill_u32(ctx, over, 2, 0u));
    A.n_obj = n_obj;
    A.max_targets = max_targets;
    A.form = form;
    A.swap = swap;
    A.chain_mode = opts ? opts->chain_mode : OBJ_ALL_HITS;
    A.max_iter = opts ? opts->max_iterations : 0;
    A.out_ht = hash_args(out_ht);
    A.overlap_ht = hash_args(overlap_ht);
    if (opts) {
        A.counts = opts->counts;
        A.offsets = opts->offsets;
        A.results = opts->results;
        A.pair_ids = opts->pair_target_ids;
        A.pair_hits = opts->pair_target_hits;
        A.max_pair = opts->max_pair_targets;
        A.iterations = opts->iterations;
        A.cand_count = opts->candidate_count;
        A.cand_overflow = opts->candidate_overflow;
        A.stats = opts->stats;
    }
    A.overflow = over;
    unsigned nb;
    RS3D_TRY(grid_blocks(n_obj, TR_BLOCK, &nb));
    // an empty target (no nodes): every ray misses; the kernel still writes counts and tracking
    const BNode* nodes = target->num_tris ? target->nodes : nullptr;
    switch (depth) {
        case 24: k_object_rays<24><<<nb, TR_BLOCK, 0, ctx->stream>>>(nodes, target->tris, A); break;
        case 32: k_object_rays<32><<<nb, TR_BLOCK, 0, ctx->stream>>>(nodes, target->tris, A); break;
        case 48: k_object_rays<48><<<nb, TR_BLOCK, 0, ctx->stream>>>(nodes, target->tris, A); break;
        default: k_object_rays<64><<<nb, TR_BLOCK, 0, ctx->stream>>>(nodes, target->tris, A); break;
    }
    RS3D_HIP(hipGetLastError());
    if (overflow_host) RS3D_TRY(read_u64(ctx, over, overflow_host));
    return RS3D_OK;
}

}  // namespace rs3d

using namespace rs3d;

static bool mesh_view_ok(const rs3d_mesh_view* m) {
    return m && (m->num_triangles == 0 || (m->vertices && m->indices && m->tri_to_obj));
}

static bool ht_ok(const rs3d_hash_table* ht) {
    return ht && ht->slots && ht->size && ht->failed && !(ht->bitwise && (ht->size & (ht->size - 1)));
}

extern "C" {

int rs3d_containment_fallback(rs3d_ctx* ctx, const rs3d_mesh_view* src, const rs3d_bvh* target,
                              int swap, int max_targets, rs3d_hash_table* ht,
                              uint64_t* overflow) {
    if (!ctx || !mesh_view_ok(src) || !target || !ht_ok(ht) || max_targets <= 0 ||
        max_targets > 4096) {
        set_error("rs3d_containment_fallback: bad arguments");
        return RS3D_E_ARG;
    }
    return object_rays_internal(ctx, src, target, OBJ_FALLBACK, swap, max_targets, nullptr, ht, nullptr,
                                overflow);
}

int rs3d_point_in_mesh_join(rs3d_ctx* ctx, const rs3d_mesh_view* src_b, const rs3d_bvh* target_a,
                            int max_targets, const rs3d_hash_table* overlap_ht,
                            rs3d_hash_table* out_ht, uint64_t* overflow) {
    if (!ctx || !mesh_view_ok(src_b) || !target_a || !ht_ok(out_ht) || !ht_ok(overlap_ht) ||
        max_targets <= 0 || max_targets > 4096) {
        set_error("rs3d_point_in_mesh_join: bad arguments");
        return RS3D_E_ARG;
    }
    return object_rays_internal(ctx, src_b, target_a, OBJ_POINT_IN_MESH, 0, max_targets, nullptr, out_ht,
                                overlap_ht, overflow);
}

int rs3d_object_rays(rs3d_ctx* ctx, const rs3d_mesh_view* src, const rs3d_bvh* target, int form,
                     int swap, int max_targets, const rs3d_object_ray_opts* opts, rs3d_hash_table* ht,
                     const rs3d_hash_table* overlap_ht, uint64_t* overflow) {
    const bool by_origin = opts && opts->origins;
    const bool two_pass = opts && (opts->counts || opts->results);
    bool ok = ctx && target && max_targets > 0 && max_targets <= 4096 && (by_origin || mesh_view_ok(src)) &&
              (form == RS3D_OBJ_FALLBACK || form == RS3D_OBJ_POINT_IN_MESH);
    if (ok && opts) {
        ok = opts->chain_mode >= RS3D_OBJ_ALL_HITS && opts->chain_mode <= RS3D_OBJ_CHAIN_STEP &&
             opts->max_iterations >= 0 && (!opts->pair_target_ids || (opts->pair_target_hits && opts->max_pair_targets > 0)) &&
             (!opts->iterations || (opts->candidate_count && opts->candidate_overflow)) &&
             (!opts->results || opts->offsets) && !(opts->counts && opts->results);
    }
    if (ok) ok = form == RS3D_OBJ_FALLBACK ? (two_pass || ht_ok(ht)) : (!two_pass && ht_ok(ht) && ht_ok(overlap_ht));
    if (!ok) {
        set_error("rs3d_object_rays: bad arguments");
        return RS3D_E_ARG;
    }
    return object_rays_internal(ctx, src, target, form, swap, max_targets, opts, two_pass ? nullptr : ht,
                                form == RS3D_OBJ_POINT_IN_MESH ? overlap_ht : nullptr, overflow);
}

}  // extern "C"

// ---- whole joins ------------------------------------------------------------------------

namespace {

// Device hash table owned by a join call; grows by 4x until every insert succeeds and
// the fill stays at or below one half.
struct OwnedTable {
    rs3d::DevBuf buf;
    rs3d_hash_table t{};
    int init(rs3d_ctx* ctx, uint64_t slots) {
        RS3D_TRY(buf.ensure(slots * sizeof(uint64_t) + 256));
        t.slots = buf.as<uint64_t>();
        t.size = slots;
        t.bitwise = 1;
        t.max_probes = 0;
        t.failed = reinterpret_cast<uint64_t*>(buf.as<uint8_t>() + slots * sizeof(uint64_t));
        return rs3d::ht_clear_internal(ctx, t.slots, t.size, t.failed);
    }
    ~OwnedTable() { buf.release(); }
};

uint64_t pow2_at_least(uint64_t v) {
    uint64_t p = 1024;
    while (p < v) p <<= 1;
    return p;
}

int finish_table(rs3d_ctx* ctx, OwnedTable& tab, rs3d::DevBuf& keys, bool* ok, uint64_t* n) {
    uint64_t failed = 0;
    RS3D_TRY(rs3d::read_u64(ctx, tab.t.failed, &failed));
    *ok = false;
    if (failed) return RS3D_OK;
    RS3D_TRY(keys.ensure(tab.t.size * sizeof(uint64_t)));
    uint64_t occ = 0;
    RS3D_TRY(rs3d::ht_compact_internal(ctx, tab.t.slots, tab.t.size, keys.as<uint64_t>(), tab.t.size, &occ));
    if (occ * 2 > tab.t.size) return RS3D_OK;
    uint64_t u = 0;
    RS3D_TRY(rs3d_sort_unique_u64(ctx, keys.as<uint64_t>(), occ, &u));
    *n = u;
    *ok = true;
    return RS3D_OK;
}

int copy_out(rs3d_ctx* ctx, const rs3d::DevBuf& keys, uint64_t n, uint64_t* out, uint64_t max_out,
             const char* what) {
    if (n > max_out) {
        rs3d::set_error("%s: %llu pairs exceed the output capacity %llu", what, (unsigned long long)n,
                        (unsigned long long)max_out);
        return RS3D_E_CAPACITY;
    }
    if (n) RS3D_HIP(hipMemcpyAsync(out, keys.ptr, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

}  // namespace

extern "C" {

int rs3d_intersection_join(rs3d_ctx* ctx, const rs3d_mesh_view* m1, const rs3d_bvh* b1,
                           const rs3d_edges* e1, const rs3d_mesh_view* m2, const rs3d_bvh* b2,
                           const rs3d_edges* e2, int overlap_max_iterations, uint64_t table_slots,
                           uint64_t* keys_out, uint64_t max_out, uint64_t* count,
                           uint64_t* overflow) {
    if (!ctx || !mesh_view_ok(m1) || !mesh_view_ok(m2) || !b1 || !b2 || !e1 || !e2 ||
        (max_out && !keys_out)) {
        set_error("rs3d_intersection_join: bad arguments");
        return RS3D_E_ARG;
    }
    if (count) *count = 0;
    if (overflow) *overflow = 0;
    const int cap = overlap_max_iterations > 0 ? overlap_max_iterations : 100;
    uint64_t slots = pow2_at_least(table_slots ? table_slots : (1ull << 20));
    OwnedTable tab;
    DevBuf keys;
    while (true) {
        RS3D_TRY(tab.init(ctx, slots));
        RS3D_TRY(edge_hash_internal(ctx, b2, e1, 0, &tab.t, 0, cap));
        RS3D_TRY(edge_hash_internal(ctx, b1, e2, 1, &tab.t, 0, cap));
        uint64_t o1 = 0, o2 = 0;
        RS3D_TRY(object_rays_internal(ctx, m1, b2, OBJ_FALLBACK, 0, 256, nullptr, &tab.t, nullptr, &o1));
        RS3D_TRY(object_rays_internal(ctx, m2, b1, OBJ_FALLBACK, 1, 256, nullptr, &tab.t, nullptr, &o2));
        bool ok = false;
        uint64_t n = 0;
        RS3D_TRY(finish_table(ctx, tab, keys, &ok, &n));
        if (ok) {
            if (count) *count = n;
            if (overflow) *overflow = o1 + o2;
            int rc = copy_out(ctx, keys, n, keys_out, max_out, "rs3d_intersection_join");
            keys.release();
            return rc;
        }
        slots <<= 2;
        if (slots > (1ull << 31)) {
            set_error("rs3d_intersection_join: hash table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
}

int rs3d_containment_join(rs3d_ctx* ctx, const rs3d_mesh_view* ma, const rs3d_bvh* ba,
                          const rs3d_edges* ea, const rs3d_mesh_view* mb, const rs3d_bvh* bb,
                          const rs3d_edges* eb, uint64_t table_slots, uint64_t* contain_out,
                          uint64_t max_contain, uint64_t* n_contain, uint64_t* overlap_out,
                          uint64_t max_overlap, uint64_t* n_overlap, uint64_t* overflow) {
    if (!ctx || !mesh_view_ok(ma) || !mesh_view_ok(mb) || !ba || !bb || !ea || !eb ||
        (max_contain && !contain_out)) {
        set_error("rs3d_containment_join: bad arguments");
        return RS3D_E_ARG;
    }
    if (n_contain) *n_contain = 0;
    if (n_overlap) *n_overlap = 0;
    if (overflow) *overflow = 0;
    uint64_t slots = pow2_at_least(table_slots ? table_slots : (1ull << 20));
    OwnedTable ov, ct;
    DevBuf ok_keys, ct_keys;
    // phase 1: overlap set (A object, B object)
    uint64_t n_ov = 0;
    while (true) {
        RS3D_TRY(ov.init(ctx, slots));
        RS3D_TRY(edge_hash_internal(ctx, ba, eb, 1, &ov.t, 1, 256));  // B edges -> A: (hit A, src B)
        RS3D_TRY(edge_hash_internal(ctx, bb, ea, 0, &ov.t, 1, 256));  // A edges -> B: (src A, hit B)
        bool ok = false;
        RS3D_TRY(finish_table(ctx, ov, ok_keys, &ok, &n_ov));
        if (ok) break;
        slots <<= 2;
        if (slots > (1ull << 31)) {
            set_error("rs3d_containment_join: overlap table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
    // phase 2: strict containment of B's first vertices in A objects
    uint64_t cslots = pow2_at_least(table_slots ? table_slots : (1ull << 20));
    uint64_t n_ct = 0, over = 0;
    while (true) {
        RS3D_TRY(ct.init(ctx, cslots));
        RS3D_TRY(object_rays_internal(ctx, mb, ba, OBJ_POINT_IN_MESH, 0, 512, nullptr, &ct.t, &ov.t, &over));
        bool ok = false;
        RS3D_TRY(finish_table(ctx, ct, ct_keys, &ok, &n_ct));
        if (ok) break;
        cslots <<= 2;
        if (cslots > (1ull << 31)) {
            set_error("rs3d_containment_join: containment table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
    if (n_contain) *n_contain = n_ct;
    if (n_overlap) *n_overlap = n_ov;
    if (overflow) *overflow = over;
    RS3D_TRY(copy_out(ctx, ct_keys, n_ct, contain_out, max_contain, "rs3d_containment_join"));
    if (overlap_out) RS3D_TRY(copy_out(ctx, ok_keys, n_ov, overlap_out, max_overlap, "rs3d_containment_join"));
    ok_keys.release();
    ct_keys.release();
    return RS3D_OK;
}

}  // extern "C"
