// tritri.hip -- exact triangle-triangle intersection: the pairwise predicate (the
// reference's TriInt / TriInt_batch_gpu, preprocess/src/tdbase_lib/geometry/TriInt.cpp:400-492,
// include/geometry.h:164-166) and an exact intersection join of two meshes built on it.
//
// Join: one lane per triangle of mesh A, taken from A's BVH triangle records (Morton
// order, so neighbouring lanes query neighbouring space). The lane's query is the
// triangle's bounding box; it walks B's BVH with a box-overlap test against the padded
// child boxes, and at the leaves tests B triangles whose exact box overlaps its own with
// the predicate. Intersecting pairs go into the pair hash table as (object of A, object of
// B) (swapped on request). Two triangles that intersect have overlapping closed boxes, so
// the box filters only remove pairs the predicate would reject.
#include "rs3d_internal.h"
#include "traverse.h"
#include "tritri.h"

namespace rs3d {

__global__ __launch_bounds__(256) void k_tri_pairs(const float* __restrict__ a, const float* __restrict__ b,
                                                   uint64_t n, uint8_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float* p = a + 9 * i;
    const float* q = b + 9 * i;
    out[i] = tri_tri_intersect(make_float3(p[0], p[1], p[2]), make_float3(p[3], p[4], p[5]),
                               make_float3(p[6], p[7], p[8]), make_float3(q[0], q[1], q[2]),
                               make_float3(q[3], q[4], q[5]), make_float3(q[6], q[7], q[8]))
                 ? 1
                 : 0;
}

// Closed overlap of the query box with a child box (lo, hi); an empty child has lo.x > hi.x.
__device__ __forceinline__ bool box_overlap(float lx, float hx, float ly, float hy, float lz, float hz, float3 qlo,
                                            float3 qhi) {
    return qlo.x <= hx && lx <= qhi.x && qlo.y <= hy && ly <= qhi.y && qlo.z <= hz && lz <= qhi.z && lx <= hx;
}

// Counters (STATS): nodes visited, leaf triangles loaded, predicate evaluations, hits.
template <int DEPTH, bool STATS>
__global__ __launch_bounds__(TR_BLOCK) void k_tri_join(const BNode* __restrict__ nodes,
                                                       const TriRec* __restrict__ tris_b,
                                                       const TriRec* __restrict__ tris_a, uint64_t n_a,
                                                       const uint32_t* __restrict__ list, OccArgs occ, int swap,
                                                       HashArgs H, unsigned long long* __restrict__ stats) {
    __shared__ int stack_mem[DEPTH * TR_BLOCK];
    int* stack = stack_mem + threadIdx.x;
    const uint64_t i = (uint64_t)blockIdx.x * TR_BLOCK + threadIdx.x;
    TraceCounters ctr;
    if (i < n_a) {  // with list: query list[i] of n_a listed ones
        float3 a0, a1, a2;
        uint32_t orig_a;
        int obj_a;
        load_tri(tris_a, list ? (int)list[i] : (int)i, a0, a1, a2, orig_a, obj_a);
        const float3 qlo = make_float3(fminf(fminf(a0.x, a1.x), a2.x), fminf(fminf(a0.y, a1.y), a2.y),
                                       fminf(fminf(a0.z, a1.z), a2.z));
        const float3 qhi = make_float3(fmaxf(fmaxf(a0.x, a1.x), a2.x), fmaxf(fmaxf(a0.y, a1.y), a2.y),
                                       fmaxf(fmaxf(a0.z, a1.z), a2.z));
        int last = -1;  // object of the last pair inserted by this lane
        int sp = 0, node = 0;
        // a pair needs A's box to meet a B triangle's box, so to cover a set occupancy bit
        const bool walk = !occ.bits || occ_box_may_hit(occ, qlo, qhi, ctr.occ_words);
        while (walk) {
            // child 0 box: x (a.x, a.y), y (a.z, a.w), z (b.x, b.y); child 1: x (b.z, b.w),
            // y (c.x, c.y), z (c.z, c.w); link = (first 0, first 1, count 0, count 1)
            const BNode nd = nodes[node];
            if (STATS) ctr.nodes++;
            const bool h0 = box_overlap(nd.a.x, nd.a.y, nd.a.z, nd.a.w, nd.b.x, nd.b.y, qlo, qhi);
            const bool h1 = box_overlap(nd.b.z, nd.b.w, nd.c.x, nd.c.y, nd.c.z, nd.c.w, qlo, qhi);
            const int4 b = make_int4(nd.link.x, nd.link.y, nd.link.z, nd.link.w);
            // Leaf children: one loop over [j, e) of child 0, then of child 1 (a single copy
            // of the predicate in the code).
            int j = (h0 && b.z > 0) ? b.x : 0, e = (h0 && b.z > 0) ? b.x + b.z : 0;
            const int j1 = (h1 && b.w > 0) ? b.y : 0, e1 = (h1 && b.w > 0) ? b.y + b.w : 0;
            bool second = false;
            while (true) {
                if (j >= e) {
                    if (second) break;
                    second = true;
                    j = j1;
                    e = e1;
                    continue;
                }
                float3 v0, v1, v2;
                uint32_t orig;
                int obj;
                load_tri(tris_b, j++, v0, v1, v2, orig, obj);
                if (STATS) ctr.tris++;
                if (obj == last) continue;
                const bool ov = fminf(fminf(v0.x, v1.x), v2.x) <= qhi.x && qlo.x <= fmaxf(fmaxf(v0.x, v1.x), v2.x) &&
                                fminf(fminf(v0.y, v1.y), v2.y) <= qhi.y && qlo.y <= fmaxf(fmaxf(v0.y, v1.y), v2.y) &&
                                fminf(fminf(v0.z, v1.z), v2.z) <= qhi.z && qlo.z <= fmaxf(fmaxf(v0.z, v1.z), v2.z);
                if (!ov) continue;
                if (STATS) ctr.restarts++;
                if (tri_tri_intersect(a0, a1, a2, v0, v1, v2)) {
                    if (STATS) ctr.hits++;
                    last = obj;
                    const uint32_t x = (uint32_t)(swap ? obj : obj_a);
                    const uint32_t y = (uint32_t)(swap ? obj_a : obj);
                    ht_insert(H, ((uint64_t)x << 32) | y);
                }
            }
            // next node without branches around the choice (descend, push, pop in one loop)
            const bool in0 = h0 && b.z == 0;
            const bool in1 = h1 && b.w == 0;
            const bool desc = in0 || in1;
            if (!desc && sp == 0) break;
            const int popv = stack[(sp > 0 ? sp - 1 : 0) * TR_BLOCK];  // read before this step's push
            if (in0 && in1) stack[sp * TR_BLOCK] = b.y;
            node = desc ? (in0 ? b.x : b.y) : popv;
            sp += (in0 && in1) ? 1 : (desc ? 0 : -1);
        }
    }
    if (STATS) flush_counters(ctr, stats);
}

// Query selection against B's occupancy bits: sel[i] = 1 where A's triangle i has a box
// covering a set bit (the only queries that can pair).
__global__ __launch_bounds__(256) void k_tri_select(const TriRec* __restrict__ tris_a, uint64_t n_a, OccArgs occ,
                                                    uint8_t* __restrict__ sel) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_a) return;
    float3 a0, a1, a2;
    uint32_t orig;
    int obj;
    load_tri(tris_a, (int)i, a0, a1, a2, orig, obj);
    const float3 lo = make_float3(fminf(fminf(a0.x, a1.x), a2.x), fminf(fminf(a0.y, a1.y), a2.y),
                                  fminf(fminf(a0.z, a1.z), a2.z));
    const float3 hi = make_float3(fmaxf(fmaxf(a0.x, a1.x), a2.x), fmaxf(fmaxf(a0.y, a1.y), a2.y),
                                  fmaxf(fmaxf(a0.z, a1.z), a2.z));
    uint64_t words = 0;
    sel[i] = occ_box_may_hit(occ, lo, hi, words) ? 1 : 0;
}

template <bool STATS>
static int launch_tri_join(rs3d_ctx* ctx, const rs3d_bvh* a, const rs3d_bvh* b, int swap, const HashArgs& H,
                           unsigned long long* stats) {
    if (a->num_tris == 0 || b->num_tris == 0) return RS3D_OK;
    if (a->num_tris > 0x7fffffffull) {
        set_error("rs3d_tri_join: more than 2^31-1 query triangles");
        return RS3D_E_ARG;
    }
    unsigned nb;
    RS3D_TRY(grid_blocks(a->num_tris, TR_BLOCK, &nb));
    int depth;
    RS3D_TRY(pick_depth(b, &depth));
    OccArgs occ;  // B's segment occupancy bits (built here on first use)
    RS3D_TRY(seg_occ_args(ctx, const_cast<rs3d_bvh*>(b), &occ));
    const bool timed = !STATS && ctx->profile;
    if (timed) {
        RS3D_TRY(prof_collect(ctx));
        RS3D_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    }
    // Large timed joins with bits: select the queries whose box covers a set bit, compact
    // them (stable: Morton order kept) and walk only those, in full waves.
    const uint32_t* list = nullptr;
    uint64_t n_q = a->num_tris;
    if (!STATS && occ.bits && a->num_tris >= (1ull << 20)) {
        const uint64_t n = a->num_tris;
        const size_t sel_bytes = (n + 255) & ~(size_t)255;
        RS3D_TRY(ctx->edge_sel.ensure(sel_bytes + n * sizeof(uint32_t)));
        uint8_t* sel = ctx->edge_sel.as<uint8_t>();
        uint32_t* idx = reinterpret_cast<uint32_t*>(sel + sel_bytes);
        unsigned nbs;
        RS3D_TRY(grid_blocks(n, 256, &nbs));
        k_tri_select<<<nbs, 256, 0, ctx->stream>>>(a->tris, n, occ, sel);
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(rs3d_compact_flags(ctx, sel, n, idx, &n_q));
        list = idx;
        occ = OccArgs{};  // tested already
        RS3D_TRY(grid_blocks(n_q ? n_q : 1, TR_BLOCK, &nb));
    }
    switch (depth) {
        case 24: k_tri_join<24, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(b->nodes, b->tris, a->tris, n_q, list, occ, swap, H, stats); break;
        case 32: k_tri_join<32, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(b->nodes, b->tris, a->tris, n_q, list, occ, swap, H, stats); break;
        case 48: k_tri_join<48, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(b->nodes, b->tris, a->tris, n_q, list, occ, swap, H, stats); break;
        default: k_tri_join<64, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(b->nodes, b->tris, a->tris, n_q, list, occ, swap, H, stats); break;
    }
    RS3D_HIP(hipGetLastError());
    if (timed) {
        RS3D_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        ctx->prof_pending = true;
    }
    return RS3D_OK;
}

static bool ht_ok(const rs3d_hash_table* ht) {
    return ht && ht->slots && ht->size && ht->failed && (!ht->bitwise || !(ht->size & (ht->size - 1)));
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_tri_intersect(rs3d_ctx* ctx, const float* a, const float* b, uint64_t n, uint8_t* out) {
    if (!ctx || (n && (!a || !b || !out))) {
        set_error("rs3d_tri_intersect: bad arguments");
        return RS3D_E_ARG;
    }
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_tri_pairs<<<nb, 256, 0, ctx->stream>>>(a, b, n, out);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int rs3d_tri_join_hash(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b, int swap,
                       rs3d_hash_table* ht) {
    if (!ctx || !bvh_a || !bvh_b || !ht_ok(ht)) {
        set_error("rs3d_tri_join_hash: bad arguments");
        return RS3D_E_ARG;
    }
    HashArgs H{reinterpret_cast<unsigned long long*>(ht->slots), ht->size, ht->bitwise, ht->max_probes,
               reinterpret_cast<unsigned long long*>(ht->failed)};
    return launch_tri_join<false>(ctx, bvh_a, bvh_b, swap, H, nullptr);
}

int rs3d_tri_join_stats(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b,
                        rs3d_tri_join_stats_t* stats) {
    if (!ctx || !bvh_a || !bvh_b || !stats) {
        set_error("rs3d_tri_join_stats: bad arguments");
        return RS3D_E_ARG;
    }
    // A table the size of the sample's pairs is not needed for counting: inserts go into a
    // one-slot table whose failures are ignored.
    RS3D_TRY(ctx->small.ensure(256));
    unsigned long long* acc = ctx->small.as<unsigned long long>();
    unsigned long long* slot = acc + 8;
    unsigned long long* failed = acc + 9;
    RS3D_HIP(hipMemsetAsync(acc, 0, 10 * sizeof(unsigned long long), ctx->stream));
    RS3D_HIP(hipMemsetAsync(slot, 0xff, sizeof(unsigned long long), ctx->stream));
    HashArgs H{slot, 1, 1, 1, failed};
    RS3D_TRY(launch_tri_join<true>(ctx, bvh_a, bvh_b, 0, H, acc));
    unsigned long long h[4];
    RS3D_HIP(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    stats->queries = bvh_a->num_tris;
    stats->nodes_visited = h[0];
    stats->tris_loaded = h[1];
    stats->hits = h[2];
    stats->predicate_tests = h[3];
    return RS3D_OK;
}

int rs3d_tri_join(rs3d_ctx* ctx, const rs3d_bvh* bvh_a, const rs3d_bvh* bvh_b, uint64_t* keys_out,
                  uint64_t max_out, uint64_t* count) {
    if (!ctx || !bvh_a || !bvh_b || (max_out && !keys_out)) {
        set_error("rs3d_tri_join: bad arguments");
        return RS3D_E_ARG;
    }
    if (count) *count = 0;
    static thread_local DevBuf slots_buf, keys_buf;
    uint64_t size = 1ull << 20, uniq = 0;
    while (true) {
        RS3D_TRY(slots_buf.ensure(size * sizeof(uint64_t) + 256));
        uint64_t* slots = slots_buf.as<uint64_t>();
        uint64_t* failed = reinterpret_cast<uint64_t*>(slots_buf.as<uint8_t>() + size * sizeof(uint64_t));
        RS3D_TRY(ht_clear_internal(ctx, slots, size, failed));
        HashArgs H{reinterpret_cast<unsigned long long*>(slots), size, 1, 0,
                   reinterpret_cast<unsigned long long*>(failed)};
        RS3D_TRY(launch_tri_join<false>(ctx, bvh_a, bvh_b, 0, H, nullptr));
        uint64_t nfail = 0, occ = 0;
        RS3D_TRY(read_u64(ctx, failed, &nfail));
        RS3D_TRY(keys_buf.ensure(size * sizeof(uint64_t)));
        if (nfail == 0) {
            RS3D_TRY(ht_compact_internal(ctx, slots, size, keys_buf.as<uint64_t>(), size, &occ));
            if (occ * 2 <= size) {
                RS3D_TRY(rs3d_sort_unique_u64(ctx, keys_buf.as<uint64_t>(), occ, &uniq));
                break;
            }
        }
        // table too full (or an insert failed): grow and run again
        size <<= 2;
        if (size > (1ull << 31)) {
            set_error("rs3d_tri_join: hash table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
    if (count) *count = uniq;
    if (uniq > max_out) {
        set_error("rs3d_tri_join: %llu pairs exceed the output capacity %llu", (unsigned long long)uniq,
                  (unsigned long long)max_out);
        return RS3D_E_CAPACITY;
    }
    if (uniq)
        RS3D_HIP(hipMemcpyAsync(keys_out, keys_buf.as<uint64_t>(), uniq * sizeof(uint64_t),
                                hipMemcpyDeviceToDevice, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

}  // extern "C"
