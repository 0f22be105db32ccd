// batch.hip -- many independent containment and overlap queries in one persistent launch.
//
// Every query is cut into 64-ray chunks (points of a containment query, edges of one
// direction of an overlap query); a device chunk table lists (query, first ray) and one
// persistent kernel of wave-sized workgroups pulls chunks from an atomic queue head until
// the table is drained, so small and large queries share the GPU without a launch each.
// Containment rays of all queries are first reordered by one radix sort on
// (query, 2-D Morton of the origin) so waves stay coherent; overlap pairs of all queries go
// into one hash table keyed (query << 48 | a << 24 | b), are compacted, radix-sorted once
// and split per query. Per ray the same programs as the single-query kernels run, so every
// result is identical to running the query alone.
#include <algorithm>
#include <cstring>
#include <vector>

#include "batch_plan.h"

namespace rs3d {

__device__ __forceinline__ uint32_t spread2_11(uint32_t v) {  // 11 bits -> every other bit
    v &= 0x7ffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// Sort keys of containment rays: (descriptor index << 22) | 22-bit Morton of (u, w).
__global__ __launch_bounds__(256) void k_batch_keys(const BatchDesc* __restrict__ qs,
                                                    const float* __restrict__ lo_sc, uint32_t nq,
                                                    const uint64_t* __restrict__ ray_base,
                                                    uint64_t total, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= total) return;
    // query of ray r: binary search over the ray bases
    uint32_t lo = 0, hi = nq;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ray_base[mid] <= r) lo = mid;
        else hi = mid;
    }
    const BatchDesc& q = qs[lo];
    const uint64_t i = r - q.ray_base;
    const int U = q.axis == 0 ? 1 : (q.axis == 1 ? 2 : 0);
    const int W = q.axis == 0 ? 2 : (q.axis == 1 ? 0 : 1);
    const float u = q.pts[3 * i + U], w = q.pts[3 * i + W];
    const float fu = fminf(fmaxf((u - lo_sc[4 * lo]) * lo_sc[4 * lo + 2], 0.0f), 2047.0f);
    const float fw = fminf(fmaxf((w - lo_sc[4 * lo + 1]) * lo_sc[4 * lo + 3], 0.0f), 2047.0f);
    keys[r] = (lo << 22) | (spread2_11((uint32_t)fu) << 1) | spread2_11((uint32_t)fw);
    vals[r] = (uint32_t)i;
}

template <int DEPTH>
__global__ __launch_bounds__(TR_BLOCK) void k_batch(const BatchDesc* __restrict__ qs,
                                                    const Chunk* __restrict__ chunks, uint32_t nchunks,
                                                    unsigned int* __restrict__ head,
                                                    const uint32_t* __restrict__ perm, HashArgs H,
                                                    unsigned long long* __restrict__ bad_ids) {
    __shared__ int stack_mem[DEPTH * TR_BLOCK];
    const int lane = threadIdx.x;
    TraceCounters ctr;
    while (true) {
        unsigned int c = 0;
        if (lane == 0) c = atomicAdd(head, 1u);
        c = __shfl(c, 0, 64);
        if (c >= nchunks) break;
        const Chunk ch = chunks[c];
        const BatchDesc q = qs[ch.desc];
        const uint64_t i = (uint64_t)ch.start + lane;
        if (i >= q.n) continue;
        if (q.kind == BQ_CONTAIN) {
            const uint64_t pi = perm ? perm[q.ray_base + i] : i;
            const float3 o = ld3f(q.pts, pi);
            int obj;
            if (q.axis == 0) obj = point_ray<0, DEPTH, false>(q.nodes, q.tris, o, stack_mem + lane, ctr);
            else if (q.axis == 1) obj = point_ray<1, DEPTH, false>(q.nodes, q.tris, o, stack_mem + lane, ctr);
            else obj = point_ray<2, DEPTH, false>(q.nodes, q.tris, o, stack_mem + lane, ctr);
            q.out[pi] = obj;
        } else if (q.kind == BQ_OBJECTS) {
            const int ft = q.first_tri[i];
            if (ft == 0x7fffffff) continue;
            const float3 o = ld3f(q.src_v, q.src_i[3 * (uint64_t)ft]);
            int* ids = q.cand_ids + i * (uint64_t)q.max_targets;
            uint8_t* par = q.cand_par + i * (uint64_t)q.max_targets;
            unsigned long long over = 0;
            const int cnt = object_ray_parity<DEPTH>(q.nodes, q.tris, o, stack_mem + lane, ids, par,
                                                     q.max_targets, over);
            if (over) atomicAdd(bad_ids + 1, over);
            const uint64_t qk = (uint64_t)q.qid << 48;
            for (int j = 0; j < cnt; ++j) {
                if (!(par[j] & 1)) continue;
                const uint32_t a = (uint32_t)(q.swap ? ids[j] : (int)i);
                const uint32_t b = (uint32_t)(q.swap ? (int)i : ids[j]);
                if ((a | b) >= (1u << 24)) {
                    atomicAdd(bad_ids, 1ull);
                    continue;
                }
                ht_insert(H, qk | ((uint64_t)a << 24) | b);
            }
        } else {
            const int nsrc = q.E.src_cnt[i];
            const int soff = q.E.src_off[i];
            const uint64_t qk = (uint64_t)q.qid << 48;
            edge_ray<DEPTH, false>(q.nodes, q.tris, ld3f(q.E.starts, i), ld3f(q.E.ends, i), CHAIN_EPS, 100,
                                   stack_mem + lane, ctr, [&](int tgt) {
                                       for (int s = 0; s < nsrc; ++s) {
                                           const int src = q.E.src_obj[soff + s];
                                           const uint32_t a = (uint32_t)(q.swap ? tgt : src);
                                           const uint32_t b = (uint32_t)(q.swap ? src : tgt);
                                           if ((a | b) >= (1u << 24)) {
                                               atomicAdd(bad_ids, 1ull);
                                               continue;
                                           }
                                           ht_insert(H, qk | ((uint64_t)a << 24) | b);
                                       }
                                   });
        }
    }
}

__global__ __launch_bounds__(256) void k_batch_split(const uint64_t* __restrict__ keys, uint64_t n,
                                                     uint64_t* __restrict__ out) {
    // batch key (q << 48 | a << 24 | b) -> pair key (a << 32 | b)
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint64_t k = keys[i];
        out[i] = (((k >> 24) & 0xffffffull) << 32) | (k & 0xffffffull);
    }
}

// ------------------------------------------------------------------------------------
// Batch engine (host): descriptors -> chunk table -> ray reorder -> persistent launch ->
// pair table compaction and sort. Used by rs3d_batch_execute and the request server.
// ------------------------------------------------------------------------------------
int BatchPlan::add_contain(const rs3d_bvh* b, const float* pts, uint64_t n, int32_t* out, int axis, int qid) {
    if (n == 0) return RS3D_OK;
    if (b->num_tris == 0) return rs3d_trace_points(ctx_, b, pts, n, nullptr, axis, RS3D_OUT_OBJ_I32, out);
    BatchItem d;
    RS3D_TRY(base(d, b, BQ_CONTAIN, qid));
    d.pts = pts;
    d.out = out;
    d.n = n;
    d.axis = axis;
    d.ray_base = contain_rays;
    contain_rays += n;
    const int U = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
    const int W = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
    const float eu = b->bmax[U] - b->bmin[U], ew = b->bmax[W] - b->bmin[W];
    d.lo_sc[0] = b->bmin[U];
    d.lo_sc[1] = b->bmin[W];
    d.lo_sc[2] = eu > 0 ? 2048.0f / eu : 0.0f;
    d.lo_sc[3] = ew > 0 ? 2048.0f / ew : 0.0f;
    items.push_back(d);
    return RS3D_OK;
}

int BatchPlan::add_edges(const rs3d_bvh* target, const rs3d_edges* e, int swap, int qid) {
    if (!e || e->num_edges == 0 || target->num_tris == 0) return RS3D_OK;
    BatchDesc d{};
    RS3D_TRY(base(d, target, BQ_EDGES, qid));
    d.E = EdgeArgs{e->starts, e->ends, e->src_offsets, e->src_counts, e->src_objects, e->num_edges};
    d.n = e->num_edges;
    d.swap = swap;
    items.push_back(d);
    any_pairs = true;
    return RS3D_OK;
}

int BatchPlan::add_objects(const rs3d_bvh* target, const rs3d_mesh_view* src, const int* first_tri,
                           int* cand_ids, uint8_t* cand_par, int max_targets, int swap, int qid) {
    if (src->num_objects == 0 || target->num_tris == 0) return RS3D_OK;
    BatchDesc d{};
    RS3D_TRY(base(d, target, BQ_OBJECTS, qid));
    d.src_v = src->vertices;
    d.src_i = src->indices;
    d.first_tri = first_tri;
    d.cand_ids = cand_ids;
    d.cand_par = cand_par;
    d.max_targets = max_targets;
    d.n = src->num_objects;
    d.swap = swap;
    items.push_back(d);
    any_pairs = true;
    return RS3D_OK;
}

int BatchPlan::base(BatchDesc& d, const rs3d_bvh* b, int kind, int qid) {
    if (qid < 0 || qid >= (1 << 15)) {
        set_error("batch: query id %d out of range", qid);
        return RS3D_E_ARG;
    }
    int dd = 64;
    RS3D_TRY(pick_depth(b, &dd));
    depth = std::max(depth, dd);
    d.nodes = b->nodes;
    d.tris = b->tris;
    d.kind = kind;
    d.qid = qid;
    num_qids = std::max(num_qids, qid + 1);
    return RS3D_OK;
}

int BatchPlan::run(DevBuf& keys_out, uint64_t* n_keys, std::vector<uint64_t>& offsets,
                   uint64_t* overflow) {
    rs3d_ctx* ctx = ctx_;
    *n_keys = 0;
    if (overflow) *overflow = 0;
    offsets.assign(num_qids + 1, 0);
    std::stable_partition(items.begin(), items.end(), [](const BatchItem& d) { return d.kind == BQ_CONTAIN; });
    uint32_t n_contain = 0;
    for (const auto& d : items) n_contain += d.kind == BQ_CONTAIN;
    if (n_contain >= (1u << 10)) {
        set_error("batch: at most 1023 containment queries per launch");
        return RS3D_E_ARG;
    }
    std::vector<BatchDesc> descs(items.begin(), items.end());
    std::vector<float> lo_sc;
    std::vector<uint64_t> ray_base(n_contain + 1, contain_rays);
    for (uint32_t k = 0; k < n_contain; ++k) {
        ray_base[k] = items[k].ray_base;
        lo_sc.insert(lo_sc.end(), items[k].lo_sc, items[k].lo_sc + 4);
    }
    std::vector<Chunk> chunks;
    for (uint32_t k = 0; k < descs.size(); ++k)
        for (uint64_t s = 0; s < descs[k].n; s += TR_BLOCK) {
            if (s > 0xffffffffull) {
                set_error("batch: query too large");
                return RS3D_E_ARG;
            }
            chunks.push_back({k, (uint32_t)s});
        }
    // big queries first: their chunks are pulled first, the small ones fill the tail
    std::stable_sort(chunks.begin(), chunks.end(),
                     [&](const Chunk& a, const Chunk& b) { return descs[a.desc].n > descs[b.desc].n; });
    const size_t o_desc = 0;
    const size_t o_chunk = (descs.size() * sizeof(BatchDesc) + 255) & ~size_t(255);
    const size_t o_losc = o_chunk + ((chunks.size() * sizeof(Chunk) + 255) & ~size_t(255));
    const size_t o_base = o_losc + ((lo_sc.size() * sizeof(float) + 255) & ~size_t(255));
    const size_t o_head = o_base + (((n_contain + 1) * sizeof(uint64_t) + 255) & ~size_t(255));
    RS3D_TRY(meta.ensure(o_head + 256));
    uint8_t* mb = meta.as<uint8_t>();
    if (!descs.empty())
        RS3D_HIP(hipMemcpyAsync(mb + o_desc, descs.data(), descs.size() * sizeof(BatchDesc), hipMemcpyHostToDevice, ctx->stream));
    if (!chunks.empty())
        RS3D_HIP(hipMemcpyAsync(mb + o_chunk, chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, ctx->stream));
    if (!lo_sc.empty())
        RS3D_HIP(hipMemcpyAsync(mb + o_losc, lo_sc.data(), lo_sc.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    RS3D_HIP(hipMemcpyAsync(mb + o_base, ray_base.data(), ray_base.size() * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
    unsigned int* head = reinterpret_cast<unsigned int*>(mb + o_head);
    unsigned long long* bad = reinterpret_cast<unsigned long long*>(mb + o_head + 64);
    // host copies of the descriptors only stay valid until the stream drains
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    const uint32_t* perm = nullptr;
    if (contain_rays > 0 && ctx->opt_ray_sort_min > 0) {
        if (contain_rays > 0xffffffffull) {
            set_error("batch: more than 2^32-1 containment rays");
            return RS3D_E_ARG;
        }
        RS3D_TRY(ctx->ray_keys.ensure(contain_rays * sizeof(uint32_t)));
        RS3D_TRY(ctx->ray_vals.ensure(contain_rays * sizeof(uint32_t)));
        unsigned nb;
        RS3D_TRY(grid_blocks(contain_rays, 256, &nb));
        k_batch_keys<<<nb, 256, 0, ctx->stream>>>(reinterpret_cast<BatchDesc*>(mb + o_desc),
                                                   reinterpret_cast<float*>(mb + o_losc), n_contain,
                                                   reinterpret_cast<uint64_t*>(mb + o_base), contain_rays,
                                                   ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>());
        RS3D_HIP(hipGetLastError());
        int qbits = 0;  // the descriptor index occupies the bits above the 22 Morton bits
        while (qbits < 10 && (1u << qbits) < n_contain) ++qbits;
        RS3D_TRY(radix_sort_u32_pairs(ctx, ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>(),
                                      contain_rays, 22 + qbits));
        perm = ctx->ray_vals.as<uint32_t>();
    }
    uint64_t slots = 1024;
    while (slots < slot_hint) slots <<= 1;
    const unsigned waves = (unsigned)std::max(ctx->num_cus, 1) * 16u;
    while (true) {
        HashArgs H{};
        if (any_pairs) {
            RS3D_TRY(table.ensure(slots * sizeof(uint64_t) + 64));
            H.slots = table.as<unsigned long long>();
            H.size = slots;
            H.bitwise = 1;
            H.max_probes = 0;
            H.failed = reinterpret_cast<unsigned long long*>(table.as<uint8_t>() + slots * sizeof(uint64_t));
            RS3D_TRY(ht_clear_internal(ctx, table.as<uint64_t>(), slots, reinterpret_cast<uint64_t*>(H.failed)));
        }
        RS3D_HIP(hipMemsetAsync(head, 0, 128, ctx->stream));
        if (!chunks.empty()) {
            const Chunk* cp = reinterpret_cast<Chunk*>(mb + o_chunk);
            const BatchDesc* dp = reinterpret_cast<BatchDesc*>(mb + o_desc);
            const unsigned grid = (unsigned)std::min<uint64_t>(waves, chunks.size());
            const uint32_t nc = (uint32_t)chunks.size();
            switch (depth) {
                case 24: k_batch<24><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, cp, nc, head, perm, H, bad); break;
                case 32: k_batch<32><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, cp, nc, head, perm, H, bad); break;
                case 48: k_batch<48><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, cp, nc, head, perm, H, bad); break;
                default: k_batch<64><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, cp, nc, head, perm, H, bad); break;
            }
            RS3D_HIP(hipGetLastError());
        }
        if (!any_pairs) {
            RS3D_HIP(hipStreamSynchronize(ctx->stream));
            return RS3D_OK;
        }
        unsigned long long hb[2] = {0, 0};
        uint64_t failed = 0;
        RS3D_TRY(read_u64(ctx, H.failed, &failed));
        RS3D_HIP(hipMemcpy(hb, bad, sizeof hb, hipMemcpyDeviceToHost));
        if (hb[0]) {
            set_error("batch: object ids of pair queries must be < 2^24");
            return RS3D_E_ARG;
        }
        if (overflow) *overflow = hb[1];
        if (failed == 0) {
            uint64_t occ = 0;
            RS3D_TRY(keys_out.ensure(slots * sizeof(uint64_t)));
            RS3D_TRY(ht_compact_internal(ctx, table.as<uint64_t>(), slots, keys_out.as<uint64_t>(), slots, &occ));
            if (occ * 2 <= slots) {
                uint64_t u = 0;
                RS3D_TRY(rs3d_sort_unique_u64(ctx, keys_out.as<uint64_t>(), occ, &u));
                std::vector<uint64_t> hk(u);
                if (u) RS3D_HIP(hipMemcpy(hk.data(), keys_out.ptr, u * sizeof(uint64_t), hipMemcpyDeviceToHost));
                for (int q = 0; q <= num_qids; ++q)
                    offsets[q] = (uint64_t)(std::lower_bound(hk.begin(), hk.end(), (uint64_t)q << 48) - hk.begin());
                *n_keys = u;
                return RS3D_OK;
            }
        }
        slots <<= 2;
        if (slots > (1ull << 31)) {
            set_error("batch: pair table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
}

int batch_keys_to_pairs(rs3d_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out) {
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_batch_split<<<nb, 256, 0, ctx->stream>>>(keys, n, out);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_batch_execute(rs3d_ctx* ctx, const rs3d_batch_query* queries, int num_queries,
                       uint64_t* pair_keys_out, uint64_t max_pairs, uint64_t* pair_offsets,
                       uint64_t* total_pairs) {
    if (!ctx || num_queries < 0 || (num_queries && !queries) || !pair_offsets ||
        (max_pairs && !pair_keys_out) || num_queries >= (1 << 15)) {
        set_error("rs3d_batch_execute: bad arguments");
        return RS3D_E_ARG;
    }
    if (total_pairs) *total_pairs = 0;
    static thread_local BatchPlan plan(nullptr);
    plan.reset(ctx);
    for (int q = 0; q < num_queries; ++q) {
        const rs3d_batch_query& Q = queries[q];
        if (Q.kind == RS3D_BATCH_CONTAIN) {
            if (!Q.bvh1 || (Q.num_points && (!Q.points || !Q.out_obj)) || Q.axis < 0 || Q.axis > 2) {
                set_error("rs3d_batch_execute: bad containment query %d", q);
                return RS3D_E_ARG;
            }
            RS3D_TRY(plan.add_contain(Q.bvh1, Q.points, Q.num_points, Q.out_obj, Q.axis, q));
        } else if (Q.kind == RS3D_BATCH_OVERLAP) {
            if (!Q.bvh1 || !Q.bvh2 || !Q.edges1 || !Q.edges2) {
                set_error("rs3d_batch_execute: bad overlap query %d", q);
                return RS3D_E_ARG;
            }
            RS3D_TRY(plan.add_edges(Q.bvh2, Q.edges1, 0, q));
            RS3D_TRY(plan.add_edges(Q.bvh1, Q.edges2, 1, q));
        } else {
            set_error("rs3d_batch_execute: unknown query kind %d", Q.kind);
            return RS3D_E_ARG;
        }
    }
    plan.num_qids = num_queries;
    static thread_local DevBuf keys;
    uint64_t n = 0;
    std::vector<uint64_t> offs;
    RS3D_TRY(plan.run(keys, &n, offs, nullptr));
    for (int q = 0; q <= num_queries; ++q) pair_offsets[q] = q < (int)offs.size() ? offs[q] : n;
    if (total_pairs) *total_pairs = n;
    if (n > max_pairs) {
        set_error("rs3d_batch_execute: %llu pairs exceed the output capacity %llu",
                  (unsigned long long)n, (unsigned long long)max_pairs);
        return RS3D_E_CAPACITY;
    }
    RS3D_TRY(batch_keys_to_pairs(ctx, keys.as<uint64_t>(), n, pair_keys_out));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

}  // extern "C"
