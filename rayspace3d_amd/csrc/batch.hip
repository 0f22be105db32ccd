// batch.hip -- many independent containment and overlap queries in one persistent launch.
//
// Every query is cut into tasks of TASK_RAYS rays (points of a containment query, edges
// of one direction of an overlap query, objects of an intersection fallback); one
// persistent kernel of wave-sized workgroups takes tasks from striped atomic queue heads
// until all are done, so small and large queries share the GPU without a launch each.
// A containment task runs the lane-refill point program; edge and object tasks run 64
// rays at a time. Containment rays of all queries are first reordered by one radix sort
// on (query, 2-D Morton of the origin) so waves stay coherent; overlap pairs of all
// queries go into one hash table keyed (query << 48 | a << 24 | b), are compacted,
// radix-sorted once and split per query. Per ray the same programs as the single-query
// kernels run, so every result is identical to running the query alone.
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

// Sort keys of containment rays: 16-bit Morton of (u, w) on the query's own grid, with the
// descriptor index in the bits above unless each query is sorted in its own segment. One
// workgroup per containment task (the first tasks of the table).
__global__ __launch_bounds__(TASK_RAYS) void k_batch_keys(const BatchDesc* __restrict__ qs,
                                                          const uint64_t* __restrict__ task_off,
                                                          const uint32_t* __restrict__ task_q,
                                                          const float* __restrict__ lo_sc,
                                                          int query_bits,
                                                          uint32_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals) {
    const uint64_t t = blockIdx.x;
    const uint32_t qa = task_q[t];
    const BatchDesc& q = qs[qa];
    if (q.grid.cells) return;  // grid queries are not reordered
    const uint64_t i = (t - task_off[qa]) * TASK_RAYS + threadIdx.x;
    if (i >= q.n) return;
    const int U = q.axis == 0 ? 1 : (q.axis == 1 ? 2 : 0);
    const int W = q.axis == 0 ? 2 : (q.axis == 1 ? 0 : 1);
    const float u = q.pts[3 * i + U], w = q.pts[3 * i + W];
    const float fu = fminf(fmaxf((u - lo_sc[4 * qa]) * lo_sc[4 * qa + 2], 0.0f), 255.0f);
    const float fw = fminf(fmaxf((w - lo_sc[4 * qa + 1]) * lo_sc[4 * qa + 3], 0.0f), 255.0f);
    keys[q.ray_base + i] = (query_bits ? (qa << 16) : 0u) | (spread2_11((uint32_t)fu) << 1) |
                           spread2_11((uint32_t)fw);
    vals[q.ray_base + i] = (uint32_t)i;
}

// task_q[t] = the query owning task t (task_off[q] <= t < task_off[q + 1]).
__global__ __launch_bounds__(256) void k_task_owner(const uint64_t* __restrict__ task_off,
                                                    uint32_t* __restrict__ task_q) {
    const uint32_t q = blockIdx.x;
    for (uint64_t t = task_off[q] + threadIdx.x; t < task_off[q + 1]; t += blockDim.x) task_q[t] = q;
}

// Persistent batch kernel over tasks [t0, t1). KINDS selects the ray programs compiled in
// (1: containment, 2: edges and object rays, 3: both) and AX the containment ray axis (-1:
// each query's own): a kernel holding only one point program needs fewer registers and
// keeps more waves resident.
template <int DEPTH, int KINDS, int AX>
__device__ __forceinline__ void batch_loop(const BatchDesc* __restrict__ qs,
                                                    const uint64_t* __restrict__ task_off,
                                                    const uint32_t* __restrict__ task_q,
                                                    uint64_t t0, uint64_t t1,
                                                    unsigned int* __restrict__ heads,
                                                    const uint32_t* __restrict__ perm, HashArgs H,
                                                    unsigned long long* __restrict__ bad_ids, int* stack_mem) {
    int* stack = stack_mem + threadIdx.x;
    const int lane = threadIdx.x;
    TraceCounters ctr;
    uint32_t stripe = blockIdx.x % NSTRIPE;
    for (int dry = 0; dry < NSTRIPE;) {
        unsigned int k = 0;
        if (lane == 0) k = atomicAdd(&heads[stripe * STRIPE_PITCH], 1u);
        k = __builtin_amdgcn_readfirstlane(k);  // wave-uniform: descriptor loads stay scalar
        const uint64_t t = t0 + (uint64_t)k * NSTRIPE + stripe;
        if (t >= t1) {
            stripe = stripe + 1 == NSTRIPE ? 0 : stripe + 1;
            ++dry;
            continue;
        }
        const uint32_t qa = task_q[t];
        const BatchDesc& q = qs[qa];
        const uint64_t lo = (t - task_off[qa]) * q.task_rays;
        const uint64_t hi = lo + q.task_rays < q.n ? lo + q.task_rays : q.n;
        if ((KINDS & 1) && q.kind == BQ_CONTAIN && q.grid.cells) {
            // the two-phase grid form with the item-parallel walk; its LDS is the (idle)
            // traversal stack memory
            const int ax = AX >= 0 ? AX : q.axis;
            AgItemLds* L = reinterpret_cast<AgItemLds*>(stack_mem);
            uint32_t cand = 0;
            GridFilter gf{};
            int32_t* out = static_cast<int32_t*>(q.out);
            if (ax == 0) points_grid_two_phase<0, false>(q.grid, q.pts, nullptr, lo, hi, RS3D_OUT_OBJ_I32, out, cand, gf, *L);
            else if (ax == 1) points_grid_two_phase<1, false>(q.grid, q.pts, nullptr, lo, hi, RS3D_OUT_OBJ_I32, out, cand, gf, *L);
            else points_grid_two_phase<2, false>(q.grid, q.pts, nullptr, lo, hi, RS3D_OUT_OBJ_I32, out, cand, gf, *L);
            continue;
        }
        if ((KINDS & 1) && q.kind == BQ_CONTAIN) {
            const uint32_t* im = perm ? perm + q.ray_base : nullptr;
            const int ax = AX >= 0 ? AX : q.axis;
            if (ax == 0)
                points_refill<0, DEPTH>(q.nodes, q.tris, q.pts, im, lo, hi, RS3D_OUT_OBJ_I32, q.out, stack);
            else if (ax == 1)
                points_refill<1, DEPTH>(q.nodes, q.tris, q.pts, im, lo, hi, RS3D_OUT_OBJ_I32, q.out, stack);
            else
                points_refill<2, DEPTH>(q.nodes, q.tris, q.pts, im, lo, hi, RS3D_OUT_OBJ_I32, q.out, stack);
            continue;
        }
        if (!(KINDS & 2)) continue;
        for (uint64_t i = lo + lane; i < hi; i += TR_BLOCK) {
            if (q.kind == BQ_OBJECTS) {
                const int ft = q.first_tri[i];
                if (ft == 0x7fffffff) continue;
                const float3 o = ld3f(q.src_v, q.src_i[3 * (uint64_t)ft]);
                int* ids = q.cand_ids + i * (uint64_t)q.max_targets;
                uint8_t* par = q.cand_par + i * (uint64_t)q.max_targets;
                unsigned long long over = 0;
                const int cnt = object_ray_parity<DEPTH>(q.nodes, q.tris, o, stack, ids, par,
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
                edge_ray<DEPTH, false>(q.nodes, q.tris, q.occ, ld3f(q.E.starts, i), ld3f(q.E.ends, i), CHAIN_EPS,
                                       100, stack, ctr, [&](int tgt) {
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
}

// Containment-only launches fit 5 waves per SIMD (91 VGPRs); the compiler is asked for
// that occupancy explicitly. Pair launches keep the default register budget.
template <int DEPTH, int AX>
__global__ __launch_bounds__(TR_BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_batch_points(
    const BatchDesc* __restrict__ qs, const uint64_t* __restrict__ task_off,
    const uint32_t* __restrict__ task_q, uint64_t t0, uint64_t t1, unsigned int* __restrict__ heads,
    const uint32_t* __restrict__ perm, HashArgs H, unsigned long long* __restrict__ bad_ids) {
    static_assert(DEPTH * TR_BLOCK * sizeof(int) >= sizeof(AgItemLds), "the grid walk's LDS is the stack memory");
    __shared__ __align__(16) int stack_mem[DEPTH * TR_BLOCK];
    batch_loop<DEPTH, 1, AX>(qs, task_off, task_q, t0, t1, heads, perm, H, bad_ids, stack_mem);
}

template <int DEPTH, int KINDS, int AX>
__global__ __launch_bounds__(TR_BLOCK) void k_batch(
    const BatchDesc* __restrict__ qs, const uint64_t* __restrict__ task_off,
    const uint32_t* __restrict__ task_q, uint64_t t0, uint64_t t1, unsigned int* __restrict__ heads,
    const uint32_t* __restrict__ perm, HashArgs H, unsigned long long* __restrict__ bad_ids) {
    static_assert(DEPTH * TR_BLOCK * sizeof(int) >= sizeof(AgItemLds), "the grid walk's LDS is the stack memory");
    __shared__ __align__(16) int stack_mem[DEPTH * TR_BLOCK];
    batch_loop<DEPTH, KINDS, AX>(qs, task_off, task_q, t0, t1, heads, perm, H, bad_ids, stack_mem);
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
    bool grid = false;
    if (n <= 0xffffffffull)
        RS3D_TRY(axis_grid_ready(ctx_, const_cast<rs3d_bvh*>(b), axis, &grid, build_structures));
    if (grid) {  // rays in the caller's order on the grid: no reorder keys for this query
        const AxisGrid* g = b->agrid[axis];
        d.grid = ag_args(g);
        items.push_back(d);
        return RS3D_OK;
    }
    d.ray_base = contain_rays;
    contain_rays += n;
    const int U = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
    const int W = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
    const float eu = b->bmax[U] - b->bmin[U], ew = b->bmax[W] - b->bmin[W];
    d.lo_sc[0] = b->bmin[U];
    d.lo_sc[1] = b->bmin[W];
    int qb = 1;  // 4^qb cells of about 64..256 rays, at most 256 x 256
    while (qb < 8 && ((uint64_t)64 << (2 * (qb + 1))) <= n) ++qb;
    const float cells = (float)(1u << qb);
    d.lo_sc[2] = eu > 0 ? cells / eu : 0.0f;
    d.lo_sc[3] = ew > 0 ? cells / ew : 0.0f;
    items.push_back(d);
    return RS3D_OK;
}

int BatchPlan::add_edges(const rs3d_bvh* target, const rs3d_edges* e, int swap, int qid) {
    if (!e || e->num_edges == 0 || target->num_tris == 0) return RS3D_OK;
    BatchDesc d{};
    RS3D_TRY(base(d, target, BQ_EDGES, qid));
    d.E = EdgeArgs{e->starts, e->ends, e->src_offsets, e->src_counts, e->src_objects, e->num_edges};
    RS3D_TRY(seg_occ_args(ctx_, const_cast<rs3d_bvh*>(target), &d.occ, build_structures));
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

// Host tables (descriptors, Morton frames, task offsets) and their upload; the device
// layout stays fixed until the next prepare, so a captured device pass stays valid.
int BatchPlan::prepare() {
    rs3d_ctx* ctx = ctx_;
    // containment first, grouped by ray axis; their rays take consecutive ranges of the
    // reordered ray list in this order
    std::stable_sort(items.begin(), items.end(), [](const BatchItem& a, const BatchItem& b) {
        const int ka = a.kind == BQ_CONTAIN ? a.axis : 3, kb = b.kind == BQ_CONTAIN ? b.axis : 3;
        return ka < kb;
    });
    // Reorder layout: when whole-chunk padding of every containment query costs at most a
    // quarter more keys, each query gets its own segment (a multiple of RS_SEG_KEYS keys,
    // padded with the largest key) and a segmented 16-bit sort replaces the global sort
    // on (query, Morton) - one radix pass less.
    // (queries on an axis grid take no part in the reorder)
    n_contain = 0;
    uint64_t padded = 0;
    for (auto& d : items)
        if (d.kind == BQ_CONTAIN) {
            ++n_contain;
            if (!d.grid.cells) padded += (d.n + RS_SEG_KEYS - 1) / RS_SEG_KEYS * RS_SEG_KEYS;
        }
    segmented = contain_rays > 0 && padded <= contain_rays + contain_rays / 4;
    uint64_t rb = 0;
    for (auto& d : items)
        if (d.kind == BQ_CONTAIN && !d.grid.cells) {
            d.ray_base = rb;
            rb += segmented ? (d.n + RS_SEG_KEYS - 1) / RS_SEG_KEYS * RS_SEG_KEYS : d.n;
        }
    sorted_rays = rb;
    for (int a = 0; a <= 3; ++a) axis_first[a] = 0;
    {
        uint32_t k = 0;
        for (int a = 0; a < 3; ++a) {
            axis_first[a] = k;
            while (k < n_contain && items[k].axis == a) ++k;
        }
        axis_first[3] = n_contain;
    }
    if (n_contain >= (1u << 10)) {
        set_error("batch: at most 1023 containment queries per launch");
        return RS3D_E_ARG;
    }
    if (sorted_rays > 0xffffffffull) {
        set_error("batch: more than 2^32-1 containment rays");
        return RS3D_E_ARG;
    }
    nq = (uint32_t)items.size();
    task_off.assign(nq + 1, 0);
    for (uint32_t k = 0; k < nq; ++k) {
        items[k].task_rays = items[k].grid.cells ? GRID_TASK_RAYS : TASK_RAYS;
        task_off[k + 1] = task_off[k] + (items[k].n + items[k].task_rays - 1) / items[k].task_rays;
    }
    ntasks = task_off[nq];
    if (task_off[n_contain] > 0x7fffffffull) {
        set_error("batch: too many containment tasks");
        return RS3D_E_ARG;
    }
    // one staging block: descriptors | Morton frames | task offsets | queue heads | flags
    const auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    // segment table of the segmented sort: per chunk {first chunk of its query, chunks}
    std::vector<uint32_t> segtab;
    if (segmented)
        for (uint32_t k = 0; k < n_contain; ++k) {
            if (items[k].grid.cells) continue;
            const uint32_t c0 = (uint32_t)(items[k].ray_base / RS_SEG_KEYS);
            const uint32_t nc = (uint32_t)((items[k].n + RS_SEG_KEYS - 1) / RS_SEG_KEYS);
            for (uint32_t c = 0; c < nc; ++c) {
                segtab.push_back(c0);
                segtab.push_back(nc);
            }
        }
    o_losc = up(nq * sizeof(BatchDesc));
    o_seg = o_losc + up(4 * (size_t)n_contain * sizeof(float));
    o_task = o_seg + up(segtab.size() * sizeof(uint32_t));
    o_head = o_task + up((nq + 1) * sizeof(uint64_t));
    o_bad = o_head + 4 * NSTRIPE * STRIPE_PITCH * sizeof(unsigned int);  // one queue per launch
    meta_bytes = o_bad + 256;
    host_meta.assign(o_head, 0);
    for (uint32_t k = 0; k < nq; ++k) {
        const BatchDesc d = items[k];
        std::memcpy(host_meta.data() + k * sizeof(BatchDesc), &d, sizeof d);
    }
    for (uint32_t k = 0; k < n_contain; ++k)
        std::memcpy(host_meta.data() + o_losc + 4 * k * sizeof(float), items[k].lo_sc, 4 * sizeof(float));
    if (!segtab.empty())
        std::memcpy(host_meta.data() + o_seg, segtab.data(), segtab.size() * sizeof(uint32_t));
    std::memcpy(host_meta.data() + o_task, task_off.data(), (nq + 1) * sizeof(uint64_t));
    RS3D_TRY(meta.ensure(meta_bytes));
    RS3D_HIP(hipMemcpyAsync(meta.ptr, host_meta.data(), o_head, hipMemcpyHostToDevice, ctx->stream));
    RS3D_TRY(task_q.ensure((ntasks + 1) * sizeof(uint32_t)));
    if (nq > 0) {
        k_task_owner<<<nq, 256, 0, ctx->stream>>>(reinterpret_cast<const uint64_t*>(meta.as<uint8_t>() + o_task),
                                                  task_q.as<uint32_t>());
        RS3D_HIP(hipGetLastError());
    }
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

template <int KINDS, int AX>
int BatchPlan::launch_batch(const BatchDesc* dp, const uint64_t* tp, uint64_t t0, uint64_t t1,
                            unsigned int* heads, const uint32_t* perm, const HashArgs& H,
                            unsigned long long* bad) {
    if (t1 <= t0) return RS3D_OK;
    rs3d_ctx* ctx = ctx_;
    const unsigned waves = (unsigned)std::max(ctx->num_cus, 1) * 20u;
    const unsigned grid = (unsigned)std::min<uint64_t>(waves, t1 - t0);
    const uint32_t* tq = task_q.as<uint32_t>();
    switch (depth) {
        case 24:
            if constexpr (KINDS == 1) k_batch_points<24, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            else k_batch<24, KINDS, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            break;
        case 32:
            if constexpr (KINDS == 1) k_batch_points<32, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            else k_batch<32, KINDS, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            break;
        case 48:
            if constexpr (KINDS == 1) k_batch_points<48, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            else k_batch<48, KINDS, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            break;
        default:
            if constexpr (KINDS == 1) k_batch_points<64, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            else k_batch<64, KINDS, AX><<<grid, TR_BLOCK, 0, ctx->stream>>>(dp, tp, tq, t0, t1, heads, perm, H, bad);
            break;
    }
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

// The device work of one execution: ray keys and reorder sort, pair-table clear, queue
// reset, persistent launch. Kernel launches, kernel fills and device copies only (no host
// synchronisation, no allocation once the buffers have their size), so it can be captured.
int BatchPlan::device_pass(uint64_t slots) {
    rs3d_ctx* ctx = ctx_;
    uint8_t* mb = meta.as<uint8_t>();
    const BatchDesc* dp = reinterpret_cast<const BatchDesc*>(mb);
    unsigned int* heads = reinterpret_cast<unsigned int*>(mb + o_head);
    unsigned long long* bad = reinterpret_cast<unsigned long long*>(mb + o_bad);
    const uint64_t* tp = reinterpret_cast<const uint64_t*>(mb + o_task);
    const uint32_t* perm = nullptr;
    if (contain_rays > 0 && ctx->opt_ray_sort_min > 0) {
        RS3D_TRY(ctx->ray_keys.ensure(sorted_rays * sizeof(uint32_t)));
        RS3D_TRY(ctx->ray_vals.ensure(sorted_rays * sizeof(uint32_t)));
        if (segmented)  // padding keys sort after every real key of their segment
            RS3D_TRY(fill_u32(ctx, ctx->ray_keys.ptr, sorted_rays, 0xffffu));
        k_batch_keys<<<(unsigned)task_off[n_contain], TASK_RAYS, 0, ctx->stream>>>(
            dp, tp, task_q.as<uint32_t>(), reinterpret_cast<const float*>(mb + o_losc),
            segmented ? 0 : 1, ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>());
        RS3D_HIP(hipGetLastError());
        uint32_t *sk = nullptr, *sv = nullptr;
        if (segmented) {
            RS3D_TRY(radix_sort_u32_pairs_to(ctx, ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>(),
                                             sorted_rays, 16, &sk, &sv,
                                             reinterpret_cast<const uint2*>(mb + o_seg)));
        } else {
            int qbits = 0;  // the descriptor index occupies the bits above the 16 Morton bits
            while (qbits < 10 && (1u << qbits) < n_contain) ++qbits;
            RS3D_TRY(radix_sort_u32_pairs_to(ctx, ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>(),
                                             sorted_rays, 16 + qbits, &sk, &sv));
        }
        perm = sv;
    }
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
    RS3D_TRY(fill_u32(ctx, heads, (meta_bytes - o_head) / sizeof(unsigned int), 0u));
    // Containment tasks come first in the task table ([0, ct)), pair tasks after. By default
    // each kind runs in its own persistent launch (the point kernel without the edge and
    // object programs keeps 5 waves per SIMD instead of 4); split_kinds = false runs every
    // task in one launch.
    // Containment queries are ordered by axis, so each axis's tasks form one range and get
    // a kernel specialised to that axis.
    const uint64_t ct = task_off[n_contain];
    if (split_kinds) {
        for (int a = 0; a < 3; ++a) {
            unsigned int* hq = heads + (size_t)a * NSTRIPE * STRIPE_PITCH;
            const uint64_t a0 = task_off[axis_first[a]], a1 = task_off[axis_first[a + 1]];
            if (a == 0) RS3D_TRY((launch_batch<1, 0>(dp, tp, a0, a1, hq, perm, H, bad)));
            if (a == 1) RS3D_TRY((launch_batch<1, 1>(dp, tp, a0, a1, hq, perm, H, bad)));
            if (a == 2) RS3D_TRY((launch_batch<1, 2>(dp, tp, a0, a1, hq, perm, H, bad)));
        }
        RS3D_TRY((launch_batch<2, 2>(dp, tp, ct, ntasks, heads + 3 * NSTRIPE * STRIPE_PITCH, perm, H, bad)));
    } else {
        RS3D_TRY((launch_batch<3, -1>(dp, tp, 0, ntasks, heads, perm, H, bad)));
    }
    return RS3D_OK;
}

// After a device pass: checks, pair-table compaction, sort/unique and per-query split.
// *grow is set when the table overflowed or is more than half full (the pass must be
// rerun with a larger table).
int BatchPlan::finish(uint64_t slots, DevBuf& keys_out, uint64_t* n_keys,
                      std::vector<uint64_t>& offsets, uint64_t* overflow, bool* grow) {
    rs3d_ctx* ctx = ctx_;
    *grow = false;
    *n_keys = 0;
    if (overflow) *overflow = 0;
    offsets.assign(num_qids + 1, 0);
    if (!any_pairs) {
        RS3D_HIP(hipStreamSynchronize(ctx->stream));
        return RS3D_OK;
    }
    unsigned long long hb[2] = {0, 0};
    uint64_t failed = 0;
    RS3D_TRY(read_u64(ctx, table.as<uint8_t>() + slots * sizeof(uint64_t), &failed));
    RS3D_HIP(hipMemcpy(hb, meta.as<uint8_t>() + o_bad, sizeof hb, hipMemcpyDeviceToHost));
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
    *grow = true;
    return RS3D_OK;
}

int BatchPlan::run(DevBuf& keys_out, uint64_t* n_keys, std::vector<uint64_t>& offsets,
                   uint64_t* overflow) {
    RS3D_TRY(prepare());
    uint64_t slots = 1024;
    while (slots < slot_hint) slots <<= 1;
    while (true) {
        RS3D_TRY(device_pass(slots));
        bool grow = false;
        RS3D_TRY(finish(slots, keys_out, n_keys, offsets, overflow, &grow));
        if (!grow) {
            slots_used = slots;
            return RS3D_OK;
        }
        slots <<= 2;
        if (slots > (1ull << 31)) {
            set_error("batch: pair table would exceed 2^31 slots");
            return RS3D_E_CAPACITY;
        }
    }
}

int BatchPlan::capture(hipGraphExec_t* exec) {
    rs3d_ctx* ctx = ctx_;
    hipGraph_t graph = nullptr;
    RS3D_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = device_pass(slots_used);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != RS3D_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) {
        set_error("batch capture: %s", hipGetErrorString(e));
        return RS3D_E_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        set_error("batch graph instantiate: %s", hipGetErrorString(ei));
        return RS3D_E_HIP;
    }
    return RS3D_OK;
}

int BatchPlan::replay(hipGraphExec_t exec, DevBuf& keys_out, uint64_t* n_keys,
                      std::vector<uint64_t>& offsets, uint64_t* overflow, bool* grow) {
    RS3D_HIP(hipGraphLaunch(exec, ctx_->stream));
    return finish(slots_used, keys_out, n_keys, offsets, overflow, grow);
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

// Adds the queries to a plan (shared by rs3d_batch_execute and rs3d_batch_create).
static int add_queries(BatchPlan& plan, const rs3d_batch_query* queries, int num_queries,
                       const char* who) {
    for (int q = 0; q < num_queries; ++q) {
        const rs3d_batch_query& Q = queries[q];
        if (Q.kind == RS3D_BATCH_CONTAIN) {
            if (!Q.bvh1 || (Q.num_points && (!Q.points || !Q.out_obj)) || Q.axis < 0 || Q.axis > 2) {
                set_error("%s: bad containment query %d", who, q);
                return RS3D_E_ARG;
            }
            RS3D_TRY(plan.add_contain(Q.bvh1, Q.points, Q.num_points, Q.out_obj, Q.axis, q));
        } else if (Q.kind == RS3D_BATCH_OVERLAP) {
            if (!Q.bvh1 || !Q.bvh2 || !Q.edges1 || !Q.edges2) {
                set_error("%s: bad overlap query %d", who, q);
                return RS3D_E_ARG;
            }
            RS3D_TRY(plan.add_edges(Q.bvh2, Q.edges1, 0, q));
            RS3D_TRY(plan.add_edges(Q.bvh1, Q.edges2, 1, q));
        } else {
            set_error("%s: unknown query kind %d", who, Q.kind);
            return RS3D_E_ARG;
        }
    }
    plan.num_qids = num_queries;
    return RS3D_OK;
}

// Per-query offsets, capacity check and (query, a, b) -> (a, b) conversion of a finished
// execution into the caller's buffers.
static int deliver_pairs(rs3d_ctx* ctx, const DevBuf& keys, uint64_t n, const std::vector<uint64_t>& offs,
                         int num_queries, uint64_t* pair_keys_out, uint64_t max_pairs,
                         uint64_t* pair_offsets, uint64_t* total_pairs, const char* who) {
    for (int q = 0; q <= num_queries; ++q) pair_offsets[q] = q < (int)offs.size() ? offs[q] : n;
    if (total_pairs) *total_pairs = n;
    if (n > max_pairs) {
        set_error("%s: %llu pairs exceed the output capacity %llu", who, (unsigned long long)n,
                  (unsigned long long)max_pairs);
        return RS3D_E_CAPACITY;
    }
    RS3D_TRY(batch_keys_to_pairs(ctx, keys.as<uint64_t>(), n, pair_keys_out));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

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
    RS3D_TRY(add_queries(plan, queries, num_queries, "rs3d_batch_execute"));
    static thread_local DevBuf keys;
    uint64_t n = 0;
    std::vector<uint64_t> offs;
    RS3D_TRY(plan.run(keys, &n, offs, nullptr));
    return deliver_pairs(ctx, keys, n, offs, num_queries, pair_keys_out, max_pairs, pair_offsets,
                         total_pairs, "rs3d_batch_execute");
}

}  // extern "C"

// A prepared batch: its own scratch (a private context on the caller's stream, so other
// calls on the caller's context cannot move buffers the graph points into), the plan, and
// the captured device pass.
struct rs3d_batch {
    rs3d_ctx* user = nullptr;
    rs3d_ctx priv;
    BatchPlan plan{nullptr};
    DevBuf keys;
    hipGraphExec_t exec = nullptr;
    int num_queries = 0;
};

static void batch_release(rs3d_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->plan.release();
    b->keys.release();
    rs3d_ctx& p = b->priv;
    for (DevBuf* d : {&p.sort_keys, &p.sort_vals, &p.sort_hist, &p.sort_csum, &p.scan_tmp, &p.tmp_a, &p.tmp_b,
                      &p.tmp_c, &p.small, &p.ray_keys, &p.ray_vals, &p.accs, &p.edge_sel})
        d->release();
    if (p.pinned) (void)hipHostFree(p.pinned);
    p.pinned = nullptr;
}

extern "C" {

int rs3d_batch_create(rs3d_ctx* ctx, const rs3d_batch_query* queries, int num_queries,
                      rs3d_batch** out) {
    if (!ctx || !out || num_queries < 0 || (num_queries && !queries) || num_queries >= (1 << 15)) {
        set_error("rs3d_batch_create: bad arguments");
        return RS3D_E_ARG;
    }
    *out = nullptr;
    RS3D_HIP(hipSetDevice(ctx->device));
    rs3d_batch* b = new rs3d_batch();
    b->user = ctx;
    b->num_queries = num_queries;
    rs3d_ctx& p = b->priv;
    p.device = ctx->device;
    p.stream = ctx->stream;
    p.tmp_pool = ctx->tmp_pool;  // borrowed: the caller's context owns it
    p.num_cus = ctx->num_cus;
    p.opt_ray_sort_min = ctx->opt_ray_sort_min;
    p.opt_rays_per_wave = ctx->opt_rays_per_wave;
    p.opt_axis_grid = ctx->opt_axis_grid;
    p.opt_grid_rays_per_wave = ctx->opt_grid_rays_per_wave;
    p.opt_grid_cell_pct = ctx->opt_grid_cell_pct;
    p.opt_grid_cell_pct_set = ctx->opt_grid_cell_pct_set;
    p.opt_seg_occupancy = ctx->opt_seg_occupancy;
    p.opt_seg_cell_pct = ctx->opt_seg_cell_pct;
    if (hipHostMalloc((void**)&p.pinned, CTX_PINNED_BYTES, hipHostMallocDefault) != hipSuccess) {
        delete b;
        set_error("rs3d_batch_create: hipHostMalloc failed");
        return RS3D_E_NOMEM;
    }
    b->plan.reset(&p);
    int rc = add_queries(b->plan, queries, num_queries, "rs3d_batch_create");
    uint64_t n = 0;
    std::vector<uint64_t> offs;
    if (rc == RS3D_OK) rc = b->plan.run(b->keys, &n, offs, nullptr);  // sizes every buffer
    if (rc == RS3D_OK) rc = b->plan.capture(&b->exec);
    if (rc != RS3D_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return RS3D_OK;
}

int rs3d_batch_run(rs3d_batch* b, uint64_t* pair_keys_out, uint64_t max_pairs,
                   uint64_t* pair_offsets, uint64_t* total_pairs) {
    if (!b || !pair_offsets || (max_pairs && !pair_keys_out)) {
        set_error("rs3d_batch_run: bad arguments");
        return RS3D_E_ARG;
    }
    if (total_pairs) *total_pairs = 0;
    RS3D_HIP(hipSetDevice(b->priv.device));
    uint64_t n = 0;
    std::vector<uint64_t> offs;
    bool grow = false;
    RS3D_TRY(b->plan.replay(b->exec, b->keys, &n, offs, nullptr, &grow));
    if (grow) {
        // more pairs than the captured table holds: run uncaptured with a growing table,
        // then capture again at the new size
        RS3D_TRY(b->plan.run(b->keys, &n, offs, nullptr));
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
        RS3D_TRY(b->plan.capture(&b->exec));
    }
    return deliver_pairs(&b->priv, b->keys, n, offs, b->num_queries, pair_keys_out, max_pairs,
                         pair_offsets, total_pairs, "rs3d_batch_run");
}

int rs3d_batch_destroy(rs3d_batch* b) {
    if (!b) return RS3D_OK;
    (void)hipSetDevice(b->priv.device);
    (void)hipStreamSynchronize(b->priv.stream);
    batch_release(b);
    delete b;
    return RS3D_OK;
}

}  // extern "C"
