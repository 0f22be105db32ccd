// trace.hip -- wavefront BVH traversal kernels (replace the OptiX raygen/closest-hit
// programs of query/src/cuda/raytracing.cu and query/src/cuda/mesh_overlap_edges.cu).
//
// One lane per ray, 64-lane workgroups, per-lane traversal stack in LDS (stride 64 so the
// lanes of a wave hit distinct banks). Each traversal collects the KBUF smallest
// (t, triangle) hits at or beyond the current chain threshold in a register-resident
// sorted buffer; the closest-hit chain of the reference is then walked inside that
// buffer. When more than KBUF hits were seen the traversal restarts from the threshold,
// which reproduces the chain exactly for any hit count. Results therefore do not depend
// on the BVH shape, leaf size or stack depth.
#include <atomic>
#include <chrono>
#include <cstring>
#include <type_traits>

#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

// A.1 point containment (query/src/cuda/raytracing.cu:22-77 with the canonical chain).
template <int AX, int DEPTH, bool STATS>
__global__ __launch_bounds__(TR_BLOCK) void k_points(const BNode* __restrict__ nodes,
                                                     const TriRec* __restrict__ tris,
                                                     const float* __restrict__ pts, uint64_t n,
                                                     const uint32_t* __restrict__ imap, int out_kind,
                                                     void* __restrict__ out,
                                                     unsigned long long* __restrict__ stats) {
    __shared__ int stack_mem[DEPTH * TR_BLOCK];
    const uint64_t i = (uint64_t)blockIdx.x * TR_BLOCK + threadIdx.x;
    TraceCounters ctr;
    if (i < n) {
        const uint64_t pi = imap ? (uint64_t)imap[i] : i;
        const int obj = point_ray<AX, DEPTH, STATS>(nodes, tris, ld3f(pts, pi), stack_mem + threadIdx.x, ctr);
        if (out_kind == RS3D_OUT_RAY_RESULT) {
            rs3d_ray_result r;
            r.ray_id = (int32_t)pi;
            r.polygon_index = obj;
            static_cast<rs3d_ray_result*>(out)[pi] = r;
        } else if (out_kind == RS3D_OUT_FLAG_U8) {
            static_cast<uint8_t*>(out)[pi] = obj >= 0 ? 1 : 0;
        } else {
            static_cast<int32_t*>(out)[pi] = obj;
        }
    }
    if (STATS) flush_counters(ctr, stats);
}

// Lane-refill form of k_points (see points_refill): wave b owns rays
// [b * per_wave, (b + 1) * per_wave) of the ray list.
template <int AX, int DEPTH>
__global__ __launch_bounds__(TR_BLOCK) void k_points_refill(const BNode* __restrict__ nodes,
                                                            const TriRec* __restrict__ tris,
                                                            const float* __restrict__ pts,
                                                            uint64_t n,
                                                            const uint32_t* __restrict__ imap,
                                                            int out_kind, void* __restrict__ out,
                                                            uint32_t per_wave) {
    __shared__ int stack_mem[DEPTH * TR_BLOCK];
    const uint64_t lo = (uint64_t)blockIdx.x * per_wave;
    const uint64_t hi = lo + per_wave < n ? lo + per_wave : n;
    points_refill<AX, DEPTH>(nodes, tris, pts, imap, lo, hi, out_kind, out, stack_mem + threadIdx.x);
}

// Filter flags of the filter-refine pipeline (the per-point test is box_filter_pass).
template <int AX>
__global__ __launch_bounds__(256) void k_filter_box(const float* __restrict__ pts, uint64_t n, FBox box,
                                                    uint8_t* __restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = box_filter_pass<AX>(ld3f(pts, i), box) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_count_nonneg(const int32_t* __restrict__ p, uint64_t n,
                                                      unsigned long long* __restrict__ cnt) {
    // 16-byte loads over the aligned body; the first `head` values (up to the first 16-byte
    // boundary) and the tail one at a time
    const uint64_t head = min(n, (uint64_t)((16u - ((uint32_t)(uintptr_t)p & 15u)) & 15u) / 4);
    const uint64_t n4 = (n - head) / 4;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t c = 0;
    const int4* p4 = reinterpret_cast<const int4*>(p + head);
    for (uint64_t i = tid; i < n4; i += stride) {
        const int4 v = p4[i];
        c += (v.x >= 0) + (v.y >= 0) + (v.z >= 0) + (v.w >= 0);
    }
    if (tid < head) c += p[tid] >= 0 ? 1 : 0;
    for (uint64_t i = head + n4 * 4 + tid; i < n; i += stride) c += p[i] >= 0 ? 1 : 0;
    const unsigned long long w = block_sum_256(c);
    if (threadIdx.x == 0 && w) atomicAdd(acc_stripe(cnt), w);
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, uint64_t n, int32_t v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------------------
// Edge rays (A.2): segment from start towards end, hits in [thr, len + 1e-6].
// ------------------------------------------------------------------------------------
// One edge ray of the A.2 join and its emissions in the mode's form (count, pair or key
// write at the edge's offset, hash insert).
template <int MODE, int DEPTH, bool STATS>
__device__ __forceinline__ void edge_one(const BNode* __restrict__ nodes, const TriRec* __restrict__ tris,
                                         const OccArgs& occ, const EdgeArgs& E, uint64_t e, int swap,
                                         int chain_rule, int max_iter, int32_t* __restrict__ counts,
                                         const int64_t* __restrict__ offsets, int64_t offset_bias,
                                         void* __restrict__ out, const HashArgs& H, int* stack, TraceCounters& ctr) {
    const int nsrc = E.src_cnt[e];
    const int soff = E.src_off[e];
    int64_t cursor = 0;
    if (MODE == EMODE_WRITE_PAIRS || MODE == EMODE_WRITE_KEYS) cursor = offsets[e] + offset_bias;
    const int hits = edge_ray<DEPTH, STATS, (DEPTH < EDGE_LDS_DEPTH ? DEPTH : EDGE_LDS_DEPTH)>(
        nodes, tris, occ, ld3f(E.starts, e), ld3f(E.ends, e), chain_rule, max_iter, stack, ctr, [&](int tgt) {
            for (int q = 0; q < nsrc; ++q) {
                const int src = E.src_obj[soff + q];
                const uint32_t a = (uint32_t)(swap ? tgt : src);
                const uint32_t b = (uint32_t)(swap ? src : tgt);
                if (MODE == EMODE_WRITE_PAIRS) {
                    rs3d_pair pr;
                    pr.object_id_mesh1 = (int32_t)a;
                    pr.object_id_mesh2 = (int32_t)b;
                    static_cast<rs3d_pair*>(out)[cursor++] = pr;
                } else if (MODE == EMODE_WRITE_KEYS) {
                    static_cast<uint64_t*>(out)[cursor++] = ((uint64_t)a << 32) | b;
                } else if (MODE == EMODE_HASH) {
                    ht_insert(H, ((uint64_t)a << 32) | b);
                }
            }
        });
    if (MODE == EMODE_COUNT) counts[e] = hits * nsrc;
}

template <bool STATS>
__device__ __forceinline__ void flush_edge_counters(const TraceCounters& ctr, unsigned long long* stats) {
    if (!STATS) return;
    flush_counters(ctr, stats);
    uint64_t w = ctr.walked, o = ctr.occ_words;
    for (int off = 32; off > 0; off >>= 1) {
        w += __shfl_xor(w, off, 64);
        o += __shfl_xor(o, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&stats[4], (unsigned long long)w);
        atomicAdd(&stats[5], (unsigned long long)o);
    }
}

// Edge rays (A.2), one per lane: segment from start towards end, hits in [thr, len + 1e-6].
// With list, lane i runs edge list[i] of n_list; otherwise edge i of E. The walk waits on
// chains of dependent node fetches: eight waves per SIMD (64 registers, a few spilled; the
// stack mostly in LDS) hide more of that latency than seven without spills.
template <int MODE, int DEPTH, bool STATS>
__global__ __launch_bounds__(TR_BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_edges(const BNode* __restrict__ nodes,
                                                    const TriRec* __restrict__ tris, OccArgs occ, EdgeArgs E,
                                                    const uint32_t* __restrict__ list, uint64_t n_list,
                                                    int swap, int chain_rule, int max_iter,
                                                    int32_t* __restrict__ counts,
                                                    const int64_t* __restrict__ offsets,
                                                    int64_t offset_bias, void* __restrict__ out,
                                                    HashArgs H, unsigned long long* __restrict__ stats) {
    __shared__ int stack_mem[(DEPTH < EDGE_LDS_DEPTH ? DEPTH : EDGE_LDS_DEPTH) * TR_BLOCK];
    const uint64_t i = (uint64_t)blockIdx.x * TR_BLOCK + threadIdx.x;
    TraceCounters ctr;
    if (i < (list ? n_list : E.num_edges))
        edge_one<MODE, DEPTH, STATS>(nodes, tris, occ, E, list ? list[i] : i, swap, chain_rule, max_iter, counts,
                                     offsets, offset_bias, out, H, stack_mem + threadIdx.x, ctr);
    flush_edge_counters<STATS>(ctr, stats);
}

// Selection and compaction of the edges to walk, in one pass: block b tests edges
// [b * SEL_BLOCK_EDGES, (b + 1) * SEL_BLOCK_EDGES) against the occupancy bits (edge_may_hit)
// and writes the indices of those that may hit, in increasing order, at positions it
// claims with one atomic add on *count. Blocks claim in any order: the list holds each
// block's stretch of the (spatially ordered) edges in order.
constexpr int SEL_EPT = 4;  // edges per thread
constexpr int SEL_BLOCK_EDGES = 256 * SEL_EPT;
__global__ __launch_bounds__(256) void k_edge_select_list(EdgeArgs E, OccArgs occ, uint32_t* __restrict__ list,
                                                          uint32_t* __restrict__ count) {
    __shared__ uint32_t wcnt[SEL_EPT * 4];
    __shared__ uint32_t base_s;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t t0 = (uint64_t)blockIdx.x * SEL_BLOCK_EDGES;
    // the block's endpoints from a block base (32-bit offsets within the block)
    const char* sb = reinterpret_cast<const char*>(E.starts + 3 * t0);
    const char* eb = reinterpret_cast<const char*>(E.ends + 3 * t0);
    auto ld = [](const char* p, uint32_t j) {
        const float* q = reinterpret_cast<const float*>(p + 12u * j);
        return make_float3(q[0], q[1], q[2]);
    };
    uint64_t ball[SEL_EPT];
#pragma unroll
    for (int k = 0; k < SEL_EPT; ++k) {
        const uint32_t j = (uint32_t)(k * 256 + tid);
        const uint64_t i = t0 + j;
        uint64_t words = 0;
        const bool f = i < E.num_edges && edge_may_hit(occ, ld(sb, j), ld(eb, j), words);
        ball[k] = __ballot(f);
        if (lane == 0) wcnt[k * 4 + w] = (uint32_t)__popcll(ball[k]);
    }
    __syncthreads();
    if (tid == 0) {  // exclusive offsets in (item, wave) order, and the block's claim
        uint32_t run = 0;
        for (int q = 0; q < SEL_EPT * 4; ++q) {
            const uint32_t c = wcnt[q];
            wcnt[q] = run;
            run += c;
        }
        base_s = run ? atomicAdd(count, run) : 0u;
    }
    __syncthreads();
    const uint32_t base = base_s;
#pragma unroll
    for (int k = 0; k < SEL_EPT; ++k)
        if ((ball[k] >> lane) & 1ull)
            list[base + wcnt[k * 4 + w] + (uint32_t)__popcll(ball[k] & lt)] =
                (uint32_t)(t0 + (uint64_t)k * 256 + tid);
}

__global__ __launch_bounds__(256) void k_sum_i32(const int32_t* __restrict__ v, uint64_t n,
                                                 unsigned long long* __restrict__ acc) {
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        s += (unsigned long long)(v[i] > 0 ? v[i] : 0);
    s = block_sum_256(s);
    if (threadIdx.x == 0 && s) atomicAdd(acc_stripe(acc), s);
}

// ------------------------------------------------------------------------------------
// Host launchers
// ------------------------------------------------------------------------------------
// The traversal pushes one child only where both children are inner nodes, and such a
// node has binary height >= 2; heights fall strictly along a path, so at most
// max_depth - 1 entries are ever pending on the stack.
int pick_depth(const rs3d_bvh* bvh, int* depth) {
    const int need = bvh->max_depth - 1;
    if (need <= 24) *depth = 24;
    else if (need <= 32) *depth = 32;
    else if (need <= 48) *depth = 48;
    else if (need <= 64) *depth = 64;
    else {
        set_error("BVH height %d needs more than the traversal stack's 64 entries", bvh->max_depth);
        return RS3D_E_DEPTH;
    }
    return RS3D_OK;
}

template <int AX, bool STATS>
static int launch_points_ax(rs3d_ctx* ctx, const rs3d_bvh* bvh, int depth, const float* pts,
                            uint64_t n, const uint32_t* imap, int out_kind, void* out,
                            unsigned long long* stats) {
    unsigned nb;
    const int64_t pw = ctx->opt_rays_per_wave;
    if (!STATS && pw >= TR_BLOCK && n <= 0xffffffffull) {
        RS3D_TRY(grid_blocks(n, (uint64_t)pw, &nb));
        const uint32_t w = (uint32_t)pw;
        switch (depth) {
            case 24: k_points_refill<AX, 24><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, w); break;
            case 32: k_points_refill<AX, 32><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, w); break;
            case 48: k_points_refill<AX, 48><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, w); break;
            default: k_points_refill<AX, 64><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, w); break;
        }
        RS3D_HIP(hipGetLastError());
        return RS3D_OK;
    }
    RS3D_TRY(grid_blocks(n, TR_BLOCK, &nb));
    switch (depth) {
        case 24: k_points<AX, 24, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, stats); break;
        case 32: k_points<AX, 32, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, stats); break;
        case 48: k_points<AX, 48, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, stats); break;
        default: k_points<AX, 64, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(bvh->nodes, bvh->tris, pts, n, imap, out_kind, out, stats); break;
    }
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

template <bool STATS>
static int launch_points(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* pts, uint64_t n,
                         const uint32_t* imap, int axis, int out_kind, void* out,
                         unsigned long long* stats) {
    int depth;
    RS3D_TRY(pick_depth(bvh, &depth));
    switch (axis) {
        case 0: return launch_points_ax<0, STATS>(ctx, bvh, depth, pts, n, imap, out_kind, out, stats);
        case 1: return launch_points_ax<1, STATS>(ctx, bvh, depth, pts, n, imap, out_kind, out, stats);
        default: return launch_points_ax<2, STATS>(ctx, bvh, depth, pts, n, imap, out_kind, out, stats);
    }
}

__global__ void k_fill_results_miss(void* out, uint64_t n, int out_kind, const uint32_t* imap) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t pi = imap ? imap[i] : i;
    if (out_kind == RS3D_OUT_RAY_RESULT) {
        rs3d_ray_result r;
        r.ray_id = (int32_t)pi;
        r.polygon_index = -1;
        static_cast<rs3d_ray_result*>(out)[pi] = r;
    } else if (out_kind == RS3D_OUT_FLAG_U8) {
        static_cast<uint8_t*>(out)[pi] = 0;
    } else {
        static_cast<int32_t*>(out)[pi] = -1;
    }
}

__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0xffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

// 2-D Morton key of each ray origin in the plane perpendicular to the ray axis, on a grid
// of (qmax + 1)^2 cells.
__global__ __launch_bounds__(256) void k_ray_keys(const float* __restrict__ pts, uint64_t n,
                                                  const uint32_t* __restrict__ imap, int axis,
                                                  float lo_u, float lo_w, float sc_u, float sc_w,
                                                  float qmax, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t pi = imap ? imap[i] : (uint32_t)i;
    const int U = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
    const int W = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
    const float u = pts[3 * (uint64_t)pi + U], w = pts[3 * (uint64_t)pi + W];
    const float fu = fminf(fmaxf((u - lo_u) * sc_u, 0.0f), qmax);
    const float fw = fminf(fmaxf((w - lo_w) * sc_w, 0.0f), qmax);
    keys[i] = (spread16((uint32_t)fu) << 1) | spread16((uint32_t)fw);
    vals[i] = pi;
}

int trace_points_internal(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* pts, uint64_t n,
                          const uint32_t* imap, int axis, int out_kind, void* out,
                          unsigned long long* stats) {
    if (n == 0) return RS3D_OK;
    if (bvh->num_tris == 0) {
        unsigned nb;
        RS3D_TRY(grid_blocks(n, 256, &nb));
        k_fill_results_miss<<<nb, 256, 0, ctx->stream>>>(out, n, out_kind, imap);
        RS3D_HIP(hipGetLastError());
        return RS3D_OK;
    }
    if (stats) return launch_points<true>(ctx, bvh, pts, n, imap, axis, out_kind, out, stats);
    // The axis grid, when one is accepted, gives the same results with one lookup per ray
    // (built on first use; the cache on the BVH is not part of its value). Its rays run in
    // the caller's order: reordering them costs more than it saves there.
    bool grid = false;
    if (n <= 0xffffffffull) RS3D_TRY(axis_grid_ready(ctx, const_cast<rs3d_bvh*>(bvh), axis, &grid));
    if (grid) {
        if (ctx->profile) {
            RS3D_TRY(prof_collect(ctx));
            RS3D_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        }
        RS3D_TRY(axis_grid_launch(ctx, const_cast<rs3d_bvh*>(bvh), axis, pts, n, imap, out_kind, out));
        if (ctx->profile) {
            RS3D_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            ctx->prof_pending = true;
        }
        return RS3D_OK;
    }
    if (ctx->opt_ray_sort_min > 0 && (int64_t)n >= ctx->opt_ray_sort_min && n <= 0xffffffffull) {
        // reorder rays along a Morton curve of their origins (identical results, coherent waves)
        RS3D_TRY(ctx->ray_keys.ensure(n * sizeof(uint32_t)));
        RS3D_TRY(ctx->ray_vals.ensure(n * sizeof(uint32_t)));
        const int U = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
        const int W = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
        const float lo_u = bvh->bmin[U], lo_w = bvh->bmin[W];
        const float eu = bvh->bmax[U] - lo_u, ew = bvh->bmax[W] - lo_w;
        // cells of about 64..256 rays: q bits per axis with 4^q ~ n / 64
        int q = 4;
        while (q < 16 && ((uint64_t)64 << (2 * (q + 1))) <= n) ++q;
        const float cells = (float)(1u << q);
        const float sc_u = eu > 0.0f ? cells / eu : 0.0f, sc_w = ew > 0.0f ? cells / ew : 0.0f;
        unsigned nb;
        RS3D_TRY(grid_blocks(n, 256, &nb));
        k_ray_keys<<<nb, 256, 0, ctx->stream>>>(pts, n, imap, axis, lo_u, lo_w, sc_u, sc_w,
                                                 cells - 1.0f, ctx->ray_keys.as<uint32_t>(),
                                                 ctx->ray_vals.as<uint32_t>());
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(radix_sort_u32_pairs(ctx, ctx->ray_keys.as<uint32_t>(), ctx->ray_vals.as<uint32_t>(),
                                      n, 2 * q));
        imap = ctx->ray_vals.as<uint32_t>();
    }
    if (ctx->profile) {
        RS3D_TRY(prof_collect(ctx));
        RS3D_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    }
    RS3D_TRY(launch_points<false>(ctx, bvh, pts, n, imap, axis, out_kind, out, nullptr));
    if (ctx->profile) {
        RS3D_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        ctx->prof_pending = true;
    }
    return RS3D_OK;
}

int trace_points_walk(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* pts, uint64_t n,
                      const uint32_t* imap, int axis, int out_kind, void* out) {
    if (n == 0 || bvh->num_tris == 0) return trace_points_internal(ctx, bvh, pts, n, imap, axis, out_kind, out, nullptr);
    return launch_points<false>(ctx, bvh, pts, n, imap, axis, out_kind, out, nullptr);
}

int fill_i32(rs3d_ctx* ctx, int32_t* p, uint64_t n, int32_t v) {
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_fill_i32<<<nb, 256, 0, ctx->stream>>>(p, n, v);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int count_nonneg(rs3d_ctx* ctx, const int32_t* p, uint64_t n, unsigned long long* acc) {
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_count_nonneg<<<std::min(nb, 2048u), 256, 0, ctx->stream>>>(p, n, acc);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

template <int MODE>
static int launch_edges(rs3d_ctx* ctx, const rs3d_bvh* bvh, const rs3d_edges* edges, int swap,
                        int32_t* counts, const int64_t* offsets, int64_t bias, void* out,
                        const HashArgs& H, int chain_rule = CHAIN_EPS, int max_iter = 100,
                        unsigned long long* stats = nullptr) {
    if (edges->num_edges == 0) return RS3D_OK;
    EdgeArgs E{edges->starts, edges->ends, edges->src_offsets, edges->src_counts, edges->src_objects,
               edges->num_edges};
    unsigned nb;
    RS3D_TRY(grid_blocks(edges->num_edges, TR_BLOCK, &nb));
    if (bvh->num_tris == 0) {
        if (MODE == EMODE_COUNT)
            RS3D_HIP(hipMemsetAsync(counts, 0, edges->num_edges * sizeof(int32_t), ctx->stream));
        return RS3D_OK;
    }
    int depth;
    RS3D_TRY(pick_depth(bvh, &depth));
    OccArgs occ;
    RS3D_TRY(seg_occ_args(ctx, const_cast<rs3d_bvh*>(bvh), &occ));
    // One ray per lane. With occupancy bits (large timed launches): a selection pass that
    // also compacts the selected edges (in stretches of SEL_BLOCK_EDGES, each in order), and
    // the walk over them alone (the counting pass keeps the one-kernel form, which counts
    // the words read).
    const uint32_t* list = nullptr;
    uint64_t n_list = 0;
    unsigned nbl = nb;
    auto launch = [&](auto depth_tag, auto stats_tag) {
        constexpr int D = decltype(depth_tag)::value;
        constexpr bool ST = decltype(stats_tag)::value;
        unsigned long long* st = ST ? stats : nullptr;
        // over a selected list the bits are tested already: the walk only looks up its
        // start node
        OccArgs walk = occ;
        if (list) walk.bits = nullptr;
        k_edges<MODE, D, ST><<<list ? nbl : nb, TR_BLOCK, 0, ctx->stream>>>(
            bvh->nodes, bvh->tris, walk, E, list, n_list, swap, chain_rule, max_iter, counts, offsets,
            bias, out, H, st);
    };
    auto launch_depth = [&](auto stats_tag) {
        switch (depth) {
            case 24: launch(std::integral_constant<int, 24>{}, stats_tag); break;
            case 32: launch(std::integral_constant<int, 32>{}, stats_tag); break;
            case 48: launch(std::integral_constant<int, 48>{}, stats_tag); break;
            default: launch(std::integral_constant<int, 64>{}, stats_tag); break;
        }
    };
    if (stats) {
        launch_depth(std::integral_constant<bool, true>{});
        RS3D_HIP(hipGetLastError());
        return RS3D_OK;
    }
    if (ctx->profile) {
        RS3D_TRY(prof_collect(ctx));
        RS3D_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
    }
    // (below 2^20 edges the selection's extra launches and host read cost more than they
    // save: the walking kernel tests the bits itself)
    if (occ.bits && edges->num_edges >= (1ull << 20) && edges->num_edges <= 0xffffffffull) {
        const uint64_t n = edges->num_edges;
        RS3D_TRY(ctx->edge_sel.ensure(256 + n * sizeof(uint32_t)));  // count, then the list
        uint32_t* cnt = ctx->edge_sel.as<uint32_t>();
        uint32_t* idx = reinterpret_cast<uint32_t*>(ctx->edge_sel.as<uint8_t>() + 256);
        RS3D_TRY(fill_u32(ctx, cnt, 1, 0u));
        unsigned nbs;
        RS3D_TRY(grid_blocks(n, SEL_BLOCK_EDGES, &nbs));
        k_edge_select_list<<<nbs, 256, 0, ctx->stream>>>(E, occ, idx, cnt);
        RS3D_HIP(hipGetLastError());
        if (MODE == EMODE_COUNT) RS3D_TRY(fill_u32(ctx, counts, n, 0u));  // unselected edges: no hits
        RS3D_HIP(hipMemcpyAsync(ctx->pinned, cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        RS3D_HIP(hipStreamSynchronize(ctx->stream));
        uint32_t h;
        memcpy(&h, ctx->pinned, sizeof h);
        n_list = h;
        list = idx;
        RS3D_TRY(grid_blocks(n_list ? n_list : 1, TR_BLOCK, &nbl));
    }
    if (!list || n_list) launch_depth(std::integral_constant<bool, false>{});
    RS3D_HIP(hipGetLastError());
    if (ctx->profile) {
        RS3D_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
        ctx->prof_pending = true;
    }
    return RS3D_OK;
}

static bool edges_ok(const rs3d_edges* e) {
    return e && (e->num_edges == 0 ||
                 (e->starts && e->ends && e->src_offsets && e->src_counts && e->src_objects));
}

int edge_count_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e, int32_t* counts) {
    HashArgs H{};
    return launch_edges<EMODE_COUNT>(ctx, t, e, 0, counts, nullptr, 0, nullptr, H);
}
// Overlap counters of rs3d_edge_profile from the per-edge counts of a counting pass (chain
// hits x sources): per source object of an edge of length >= 1e-6 one trace call with
// min(hits + 1, cap) iterations (trace_edge_multi_hits, query/src/cuda/mesh_intersection.cu:
// 16-91); the sums are combined per block before one atomic per counter.
__global__ __launch_bounds__(256) void k_edge_profile(const float* __restrict__ starts, const float* __restrict__ ends,
                                                      const int32_t* __restrict__ src_cnt,
                                                      const int32_t* __restrict__ counts, uint64_t n, int cap,
                                                      rs3d_intersection_stats* __restrict__ st) {
    const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long calls = 0, iters = 0, hits = 0;
    uint32_t mx = 0;
    if (e < n) {
        const float3 a = ld3f(starts, e), b = ld3f(ends, e);
        const float dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        const int sc = src_cnt[e];
        if (len >= 1e-6f && sc > 0) {
            const uint32_t h = (uint32_t)(counts[e] / sc);
            const uint32_t it = h + 1 < (uint32_t)cap ? h + 1 : (uint32_t)cap;
            calls = (unsigned long long)sc;
            iters = (unsigned long long)it * sc;
            hits = (unsigned long long)counts[e];
            mx = it;
        }
    }
    const unsigned long long c_sum = block_sum_256(calls);
    __syncthreads();
    const unsigned long long i_sum = block_sum_256(iters);
    __syncthreads();
    const unsigned long long h_sum = block_sum_256(hits);
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
    if ((threadIdx.x & 63) == 0 && mx) atomicMax(&st->overlap_max_iterations_per_trace, mx);
    if (threadIdx.x == 0) {
        if (c_sum) atomicAdd(reinterpret_cast<unsigned long long*>(&st->overlap_trace_calls), c_sum);
        if (i_sum) atomicAdd(reinterpret_cast<unsigned long long*>(&st->overlap_iterations_total), i_sum);
        if (h_sum) atomicAdd(reinterpret_cast<unsigned long long*>(&st->overlap_hits_total), h_sum);
    }
}

int edge_write_keys_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e,
                             const int64_t* offsets, int64_t bias, int swap, uint64_t* out) {
    HashArgs H{};
    return launch_edges<EMODE_WRITE_KEYS>(ctx, t, e, swap, nullptr, offsets, bias, out, H);
}
int edge_hash_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e, int swap,
                       const rs3d_hash_table* ht, int chain_rule, int max_iter) {
    HashArgs H{reinterpret_cast<unsigned long long*>(ht->slots), ht->size, ht->bitwise,
               ht->max_probes, reinterpret_cast<unsigned long long*>(ht->failed)};
    return launch_edges<EMODE_HASH>(ctx, t, e, swap, nullptr, nullptr, 0, nullptr, H, chain_rule,
                                    max_iter);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_trace_points(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                      const uint32_t* index_map, int axis, int out_kind, void* out) {
    if (!ctx || !bvh || (n && (!points || !out)) || axis < 0 || axis > 2 || out_kind < 0 ||
        out_kind > 2) {
        set_error("rs3d_trace_points: bad arguments");
        return RS3D_E_ARG;
    }
    return trace_points_internal(ctx, bvh, points, n, index_map, axis, out_kind, out, nullptr);
}

int rs3d_trace_points_stats(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                            int axis, rs3d_trace_stats* stats) {
    if (!ctx || !bvh || !stats || (n && !points) || axis < 0 || axis > 2) {
        set_error("rs3d_trace_points_stats: bad arguments");
        return RS3D_E_ARG;
    }
    memset(stats, 0, sizeof *stats);
    stats->rays = n;
    if (n == 0 || bvh->num_tris == 0) return RS3D_OK;
    RS3D_TRY(ctx->small.ensure(256));
    RS3D_TRY(ctx->tmp_a.ensure(n));
    unsigned long long* dstats = ctx->small.as<unsigned long long>() + 16;
    RS3D_HIP(hipMemsetAsync(dstats, 0, 4 * sizeof(unsigned long long), ctx->stream));
    RS3D_TRY(trace_points_internal(ctx, bvh, points, n, nullptr, axis, RS3D_OUT_FLAG_U8,
                                   ctx->tmp_a.ptr, dstats));
    unsigned long long h[4];
    RS3D_HIP(hipMemcpyAsync(h, dstats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    stats->nodes_visited = h[0];
    stats->tris_tested = h[1];
    stats->hits = h[2];
    stats->restarts = h[3];
    return RS3D_OK;
}

int rs3d_edge_stats(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges,
                    rs3d_edge_stats_t* stats) {
    if (!ctx || !target || !stats || !edges_ok(edges)) {
        set_error("rs3d_edge_stats: bad arguments");
        return RS3D_E_ARG;
    }
    memset(stats, 0, sizeof *stats);
    const uint64_t n = edges->num_edges;
    stats->rays = n;
    if (n == 0 || target->num_tris == 0) return RS3D_OK;
    RS3D_TRY(ctx->small.ensure(256));
    RS3D_TRY(ctx->tmp_a.ensure(n * sizeof(int32_t)));
    unsigned long long* dstats = ctx->small.as<unsigned long long>() + 16;
    RS3D_HIP(hipMemsetAsync(dstats, 0, 6 * sizeof(unsigned long long), ctx->stream));
    HashArgs H{};
    int32_t* counts = ctx->tmp_a.as<int32_t>();
    RS3D_TRY(launch_edges<EMODE_COUNT>(ctx, target, edges, 0, counts, nullptr, 0, nullptr, H,
                                       CHAIN_EPS, 100, dstats));
    unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    unsigned long long *emits, *refs;
    RS3D_TRY(acc_slot(ctx, 0, 0, &emits));
    RS3D_TRY(acc_slot(ctx, 1, 0, &refs));
    k_sum_i32<<<nb, 256, 0, ctx->stream>>>(counts, n, emits);
    k_sum_i32<<<nb, 256, 0, ctx->stream>>>(edges->src_counts, n, refs);
    RS3D_HIP(hipGetLastError());
    unsigned long long h[6];
    RS3D_HIP(hipMemcpyAsync(h, dstats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    uint64_t sums[2];
    RS3D_TRY(acc_read_sum(ctx, emits, &sums[0]));
    RS3D_TRY(acc_read_sum(ctx, refs, &sums[1]));
    stats->nodes_visited = h[0];
    stats->tris_tested = h[1];
    stats->hits = h[2];
    stats->restarts = h[3];
    stats->rays_walked = h[4];
    stats->occ_words = h[5];
    stats->pair_emits = sums[0];
    stats->source_refs = sums[1];
    return RS3D_OK;
}

int rs3d_filter_box(rs3d_ctx* ctx, const float* box, const float* points, uint64_t n, int axis,
                    uint8_t* flags) {
    if (!ctx || !box || (n && (!points || !flags)) || axis < 0 || axis > 2) {
        set_error("rs3d_filter_box: bad arguments");
        return RS3D_E_ARG;
    }
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    const FBox fb{box[0], box[1], box[2], box[3], box[4], box[5]};
    if (axis == 0)
        k_filter_box<0><<<nb, 256, 0, ctx->stream>>>(points, n, fb, flags);
    else if (axis == 1)
        k_filter_box<1><<<nb, 256, 0, ctx->stream>>>(points, n, fb, flags);
    else
        k_filter_box<2><<<nb, 256, 0, ctx->stream>>>(points, n, fb, flags);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

}  // extern "C"

namespace rs3d {
// The grid filter-refine counts of a context: ACC_STRIPES stripes, zeroed once here and left
// zeroed by every launch, and the host-coherent word the launch writes.
int fr_counts_ready(rs3d_ctx* ctx) {
    const size_t bytes = (size_t)ACC_STRIPES * ACC_PITCH * sizeof(unsigned long long);
    if (ctx->fr_total && !ctx->fr_dirty) return RS3D_OK;
    // first use, or a call that stopped between its launch and the finisher: zero again
    RS3D_TRY(ctx->fr_acc.ensure(bytes));
    RS3D_HIP(hipMemsetAsync(ctx->fr_acc.ptr, 0, bytes, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    ctx->fr_dirty = false;
    if (ctx->fr_total) return RS3D_OK;
    void* h = nullptr;
    const hipError_t e = hipHostMalloc(&h, 256, hipHostMallocCoherent | hipHostMallocMapped);
    if (e != hipSuccess) {
        set_error("hipHostMalloc of the filter-refine counts failed: %s", hipGetErrorString(e));
        return RS3D_E_NOMEM;
    }
    ctx->fr_total = static_cast<unsigned long long*>(h);
    *ctx->fr_total = 0;
    return RS3D_OK;
}

// Waits for the counts of a filter-refine call, whose word the host set to all ones before
// the launch: polls it for up to 2 ms (a short launch is seen as soon as its finisher's one
// 8-byte store lands, without the stream synchronize's wake-up), then falls back to the
// stream synchronize, which also reports a launch that failed. (All ones are also the counts
// of 2^32 - 1 candidates all inside: such a call waits the 2 ms and then synchronizes.)
int fr_wait(rs3d_ctx* ctx) {
    volatile unsigned long long* w = ctx->fr_total;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned i = 1;; ++i) {
        if (*w != ~0ull) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return RS3D_OK;
        }
        if ((i & 1023) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
            break;
    }
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}
}  // namespace rs3d

extern "C" {

int rs3d_filter_refine(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                       int axis, int32_t* out_obj, uint64_t* candidate_count,
                       uint64_t* inside_count) {
    if (!ctx || !bvh || (n && (!points || !out_obj)) || axis < 0 || axis > 2 ||
        n > 0xffffffffull) {
        set_error("rs3d_filter_refine: bad arguments");
        return RS3D_E_ARG;
    }
    if (candidate_count) *candidate_count = 0;
    if (inside_count) *inside_count = 0;
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    const float box[6] = {bvh->vmin[0], bvh->vmin[1], bvh->vmin[2],
                          bvh->vmax[0], bvh->vmax[1], bvh->vmax[2]};
    bool grid = false;
    if (bvh->num_tris) RS3D_TRY(axis_grid_ready(ctx, const_cast<rs3d_bvh*>(bvh), axis, &grid));
    if (grid) {
        // one pass: filter, trace and count per point (same answers as the three phases);
        // the launch leaves its counts in fr_total and its stripes zeroed for the next call
        RS3D_TRY(fr_counts_ready(ctx));
        unsigned long long* cnt = ctx->fr_acc.as<unsigned long long>();
        if (ctx->profile) {
            RS3D_TRY(prof_collect(ctx));
            RS3D_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
        }
        ctx->fr_dirty = true;
        RS3D_TRY(axis_grid_filter_refine(ctx, const_cast<rs3d_bvh*>(bvh), axis, points, n, box, out_obj, cnt));
        if (ctx->profile) {
            RS3D_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
            ctx->prof_pending = true;
        }
        *reinterpret_cast<volatile unsigned long long*>(ctx->fr_total) = ~0ull;
        RS3D_TRY(axis_grid_counts_finish(ctx, cnt, ctx->fr_total));
        ctx->fr_dirty = false;
        RS3D_TRY(fr_wait(ctx));
        const uint64_t h = *reinterpret_cast<volatile unsigned long long*>(ctx->fr_total);
        if (candidate_count) *candidate_count = h >> 32;
        if (inside_count) *inside_count = h & 0xffffffffull;
        return RS3D_OK;
    }
    k_fill_i32<<<nb, 256, 0, ctx->stream>>>(out_obj, n, -1);
    RS3D_HIP(hipGetLastError());
    if (bvh->num_tris == 0) return RS3D_OK;
    // filter flags live in tmp_c; compact_flags uses tmp_b for positions
    RS3D_TRY(ctx->tmp_c.ensure(n + n * sizeof(uint32_t) + 256));
    uint8_t* flags = ctx->tmp_c.as<uint8_t>();
    uint32_t* cand = reinterpret_cast<uint32_t*>(ctx->tmp_c.as<uint8_t>() + ((n + 255) & ~255ull));
    RS3D_TRY(rs3d_filter_box(ctx, box, points, n, axis, flags));
    uint64_t nc = 0;
    RS3D_TRY(rs3d_compact_flags(ctx, flags, n, cand, &nc));
    if (candidate_count) *candidate_count = nc;
    if (nc == 0) return RS3D_OK;
    RS3D_TRY(trace_points_internal(ctx, bvh, points, nc, cand, axis, RS3D_OUT_OBJ_I32, out_obj, nullptr));
    if (inside_count) {
        unsigned long long* cnt;
        RS3D_TRY(acc_slot(ctx, 0, 0, &cnt));
        k_count_nonneg<<<std::min(nb, 2048u), 256, 0, ctx->stream>>>(out_obj, n, cnt);
        RS3D_HIP(hipGetLastError());
        uint64_t h = 0;
        RS3D_TRY(acc_read_sum(ctx, cnt, &h));
        *inside_count = h;
    }
    return RS3D_OK;
}

int rs3d_edge_count(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int32_t* counts) {
    if (!ctx || !target || !edges_ok(edges) || (edges->num_edges && !counts)) {
        set_error("rs3d_edge_count: bad arguments");
        return RS3D_E_ARG;
    }
    return edge_count_internal(ctx, target, edges, counts);
}

int rs3d_edge_write(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges,
                    const int64_t* offsets, int swap, rs3d_pair* out) {
    if (!ctx || !target || !edges_ok(edges) || (edges->num_edges && (!offsets || !out))) {
        set_error("rs3d_edge_write: bad arguments");
        return RS3D_E_ARG;
    }
    HashArgs H{};
    return launch_edges<EMODE_WRITE_PAIRS>(ctx, target, edges, swap, nullptr, offsets, 0, out, H);
}

int rs3d_edge_hash(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int swap,
                   rs3d_hash_table* ht) {
    if (!ctx || !target || !edges_ok(edges) || !ht || !ht->slots || ht->size == 0 || !ht->failed ||
        (ht->bitwise && (ht->size & (ht->size - 1)))) {
        set_error("rs3d_edge_hash: bad arguments");
        return RS3D_E_ARG;
    }
    return edge_hash_internal(ctx, target, edges, swap, ht, CHAIN_EPS, 100);
}

int rs3d_edge_count_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                       int max_iterations, int32_t* counts) {
    if (!ctx || !target || !edges_ok(edges) || (edges->num_edges && !counts) || chain_rule < 0 ||
        chain_rule > 1 || max_iterations <= 0) {
        set_error("rs3d_edge_count_ex: bad arguments");
        return RS3D_E_ARG;
    }
    HashArgs H{};
    return launch_edges<EMODE_COUNT>(ctx, target, edges, 0, counts, nullptr, 0, nullptr, H, chain_rule,
                                     max_iterations);
}

int rs3d_edge_write_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                       int max_iterations, const int64_t* offsets, int swap, rs3d_pair* out) {
    if (!ctx || !target || !edges_ok(edges) || (edges->num_edges && (!offsets || !out)) || chain_rule < 0 ||
        chain_rule > 1 || max_iterations <= 0) {
        set_error("rs3d_edge_write_ex: bad arguments");
        return RS3D_E_ARG;
    }
    HashArgs H{};
    return launch_edges<EMODE_WRITE_PAIRS>(ctx, target, edges, swap, nullptr, offsets, 0, out, H, chain_rule,
                                           max_iterations);
}

int rs3d_edge_profile(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int chain_rule,
                      int max_iterations, rs3d_intersection_stats* stats) {
    if (!ctx || !target || !edges_ok(edges) || !stats || chain_rule < 0 || chain_rule > 1 || max_iterations <= 0) {
        set_error("rs3d_edge_profile: bad arguments");
        return RS3D_E_ARG;
    }
    const uint64_t n = edges->num_edges;
    if (n == 0) return RS3D_OK;
    static thread_local DevBuf counts;
    RS3D_TRY(counts.ensure(n * sizeof(int32_t)));
    HashArgs H{};
    RS3D_TRY(launch_edges<EMODE_COUNT>(ctx, target, edges, 0, counts.as<int32_t>(), nullptr, 0, nullptr, H,
                                       chain_rule, max_iterations));
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_edge_profile<<<nb, 256, 0, ctx->stream>>>(edges->starts, edges->ends, edges->src_counts,
                                                counts.as<int32_t>(), n, max_iterations, stats);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

int rs3d_edge_hash_ex(rs3d_ctx* ctx, const rs3d_bvh* target, const rs3d_edges* edges, int swap,
                      int chain_rule, int max_iterations, rs3d_hash_table* ht) {
    if (!ctx || !target || !edges_ok(edges) || !ht || !ht->slots || ht->size == 0 || !ht->failed ||
        (ht->bitwise && (ht->size & (ht->size - 1))) || chain_rule < 0 || chain_rule > 1 ||
        max_iterations <= 0) {
        set_error("rs3d_edge_hash_ex: bad arguments");
        return RS3D_E_ARG;
    }
    return edge_hash_internal(ctx, target, edges, swap, ht, chain_rule, max_iterations);
}

}  // extern "C"
