// traverse.h -- device-side traversal building blocks shared by the ray kernels:
// hit buffer, node/triangle fetch, all-hits traversal for axis rays and segments, and
// the pair hash insert of the reference's open-addressing table.
#pragma once

#include "rs3d_internal.h"

namespace rs3d {

constexpr int TR_BLOCK = 64;
constexpr int KBUF = 8;
// Hit buffer of the axis-grid walk: a grid ray meets few triangles (a sparse scene's rays
// cross a handful of surfaces), so a smaller buffer saves registers; more hits restart.
constexpr int KBUF_GRID = 4;
// Hit buffer of a segment walk: one hit. A walked edge meets 0.1 triangles on average, and an
// edge with k hits takes k passes; the registers saved buy a seventh wave per SIMD.
constexpr int KBUF_EDGE = 1;
// LDS part of a segment walk's stack (entries per lane).
constexpr int EDGE_LDS_DEPTH = 16;

__device__ __forceinline__ float3 ld3f(const float* p, uint64_t i) {
    return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

// Sorted insertion of (t, tri, obj) into the N smallest, ordered by (t, tri).
template <int N>
__device__ __forceinline__ void kbuf_insert(float (&bt)[N], uint32_t (&bi)[N], int (&bo)[N], float t,
                                            uint32_t tri, int obj) {
    float ct = t;
    uint32_t ci = tri;
    int co = obj;
#pragma unroll
    for (int s = 0; s < N; ++s) {
        const bool less = ct < bt[s] || (ct == bt[s] && ci < bi[s]);
        if (less) {
            const float tt = bt[s];
            const uint32_t ti = bi[s];
            const int to = bo[s];
            bt[s] = ct;
            bi[s] = ci;
            bo[s] = co;
            ct = tt;
            ci = ti;
            co = to;
        }
    }
}

template <int N>
__device__ __forceinline__ void kbuf_clear(float (&bt)[N], uint32_t (&bi)[N], int (&bo)[N]) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
        bt[s] = INFINITY;
        bi[s] = 0xffffffffu;
        bo[s] = -1;
    }
}

// ------------------------------------------------------------------------------------
// Dataset-box filter of the filter-refine pipeline: parity of the +AX ray against the
// 12-triangle box mesh of query/src/geometry/BoundingBox.cpp:26-49 (same vertex order and
// triangulation), the chain of the point kernels over its hits.
// Only the 4 triangles of the two faces perpendicular to AX are tested. A triangle of any
// other face never counts: its three vertices share one sheared coordinate a (the face's
// plane; the shear of an axis ray is +0), and two of its three corners share the other one,
// so the edge function of that pair is fl(a*y) - fl(y*a) = 0. That forces the double
// recomputation, in which the remaining two are fl(a*p - a*q) and fl(a*q - a*p), exact
// negatives under round-to-nearest: mixed strict signs (miss) or all zero (det = 0, miss).
// Overflowed or NaN coordinates give t = NaN or inf, which the t window rejects, as it would
// on the full mesh. Hits keep their triangle index for ties.
// ------------------------------------------------------------------------------------
struct FBox {
    float x0, y0, z0, x1, y1, z1;
};

template <int AX>
__host__ __device__ __forceinline__ bool box_filter_full(const float3 o, const FBox& b) {
    const float3 bv[8] = {make_float3(b.x0, b.y0, b.z0), make_float3(b.x1, b.y0, b.z0),
                          make_float3(b.x1, b.y1, b.z0), make_float3(b.x0, b.y1, b.z0),
                          make_float3(b.x0, b.y0, b.z1), make_float3(b.x1, b.y0, b.z1),
                          make_float3(b.x1, b.y1, b.z1), make_float3(b.x0, b.y1, b.z1)};
    const int tri[12][3] = {{0, 2, 1}, {0, 3, 2}, {4, 5, 6}, {4, 6, 7}, {0, 1, 5}, {0, 5, 4},
                            {3, 6, 2}, {3, 7, 6}, {0, 4, 7}, {0, 7, 3}, {1, 2, 6}, {1, 6, 5}};
    // faces perpendicular to AX: z -> triangles 0-3, y -> 4-7, x -> 8-11
    constexpr int K0 = AX == 2 ? 0 : (AX == 1 ? 4 : 8);
    constexpr int NB = 4;
    float bt[NB];
    uint32_t bi[NB];
#pragma unroll
    for (int s = 0; s < NB; ++s) {
        bt[s] = INFINITY;
        bi[s] = 0xffffffffu;
    }
#pragma unroll
    for (int k = K0; k < K0 + 4; ++k) {
        float t;
        if (tri_hit_axis<AX>(o, bv[tri[k][0]], bv[tri[k][1]], bv[tri[k][2]], t) && t >= 1e-6f && t < 1e16f) {
            float ct = t;
            uint32_t ci = (uint32_t)k;
#pragma unroll
            for (int s = 0; s < NB; ++s) {
                const bool less = ct < bt[s] || (ct == bt[s] && ci < bi[s]);
                if (less) {
                    const float tt = bt[s];
                    const uint32_t ti = bi[s];
                    bt[s] = ct;
                    bi[s] = ci;
                    ct = tt;
                    ci = ti;
                }
            }
        }
    }
    float thr = 1e-6f;
    int count = 0;
#pragma unroll
    for (int s = 0; s < NB; ++s) {
        if (bt[s] >= thr && bt[s] < INFINITY) {
            ++count;
            float nthr = bt[s] + 1e-6f;
            if (nthr <= bt[s]) nthr = nextafterf(bt[s], INFINITY);
            thr = nthr;
        }
    }
    return (count & 1) != 0;
}

// The same answer without the triangle tests for the points that are clearly placed: with
// m = 2^-16 of the largest magnitude among the box and the point, plus 1e-5, a point whose
// in-plane coordinates lie at least m inside the face square, whose distance from the
// faces' common diagonal (corner lo-lo to hi-hi in the plane) is at least 2^-14 mag^2 in the
// un-normalised edge function, and whose axis coordinate is at least m from both faces
// gets exactly the hits the full test would give: every edge function of the face
// triangles then has its exact sign (the rounding of the differences, products and sum is
// below 2^-19 mag^2, and the side edges' signs are exact by monotone rounding), so one
// triangle per face is hit, and t = fl(T / det) lies within 2^-21 relative of the
// distance to that face, which is at least m: inside the box the top face is hit at
// t >= 1e-6 and the bottom face behind the origin, one hit (mag < 1e12 keeps the edge
// functions, T = U*Az + V*Bz + W*Cz and t = T / det below overflow and t below 1e16);
// above the box no hit; below it two hits at least the box height (> m) apart, which the
// chain counts twice. A point at least m outside the square in one coordinate and off the
// diagonal's line by the same margin misses every triangle (a side edge or the diagonal
// separates it, with exact sign). Everything else -- near a face, an edge or the diagonal,
// flat or huge boxes, non-finite values -- takes the full test.
template <int AX>
__host__ __device__ __forceinline__ bool box_filter_pass(const float3 o, const FBox& b) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float lo[3] = {b.x0, b.y0, b.z0}, hi[3] = {b.x1, b.y1, b.z1};
    const float ulo = lo[U], uhi = hi[U], wlo = lo[W], whi = hi[W], alo = lo[AX], ahi = hi[AX];
    const float u = get3(o, U), w = get3(o, W), a = get3(o, AX);
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(ulo), fabsf(uhi)), fmaxf(fabsf(wlo), fabsf(whi))),
                            fmaxf(fmaxf(fabsf(alo), fabsf(ahi)), fmaxf(fmaxf(fabsf(u), fabsf(w)), fabsf(a))));
    if (mag < 1e12f) {
        const float m = mag * 1.52587890625e-5f + 1e-5f;  // 2^-16 mag + 1e-5
        const float eu = uhi - ulo, ew = whi - wlo;
        const float dg = (u - ulo) * ew - (w - wlo) * eu;
        const bool off_diag = fabsf(dg) > mag * mag * 6.103515625e-5f;  // 2^-14 mag^2
        if (off_diag && eu > 2.0f * m && ew > 2.0f * m && ahi - alo > 2.0f * m) {
            const bool in_sq = u > ulo + m && u < uhi - m && w > wlo + m && w < whi - m;
            const bool out_sq = u < ulo - m || u > uhi + m || w < wlo - m || w > whi + m;
            if (out_sq) return false;
            if (in_sq) {
                if (a > alo + m && a < ahi - m) return true;
                if (a > ahi + m || a < alo - m) return false;
            }
        }
    }
    return box_filter_full<AX>(o, b);
}

struct NodeBoxes {
    float lo0[3], hi0[3], lo1[3], hi1[3];
    int l0, l1, n0, n1;
};

__device__ __forceinline__ NodeBoxes load_node(const BNode* __restrict__ nodes, int idx) {
    const BNode nd = nodes[idx];
    NodeBoxes b;
    b.lo0[0] = nd.a.x; b.hi0[0] = nd.a.y; b.lo0[1] = nd.a.z; b.hi0[1] = nd.a.w;
    b.lo0[2] = nd.b.x; b.hi0[2] = nd.b.y; b.lo1[0] = nd.b.z; b.hi1[0] = nd.b.w;
    b.lo1[1] = nd.c.x; b.hi1[1] = nd.c.y; b.lo1[2] = nd.c.z; b.hi1[2] = nd.c.w;
    b.l0 = nd.link.x; b.l1 = nd.link.y; b.n0 = nd.link.z; b.n1 = nd.link.w;
    return b;
}

__device__ __forceinline__ void load_tri(const TriRec* __restrict__ tris, int j, float3& v0, float3& v1,
                                         float3& v2, uint32_t& orig, int& obj) {
    const TriRec r = tris[j];
    v0 = make_float3(r.v0.x, r.v0.y, r.v0.z);
    v1 = make_float3(r.v1.x, r.v1.y, r.v1.z);
    v2 = make_float3(r.v2.x, r.v2.y, r.v2.z);
    orig = (uint32_t)__float_as_int(r.v0.w);
    obj = __float_as_int(r.v1.w);
}

struct TraceCounters {
    uint64_t nodes = 0, tris = 0, hits = 0, restarts = 0;
    uint64_t walked = 0, occ_words = 0;  // segment rays: walks after the occupancy test, words read
};

// ------------------------------------------------------------------------------------
// Segment occupancy bits of a target BVH: one bit per cube cell of a uniform 3-D grid over
// its padded bounds, set where some padded triangle box (the BVH's leaf box) overlaps the
// cell. A segment whose (slightly widened) bounding box covers no set bit cannot hit any
// triangle: a reported hit lies, up to rounding far below the widening, on its triangle,
// inside the padded box, and on the segment, inside the segment's box, so the two boxes
// meet and so do their cell ranges (cells come from one monotone clamped map for both).
// ------------------------------------------------------------------------------------
struct OccArgs {
    const uint32_t* bits;  // null: no occupancy test
    float org_x, org_y, org_z, inv;
    int nx, ny, nz;
    // start nodes (null: every walk starts at the root): for block (bx, by, bz) of
    // OCC_BLOCK^3 cells, the deepest BVH node whose subtree holds every triangle whose padded
    // box overlaps a cell of the 2x2x2 blocks from (bx, by, bz) on
    const uint32_t* snode;
    int bnx, bny;
};

constexpr int OCC_BLOCK = 4;

__device__ __forceinline__ int occ_cell(float x, float org, float inv, int n) {
    const float f = fminf(fmaxf((x - org) * inv, 0.0f), (float)(n - 1));
    return (int)f;
}

// Whether the box [lo, hi] covers a set bit, i.e. may meet a padded triangle box of the
// bits' BVH (true without bits). Boxes over more than 64 cells are not tested (answer true).
// words: occupancy words read. With start, a walk for anything inside the box may begin at *start (left
// unchanged, the root, when the box's cells do not fit one block's 2x2x2 neighbourhood):
// every triangle whose padded box meets the box overlaps one of the box's cells.
__device__ __forceinline__ bool occ_box_may_hit(const OccArgs& g, const float3 lo, const float3 hi,
                                                uint64_t& words, int* start = nullptr) {
    const int x0 = occ_cell(lo.x, g.org_x, g.inv, g.nx), x1 = occ_cell(hi.x, g.org_x, g.inv, g.nx);
    const int y0 = occ_cell(lo.y, g.org_y, g.inv, g.ny), y1 = occ_cell(hi.y, g.org_y, g.inv, g.ny);
    const int z0 = occ_cell(lo.z, g.org_z, g.inv, g.nz), z1 = occ_cell(hi.z, g.org_z, g.inv, g.nz);
    const int ex = x1 - x0, ey = y1 - y0 + 1, ez = z1 - z0 + 1;
    if ((ex + 1) * ey * ez > 64) return true;
    bool any = !g.bits;  // without bits (a walk over selected segments): start node only
    if (g.bits && ex >= 0 && ey > 0 && ez > 0) {
        // The box's cells of one (y, z) row are consecutive bits: one or two words per row
        // (three or more only for a row of over 32 cells). Rows go in groups of four whose
        // loads are all issued before the test; padding rows repeat the group's first row.
        // (cell numbers fit 32 bits: a grid holds at most 2^30 cells)
        const int rows = ey * ez;
        const uint32_t nx = (uint32_t)g.nx;
        const uint32_t wrap_step = (uint32_t)g.ny * nx - (uint32_t)(ey - 1) * nx;  // last y of a plane -> next z
        uint32_t cb = ((uint32_t)z0 * (uint32_t)g.ny + (uint32_t)y0) * nx + (uint32_t)x0;  // group's first row
        int yb = y0;
        for (int r0 = 0; r0 < rows && !any; r0 += 4) {
            uint32_t acc = 0;
            uint32_t c = cb;
            int y = yb;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool real = r0 + k < rows;
                const uint32_t c0 = real ? c : cb;
                const uint32_t c1 = c0 + (uint32_t)ex;
                const uint32_t w0 = c0 >> 5, w1 = c1 >> 5;
                const uint32_t mlo = ~0u << (c0 & 31), mhi = ~0u >> (31 - (c1 & 31));
                // (byte offsets below 2^27: 32-bit, so the loads take the scalar base)
                const uint32_t a = *(const uint32_t*)((const char*)g.bits + (w0 << 2));
                const uint32_t b = *(const uint32_t*)((const char*)g.bits + (w1 << 2));
                acc |= w0 == w1 ? (a & mlo & mhi) : ((a & mlo) | (b & mhi));
                for (uint32_t w = w0 + 1; w < w1; ++w) acc |= g.bits[w];
                if (real) words += w1 - w0 + 1;
                const bool wrap = ++y > y1;
                c += wrap ? wrap_step : nx;
                y = wrap ? y0 : y;
            }
            cb = c;
            yb = y;
            any = acc != 0;
        }
    }
    if (any && start && g.snode) {
        const int bx = x0 / OCC_BLOCK, by = y0 / OCC_BLOCK, bz = z0 / OCC_BLOCK;
        if (x1 < (bx + 2) * OCC_BLOCK && y1 < (by + 2) * OCC_BLOCK && z1 < (bz + 2) * OCC_BLOCK)
            *start = (int)g.snode[((uint64_t)bz * g.bny + by) * g.bnx + bx];
    }
    return any;
}

// Whether segment a-b may hit a triangle of the bits' BVH: its box, widened by 2^-14 of the
// coordinates' magnitude plus pad on each axis, through occ_box_may_hit.
__device__ __forceinline__ bool occ_segment_may_hit(const OccArgs& g, const float3 a, const float3 b, float pad,
                                                    uint64_t& words, int* start = nullptr) {
    const float wa = (fabsf(a.x) + fabsf(b.x)) * 6.103515625e-5f + pad;  // 2^-14
    const float wb = (fabsf(a.y) + fabsf(b.y)) * 6.103515625e-5f + pad;
    const float wc = (fabsf(a.z) + fabsf(b.z)) * 6.103515625e-5f + pad;
    return occ_box_may_hit(g, make_float3(fminf(a.x, b.x) - wa, fminf(a.y, b.y) - wb, fminf(a.z, b.z) - wc),
                           make_float3(fmaxf(a.x, b.x) + wa, fmaxf(a.y, b.y) + wb, fmaxf(a.z, b.z) + wc), words,
                           start);
}

// ------------------------------------------------------------------------------------
// All-hits traversal for an axis-aligned ray (+AX). Hits with thr <= t < tmax_excl are
// offered to the buffer; returns the number of such hits.
// ------------------------------------------------------------------------------------
template <int AX, int DEPTH, bool STATS>
__device__ __forceinline__ uint32_t collect_axis(const BNode* __restrict__ nodes,
                                                 const TriRec* __restrict__ tris, const float3 o,
                                                 float thr, uint32_t thr_tri, float tmax_excl,
                                                 int* stack,
                                                 float (&bt)[KBUF], uint32_t (&bi)[KBUF],
                                                 int (&bo)[KBUF], TraceCounters& ctr) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float ou = get3(o, U), ow = get3(o, W), oa = get3(o, AX);
    uint32_t total = 0;
    int sp = 0;
    int node = 0;
    auto leaf = [&](int first, int cnt) {
        for (int j = first; j < first + cnt; ++j) {
            float3 v0, v1, v2;
            uint32_t orig;
            int obj;
            load_tri(tris, j, v0, v1, v2, orig, obj);
            if (STATS) ctr.tris++;
            float t;
            if (tri_hit_axis<AX>(o, v0, v1, v2, t) && (t > thr || (t == thr && orig >= thr_tri)) && t < tmax_excl) {
                ++total;
                if (STATS) ctr.hits++;
                if (t < bt[KBUF - 1] || (t == bt[KBUF - 1] && orig < bi[KBUF - 1]))
                    kbuf_insert(bt, bi, bo, t, orig, obj);
            }
        }
    };
    while (true) {
        const NodeBoxes b = load_node(nodes, node);
        if (STATS) ctr.nodes++;
        const int l0 = b.l0, l1 = b.l1, n0 = b.n0, n1 = b.n1;
        // A triangle entirely behind the origin (all vertices <= o along AX) yields t <= 0,
        // so boxes are culled only on the perpendicular slab and on hi < o along AX.
        const bool h0 = ou >= b.lo0[U] && ou <= b.hi0[U] && ow >= b.lo0[W] && ow <= b.hi0[W] &&
                        b.hi0[AX] >= oa;
        const bool h1 = ou >= b.lo1[U] && ou <= b.hi1[U] && ow >= b.lo1[W] && ow <= b.hi1[W] &&
                        b.hi1[AX] >= oa;
        if (h0 && n0 > 0) leaf(l0, n0);
        if (h1 && n1 > 0) leaf(l1, n1);
        // next node without branches around the choice (as in collect_segment): the nearer
        // child along AX first when both are inner nodes, the other pushed; else pop
        const bool in0 = h0 && n0 == 0;
        const bool in1 = h1 && n1 == 0;
        const bool desc = in0 || in1;
        if (!desc && sp == 0) break;
        const int popv = stack[(sp > 0 ? sp - 1 : 0) * TR_BLOCK];  // read before this step's push
        const bool both = in0 && in1;
        const bool first0 = b.lo0[AX] <= b.lo1[AX];
        if (both) stack[sp * TR_BLOCK] = first0 ? l1 : l0;
        node = both ? (first0 ? l0 : l1) : (desc ? (in0 ? l0 : l1) : popv);
        sp += both ? 1 : (desc ? 0 : -1);
    }
    return total;
}

__device__ __forceinline__ void flush_counters(const TraceCounters& c, unsigned long long* stats) {
    uint64_t v[4] = {c.nodes, c.tris, c.hits, c.restarts};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint64_t x = v[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[k] = x;
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) atomicAdd(&stats[k], (unsigned long long)v[k]);
    }
}

enum { EMODE_COUNT = 0, EMODE_WRITE_PAIRS = 1, EMODE_WRITE_KEYS = 2, EMODE_HASH = 3 };

struct EdgeArgs {
    const float* starts;
    const float* ends;
    const int32_t* src_off;
    const int32_t* src_cnt;
    const int32_t* src_obj;
    uint64_t num_edges;
};

struct HashArgs {
    unsigned long long* slots;
    uint64_t size;
    int bitwise;
    int max_probes;
    unsigned long long* failed;
};

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// Insert of the reference's open-addressing table (query/src/cuda/mesh_overlap_edges.cu:62-104):
// fmix64 hash, linear probing, CAS from the all-ones empty key. A read before the CAS
// skips the atomic when the slot already holds the key or another key.
__device__ __forceinline__ void ht_insert(const HashArgs& h, uint64_t key) {
    const uint64_t k = fmix64(key);
    uint64_t s = h.bitwise ? (k & (h.size - 1)) : (k % h.size);
    const uint64_t probes = h.max_probes > 0 ? (uint64_t)h.max_probes : h.size;
    for (uint64_t p = 0; p < probes; ++p) {
        const unsigned long long cur = __hip_atomic_load(&h.slots[s], __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
        if (cur == (unsigned long long)key) return;
        if (cur == ~0ull) {
            const unsigned long long old = atomicCAS(&h.slots[s], ~0ull, (unsigned long long)key);
            if (old == ~0ull || old == (unsigned long long)key) return;
        }
        s = (s + 1 == h.size) ? 0 : s + 1;
    }
    atomicAdd(h.failed, 1ull);
}

// Membership test in a table filled by ht_insert (the lookup of
// query/src/cuda/mesh_containment.cu:56-68, without its 1000-probe cut-off).
__device__ __forceinline__ bool ht_contains(const HashArgs& h, uint64_t key) {
    const uint64_t k = fmix64(key);
    uint64_t s = h.bitwise ? (k & (h.size - 1)) : (k % h.size);
    for (uint64_t p = 0; p < h.size; ++p) {
        const unsigned long long cur = h.slots[s];
        if (cur == (unsigned long long)key) return true;
        if (cur == ~0ull) return false;
        s = (s + 1 == h.size) ? 0 : s + 1;
    }
    return false;
}

// One node visit of a segment walk (ray o + t d, inv = 1/d, rp its shear) over hits in
// [thr, tmax]: tests both children of node, offers the hits of leaf children to the
// buffer (total counts them) and moves node to the next node to visit. Returns false when
// the pass is complete. The stack's first DL entries live in LDS
// (stride TR_BLOCK), deeper ones in deep, a per-lane private array: a walk from a start
// node rarely nests deep, and a small LDS stack leaves room for more waves per CU.
// The computed t of any reported hit lies within the kz extent of its triangle's vertices
// (scaled by Sz) up to a few ulps; boxes are tested against [lo_t, hi_t] (thr and tmax with
// a relative margin) on that slab, and against the full line on all three slabs.
template <int DEPTH, int DL>
using SegDeep = int[DEPTH > DL ? DEPTH - DL : 1];

template <int DEPTH, bool STATS, int KB, int DL>
__device__ __forceinline__ bool seg_visit(const BNode* __restrict__ nodes, const TriRec* __restrict__ tris,
                                          const float3 o, const float3 inv, const RayPre& rp, float thr,
                                          float lo_t, float tmax, float hi_t, int& node, int& sp, int* stack,
                                          SegDeep<DEPTH, DL>& deep, float (&bt)[KB], uint32_t (&bi)[KB],
                                          int (&bo)[KB], uint32_t& total, TraceCounters& ctr) {
    // Box test of one child (its six bounds as values: no array or pointer selects, which
    // led the compiler to keep node fields in scratch memory). The kz slab's values
    // (bound - o[kz]) * Sz are the slab values of axis kz, bit for bit: Sz and inv[kz] are
    // both 1.0f / d[kz], and d[kz] is never 0 (|d[kz]| is the largest component).
    auto box = [&](float lx, float ly, float lz, float hx, float hy, float hz) -> bool {
        const float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
        const float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
        const float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
        const float nx = fminf(tx0, tx1), ny = fminf(ty0, ty1), nz = fminf(tz0, tz1);
        const float fx = fmaxf(tx0, tx1), fy = fmaxf(ty0, ty1), fz = fmaxf(tz0, tz1);
        const float tn = fmaxf(fmaxf(nx, ny), nz);
        const float tf = fminf(fminf(fx, fy), fz);
        const float kmin = rp.kz == 0 ? nx : (rp.kz == 1 ? ny : nz);
        const float kmax = rp.kz == 0 ? fx : (rp.kz == 1 ? fy : fz);
        return (tn <= tf + fabsf(tf) * 9.5367431640625e-7f) && kmax >= lo_t && kmin <= hi_t && lx <= hx;
    };
    auto leaf = [&](int first, int cnt) {
        for (int j = first; j < first + cnt; ++j) {
            float3 v0, v1, v2;
            uint32_t orig;
            int obj;
            load_tri(tris, j, v0, v1, v2, orig, obj);
            if (STATS) ctr.tris++;
            float t;
            if (tri_hit(o, rp, v0, v1, v2, t) && t >= thr && t <= tmax) {
                ++total;
                if (STATS) ctr.hits++;
                if (t < bt[KB - 1] || (t == bt[KB - 1] && orig < bi[KB - 1])) kbuf_insert(bt, bi, bo, t, orig, obj);
            }
        }
    };
    const NodeBoxes b = load_node(nodes, node);
    if (STATS) ctr.nodes++;
    const int l0 = b.l0, l1 = b.l1, n0 = b.n0, n1 = b.n1;
    const bool h0 = box(b.lo0[0], b.lo0[1], b.lo0[2], b.hi0[0], b.hi0[1], b.hi0[2]);
    const bool h1 = box(b.lo1[0], b.lo1[1], b.lo1[2], b.hi1[0], b.hi1[1], b.hi1[2]);
    if (h0 && n0 > 0) leaf(l0, n0);
    if (h1 && n1 > 0) leaf(l1, n1);
    // Next node without branches around the choice: descend (pushing the second child when
    // both are inner nodes) or pop. Lanes that descend and lanes that pop then stay in one
    // loop; a branch per case lets the compiler nest a descend loop inside a pop loop,
    // where a lane that must pop waits for every lane of the wave to pop.
    const bool in0 = h0 && n0 == 0;
    const bool in1 = h1 && n1 == 0;
    const bool desc = in0 || in1;
    if (!desc && sp == 0) return false;
    const int top = sp - 1;  // the entry a pop takes (read before this step's push)
    int popv = stack[(DL >= DEPTH || top < DL ? (top > 0 ? top : 0) : DL - 1) * TR_BLOCK];
    if (DL < DEPTH && !desc && top >= DL) popv = *(volatile int*)&deep[top - DL];
    if (in0 && in1) {
        if (DL >= DEPTH || sp < DL) stack[sp * TR_BLOCK] = l1;
        else deep[sp - DL] = l1;
    }
    node = desc ? (in0 ? l0 : l1) : popv;
    sp += (in0 && in1) ? 1 : (desc ? 0 : -1);
    return true;
}

// Margins of the kz-slab test for hits in [thr, tmax].
__device__ __forceinline__ float seg_lo_t(float thr) { return thr - fabsf(thr) * 9.5367431640625e-7f; }  // 2^-20
__device__ __forceinline__ float seg_hi_t(float tmax) { return tmax + fabsf(tmax) * 9.5367431640625e-7f; }

// All hits of one pass of a segment walk at or above thr, from node root down, into the
// buffer; returns their number.
template <int DEPTH, bool STATS, int KB, int DL>
__device__ __forceinline__ uint32_t collect_segment(const BNode* __restrict__ nodes,
                                                    const TriRec* __restrict__ tris, const float3 o,
                                                    const float3 inv, const RayPre& rp, float thr,
                                                    float tmax, int* stack, float (&bt)[KB],
                                                    uint32_t (&bi)[KB], int (&bo)[KB],
                                                    TraceCounters& ctr, int root = 0) {
    const float lo_t = seg_lo_t(thr), hi_t = seg_hi_t(tmax);
    uint32_t total = 0;
    int sp = 0;
    int node = root;
    SegDeep<DEPTH, DL> deep;
    while (seg_visit<DEPTH, STATS, KB, DL>(nodes, tris, o, inv, rp, thr, lo_t, tmax, hi_t, node, sp, stack,
                                           deep, bt, bi, bo, total, ctr)) {
    }
    return total;
}


// ------------------------------------------------------------------------------------
// Per-ray programs shared by the single-query kernels and the batched persistent kernel.
// ------------------------------------------------------------------------------------

// A.1 point containment (query/src/cuda/raytracing.cu:22-77 with the canonical chain):
// returns the object of the first chain hit when the chain length is odd, else -1.
template <int AX, int DEPTH, bool STATS>
__device__ __forceinline__ int point_ray(const BNode* __restrict__ nodes,
                                         const TriRec* __restrict__ tris, const float3 o,
                                         int* stack, TraceCounters& ctr) {
    const float eps = 1e-6f, tmax = 1e16f;
    float thr = eps;
    int count = 0;
    int first_obj = -1;
    float bt[KBUF];
    uint32_t bi[KBUF];
    int bo[KBUF];
    while (true) {
        kbuf_clear(bt, bi, bo);
        const uint32_t total =
            collect_axis<AX, DEPTH, STATS>(nodes, tris, o, thr, 0u, tmax, stack, bt, bi, bo, ctr);
#pragma unroll
        for (int k = 0; k < KBUF; ++k) {
            if ((uint32_t)k < total && count < 10000 && bt[k] >= thr) {
                if (count == 0) first_obj = bo[k];
                ++count;
                float nthr = bt[k] + eps;
                if (nthr <= bt[k]) nthr = nextafterf(bt[k], INFINITY);
                thr = nthr;
            }
        }
        if (total <= (uint32_t)KBUF || count >= 10000) break;
        if (STATS) ctr.restarts++;
    }
    return (count & 1) ? first_obj : -1;
}

// Chain rules: CHAIN_EPS (query/src/cuda/mesh_overlap_edges.cu:173-177): next threshold
// fl(t + 1e-6), or t + 2e-4 when the add is absorbed. CHAIN_NEXT
// (query/src/cuda/mesh_containment.cu:150-160): next threshold nextafterf(t, tmax).
enum { CHAIN_EPS = 0, CHAIN_NEXT = 1 };

// A.2 edge ray from s towards en: calls emit(target object) once per chain hit; returns
// the number of chain hits.
// Whether edge s-en may have chain hits: its length reaches 1e-6 and, with occupancy bits,
// the box of s-en covers a set bit. Every hit lies on s + t d, d the unit direction, t in
// [1e-6, len + 1e-6]: within 1e-6 beyond en on each axis, which the absolute pad of 2e-6
// covers; the rounding of the computed hit, of d and of en is covered by the 2^-14 relative
// widening (and the padded triangle boxes of the bits).
__device__ __forceinline__ bool edge_may_hit(const OccArgs& occ, const float3 s, const float3 en, uint64_t& words,
                                             int* start = nullptr) {
    const float eps = 1e-6f;
    const float dx = en.x - s.x, dy = en.y - s.y, dz = en.z - s.z;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    if (!(len >= eps)) return false;
    if (!occ.bits && !(start && occ.snode)) return true;
    return occ_segment_may_hit(occ, s, en, 2e-6f, words, start);
}

// The chain of a segment walk inside the sorted buffer of one pass (total hits at or above
// thr were offered): the members' buffer positions as a bit mask, in order; iter counts
// them and thr moves past each by the chain rule.
template <int KB>
__device__ __forceinline__ uint32_t seg_chain(const float (&bt)[KB], uint32_t total, int chain_rule, int max_iter,
                                              float tmax, int& iter, float& thr) {
    const float eps = 1e-6f;
    uint32_t chain = 0;
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        if ((uint32_t)k < total && iter < max_iter && thr <= tmax && bt[k] >= thr) {
            chain |= 1u << k;
            ++iter;
            float nthr;
            if (chain_rule == CHAIN_EPS) {
                nthr = bt[k] + eps;
                if (nthr <= bt[k]) nthr = bt[k] + 2e-4f;
            } else {
                nthr = nextafterf(bt[k], tmax);
                if (nthr <= bt[k]) nthr = bt[k] + eps;
            }
            thr = nthr;
        }
    }
    return chain;
}

// bo[k] for a k known only at run time (selects, so the buffer stays in registers).
template <int KB>
__device__ __forceinline__ int kbuf_at(const int (&bo)[KB], int k) {
    int v = bo[0];
#pragma unroll
    for (int q = 1; q < KB; ++q)
        if (k == q) v = bo[q];
    return v;
}

template <int DEPTH, bool STATS, int DL = DEPTH, typename Emit>
__device__ __forceinline__ int edge_ray(const BNode* __restrict__ nodes,
                                        const TriRec* __restrict__ tris, const OccArgs& occ,
                                        const float3 s, const float3 en, int chain_rule, int max_iter,
                                        int* stack, TraceCounters& ctr, Emit&& emit) {
    int root = 0;  // the walk starts below the root when the occupancy blocks allow it
    if (!edge_may_hit(occ, s, en, ctr.occ_words, &root)) return 0;
    if (STATS) ctr.walked++;
    const float eps = 1e-6f;
    const float dx = en.x - s.x, dy = en.y - s.y, dz = en.z - s.z;
    const float len = sqrtf(dx * dx + dy * dy + dz * dz);
    const float3 d = make_float3(dx / len, dy / len, dz / len);
    const float tmax = len + eps;
    const RayPre rp = ray_pre(d);
    const float3 inv = make_float3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    float thr = eps;
    int iter = 0;
    float bt[KBUF_EDGE];
    uint32_t bi[KBUF_EDGE];
    int bo[KBUF_EDGE];
    while (true) {
        kbuf_clear(bt, bi, bo);
        const uint32_t total =
            collect_segment<DEPTH, STATS, KBUF_EDGE, DL>(nodes, tris, s, inv, rp, thr, tmax, stack, bt, bi, bo, ctr,
                                                         root);
        // the chain over the buffer first, then one emission site for its members in order
        for (uint32_t chain = seg_chain(bt, total, chain_rule, max_iter, tmax, iter, thr); chain; chain &= chain - 1)
            emit(kbuf_at(bo, __builtin_ctz(chain)));
        if (total <= (uint32_t)KBUF_EDGE || iter >= max_iter || thr > tmax) break;
        if (STATS) ctr.restarts++;
    }
    return iter;
}

// Work counters of one object ray (the containment fields of the reference's
// MeshIntersectionProfilingStats, query/src/cuda/mesh_intersection.h:7-22).
struct ObjRayCtr {
    uint32_t iters = 0, hits = 0, suppressed = 0, additions = 0, toggles = 0, overflow = 0;
};

// One hit on target object obj: toggles its parity (times times) in the candidate list
// (ids, par, hit counts hc when given), or adds it while fewer than max_targets are
// listed, or counts an overflow.
__device__ __forceinline__ void cand_hit(int* ids, uint8_t* par, uint32_t* hc, int& count, int max_targets,
                                         int obj, uint32_t times, ObjRayCtr& c) {
    for (int j = 0; j < count; ++j)
        if (ids[j] == obj) {
            par[j] ^= (uint8_t)(times & 1u);
            if (hc) hc[j] += times;
            c.toggles += times;
            return;
        }
    if (count < max_targets) {
        ids[count] = obj;
        par[count] = (uint8_t)(times & 1u);
        if (hc) hc[count] = times;
        ++count;
        c.additions += 1;
        c.toggles += times - 1;
    } else {
        c.overflow += times;
    }
}

enum { OBJ_ALL_HITS = 0, OBJ_CHAIN_NEXTAFTER = 1, OBJ_CHAIN_STEP = 2 };

// A.3/A.4 object ray: +z from o, parity per target object in (ids, par) (at most
// max_targets distinct objects, further new objects counted as overflow). Returns the
// number of distinct objects recorded. Forms:
//  OBJ_ALL_HITS: every hit in [1e-4, 1e10] in (t, triangle) order; a hit of the same object
//    as the previous hit within 1e-5 of it is skipped (the any-hit forms,
//    query/src/cuda/mesh_intersection_anyhit.cu:194-239, mesh_containment.cu:196-256);
//  OBJ_CHAIN_NEXTAFTER: one closest hit per iteration over [thr, 1e10], thr from 1e-4 and
//    then nextafterf(t, 1e10), at most max_iter iterations (default 2048), a repeat of the
//    previous triangle within 1e-5 skipped (query/src/cuda/mesh_intersection.cu:134-230);
//  OBJ_CHAIN_STEP: one closest hit per iteration over [thr, 1e10), thr from 1e-4 and then
//    fl(t + 1e-4), at most max_iter iterations (default 2000); when that add is absorbed
//    every remaining iteration finds the same hit again (query/src/cuda/mesh_containment.cu:336-366).
// The closest hit of an iteration is the smallest (t, triangle index) at or above thr. The
// walk keeps the sorted hit buffer of the traversal kernels and refills it from the
// current threshold, so no hit is lost for any count.
template <int DEPTH>
__device__ __forceinline__ int object_ray_walk(const BNode* __restrict__ nodes,
                                               const TriRec* __restrict__ tris, const float3 o,
                                               int* stack, int* ids, uint8_t* par, uint32_t* hc,
                                               int max_targets, int mode, int max_iter, ObjRayCtr& c) {
    int count = 0;
    float bt[KBUF];
    uint32_t bi[KBUF];
    int bo[KBUF];
    TraceCounters ctr;
    if (mode == OBJ_ALL_HITS) {
        float lo_t = 1e-4f;
        uint32_t lo_tri = 0;
        const float tmax_excl = nextafterf(1e10f, INFINITY);
        bool has_last = false;
        int last_obj = -1;
        float last_t = 0.0f;
        c.iters = 1;
        while (true) {
            kbuf_clear(bt, bi, bo);
            const uint32_t total = collect_axis<2, DEPTH, false>(nodes, tris, o, lo_t, lo_tri, tmax_excl,
                                                                 stack, bt, bi, bo, ctr);
#pragma unroll
            for (int k = 0; k < KBUF; ++k) {
                if ((uint32_t)k < total) {
                    const float t = bt[k];
                    const int obj = bo[k];
                    const bool same = has_last && obj == last_obj && fabsf(t - last_t) <= 1e-5f;
                    has_last = true;
                    last_obj = obj;
                    last_t = t;
                    c.hits++;
                    if (same) c.suppressed++;
                    else cand_hit(ids, par, hc, count, max_targets, obj, 1u, c);
                }
            }
            if (total <= (uint32_t)KBUF) break;
            lo_t = bt[KBUF - 1];
            lo_tri = bi[KBUF - 1] + 1;
        }
        return count;
    }
    const float tmax = 1e10f;
    // window: t <= 1e10 for the nextafter loop; t < 1e10 for the step loop (a hit at 1e10
    // ends it like a miss)
    const float tmax_excl = mode == OBJ_CHAIN_NEXTAFTER ? nextafterf(tmax, INFINITY) : tmax;
    const uint32_t cap = max_iter > 0 ? (uint32_t)max_iter : (mode == OBJ_CHAIN_NEXTAFTER ? 2048u : 2000u);
    float thr = 1e-4f, last_t = -1.0f;
    uint32_t last_tri = 0xffffffffu;
    bool done = false;
    while (!done) {
        kbuf_clear(bt, bi, bo);
        const uint32_t total = collect_axis<2, DEPTH, false>(nodes, tris, o, thr, 0u, tmax_excl, stack,
                                                             bt, bi, bo, ctr);
        const uint32_t nb = total < (uint32_t)KBUF ? total : (uint32_t)KBUF;
        uint32_t k = 0;
        while (true) {
            if (c.iters >= cap) {
                done = true;
                break;
            }
            if (mode == OBJ_CHAIN_NEXTAFTER && thr > tmax) {
                c.iters++;
                done = true;
                break;
            }
            while (k < nb && bt[k] < thr) ++k;
            if (k == nb) {
                if (total > (uint32_t)KBUF) break;  // more hits lie beyond the buffer: refill
                c.iters++;                          // the iteration whose trace misses
                done = true;
                break;
            }
            c.iters++;
            c.hits++;
            const float t = bt[k];
            const uint32_t tri = bi[k];
            const int obj = bo[k];
            if (mode == OBJ_CHAIN_NEXTAFTER) {
                const bool same = tri == last_tri && last_t >= 0.0f && fabsf(t - last_t) <= 1e-5f;
                if (same) c.suppressed++;
                else cand_hit(ids, par, hc, count, max_targets, obj, 1u, c);
                last_t = t;
                last_tri = tri;
                float nx = nextafterf(t, tmax);
                if (nx <= t) nx = t + 1e-6f;
                if (nx <= thr) nx = nextafterf(thr, tmax);
                if (nx <= thr) {
                    done = true;
                    break;
                }
                thr = nx;
            } else {
                const float nx = t + 1e-4f;
                if (nx <= t) {
                    // absorbed: this hit is found again on each remaining iteration
                    const uint32_t r = cap - c.iters + 1;
                    c.hits += r - 1;
                    c.iters = cap;
                    cand_hit(ids, par, hc, count, max_targets, obj, r, c);
                    done = true;
                    break;
                }
                cand_hit(ids, par, hc, count, max_targets, obj, 1u, c);
                thr = nx;
            }
            ++k;
        }
    }
    return count;
}

// The all-hits object ray without per-candidate hit counts (the batch engine's form).
template <int DEPTH>
__device__ __forceinline__ int object_ray_parity(const BNode* __restrict__ nodes,
                                                 const TriRec* __restrict__ tris, const float3 o,
                                                 int* stack, int* ids, uint8_t* par,
                                                 int max_targets, unsigned long long& over) {
    ObjRayCtr c;
    const int n = object_ray_walk<DEPTH>(nodes, tris, o, stack, ids, par, nullptr, max_targets, OBJ_ALL_HITS, 0, c);
    over += c.overflow;
    return n;
}


// Lane-refill point containment (the program of point_ray) over rays [lo, hi) of a ray
// list: ray r is point imap[r] (r itself without imap). The wave hands the next ray of
// [lo, hi) to any lane whose ray has finished, in the same loop iteration, so lanes do not
// idle behind the longest ray of the wave; each ray still runs the same traversal order,
// hit buffer, restarts and chain as point_ray, so the results are identical. Writes the
// result for point p to out[p] in the given output kind. All 64 lanes must call it with
// the same lo and hi.
template <int AX, int DEPTH>
__device__ __forceinline__ void points_refill(const BNode* __restrict__ nodes,
                                              const TriRec* __restrict__ tris,
                                              const float* __restrict__ pts,
                                              const uint32_t* __restrict__ imap, uint64_t lo,
                                              uint64_t hi, int out_kind, void* __restrict__ out,
                                              int* stack) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float eps = 1e-6f, tmax = 1e16f;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint64_t run_hi = hi;
    uint64_t next = lo;  // wave-uniform: first ray of the run not handed out yet
    bool active = false;
    uint32_t pi = 0;
    float3 o = make_float3(0.0f, 0.0f, 0.0f);
    float thr = eps;
    int count = 0, first_obj = -1, node = 0, sp = 0;
    uint32_t total = 0;
    float bt[KBUF];
    uint32_t bi[KBUF];
    int bo[KBUF];
    kbuf_clear(bt, bi, bo);
    while (true) {
        const uint64_t idle = __ballot(!active);
        if (idle && next < run_hi) {
            const uint64_t r = next + (uint64_t)__popcll(idle & lt_mask);
            if (!active && r < run_hi) {
                pi = imap ? imap[r] : (uint32_t)r;
                o = ld3f(pts, pi);
                thr = eps;
                count = 0;
                first_obj = -1;
                total = 0;
                node = 0;
                sp = 0;
                kbuf_clear(bt, bi, bo);
                active = true;
            }
            next += (uint64_t)__popcll(idle);
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;
        const float ou = get3(o, U), ow = get3(o, W), oa = get3(o, AX);
        const NodeBoxes b = load_node(nodes, node);
        const bool h0 = ou >= b.lo0[U] && ou <= b.hi0[U] && ow >= b.lo0[W] && ow <= b.hi0[W] &&
                        b.hi0[AX] >= oa;
        const bool h1 = ou >= b.lo1[U] && ou <= b.hi1[U] && ow >= b.lo1[W] && ow <= b.hi1[W] &&
                        b.hi1[AX] >= oa;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bool hit = c == 0 ? h0 : h1;
            const int cnt = c == 0 ? b.n0 : b.n1;
            const int first = c == 0 ? b.l0 : b.l1;
            if (hit && cnt > 0) {
                for (int j = first; j < first + cnt; ++j) {
                    float3 v0, v1, v2;
                    uint32_t orig;
                    int obj;
                    load_tri(tris, j, v0, v1, v2, orig, obj);
                    float t;
                    if (tri_hit_axis<AX>(o, v0, v1, v2, t) && t >= thr && t < tmax) {
                        ++total;
                        if (t < bt[KBUF - 1] || (t == bt[KBUF - 1] && orig < bi[KBUF - 1]))
                            kbuf_insert(bt, bi, bo, t, orig, obj);
                    }
                }
            }
        }
        const bool in0 = h0 && b.n0 == 0;
        const bool in1 = h1 && b.n1 == 0;
        if (in0 && in1) {
            const bool first0 = b.lo0[AX] <= b.lo1[AX];
            stack[sp * TR_BLOCK] = first0 ? b.l1 : b.l0;
            ++sp;
            node = first0 ? b.l0 : b.l1;
            continue;
        }
        if (in0 || in1) {
            node = in0 ? b.l0 : b.l1;
            continue;
        }
        if (sp > 0) {
            --sp;
            node = stack[sp * TR_BLOCK];
            continue;
        }
        // traversal pass complete: walk the chain inside the buffer (as point_ray)
#pragma unroll
        for (int k = 0; k < KBUF; ++k) {
            if ((uint32_t)k < total && count < 10000 && bt[k] >= thr) {
                if (count == 0) first_obj = bo[k];
                ++count;
                float nthr = bt[k] + eps;
                if (nthr <= bt[k]) nthr = nextafterf(bt[k], INFINITY);
                thr = nthr;
            }
        }
        if (total > (uint32_t)KBUF && count < 10000) {
            kbuf_clear(bt, bi, bo);
            total = 0;
            node = 0;
            continue;
        }
        const int res = (count & 1) ? first_obj : -1;
        if (out_kind == RS3D_OUT_RAY_RESULT) {
            rs3d_ray_result rr;
            rr.ray_id = (int32_t)pi;
            rr.polygon_index = res;
            static_cast<rs3d_ray_result*>(out)[pi] = rr;
        } else if (out_kind == RS3D_OUT_FLAG_U8) {
            static_cast<uint8_t*>(out)[pi] = res >= 0 ? 1 : 0;
        } else {
            static_cast<int32_t*>(out)[pi] = res;
        }
        active = false;
    }
}

}  // namespace rs3d
