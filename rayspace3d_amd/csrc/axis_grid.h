// axis_grid.h -- the axis grid of a BVH (a 2-D grid of padded triangle boxes for rays
// along one axis) and the per-ray program on it, shared by the single-query kernel
// (axis_grid.hip) and the batch engine.
#pragma once

#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

// Cells are numbered in tiles of 8 (u) x 4 (w): cell (cu, cw) is
// ((cw / 4) * ntu + cu / 8) * 32 + (cw % 4) * 8 + cu % 8, with ntu = ceil(nu / 8), so one
// 32-bit occupancy word covers one tile and a tile's cells are consecutive in the cell
// table and the records. Occupancy has two levels: tsum[g] holds, for tiles 32g..32g+31, a
// bit per tile that lists records (.x) and the number of such tiles before the group (.y);
// occ holds the cell bits of those tiles only, in tile order (rank .y + popcount of the
// lower bits of .x). A point in an empty tile is answered from tsum alone, which is small
// enough (8 B per 1024 cells) to stay in L2 while records stream through it.
struct AxisGrid {
    int axis = 2, nu = 1, nw = 1, ntu = 1;
    float lo_u = 0.0f, lo_w = 0.0f, inv_u = 0.0f, inv_w = 0.0f;
    uint64_t entries = 0;
    uint2* cells = nullptr;     // tiled cells + 1 entries: first record, box top of that record
    uint2* tsum = nullptr;      // per 32 tiles: occupied-tile bits, occupied tiles before
    uint32_t* occ = nullptr;    // per occupied tile: its 32 cell bits
    float4* rec = nullptr;      // 4 float4 per record
    size_t bytes = 0;
    double expected = 0.0;
    double build_ms = 0.0;
};

struct AGArgs {
    const uint2* cells;
    const uint2* tsum;
    const uint32_t* occ;
    const float4* rec;
    float lo_u, lo_w, inv_u, inv_w;
    int nu, nw, ntu;
};

inline AGArgs ag_args(const AxisGrid* g) {
    return AGArgs{g->cells, g->tsum, g->occ, g->rec, g->lo_u, g->lo_w, g->inv_u, g->inv_w, g->nu, g->nw, g->ntu};
}

__host__ __device__ __forceinline__ uint32_t ag_cell_id(int cu, int cw, int ntu) {
    return (((uint32_t)cw >> 2) * (uint32_t)ntu + ((uint32_t)cu >> 3)) * 32u + (((uint32_t)cw & 3u) << 3) +
           ((uint32_t)cu & 7u);
}

// Rank of tile t among the occupied tiles, given its group's tsum entry; -1 when t is empty.
__device__ __forceinline__ int64_t ag_tile_rank(uint2 g, uint32_t t) {
    const uint32_t b = t & 31u;
    if (!((g.x >> b) & 1u)) return -1;
    return (int64_t)g.y + __popc(g.x & ((1u << b) - 1u));
}

// Cell coordinate along one axis: monotone in x (clamped), NaN goes to cell 0.
__device__ __forceinline__ int ag_cell(float x, float lo, float inv, int n) {
    const float f = fminf(fmaxf((x - lo) * inv, 0.0f), (float)(n - 1));
    return (int)f;
}

// The dataset-box filter of a filter-refine grid run (points that fail box_filter_pass are
// answered -1 without a walk) and the lane's count of points found inside.
struct GridFilter {
    FBox box;
    uint32_t inside;
};

__device__ __forceinline__ void grid_write_result(int out_kind, void* __restrict__ out, uint32_t pi, int res) {
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
}

// Point rays [lo, run_hi) of a ray list on the grid, lane-refill form: the 64 lanes of the
// calling wave take rays from the run as they finish. The per-ray program (hit buffer,
// restarts, chain, output) is the BVH kernels'; only the candidate walk differs: a lane
// reads two records per step (the second is in flight while the first is tested). With
// gf, the lane counts the points it finds inside. With PRE, the list holds points whose
// record range was looked up already: qrange[r] is ray imap[r]'s (first, end) record.
template <int AX, bool STATS, bool PRE = false>
__device__ __forceinline__ void points_grid_run(const AGArgs& G, const float* __restrict__ pts,
                                                const uint32_t* __restrict__ imap, uint64_t lo, uint64_t run_hi,
                                                int out_kind, void* __restrict__ out, TraceCounters& ctr,
                                                GridFilter* gf = nullptr, const uint2* qrange = nullptr) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float eps = 1e-6f, tmax = 1e16f;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t next = lo;
    bool active = false;
    uint32_t pi = 0, j = 0, je = 0, js = 0;
    float3 o = make_float3(0.0f, 0.0f, 0.0f);
    float thr = eps;
    int count = 0, first_obj = -1;
    uint32_t total = 0;
    float bt[KBUF_GRID];
    uint32_t bi[KBUF_GRID];
    int bo[KBUF_GRID];
    kbuf_clear(bt, bi, bo);
    while (true) {
        const uint64_t idle = __ballot(!active);
        if (idle && next < run_hi) {
            const uint64_t r = next + (uint64_t)__popcll(idle & lt_mask);
            if (!active && r < run_hi) {
                pi = imap ? imap[r] : (uint32_t)r;
                o = ld3f(pts, pi);
                if (PRE) {
                    const uint2 rg = qrange[r];
                    js = rg.x;
                    je = rg.y;
                } else {
                    const int cu = ag_cell(get3(o, U), G.lo_u, G.inv_u, G.nu);
                    const int cw = ag_cell(get3(o, W), G.lo_w, G.inv_w, G.nw);
                    const uint32_t c = ag_cell_id(cu, cw, G.ntu);
                    // most rays of a sparse scene fall in empty tiles or cells: the occupancy
                    // levels answer those without the range fetch
                    js = je = 0;
                    const int64_t tr = ag_tile_rank(G.tsum[c >> 10], c >> 5);
                    if (tr >= 0 && ((G.occ[tr] >> (c & 31)) & 1u)) {
                        const uint2 c0 = G.cells[c], c1 = G.cells[c + 1];
                        if (STATS) ctr.nodes++;
                        // the first record's box top decides as the walk's exit test would
                        if (!(__uint_as_float(c0.y) < get3(o, AX))) {
                            js = c0.x;
                            je = c1.x;
                        }
                    }
                }
                j = js;
                thr = eps;
                count = 0;
                first_obj = -1;
                total = 0;
                kbuf_clear(bt, bi, bo);
                active = true;
            }
            next += (uint64_t)__popcll(idle);
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;
        const float ou = get3(o, U), ow = get3(o, W), oa = get3(o, AX);
        if (j < je) {
            const bool two = j + 1 < je;
            const uint64_t b0 = 4 * (uint64_t)j, b1 = 4 * (uint64_t)(two ? j + 1 : j);
            const float4 e0 = G.rec[b0], e1 = G.rec[b0 + 1], e2 = G.rec[b0 + 2], e3 = G.rec[b0 + 3];
            const float4 f0 = G.rec[b1], f1 = G.rec[b1 + 1], f2 = G.rec[b1 + 2], f3 = G.rec[b1 + 3];
            // one record: box test in the plane, then the triangle
            auto test = [&](const float4& a0, const float4& a1, const float4& a2, const float4& a3) {
                if (ou >= a3.x && ou <= a3.y && ow >= a3.z && ow <= a3.w) {
                    float t;
                    const uint32_t orig = (uint32_t)__float_as_int(a0.w);
                    if (tri_hit_axis<AX>(o, make_float3(a0.x, a0.y, a0.z), make_float3(a1.x, a1.y, a1.z),
                                         make_float3(a2.x, a2.y, a2.z), t) &&
                        t >= thr && t < tmax) {
                        ++total;
                        if (STATS) ctr.hits++;
                        if (t < bt[KBUF_GRID - 1] || (t == bt[KBUF_GRID - 1] && orig < bi[KBUF_GRID - 1]))
                            kbuf_insert(bt, bi, bo, t, orig, __float_as_int(a1.w));
                    }
                }
            };
            // records are ordered by box top: from the first that ends below the origin on,
            // none can be hit
            if (STATS) ctr.tris++;
            if (e2.w < oa) {
                j = je;
                continue;
            }
            test(e0, e1, e2, e3);
            if (!two) {
                ++j;
                continue;
            }
            if (STATS) ctr.tris++;
            if (f2.w < oa) {
                j = je;
                continue;
            }
            test(f0, f1, f2, f3);
            j += 2;
            continue;
        }
        // list complete: walk the chain inside the buffer (as the BVH kernels do)
#pragma unroll
        for (int k = 0; k < KBUF_GRID; ++k) {
            if ((uint32_t)k < total && count < 10000 && bt[k] >= thr) {
                if (count == 0) first_obj = bo[k];
                ++count;
                float nthr = bt[k] + eps;
                if (nthr <= bt[k]) nthr = nextafterf(bt[k], INFINITY);
                thr = nthr;
            }
        }
        if (total > (uint32_t)KBUF_GRID && count < 10000) {
            kbuf_clear(bt, bi, bo);
            total = 0;
            j = js;
            if (STATS) ctr.restarts++;
            continue;
        }
        const int res = (count & 1) ? first_obj : -1;
        grid_write_result(out_kind, out, pi, res);
        if (gf && res >= 0) gf->inside++;
        active = false;
    }
}

// Queue capacity of the two-phase grid form (points; 6 KB of LDS with their ranges).
constexpr uint32_t AG_QCAP = 512;

// Point rays [lo, hi) on the grid in two alternating phases. Phase 1 streams the points 256
// at a time (four independent loads per lane in flight): origin, with FILT the dataset-box
// filter of gf.box, cell, occupancy bit and the cell-table entries; points that fail the
// filter, fall in an empty cell or lie above the cell's highest box top are answered -1 at
// once, the others are queued (qpi: point, qrg: record range; AG_QCAP entries of LDS).
// Ray r of the list is point imap[r] (r itself without imap); answers go out as out_kind.
// Phase 2 walks the queue with lane refill (points_grid_run, PRE) whenever phase 1 could
// not take another 256 points, and at the end. cand counts filter passes (FILT), gf.inside
// the points found inside. The whole wave calls it with the same lo, hi.
template <int AX, bool FILT>
__device__ __forceinline__ void points_grid_two_phase(const AGArgs& G, const float* __restrict__ pts,
                                                      const uint32_t* __restrict__ imap, uint64_t lo, uint64_t hi,
                                                      int out_kind, void* __restrict__ out, uint32_t* qpi,
                                                      uint2* qrg, uint32_t& cand, GridFilter& gf,
                                                      TraceCounters& ctr) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t base = lo;
    while (true) {
        uint32_t nq = 0;
        for (; base < hi && nq + 4 * 64 <= AG_QCAP; base += 4 * 64) {
            float3 o[4];
            bool ok[4];
            uint32_t cell[4], word[4], pid[4];
            uint2 rg[4], ts[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t r = base + 64 * k + lane;
                ok[k] = r < hi;
                pid[k] = ok[k] ? (imap ? imap[r] : (uint32_t)r) : 0u;
                o[k] = ok[k] ? ld3f(pts, pid[k]) : make_float3(0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (FILT) {
                    ok[k] = ok[k] && box_filter_pass<AX>(o[k], gf.box);
                    cand += ok[k] ? 1u : 0u;
                }
                const int cu = ag_cell(get3(o[k], U), G.lo_u, G.inv_u, G.nu);
                const int cw = ag_cell(get3(o[k], W), G.lo_w, G.inv_w, G.nw);
                cell[k] = ag_cell_id(cu, cw, G.ntu);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) ts[k] = ok[k] ? G.tsum[cell[k] >> 10] : make_uint2(0u, 0u);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t tr = ag_tile_rank(ts[k], cell[k] >> 5);
                word[k] = tr >= 0 ? G.occ[tr] : 0u;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ok[k] = (word[k] >> (cell[k] & 31)) & 1u;
                uint2 c0 = make_uint2(0u, 0u), c1 = make_uint2(0u, 0u);
                if (ok[k]) {
                    c0 = G.cells[cell[k]];
                    c1 = G.cells[cell[k] + 1];
                }
                // a point above the cell's highest box top has nothing to test (the walk's
                // exit test on the first record)
                ok[k] = ok[k] && !(__uint_as_float(c0.y) < get3(o[k], AX));
                rg[k] = make_uint2(c0.x, c1.x);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t r = base + 64 * k + lane;
                if (r < hi && !ok[k]) grid_write_result(out_kind, out, pid[k], -1);
                const uint64_t m = __ballot(ok[k]);
                if (ok[k]) {
                    const uint32_t slot = nq + (uint32_t)__popcll(m & lt_mask);
                    qpi[slot] = pid[k];
                    qrg[slot] = rg[k];
                }
                nq += (uint32_t)__popcll(m);
            }
        }
        __syncthreads();
        points_grid_run<AX, false, true>(G, pts, qpi, 0, nq, out_kind, out, ctr, &gf, qrg);
        __syncthreads();
        if (base >= hi) break;
    }
}

}  // namespace rs3d
