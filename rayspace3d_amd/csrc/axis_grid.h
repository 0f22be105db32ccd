// axis_grid.h -- the axis grid of a BVH (a 2-D grid of padded triangle boxes for rays
// along one axis) and the per-ray program on it, shared by the single-query kernel
// (axis_grid.hip) and the batch engine.
#pragma once

#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

// Cells are numbered in tiles of 8 (u) x 4 (w): cell (cu, cw) is
// ((cw / 4) * ntu + cu / 8) * 32 + (cw % 4) * 8 + cu % 8, with ntu = ceil(nu / 8), so one
// 32-bit occupancy word covers one tile and a tile's cells are consecutive in the records.
// Occupancy has two levels: tsum[g] holds, for tiles 32g..32g+31, a bit per tile that lists
// records (.x) and the number of such tiles before the group (.y); occ holds, for those
// tiles only, in tile order (rank .y + popcount of the lower bits of .x): the tile's cell
// bits, the highest top of its records (the largest first-record top of its cells; +inf
// when one is NaN) and the rank of its first occupied cell. A point in an empty tile is
// answered from tsum alone, which is small enough (8 B per 1024 cells) to stay in L2 while
// records stream through it; a point above its tile's highest top has nothing to test in
// any of its cells and is answered without the cell-table read. The cell table lists the
// occupied cells only (rank = occ .z + popcount of the lower cell bits), a quarter of the
// dense table's size or less, which keeps it in the caches far better.
struct AxisGrid {
    int axis = 2, nu = 1, nw = 1, ntu = 1;
    float lo_u = 0.0f, lo_w = 0.0f, inv_u = 0.0f, inv_w = 0.0f;
    uint64_t entries = 0;
    uint2* cells = nullptr;     // occupied cells + 1 entries: first record, top of that record
    uint2* tsum = nullptr;      // per 32 tiles: occupied-tile bits, occupied tiles before
    uint4* occ = nullptr;       // per occupied tile: cell bits, highest top (bits), first cell rank
    float* rec = nullptr;       // AG_RECF floats per record
    int32_t* obj = nullptr;     // object of each original triangle
    size_t bytes = 0;
    double expected = 0.0;
    double build_ms = 0.0;
};

// A record is ten floats: v0, v1, v2 and the original triangle index (its bits) -- 40
// bytes, so a record pair at an even index is 80 bytes, five 16-byte loads. A list is
// ordered by its triangles' largest vertex coordinate along the axis ("top"), largest
// first: a ray whose origin lies above a triangle's top cannot hit it (every vertex is at or
// below the origin, so T <= 0 and t <= 0 in tri_hit_axis), hence nothing from the first
// such record on. The object of a triangle is looked up (obj) only for a point found inside.
constexpr int AG_RECF = 10;

struct AgRec {
    float3 v0, v1, v2;
    uint32_t orig;
};

struct AGArgs {
    const uint2* cells;
    const uint2* tsum;
    const uint4* occ;
    const float* rec;
    const int32_t* obj;
    float lo_u, lo_w, inv_u, inv_w;
    int nu, nw, ntu;
};

inline AGArgs ag_args(const AxisGrid* g) {
    return AGArgs{g->cells, g->tsum, g->occ, g->rec, g->obj, g->lo_u, g->lo_w, g->inv_u, g->inv_w, g->nu, g->nw, g->ntu};
}

__host__ __device__ __forceinline__ uint32_t ag_cell_id(int cu, int cw, int ntu) {
    return (((uint32_t)cw >> 2) * (uint32_t)ntu + ((uint32_t)cu >> 3)) * 32u + (((uint32_t)cw & 3u) << 3) +
           ((uint32_t)cu & 7u);
}

// Pins a loaded value in a register at this point. The grid kernels issue the loads of a
// wave's four rays together and use the values only for some lanes; without the pin the
// compiler moves each load into the branch that uses it and waits for it there, which
// serialises the four loads. (An empty statement: it emits no instruction.)
__device__ __forceinline__ void ag_pin(uint32_t& x) { asm volatile("" : "+v"(x)); }

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

// Record j (8-byte aligned: five 8-byte loads).
__device__ __forceinline__ AgRec ag_load_one(const float* __restrict__ rec, uint64_t j) {
    const float2* p = reinterpret_cast<const float2*>(rec + AG_RECF * j);
    const float2 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
    AgRec r;
    r.v0 = make_float3(a.x, a.y, b.x);
    r.v1 = make_float3(b.y, c.x, c.y);
    r.v2 = make_float3(d.x, d.y, e.x);
    r.orig = __float_as_uint(e.y);
    return r;
}

// Records j and j + 1 for an even j (16-byte aligned: five 16-byte loads).
__device__ __forceinline__ void ag_load_pair(const float* __restrict__ rec, uint64_t j, AgRec& r0, AgRec& r1) {
    const float4* p = reinterpret_cast<const float4*>(rec + AG_RECF * j);
    const float4 a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
    r0.v0 = make_float3(a.x, a.y, a.z);
    r0.v1 = make_float3(a.w, b.x, b.y);
    r0.v2 = make_float3(b.z, b.w, c.x);
    r0.orig = __float_as_uint(c.y);
    r1.v0 = make_float3(c.z, c.w, d.x);
    r1.v1 = make_float3(d.y, d.z, d.w);
    r1.v2 = make_float3(e.x, e.y, e.z);
    r1.orig = __float_as_uint(e.w);
}

// The record's top: its largest vertex coordinate along the axis (NaN only where every
// coordinate is NaN, as in the pad record).
template <int AX>
__device__ __forceinline__ float ag_top(const AgRec& r) {
    return fmaxf(fmaxf(get3(r.v0, AX), get3(r.v1, AX)), get3(r.v2, AX));
}

// One grid record against a point ray: the triangle reaches the origin's height and is hit
// at t in [thr, 1e16). (The record keeps no box: the padded box contains every hit, so
// testing it first could only skip triangle tests.)
template <int AX>
__device__ __forceinline__ bool grid_rec_hit(const float3 o, const AgRec& r, float thr, float& t) {
    if (ag_top<AX>(r) < get3(o, AX)) return false;
    return tri_hit_axis<AX>(o, r.v0, r.v1, r.v2, t) && t >= thr && t < 1e16f;
}

// The answer of a point whose chain counted `count` hits, the first on original triangle
// first_orig: that triangle's object when the count is odd, else -1.
__device__ __forceinline__ int ag_answer(const AGArgs& G, int count, int first_orig) {
    return (count & 1) ? G.obj[(uint32_t)first_orig] : -1;
}

// Where the item-parallel form puts a point's answer: the session's LDS answers when it
// buffers them (points sbase.. of a contiguous run), else the output.
struct GridSink {
    int out_kind;
    void* out;
    int32_t* res;
    uint32_t sbase;
    __device__ __forceinline__ void put(uint32_t pid, int r) const {
        if (res) res[pid - sbase] = r;
        else grid_write_result(out_kind, out, pid, r);
    }
};

// Point rays [lo, run_hi) of a ray list on the grid, lane-refill form: the 64 lanes of the
// calling wave take rays from the run as they finish. The per-ray program (hit buffer,
// restarts, chain, output) is the BVH kernels'; only the candidate walk differs: a lane
// reads two records per step (the second is in flight while the first is tested). This
// sequential form is the counting pass (STATS: cell-table reads, records read, hits,
// restarts) that gives the grid launches' algorithmic bytes.
template <int AX, bool STATS>
__device__ __forceinline__ void points_grid_run(const AGArgs& G, const float* __restrict__ pts,
                                                const uint32_t* __restrict__ imap, uint64_t lo, uint64_t run_hi,
                                                int out_kind, void* __restrict__ out, TraceCounters& ctr) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float eps = 1e-6f;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint64_t next = lo;
    bool active = false;
    uint32_t pi = 0, j = 0, je = 0, js = 0;
    float3 o = make_float3(0.0f, 0.0f, 0.0f);
    float thr = eps;
    int count = 0, first_orig = -1;
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
                const int cu = ag_cell(get3(o, U), G.lo_u, G.inv_u, G.nu);
                const int cw = ag_cell(get3(o, W), G.lo_w, G.inv_w, G.nw);
                const uint32_t c = ag_cell_id(cu, cw, G.ntu);
                // most rays of a sparse scene fall in empty tiles or cells: the occupancy
                // levels answer those without the range fetch
                js = je = 0;
                const int64_t tr = ag_tile_rank(G.tsum[c >> 10], c >> 5);
                const uint4 oc = tr >= 0 ? G.occ[tr] : make_uint4(0u, 0u, 0u, 0u);
                if (((oc.x >> (c & 31)) & 1u) && !(__uint_as_float(oc.y) < get3(o, AX))) {
                    const uint32_t rk = oc.z + (uint32_t)__popc(oc.x & ((1u << (c & 31)) - 1u));
                    const uint2 c0 = G.cells[rk], c1 = G.cells[rk + 1];
                    if (STATS) ctr.nodes++;
                    // the first record's box top decides as the walk's exit test would
                    if (!(__uint_as_float(c0.y) < get3(o, AX))) {
                        js = c0.x;
                        je = c1.x;
                    }
                }
                j = js;
                thr = eps;
                count = 0;
                first_orig = -1;
                total = 0;
                kbuf_clear(bt, bi, bo);
                active = true;
            }
            next += (uint64_t)__popcll(idle);
        }
        if (__ballot(active) == 0) break;
        if (!active) continue;
        const float oa = get3(o, AX);
        if (j < je) {
            // (lists have even lengths and start at even records: j is even here)
            const bool two = j + 1 < je;
            AgRec e, f;
            if (two) ag_load_pair(G.rec, j, e, f);
            else e = f = ag_load_one(G.rec, j);
            // one record (the buffer keeps the original triangle as the hit's tag)
            auto test = [&](const AgRec& r) {
                float t;
                if (grid_rec_hit<AX>(o, r, thr, t)) {
                    ++total;
                    if (STATS) ctr.hits++;
                    if (t < bt[KBUF_GRID - 1] || (t == bt[KBUF_GRID - 1] && r.orig < bi[KBUF_GRID - 1]))
                        kbuf_insert(bt, bi, bo, t, r.orig, (int)r.orig);
                }
            };
            // records are ordered by top: from the first below the origin on, none can be hit
            if (STATS) ctr.tris++;
            if (ag_top<AX>(e) < oa) {
                j = je;
                continue;
            }
            test(e);
            if (!two) {
                ++j;
                continue;
            }
            if (STATS) ctr.tris++;
            if (ag_top<AX>(f) < oa) {
                j = je;
                continue;
            }
            test(f);
            j += 2;
            continue;
        }
        // list complete: walk the chain inside the buffer (as the BVH kernels do)
#pragma unroll
        for (int k = 0; k < KBUF_GRID; ++k) {
            if ((uint32_t)k < total && count < 10000 && bt[k] >= thr) {
                if (count == 0) first_orig = bo[k];
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
        grid_write_result(out_kind, out, pi, ag_answer(G, count, first_orig));
        active = false;
    }
}

// One point's walk, one record per step, with the hit buffer and its restarts: the
// sequential walk's result for list [js, je) (used for the few entries the item-parallel
// form leaves).
template <int AX>
__device__ __forceinline__ void grid_walk_one(const AGArgs& G, const float3 o, uint32_t pid, uint32_t js,
                                              uint32_t je, const GridSink& sink, GridFilter* gf) {
    const float eps = 1e-6f;
    const float oa = get3(o, AX);
    float thr = eps;
    int count = 0, first_orig = -1;
    float bt[KBUF_GRID];
    uint32_t bi[KBUF_GRID];
    int bo[KBUF_GRID];
    while (true) {
        kbuf_clear(bt, bi, bo);
        uint32_t total = 0;
        for (uint32_t j = js; j < je; ++j) {
            const AgRec r = ag_load_one(G.rec, j);
            if (ag_top<AX>(r) < oa) break;
            float t;
            if (grid_rec_hit<AX>(o, r, thr, t)) {
                ++total;
                if (t < bt[KBUF_GRID - 1] || (t == bt[KBUF_GRID - 1] && r.orig < bi[KBUF_GRID - 1]))
                    kbuf_insert(bt, bi, bo, t, r.orig, (int)r.orig);
            }
        }
#pragma unroll
        for (int k = 0; k < KBUF_GRID; ++k) {
            if ((uint32_t)k < total && count < 10000 && bt[k] >= thr) {
                if (count == 0) first_orig = bo[k];
                ++count;
                float nthr = bt[k] + eps;
                if (nthr <= bt[k]) nthr = nextafterf(bt[k], INFINITY);
                thr = nthr;
            }
        }
        if (!(total > (uint32_t)KBUF_GRID && count < 10000)) break;
    }
    const int res = ag_answer(G, count, first_orig);
    sink.put(pid, res);
    if (gf && res >= 0) gf->inside++;
}

// Item-parallel form of the walk (used by the launches that give it the LDS below): the
// queued points' lists are cut into record pairs ("items", entry-major); each lane takes an
// item, tests its two records against its point (every record of the list: those past the
// walk's exit test fail the same box-top test) and appends the hits to an LDS list in item
// order, so each entry's hits are consecutive and the list is ordered by entry. One lane per
// entry then runs the chain on its hits, taking them in (t, triangle) order by selection:
// the hits the sequential walk's buffer would take in turn, so the result is the same.
// Entries with more than AG_CHAIN_MAX hits, and those from the first whose hits did not
// all fit the list on, are walked by grid_walk_one afterwards. A wave's items are
// independent loads, so no lane idles behind another's long list. An entry carries its
// point's origin, so the items read it from LDS and not again from memory.
constexpr uint32_t AG_SESSION = 256;    // rays of one session (four per lane)
constexpr uint32_t AG_IQCAP = 128;      // queue entries of one items pass
constexpr uint32_t AG_HCAP = 128;       // hits of one items pass
constexpr uint32_t AG_CHAIN_MAX = 16;   // hits one entry's chain takes from the list
constexpr int32_t AG_DEFER = INT32_MIN; // res mark: queued by a later pass of the session

struct AgItemLds {
    uint32_t pi[AG_IQCAP];     // point
    uint32_t first[AG_IQCAP];  // first record of its list (even)
    uint32_t pre[AG_IQCAP];    // record pairs of the list, then the pairs before it
    float ox[AG_IQCAP];        // the point's origin
    float oy[AG_IQCAP];
    float oz[AG_IQCAP];
    float ht[AG_HCAP];
    uint32_t ho[AG_HCAP];      // original triangle
    uint16_t hq[AG_HCAP];      // entry
    int32_t res[AG_SESSION];   // per ray of the session: its answer when the session buffers
                               // them, else -1; AG_DEFER while it waits for a later pass
};

__device__ __forceinline__ uint32_t ag_wave_incl(uint32_t v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}

// First index in hq[0, n) whose entry is at least q (hq is non-decreasing).
__device__ __forceinline__ uint32_t ag_hit_lb(const uint16_t* hq, uint32_t n, uint32_t q) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (hq[mid] < q) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int AX>
__device__ __forceinline__ void points_grid_items(const AGArgs& G, uint32_t nq, AgItemLds& L, const GridSink& sink,
                                                  GridFilter* gf) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const float eps = 1e-6f;
    uint32_t items = 0;
    for (uint32_t b = 0; b < nq; b += 64) {
        const uint32_t q = b + lane;
        const uint32_t n = q < nq ? L.pre[q] : 0u;
        const uint32_t inc = ag_wave_incl(n, lane);
        if (q < nq) L.pre[q] = items + inc - n;
        items += __shfl(inc, 63, 64);
    }
    __syncthreads();
    uint32_t hpos = 0, qdrop = nq;
    for (uint32_t b = 0; b < items; b += 64) {
        const uint32_t i = b + lane;
        bool h0 = false, h1 = false;
        float t0 = 0.0f, t1 = 0.0f;
        uint32_t q = 0;
        AgRec e, f;
        if (i < items) {
            // the last entry whose first item is at or before i
            uint32_t lo = 0, hi = nq;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (L.pre[mid] <= i) lo = mid;
                else hi = mid;
            }
            q = lo;
            const uint64_t j = (uint64_t)L.first[q] + 2 * (uint64_t)(i - L.pre[q]);
            const float3 o = make_float3(L.ox[q], L.oy[q], L.oz[q]);
            ag_load_pair(G.rec, j, e, f);
            h0 = grid_rec_hit<AX>(o, e, eps, t0);
            h1 = grid_rec_hit<AX>(o, f, eps, t1);
        }
        const uint64_t m0 = __ballot(h0), m1 = __ballot(h1);
        const uint32_t p0 = hpos + (uint32_t)__popcll(m0 & lt_mask) + (uint32_t)__popcll(m1 & lt_mask);
        const uint32_t p1 = p0 + (h0 ? 1u : 0u);
        if (h0 && p0 < AG_HCAP) {
            L.ht[p0] = t0;
            L.ho[p0] = e.orig;
            L.hq[p0] = (uint16_t)q;
        }
        if (h1 && p1 < AG_HCAP) {
            L.ht[p1] = t1;
            L.ho[p1] = f.orig;
            L.hq[p1] = (uint16_t)q;
        }
        // the entry of the first hit that did not fit
        const uint64_t md = __ballot((h0 && p0 == AG_HCAP) || (h1 && p1 == AG_HCAP));
        if (md) qdrop = __shfl(q, __ffsll((unsigned long long)md) - 1, 64);
        hpos += (uint32_t)(__popcll(m0) + __popcll(m1));
    }
    const uint32_t nh = hpos < AG_HCAP ? hpos : AG_HCAP;
    __syncthreads();
    // the chain per entry; entries it cannot finish move to the front of pi / first / pre
    // (pre then holds the end of the list)
    uint32_t nfb = 0;
    for (uint32_t b = 0; b < nq; b += 64) {
        const uint32_t q = b + lane;
        const bool valid = q < nq;
        uint32_t pid = 0, js = 0, je = 0, s0 = 0, s1 = 0;
        if (valid) {
            pid = L.pi[q];
            js = L.first[q];
            je = js + 2 * ((q + 1 < nq ? L.pre[q + 1] : items) - L.pre[q]);
            s0 = ag_hit_lb(L.hq, nh, q);
            s1 = ag_hit_lb(L.hq, nh, q + 1);
        }
        const bool fb = valid && (q >= qdrop || s1 - s0 > AG_CHAIN_MAX);
        if (valid && !fb) {
            float thr = eps;
            int count = 0, first_orig = -1;
            while (true) {
                // the smallest (t, triangle) at or above thr
                float bt = INFINITY;
                uint32_t bi = 0xffffffffu;
                bool found = false;
                for (uint32_t k = s0; k < s1; ++k) {
                    const float t = L.ht[k];
                    if (t >= thr) {
                        const uint32_t tri = L.ho[k];
                        if (!found || t < bt || (t == bt && tri < bi)) {
                            bt = t;
                            bi = tri;
                            found = true;
                        }
                    }
                }
                if (!found) break;
                if (count == 0) first_orig = (int)bi;
                ++count;
                float nthr = bt + eps;
                if (nthr <= bt) nthr = nextafterf(bt, INFINITY);
                thr = nthr;
            }
            const int res = ag_answer(G, count, first_orig);
            sink.put(pid, res);
            if (gf && res >= 0) gf->inside++;
        }
        const uint64_t mf = __ballot(fb);
        if (fb) {
            // (every lane's reads of this batch precede its writes, and later batches' entries
            // lie beyond every slot written so far)
            const uint32_t slot = nfb + (uint32_t)__popcll(mf & lt_mask);
            const float ox = L.ox[q], oy = L.oy[q], oz = L.oz[q];
            L.pi[slot] = pid;
            L.first[slot] = js;
            L.pre[slot] = je;
            L.ox[slot] = ox;
            L.oy[slot] = oy;
            L.oz[slot] = oz;
        }
        nfb += (uint32_t)__popcll(mf);
    }
    if (nfb) {
        __syncthreads();
        for (uint32_t k = lane; k < nfb; k += 64)
            grid_walk_one<AX>(G, make_float3(L.ox[k], L.oy[k], L.oz[k]), L.pi[k], L.first[k], L.pre[k], sink, gf);
    }
}

// Point rays [lo, hi) on the grid in sessions of AG_SESSION rays (four per lane). Phase 1
// loads, four independent loads per lane in flight: origin, with FILT the dataset-box
// filter of gf.box, cell, tile summary, occupancy word (cell bits, tile box top) and, for an
// occupied cell below the tile's top, the cell-table entries; rays that fail the filter,
// fall in an empty tile or cell, or lie above the tile's or the cell's highest box top are
// answered -1 at once, the others are queued in L with their origin and record range.
// Phase 2 walks the queue (points_grid_items). A session whose queue would exceed
// AG_IQCAP marks the rest AG_DEFER and queues them in further passes (each pass reloads only
// its marked rays). Ray r of the list is point imap[r] (r itself without imap); answers go
// out as out_kind, over a contiguous run (no imap) gathered in L.res and written as whole
// lines at the session's end. cand counts filter passes (FILT), gf.inside the points found
// inside. The whole wave calls it with the same lo, hi.
template <int AX, bool FILT>
__device__ __forceinline__ void points_grid_two_phase(const AGArgs& G, const float* __restrict__ pts,
                                                      const uint32_t* __restrict__ imap, uint64_t lo, uint64_t hi,
                                                      int out_kind, void* __restrict__ out, uint32_t& cand,
                                                      GridFilter& gf, AgItemLds& L) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const bool buf = imap == nullptr;
    for (uint64_t sbase = lo; sbase < hi; sbase += AG_SESSION) {
        const uint64_t send = hi - sbase > AG_SESSION ? sbase + AG_SESSION : hi;
        const GridSink sink{out_kind, out, buf ? L.res : nullptr, (uint32_t)sbase};
        bool first_pass = true;
        while (true) {
            float3 o[4];
            bool act[4], ok[4];
            uint32_t cell[4], pid[4];
            uint2 ts[4];
            uint4 oc[4];
            // Every load below is issued by every lane, inactive ones at an index known to be
            // valid (the session's first ray, tile group and occupied tile 0, cell 0) with the
            // value then ignored: a load under a branch is waited for inside that branch, which
            // would serialise the four rays' loads of each step.
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t r = sbase + 64 * k + lane;
                act[k] = r < send && (first_pass || L.res[r - sbase] == AG_DEFER);
                const uint64_t rr = act[k] ? r : sbase;
                pid[k] = imap ? imap[rr] : (uint32_t)rr;
                o[k] = ld3f(pts, pid[k]);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                // a deferred ray passed the filter in the first pass (and was counted there)
                ok[k] = act[k] && (!FILT || !first_pass || box_filter_pass<AX>(o[k], gf.box));
                if (FILT && first_pass) cand += ok[k] ? 1u : 0u;
                const int cu = ag_cell(get3(o[k], U), G.lo_u, G.inv_u, G.nu);
                const int cw = ag_cell(get3(o[k], W), G.lo_w, G.inv_w, G.nw);
                cell[k] = ag_cell_id(cu, cw, G.ntu);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) ts[k] = G.tsum[cell[k] >> 10];  // (cells are clamped into the grid)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ag_pin(ts[k].x);
                ag_pin(ts[k].y);
                if (!ok[k]) ts[k] = make_uint2(0u, 0u);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t tr = ag_tile_rank(ts[k], cell[k] >> 5);
                oc[k] = G.occ[tr >= 0 ? tr : 0];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ag_pin(oc[k].x);
                ag_pin(oc[k].y);
                ag_pin(oc[k].z);
                if (ag_tile_rank(ts[k], cell[k] >> 5) < 0) oc[k] = make_uint4(0u, 0u, 0u, 0u);
            }
            // queue the points in an occupied cell that do not lie above the tile's highest box
            // top, with their origins (which are then no longer held in registers)
            uint32_t nq = 0, slot[4];
            float oa[4];
            bool deferred = false;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                oa[k] = get3(o[k], AX);
                ok[k] = ((oc[k].x >> (cell[k] & 31)) & 1u) && !(__uint_as_float(oc[k].y) < oa[k]);
                const uint32_t rs = 64u * (uint32_t)k + (uint32_t)lane;
                if (act[k] && !ok[k]) {
                    L.res[rs] = -1;
                    if (!buf) grid_write_result(out_kind, out, pid[k], -1);
                }
                const uint64_t m = __ballot(ok[k]);
                slot[k] = nq + (uint32_t)__popcll(m & lt_mask);
                if (ok[k]) {
                    if (slot[k] < AG_IQCAP) {
                        L.pi[slot[k]] = pid[k];
                        L.ox[slot[k]] = o[k].x;
                        L.oy[slot[k]] = o[k].y;
                        L.oz[slot[k]] = o[k].z;
                        L.res[rs] = -1;  // the walk writes the answer (buffered or out)
                    } else {
                        L.res[rs] = AG_DEFER;
                        deferred = true;
                        ok[k] = false;
                    }
                }
                nq += (uint32_t)__popcll(m);
            }
            // their record ranges; a point above its cell's highest box top (the walk's exit
            // test on the first record) gets an empty list
            uint2 c0[4];
            uint32_t c1[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                // the cell's rank among the occupied cells (0: a valid entry for the others)
                const uint32_t c = ok[k] ? oc[k].z + (uint32_t)__popc(oc[k].x & ((1u << (cell[k] & 31)) - 1u)) : 0u;
                c0[k] = G.cells[c];
                c1[k] = G.cells[c + 1].x;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                ag_pin(c0[k].x);
                ag_pin(c0[k].y);
                ag_pin(c1[k]);
                if (ok[k]) {
                    const bool below = !(__uint_as_float(c0[k].y) < oa[k]);
                    L.first[slot[k]] = c0[k].x;
                    L.pre[slot[k]] = below ? (c1[k] - c0[k].x) >> 1 : 0u;
                }
            }
            __syncthreads();
            points_grid_items<AX>(G, nq < AG_IQCAP ? nq : AG_IQCAP, L, sink, &gf);
            __syncthreads();
            if (!__ballot(deferred)) break;
            first_pass = false;
        }
        if (buf) {
            const uint32_t cnt = (uint32_t)(send - sbase);
            for (uint32_t k = lane; k < cnt; k += 64) grid_write_result(out_kind, out, (uint32_t)sbase + k, L.res[k]);
            __syncthreads();
        }
    }
}

}  // namespace rs3d
