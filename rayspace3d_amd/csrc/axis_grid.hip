// axis_grid.hip -- 2-D grid of padded triangle boxes for axis-aligned point rays.
//
// Point containment casts rays along one axis, so every ray is a vertical half-line over a
// point of the perpendicular plane. The grid covers that plane (the BVH's padded bounds)
// with square cells; cell c lists every triangle whose padded box -- the BVH's leaf box,
// computed by the same prim_padded_box -- overlaps c in the plane and whose projected
// footprint, widened by the fp32 test's error bound, meets c (ag_tri_meets_cell), as a
// self-contained 40-byte record (vertices and original index), ordered by the triangle's
// top (its largest vertex coordinate along the axis), largest first. A ray reads its cell's
// entry in the cell table (first record, and that record's top, so that a ray above it
// reads no record at all) and tests the triangles of the records that reach its origin.
// They include every triangle the fp32 test can report for a ray of the cell: the padding
// guarantees that no triangle is hit from outside its box, the footprint test that none is
// hit from a cell its widened footprint misses. So the hits are the BVH walk's, and the
// hit buffer, restarts and chain are the BVH kernel's: the results are the same. The gain
// is the walk: one range lookup and a few records instead of ~30 dependent node fetches.
// The object of a point's answer comes from a per-triangle map.
//
// Build (per BVH and axis, once): the cell edge is a fraction of the triangles' mean padded
// extent in the plane (RS3D_OPT_GRID_CELL_PCT, widened while the records exceed their
// budget); records per triangle from the cells of its box that its footprint meets, budget
// and list-length checks (rejected grids leave the launches on the BVH), one
// (cell << 32 | descending top) key per record, per-cell counts and their scan as the cell
// ranges, a bucket and a rank pass that order each list, and a gather of the records.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "axis_grid.h"

namespace rs3d {

void axis_grid_free(AxisGrid* g) {
    if (!g) return;
    if (g->cells) (void)hipFree(g->cells);
    if (g->occ) (void)hipFree(g->occ);
    if (g->tsum) (void)hipFree(g->tsum);
    if (g->rec) (void)hipFree(g->rec);
    if (g->obj) (void)hipFree(g->obj);
    delete g;
}

constexpr uint64_t AG_MAX_CELLS = 1ull << 26;
constexpr int AG_MAX_DIM = 32768;
constexpr double AG_MAX_UNIFORM_LIST = 64.0;    // records per cell, averaged over cells
constexpr double AG_MAX_WEIGHTED_LIST = 256.0;  // averaged over records (dense spots)
constexpr uint64_t AG_SMALL_RECORDS = 1ull << 19;  // grids this small take finer cells
constexpr double AG_MIN_CELL_PCT = 10.0;
// Triangles whose box spans more cells than this are listed in every cell of the box (the
// footprint test would cost more than it saves).
constexpr uint64_t AG_FOOTPRINT_MAX_SPAN = 4096;

// Unsigned key whose order is the float order.
__device__ __forceinline__ uint32_t ag_ord(float f) {
    const uint32_t u = __float_as_uint(f);
    return u ^ ((u >> 31) ? 0xffffffffu : 0x80000000u);
}

template <int AX>
__device__ __forceinline__ void ag_cells(const TriRec& r, const AGArgs& G, float (&plo)[3], float (&phi)[3],
                                         int& cu0, int& cu1, int& cw0, int& cw1) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    prim_padded_box(make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                    make_float3(r.v2.x, r.v2.y, r.v2.z), plo, phi);
    cu0 = ag_cell(plo[U], G.lo_u, G.inv_u, G.nu);
    cu1 = ag_cell(phi[U], G.lo_u, G.inv_u, G.nu);
    cw0 = ag_cell(plo[W], G.lo_w, G.inv_w, G.nw);
    cw1 = ag_cell(phi[W], G.lo_w, G.inv_w, G.nw);
}

// Whether the square of cell (cu, cw) can hold a ray that the fp32 axis triangle test
// (tri_hit_axis) reports for the triangle: false only when the square lies strictly outside
// one edge line of the triangle's projection by more than that edge's error bound, so
// listing the triangle only in cells that pass loses no hit.
//  * The square is widened by 2^-7 of a cell: the cell mapping ag_cell rounds (x - lo) and
//    its product with inv, a relative error under 2^-23 of at most 32,768 cells (2^-8 of a
//    cell). Boundary cells, which take every ray beyond the grid, are never pruned.
//  * The test decides the side of an edge BC by the sign of fl(Cx By) - fl(Cy Bx), C and B
//    the vertices less the origin (each rounded once): a wrong sign needs |U| <= 6 2^-24
//    |C - o| |B - o|, i.e. a distance under 2^-21.4 |C - o| |B - o| / |BC| from the line.
//    A pruned square is inside the padded box, so |C - o|, |B - o| <= D = Lmax + 4 pad
//    (Lmax the longest projected edge, pad the box padding): the edge's bound is taken as
//    2^-16 D^2 / |BC|. Edges shorter than Lmax / 256, and degenerate projections, prune
//    nothing. All of it in double.
template <int AX>
__device__ __forceinline__ bool ag_tri_meets_cell(const TriRec& r, int cu, int cw, const AGArgs& G) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    if (cu <= 0 || cw <= 0 || cu >= G.nu - 1 || cw >= G.nw - 1) return true;
    const float3 v[3] = {make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                         make_float3(r.v2.x, r.v2.y, r.v2.z)};
    double pu[3], pw[3];
    float mag = 0.0f, ext = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        pu[k] = (double)get3(v[k], U);
        pw[k] = (double)get3(v[k], W);
    }
    {
        const float lo[3] = {fminf(fminf(v[0].x, v[1].x), v[2].x), fminf(fminf(v[0].y, v[1].y), v[2].y),
                             fminf(fminf(v[0].z, v[1].z), v[2].z)};
        const float hi[3] = {fmaxf(fmaxf(v[0].x, v[1].x), v[2].x), fmaxf(fmaxf(v[0].y, v[1].y), v[2].y),
                             fmaxf(fmaxf(v[0].z, v[1].z), v[2].z)};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ext = fmaxf(ext, hi[k] - lo[k]);
            mag = fmaxf(mag, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
        }
    }
    const double area2 = (pu[1] - pu[0]) * (pw[2] - pw[0]) - (pw[1] - pw[0]) * (pu[2] - pu[0]);
    double l2max = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double du = pu[(k + 1) % 3] - pu[k], dw = pw[(k + 1) % 3] - pw[k];
        l2max = fmax(l2max, du * du + dw * dw);
    }
    if (!(fabs(area2) > l2max * 0x1p-40) || !(l2max < 1e300)) return true;  // degenerate or not finite
    const double d = sqrt(l2max) + 4.0 * (double)box_pad(mag, ext);
    const double hu = 1.0 / (double)G.inv_u, hw = 1.0 / (double)G.inv_w;
    const double u0 = (double)G.lo_u + cu * hu - hu * 0x1p-7, u1 = (double)G.lo_u + (cu + 1) * hu + hu * 0x1p-7;
    const double w0 = (double)G.lo_w + cw * hw - hw * 0x1p-7, w1 = (double)G.lo_w + (cw + 1) * hw + hw * 0x1p-7;
    const double sgn = area2 > 0.0 ? 1.0 : -1.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double au = pu[k], aw = pw[k];
        const double eu = pu[(k + 1) % 3] - au, ew = pw[(k + 1) % 3] - aw;
        const double len2 = eu * eu + ew * ew;
        if (len2 < l2max * 0x1p-16) continue;
        // the corner farthest to the triangle's side of the line (sgn * cross(e, q - a) > 0 inside)
        const double c00 = sgn * (eu * (w0 - aw) - ew * (u0 - au)), c10 = sgn * (eu * (w0 - aw) - ew * (u1 - au));
        const double c01 = sgn * (eu * (w1 - aw) - ew * (u0 - au)), c11 = sgn * (eu * (w1 - aw) - ew * (u1 - au));
        const double best = fmax(fmax(c00, c10), fmax(c01, c11));
        // best / |e| is the corner's distance from the line; the bound 2^-16 d^2 / |e| times |e|
        if (best < -0x1p-16 * d * d) return false;
    }
    return true;
}

// Sum of the triangles' padded extents in the plane (the larger of the two), for the cell size.
template <int AX>
__global__ __launch_bounds__(256) void k_ag_size(const TriRec* __restrict__ tris, uint64_t n,
                                                 unsigned long long* __restrict__ sum) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    double v = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const TriRec r = tris[i];
        float plo[3], phi[3];
        prim_padded_box(make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                        make_float3(r.v2.x, r.v2.y, r.v2.z), plo, phi);
        const float e = fmaxf(phi[U] - plo[U], phi[W] - plo[W]);
        if (e < 3.0e38f) v += (double)e;  // infinite or NaN extents do not set the cell size
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(reinterpret_cast<double*>(acc_stripe(sum)), part[0] + part[1] + part[2] + part[3]);
}

template <int AX>
__global__ __launch_bounds__(256) void k_ag_span(const TriRec* __restrict__ tris, uint64_t n, AGArgs G,
                                                 uint32_t* __restrict__ span, unsigned long long* __restrict__ total) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (i < n) {
        float plo[3], phi[3];
        int cu0, cu1, cw0, cw1;
        const TriRec r = tris[i];
        ag_cells<AX>(r, G, plo, phi, cu0, cu1, cw0, cw1);
        v = (unsigned long long)(cu1 - cu0 + 1) * (unsigned long long)(cw1 - cw0 + 1);
        if (v <= AG_FOOTPRINT_MAX_SPAN) {  // the cells of the box that the footprint meets
            v = 0;
            for (int cw = cw0; cw <= cw1; ++cw)
                for (int cu = cu0; cu <= cu1; ++cu) v += ag_tri_meets_cell<AX>(r, cu, cw, G) ? 1 : 0;
        }
        span[i] = v > 0xffffffffull ? 0xffffffffu : (uint32_t)v;
    }
    v = block_sum_256(v);
    if (threadIdx.x == 0 && v) atomicAdd(acc_stripe(total), v);
}

template <int AX>
__global__ __launch_bounds__(256) void k_ag_emit(const TriRec* __restrict__ tris, uint64_t n, AGArgs G,
                                                 const uint32_t* __restrict__ off, uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals, uint32_t* __restrict__ counts) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float plo[3], phi[3];
    int cu0, cu1, cw0, cw1;
    const TriRec r = tris[i];
    ag_cells<AX>(r, G, plo, phi, cu0, cu1, cw0, cw1);
    const float top = fmaxf(fmaxf(get3(make_float3(r.v0.x, r.v0.y, r.v0.z), AX), get3(make_float3(r.v1.x, r.v1.y, r.v1.z), AX)),
                            get3(make_float3(r.v2.x, r.v2.y, r.v2.z), AX));
    const uint64_t low = 0xffffffffu - ag_ord(top);  // the record's top (ag_top), descending
    uint64_t k = off[i];
    const bool test = (uint64_t)(cu1 - cu0 + 1) * (uint64_t)(cw1 - cw0 + 1) <= AG_FOOTPRINT_MAX_SPAN;
    for (int cw = cw0; cw <= cw1; ++cw)
        for (int cu = cu0; cu <= cu1; ++cu) {
            if (test && !ag_tri_meets_cell<AX>(r, cu, cw, G)) continue;  // as counted by k_ag_span
            const uint64_t cell = ag_cell_id(cu, cw, G.ntu);
            keys[k] = (cell << 32) | low;
            vals[k] = (uint32_t)i;
            ++k;
            atomicAdd(&counts[cell], 1u);
        }
}

// Sum of count^2 over the cells (the expected list length times the record count).
__global__ __launch_bounds__(256) void k_ag_sq(const uint32_t* __restrict__ counts, uint64_t nc,
                                               unsigned long long* __restrict__ acc) {
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long c = counts[i];
        s += c * c;
    }
    s = block_sum_256(s);
    if (threadIdx.x == 0 && s) atomicAdd(acc_stripe(acc), s);
}

// Record i of the sorted (cell, top) keys goes to its cell's padded list: pstart[cell] +
// its rank in the cell (i - start[cell]).
__global__ __launch_bounds__(256) void k_ag_gather(const TriRec* __restrict__ tris, const uint32_t* __restrict__ vals,
                                                   const uint64_t* __restrict__ keys, const uint32_t* __restrict__ start,
                                                   const uint32_t* __restrict__ pstart, uint64_t m,
                                                   float* __restrict__ rec) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t cell = (uint32_t)(keys[i] >> 32);
    const uint64_t d = (uint64_t)pstart[cell] + (i - start[cell]);
    const TriRec r = tris[vals[i]];
    float2* p = reinterpret_cast<float2*>(rec + AG_RECF * d);
    p[0] = make_float2(r.v0.x, r.v0.y);
    p[1] = make_float2(r.v0.z, r.v1.x);
    p[2] = make_float2(r.v1.y, r.v1.z);
    p[3] = make_float2(r.v2.x, r.v2.y);
    p[4] = make_float2(r.v2.z, r.v0.w);  // .w of v0: the original triangle index
}

// Sort of the (cell << 32 | descending top, triangle) records by cell, then descending top,
// then triangle (the order of a stable sort of the keys in emission order), in two passes:
// k_ag_bucket places every record in its cell's range (in any order), k_ag_rank moves each
// to its rank among its cell's records. Lists are short (their mean weighted by length is
// bounded by the acceptance checks), so a rank costs one pass over the list.
__global__ __launch_bounds__(256) void k_ag_bucket(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                   uint64_t m, const uint32_t* __restrict__ start,
                                                   uint32_t* __restrict__ cursor, uint32_t* __restrict__ bkey,
                                                   uint32_t* __restrict__ bval, uint32_t* __restrict__ bcell) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = keys[i];
    const uint32_t cell = (uint32_t)(k >> 32);
    const uint32_t pos = start[cell] + atomicAdd(&cursor[cell], 1u);
    bkey[pos] = (uint32_t)k;
    bval[pos] = vals[i];
    bcell[pos] = cell;
}

__global__ __launch_bounds__(256) void k_ag_rank(const uint32_t* __restrict__ bkey, const uint32_t* __restrict__ bval,
                                                 const uint32_t* __restrict__ bcell, uint64_t m,
                                                 const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ counts, uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t cell = bcell[i], key = bkey[i], val = bval[i];
    const uint32_t s0 = start[cell], cnt = counts[cell];
    uint32_t rank = 0;
    for (uint32_t j = s0; j < s0 + cnt; ++j) {
        const uint32_t kj = bkey[j], vj = bval[j];
        rank += (kj < key || (kj == key && vj < val)) ? 1u : 0u;
    }
    keys[s0 + rank] = ((uint64_t)cell << 32) | key;
    vals[s0 + rank] = val;
}

// The object of each original triangle (the records keep only the triangle's index).
__global__ __launch_bounds__(256) void k_ag_objmap(const TriRec* __restrict__ tris, uint64_t n,
                                                   int32_t* __restrict__ obj) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TriRec r = tris[i];
    const uint32_t orig = __float_as_uint(r.v0.w);
    if (orig < n) obj[orig] = __float_as_int(r.v1.w);
}

// Lists padded to an even number of records (a walk step, or an item, is a record pair):
// pcounts = counts rounded up to even; the pad is a sentinel record (NaN vertices, original
// index 0xffffffff) after a cell's last record, whose triangle test fails for every ray (t
// is NaN), so no chain can count it.
__global__ __launch_bounds__(256) void k_ag_pcounts(const uint32_t* __restrict__ counts, uint64_t nc,
                                                    uint32_t* __restrict__ pcounts,
                                                    unsigned long long* __restrict__ odd) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = c < nc ? counts[c] : 0u;
    if (c < nc) pcounts[c] = k + (k & 1u);
    const unsigned long long o = block_sum_256(k & 1u);
    if (threadIdx.x == 0 && o) atomicAdd(acc_stripe(odd), o);
}

__global__ __launch_bounds__(256) void k_ag_pad(const uint32_t* __restrict__ counts,
                                                const uint32_t* __restrict__ pstart, uint64_t nc,
                                                float* __restrict__ rec) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc || !(counts[c] & 1u)) return;
    const uint64_t d = (uint64_t)pstart[c] + counts[c];
    const float nan = __int_as_float(0x7fc00000);
    float2* p = reinterpret_cast<float2*>(rec + AG_RECF * d);
    for (int k = 0; k < 4; ++k) p[k] = make_float2(nan, nan);
    p[4] = make_float2(nan, __int_as_float(-1));
}

// Occupancy words from the counts: one cell per lane, a wave's ballot is the words of its
// two tiles (nc is a multiple of 32).
__global__ __launch_bounds__(256) void k_ag_occ(const uint32_t* __restrict__ counts, uint64_t nc,
                                                uint32_t* __restrict__ words) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(c < nc && counts[c] != 0u);
    if ((threadIdx.x & 31) == 0 && c < nc) words[c / 32] = (uint32_t)(m >> (threadIdx.x & 32));
}

// Tile summaries: one tile per lane; lane 32g of the ballot writes group g's tile bits and
// their count (the scan of the counts gives each group's rank base).
__global__ __launch_bounds__(256) void k_ag_tsum(const uint32_t* __restrict__ words, uint64_t nt,
                                                 uint2* __restrict__ tsum, uint32_t* __restrict__ gcount) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t m = __ballot(t < nt && words[t] != 0u);
    if ((threadIdx.x & 31) == 0 && t < nt) {
        const uint32_t b = (uint32_t)(m >> (threadIdx.x & 32));
        tsum[t / 32] = make_uint2(b, 0u);
        gcount[t / 32] = (uint32_t)__popc(b);
    }
}

// Compact entries of the occupied tiles at their ranks: the cell bits, the highest top of the
// tile's records (the cells' first-record tops, from the cell table: +inf when one is NaN,
// so that no point is ever skipped on it) and the rank of the tile's first occupied cell
// (crank of its first cell); and each group's rank base.
__global__ __launch_bounds__(256) void k_ag_fine(const uint32_t* __restrict__ words, uint64_t nt,
                                                 const uint32_t* __restrict__ gbase, const uint32_t* __restrict__ crank,
                                                 const uint2* __restrict__ cells, uint2* __restrict__ tsum,
                                                 uint4* __restrict__ occ) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const uint32_t base = gbase[t / 32];
    if ((t & 31) == 0) tsum[t / 32].y = base;
    const uint32_t w = words[t];
    if (!w) return;
    const uint32_t cb = crank[32 * t];
    float top = -INFINITY;
    for (uint32_t k = 0; k < (uint32_t)__popc(w); ++k) {
        const float c = __uint_as_float(cells[cb + k].y);
        top = c != c ? INFINITY : (c > top ? c : top);
    }
    occ[base + __popc(tsum[t / 32].x & ((1u << (t & 31)) - 1u))] = make_uint4(w, __float_as_uint(top), cb, 0u);
}

// Occupied-cell flags (for their ranks) and their number.
__global__ __launch_bounds__(256) void k_ag_oflag(const uint32_t* __restrict__ counts, uint64_t nc,
                                                  uint32_t* __restrict__ flag, unsigned long long* __restrict__ total) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t f = c < nc && counts[c] != 0u ? 1u : 0u;
    if (c < nc) flag[c] = f;
    const unsigned long long v = block_sum_256(f);
    if (threadIdx.x == 0 && v) atomicAdd(acc_stripe(total), v);
}

// Cell table over the occupied cells, in cell order (occupied cell c at crank[c]): (first
// record, top of the first record -- the largest of the cell), and (record count, -inf)
// after the last; an occupied cell's list ends where the next one's starts.
template <int AX>
__global__ __launch_bounds__(256) void k_ag_cells(const uint32_t* __restrict__ start,
                                                  const uint32_t* __restrict__ counts,
                                                  const uint32_t* __restrict__ crank, const float* __restrict__ rec,
                                                  uint64_t nc, uint32_t nocc, uint32_t total,
                                                  uint2* __restrict__ cells) {
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) {
        cells[nocc] = make_uint2(total, __float_as_uint(-INFINITY));
        cells[nocc + 1] = make_uint2(total, __float_as_uint(-INFINITY));
    }
    if (c >= nc || !counts[c]) return;
    const uint32_t s = start[c];
    cells[crank[c]] = make_uint2(s, __float_as_uint(ag_top<AX>(ag_load_one(rec, s))));
}

// Point rays on the grid: wave b owns rays [b * per_wave, ...), in the two-phase form; the
// counting pass (STATS: cell lookups, records read, hits, restarts) walks them in one.
// Held to 80 VGPRs (6 waves per SIMD, a few bytes of spills): 1 M points on a small grid
// ran 194 us against 205 us at the 83-88 the compiler chooses (5 waves).
template <int AX, bool STATS>
__global__ __launch_bounds__(TR_BLOCK) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_points_grid(
    AGArgs G, const float* __restrict__ pts, uint64_t n, const uint32_t* __restrict__ imap, int out_kind,
    void* __restrict__ out, uint32_t per_wave, unsigned long long* __restrict__ stats) {
    const uint64_t lo = (uint64_t)blockIdx.x * per_wave;
    const uint64_t hi = lo + per_wave < n ? lo + per_wave : n;
    TraceCounters ctr;
    if (STATS) {
        points_grid_run<AX, true>(G, pts, imap, lo, hi, out_kind, out, ctr);
        flush_counters(ctr, stats);
    } else {
        __shared__ AgItemLds lds;
        uint32_t cand = 0;
        GridFilter gf{};
        points_grid_two_phase<AX, false>(G, pts, imap, lo, hi, out_kind, out, cand, gf, lds);
    }
}

// Filter-refine in one pass: one wave per run of per_wave points, in the two-phase grid
// form with the dataset-box filter. counts (striped accumulator): candidates in the high 32
// bits, points inside in the low 32 (n < 2^32, so no carry).
template <int AX>
__global__ __launch_bounds__(TR_BLOCK) void k_points_grid_filtered(AGArgs G, const float* __restrict__ pts,
                                                                   uint64_t n, int32_t* __restrict__ out,
                                                                   uint32_t per_wave, FBox box,
                                                                   unsigned long long* __restrict__ counts) {
    __shared__ AgItemLds lds;
    const uint64_t lo = (uint64_t)blockIdx.x * per_wave;
    const uint64_t hi = lo + per_wave < n ? lo + per_wave : n;
    const int lane = threadIdx.x & 63;
    uint32_t cand = 0;
    GridFilter gf{box, 0u};
    points_grid_two_phase<AX, true>(G, pts, nullptr, lo, hi, RS3D_OUT_OBJ_I32, out, cand, gf, lds);
    unsigned long long v = ((unsigned long long)cand << 32) | gf.inside;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0 && v) atomicAdd(acc_stripe(counts), v);
}

// After a filter-refine launch: the stripes' sum into *total (host-coherent memory) and the
// stripes zeroed for the next launch (one wave; the kernel boundary orders it after the
// launch's atomics).
__global__ __launch_bounds__(64) void k_fr_finish(unsigned long long* __restrict__ counts,
                                                  unsigned long long* __restrict__ total) {
    const int lane = threadIdx.x & 63;
    unsigned long long s = counts[(size_t)lane * ACC_PITCH];
    counts[(size_t)lane * ACC_PITCH] = 0ull;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    // one aligned 8-byte store over the all-ones word the host left there, which it polls
    if (lane == 0) *reinterpret_cast<volatile unsigned long long*>(total) = s;
}

namespace {

// Build temporaries, from the context's stream-ordered pool when it has one: freed on the
// build's stream into the pool, which keeps them for the next build instead of unmapping
// them (hipFree of the temporaries took ~2 ms of a 9 ms build).
struct AgTmp {
    rs3d_ctx* ctx;
    std::vector<void*> ptrs;
    explicit AgTmp(rs3d_ctx* c) : ctx(c) {}
    ~AgTmp() {
        for (void* p : ptrs) (void)(ctx->tmp_pool ? hipFreeAsync(p, ctx->stream) : hipFree(p));
    }
    template <typename T>
    int get(T** out, size_t count) {
        void* p = nullptr;
        const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
        const hipError_t e = ctx->tmp_pool ? hipMallocFromPoolAsync(&p, bytes, ctx->tmp_pool, ctx->stream)
                                           : hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            set_error("axis grid: hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
            return RS3D_E_NOMEM;
        }
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return RS3D_OK;
    }
};

template <int AX>
int build_ax(rs3d_ctx* ctx, rs3d_bvh* bvh, AxisGrid** out) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    *out = nullptr;
    const uint64_t n = bvh->num_tris;
    hipStream_t s = ctx->stream;
    AGArgs G{};
    const double lo_u = bvh->bmin[U], lo_w = bvh->bmin[W];
    const double eu = (double)bvh->bmax[U] - lo_u, ew = (double)bvh->bmax[W] - lo_w;
    // Cell edge: the mean padded extent of the triangles (each then spans ~4 cells), but no
    // more than 4 cells per triangle, 2^26 cells or 32768 cells along an axis.
    double size_sum = 0.0;
    {
        unsigned long long* dsum;
        RS3D_TRY(acc_slot(ctx, 0, 0, &dsum));  // 0 bits = +0.0
        unsigned nbs;
        RS3D_TRY(grid_blocks(n, 256, &nbs));
        k_ag_size<AX><<<nbs < 1024 ? nbs : 1024, 256, 0, s>>>(bvh->tris, n, dsum);
        RS3D_HIP(hipGetLastError());
        uint64_t st[ACC_STRIPES];
        RS3D_TRY(acc_read(ctx, dsum, st));
        for (int i = 0; i < ACC_STRIPES; ++i) {
            double d;
            memcpy(&d, &st[i], sizeof d);
            size_sum += d;
        }
    }
    // Cells smaller than the mean extent shorten the lists a ray reads (fewer boxes that
    // overlap its cell but not its line) at the cost of more records: the requested edge
    // (a percentage of the mean extent) is widened by 1.5x steps until the records fit the
    // budget of their edge (16 per triangle below the mean extent, 6 at or above it).
    const double mean_ext = size_sum / (double)n;
    const double max_cells = std::min<double>(16.0 * (double)n + 1024.0, (double)AG_MAX_CELLS);
    double h_min = 0.0;
    if (eu > 0 && ew > 0) h_min = std::sqrt(eu * ew / max_cells);
    h_min = std::max({h_min, eu / AG_MAX_DIM, ew / AG_MAX_DIM});
    double h = std::max(mean_ext * (double)ctx->opt_grid_cell_pct / 100.0, h_min);
    AgTmp tmp(ctx);
    uint32_t *span, *off;
    RS3D_TRY(tmp.get(&span, n));
    RS3D_TRY(tmp.get(&off, n));
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    uint64_t total = 0;
    int nu = 1, nw = 1;
    // the grid of cell edge e: its geometry into G, nu, nw and the records per triangle
    // into span (total: their sum)
    auto span_pass = [&](double e) -> int {
        nu = nw = 1;
        if (e > 0 && std::isfinite(e)) {
            nu = (int)std::min<double>(std::max<double>(std::ceil(eu / e), 1.0), AG_MAX_DIM);
            nw = (int)std::min<double>(std::max<double>(std::ceil(ew / e), 1.0), AG_MAX_DIM);
        }
        G.lo_u = (float)lo_u;
        G.lo_w = (float)lo_w;
        G.inv_u = eu > 0 ? (float)(nu / eu) : 0.0f;
        G.inv_w = ew > 0 ? (float)(nw / ew) : 0.0f;
        G.nu = nu;
        G.nw = nw;
        G.ntu = (nu + 7) / 8;
        unsigned long long* acc;
        RS3D_TRY(acc_slot(ctx, 0, 0, &acc));
        k_ag_span<AX><<<nb, 256, 0, s>>>(bvh->tris, n, G, span, acc);
        RS3D_HIP(hipGetLastError());
        return acc_read_sum(ctx, acc, &total);
    };
    for (;;) {
        RS3D_TRY(span_pass(h));
        const bool fine = h < mean_ext;
        const uint64_t budget = (fine ? 16 : 6) * n + (1ull << 20);
        if (total <= budget && total <= 0xffffffffull) break;
        if (!fine || !std::isfinite(h) || !(h > 0)) return RS3D_OK;  // rejected: too many records
        h = std::min(h * 1.5, std::max(mean_ext, h_min));
    }
    // A small grid (under AG_SMALL_RECORDS records, a few MB: it stays in the caches) takes
    // finer cells, in 1.5x steps down to AG_MIN_CELL_PCT of the mean extent, while it stays
    // that small: its rays then read shorter lists (one 20,480-triangle sphere, 1 M points:
    // 6.1 M records read at 35 %, 4.5 M at 20 %). Not for an edge the caller set explicitly.
    if (!ctx->opt_grid_cell_pct_set && total < AG_SMALL_RECORDS) {
        const double h_floor = std::max(mean_ext * AG_MIN_CELL_PCT / 100.0, h_min);
        while (h / 1.5 >= h_floor) {
            const uint64_t prev = total;
            RS3D_TRY(span_pass(h / 1.5));
            if (total > AG_SMALL_RECORDS) {
                RS3D_TRY(span_pass(h));  // back to the last edge that fit
                if (total != prev) {
                    set_error("axis grid: span total changed between passes");
                    return RS3D_E_HIP;
                }
                break;
            }
            h /= 1.5;
        }
    }
    if ((double)total / ((double)nu * nw) > AG_MAX_UNIFORM_LIST) return RS3D_OK;  // rejected: lists too long
    const uint64_t nt = (uint64_t)G.ntu * (uint64_t)((nw + 3) / 4), nc = 32 * nt, ng = (nt + 31) / 32;
    RS3D_TRY(scan_u32_exclusive(ctx, span, off, n));
    uint64_t* keys;
    uint32_t *vals, *counts;
    RS3D_TRY(tmp.get(&keys, total));
    RS3D_TRY(tmp.get(&vals, total));
    RS3D_TRY(tmp.get(&counts, nc));
    RS3D_HIP(hipMemsetAsync(counts, 0, nc * sizeof(uint32_t), s));
    k_ag_emit<AX><<<nb, 256, 0, s>>>(bvh->tris, n, G, off, keys, vals, counts);
    unsigned nbc;
    RS3D_TRY(grid_blocks(nc, 256, &nbc));
    unsigned long long* acc_sq;
    RS3D_TRY(acc_slot(ctx, 1, 0, &acc_sq));
    k_ag_sq<<<nbc < 1024 ? nbc : 1024, 256, 0, s>>>(counts, nc, acc_sq);
    RS3D_HIP(hipGetLastError());
    uint64_t sq = 0;
    RS3D_TRY(acc_read_sum(ctx, acc_sq, &sq));
    const double weighted = total ? (double)sq / (double)total : 0.0;
    if (weighted > AG_MAX_WEIGHTED_LIST) return RS3D_OK;  // rejected: dense spots too long
    const double expected = (double)total / ((double)nu * nw);
    // padded lists (even record counts; see k_ag_pcounts)
    uint32_t *pcounts, *pstart, *start;
    RS3D_TRY(tmp.get(&pcounts, nc));
    RS3D_TRY(tmp.get(&pstart, nc));
    RS3D_TRY(tmp.get(&start, nc));
    uint64_t odd = 0;
    {
        unsigned long long* acc_odd;
        RS3D_TRY(acc_slot(ctx, 2, 0, &acc_odd));
        k_ag_pcounts<<<nbc, 256, 0, s>>>(counts, nc, pcounts, acc_odd);
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(acc_read_sum(ctx, acc_odd, &odd));
    }
    const uint64_t ptotal = total + odd;
    if (ptotal > 0xffffffffull) return RS3D_OK;  // rejected: too many records
    uint32_t *words, *gcount, *gbase, *oflag, *crank;
    RS3D_TRY(tmp.get(&words, nt));
    RS3D_TRY(tmp.get(&gcount, ng));
    RS3D_TRY(tmp.get(&gbase, ng));
    RS3D_TRY(tmp.get(&oflag, nc));
    RS3D_TRY(tmp.get(&crank, nc));
    uint64_t nocc = 0;
    {
        unsigned long long* acc_occ;
        RS3D_TRY(acc_slot(ctx, 3, 0, &acc_occ));
        k_ag_oflag<<<nbc, 256, 0, s>>>(counts, nc, oflag, acc_occ);
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(acc_read_sum(ctx, acc_occ, &nocc));
    }
    AxisGrid* g = new AxisGrid();
    g->axis = AX;
    g->nu = nu;
    g->nw = nw;
    g->ntu = G.ntu;
    g->lo_u = G.lo_u;
    g->lo_w = G.lo_w;
    g->inv_u = G.inv_u;
    g->inv_w = G.inv_w;
    g->entries = total;
    g->expected = expected;
    hipError_t e;
    // (one entry more than the table: a lane with no occupied cell reads entries 0 and 1)
    if ((e = hipMalloc(&g->cells, (nocc + 2) * sizeof(uint2))) != hipSuccess ||
        (e = hipMalloc(&g->occ, nt * sizeof(uint4))) != hipSuccess ||
        (e = hipMalloc(&g->tsum, ng * sizeof(uint2))) != hipSuccess ||
        (e = hipMalloc(&g->rec, std::max<uint64_t>(ptotal, 1) * AG_RECF * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&g->obj, n * sizeof(int32_t))) != hipSuccess) {
        set_error("axis grid: hipMalloc failed: %s", hipGetErrorString(e));
        axis_grid_free(g);
        return RS3D_E_NOMEM;
    }
    g->bytes = (nocc + 2) * sizeof(uint2) + nt * sizeof(uint4) + ng * sizeof(uint2) +
               ptotal * AG_RECF * sizeof(float) + n * sizeof(int32_t);
    const int rc = [&]() -> int {
        RS3D_TRY(scan_u32_exclusive(ctx, counts, start, nc));
        RS3D_TRY(scan_u32_exclusive(ctx, pcounts, pstart, nc));
        RS3D_TRY(scan_u32_exclusive(ctx, oflag, crank, nc));
        k_ag_occ<<<nbc, 256, 0, s>>>(counts, nc, words);
        unsigned nbt;
        RS3D_TRY(grid_blocks(nt, 256, &nbt));
        k_ag_tsum<<<nbt, 256, 0, s>>>(words, nt, g->tsum, gcount);
        RS3D_TRY(scan_u32_exclusive(ctx, gcount, gbase, ng));
        if (total) {
            unsigned nbe;
            RS3D_TRY(grid_blocks(total, 256, &nbe));
            uint32_t *cursor, *bkey, *bval, *bcell;
            RS3D_TRY(tmp.get(&cursor, nc));
            RS3D_TRY(tmp.get(&bkey, total));
            RS3D_TRY(tmp.get(&bval, total));
            RS3D_TRY(tmp.get(&bcell, total));
            RS3D_TRY(fill_u32(ctx, cursor, nc, 0u));
            k_ag_bucket<<<nbe, 256, 0, s>>>(keys, vals, total, start, cursor, bkey, bval, bcell);
            k_ag_rank<<<nbe, 256, 0, s>>>(bkey, bval, bcell, total, start, counts, keys, vals);
            k_ag_gather<<<nbe, 256, 0, s>>>(bvh->tris, vals, keys, start, pstart, total, g->rec);
            k_ag_pad<<<nbc, 256, 0, s>>>(counts, pstart, nc, g->rec);
        }
        k_ag_cells<AX><<<nbc, 256, 0, s>>>(pstart, counts, crank, g->rec, nc, (uint32_t)nocc, (uint32_t)ptotal,
                                           g->cells);
        k_ag_objmap<<<nb, 256, 0, s>>>(bvh->tris, n, g->obj);
        k_ag_fine<<<nbt, 256, 0, s>>>(words, nt, gbase, crank, g->cells, g->tsum, g->occ);
        RS3D_HIP(hipGetLastError());
        RS3D_HIP(hipStreamSynchronize(s));
        return RS3D_OK;
    }();
    if (rc != RS3D_OK) {
        axis_grid_free(g);
        return rc;
    }
    *out = g;
    return RS3D_OK;
}

int ensure_grid(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis) {
    if (bvh->agrid_state[axis] != 0) return RS3D_OK;
    if (bvh->num_tris == 0 || bvh->num_tris > 0x7fffffffull) {
        bvh->agrid_state[axis] = 2;
        return RS3D_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    AxisGrid* g = nullptr;
    int rc;
    switch (axis) {
        case 0: rc = build_ax<0>(ctx, bvh, &g); break;
        case 1: rc = build_ax<1>(ctx, bvh, &g); break;
        default: rc = build_ax<2>(ctx, bvh, &g); break;
    }
    if (rc == RS3D_E_NOMEM) {
        // the grid only speeds launches up: without memory for it they walk the BVH (same
        // results); state 3 is not retried by launches, only by rs3d_axis_grid_prepare
        (void)hipGetLastError();
        bvh->agrid_state[axis] = 3;
        return RS3D_OK;
    }
    if (rc != RS3D_OK) return rc;
    if (g) g->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    bvh->agrid[axis] = g;
    bvh->agrid_state[axis] = g ? 1 : 2;
    return RS3D_OK;
}

// Run length of a grid launch over n rays: the context's RS3D_OPT_GRID_RAYS_PER_WAVE,
// halved (down to 64) while the launch would have fewer than 16,384 waves -- about two
// waves per resident wave slot of the chip (256 CUs x 28 at this kernel's registers), which
// small launches need to hide the latency of their dependent loads (1 M points: 230 us at
// 256 per wave, 196 us at 64).
uint32_t grid_run_length(const rs3d_ctx* ctx, uint64_t n) {
    uint64_t pw = (uint64_t)ctx->opt_grid_rays_per_wave;
    while (pw > 64 && n / pw < 16384) pw = pw / 2 > 64 ? pw / 2 : 64;
    return (uint32_t)pw;
}

template <int AX, bool STATS>
int launch_grid_ax(rs3d_ctx* ctx, const AxisGrid* g, const float* pts, uint64_t n, const uint32_t* imap,
                   int out_kind, void* out, unsigned long long* stats) {
    const AGArgs G = ag_args(g);
    const uint32_t pw = grid_run_length(ctx, n);
    unsigned nb;
    RS3D_TRY(grid_blocks(n, pw, &nb));
    k_points_grid<AX, STATS><<<nb, TR_BLOCK, 0, ctx->stream>>>(G, pts, n, imap, out_kind, out, pw, stats);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

template <bool STATS>
int launch_grid(rs3d_ctx* ctx, const AxisGrid* g, const float* pts, uint64_t n, const uint32_t* imap,
                int out_kind, void* out, unsigned long long* stats) {
    switch (g->axis) {
        case 0: return launch_grid_ax<0, STATS>(ctx, g, pts, n, imap, out_kind, out, stats);
        case 1: return launch_grid_ax<1, STATS>(ctx, g, pts, n, imap, out_kind, out, stats);
        default: return launch_grid_ax<2, STATS>(ctx, g, pts, n, imap, out_kind, out, stats);
    }
}

template <int AX>
int launch_grid_filtered_ax(rs3d_ctx* ctx, const AxisGrid* g, const float* pts, uint64_t n, const FBox& box,
                            int32_t* out, unsigned long long* counts) {
    const AGArgs G = ag_args(g);
    const uint32_t pw = grid_run_length(ctx, n);
    unsigned nb;
    RS3D_TRY(grid_blocks(n, pw, &nb));
    k_points_grid_filtered<AX><<<nb, TR_BLOCK, 0, ctx->stream>>>(G, pts, n, out, pw, box, counts);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

}  // namespace

int axis_grid_filter_refine(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, const float* pts, uint64_t n,
                            const float* box, int32_t* out, unsigned long long* counts) {
    const AxisGrid* g = axis >= 0 && axis <= 2 ? bvh->agrid[axis] : nullptr;
    if (!g || n > 0xffffffffull) {
        set_error("axis grid launch without an accepted grid");
        return RS3D_E_ARG;
    }
    const FBox fb{box[0], box[1], box[2], box[3], box[4], box[5]};
    switch (axis) {
        case 0: return launch_grid_filtered_ax<0>(ctx, g, pts, n, fb, out, counts);
        case 1: return launch_grid_filtered_ax<1>(ctx, g, pts, n, fb, out, counts);
        default: return launch_grid_filtered_ax<2>(ctx, g, pts, n, fb, out, counts);
    }
}

int axis_grid_counts_finish(rs3d_ctx* ctx, unsigned long long* counts, unsigned long long* total) {
    k_fr_finish<<<1, 64, 0, ctx->stream>>>(counts, total);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int axis_grid_ready(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, bool* ready, bool build) {
    *ready = false;
    if (!ctx->opt_axis_grid || axis < 0 || axis > 2) return RS3D_OK;
    if (build) RS3D_TRY(ensure_grid(ctx, bvh, axis));
    *ready = bvh->agrid_state[axis] == 1 && bvh->agrid[axis] != nullptr;
    return RS3D_OK;
}

uint64_t axis_grid_bytes_bound(const rs3d_bvh* bvh) {
    // Peak device bytes of a build: records are at most 16 per triangle + 2^20 (the budget
    // the build widens its cells to meet), 40 B each plus up to 1/8 for the padding to even
    // lists, and while the build sorts them 24 B more each (keys, values and the sort's
    // alternate buffers); cells are within 16 per triangle + 1024 and the padding of the
    // edge tiles (under 2^19 at 32768 cells a side), at most 8 B in the table, 24 B of
    // counts, offsets and ranks during the build and under 1 B of tile words, tile entries
    // and summaries; per triangle 8 B of spans and offsets and 4 B of its object.
    const uint64_t n = bvh->num_tris;
    const uint64_t recs = 16 * n + (1ull << 20), cells = 16 * n + 1024 + (1ull << 19);
    return recs * (45 + 24) + cells * 33 + n * 12 + (64ull << 20);
}

int axis_grid_launch(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, const float* pts, uint64_t n, const uint32_t* imap,
                     int out_kind, void* out) {
    const AxisGrid* g = bvh->agrid[axis];
    if (!g || n > 0xffffffffull) {
        set_error("axis grid launch without an accepted grid");
        return RS3D_E_ARG;
    }
    return launch_grid<false>(ctx, g, pts, n, imap, out_kind, out, nullptr);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_axis_grid_prepare(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, rs3d_axis_grid_info* info) {
    if (!ctx || !bvh || axis < 0 || axis > 2) {
        set_error("rs3d_axis_grid_prepare: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(ctx->device));
    if (bvh->agrid_state[axis] == 3) bvh->agrid_state[axis] = 0;  // refused for memory: try again
    RS3D_TRY(ensure_grid(ctx, bvh, axis));
    return info ? rs3d_axis_grid_query(bvh, axis, info) : RS3D_OK;
}

int rs3d_axis_grid_query(const rs3d_bvh* bvh, int axis, rs3d_axis_grid_info* info) {
    if (!bvh || axis < 0 || axis > 2 || !info) {
        set_error("rs3d_axis_grid_query: bad arguments");
        return RS3D_E_ARG;
    }
    {
        memset(info, 0, sizeof *info);
        info->state = bvh->agrid_state[axis];
        const AxisGrid* g = bvh->agrid[axis];
        if (g) {
            info->cells_u = g->nu;
            info->cells_w = g->nw;
            info->entries = g->entries;
            info->expected_list = g->expected;
            info->device_bytes = g->bytes;
            info->build_ms = g->build_ms;
        }
    }
    return RS3D_OK;
}

int rs3d_axis_grid_stats(rs3d_ctx* ctx, rs3d_bvh* bvh, const float* points, uint64_t n, int axis,
                         rs3d_trace_stats* stats) {
    if (!ctx || !bvh || !stats || (n && !points) || axis < 0 || axis > 2 || n > 0xffffffffull) {
        set_error("rs3d_axis_grid_stats: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ensure_grid(ctx, bvh, axis));
    const AxisGrid* g = bvh->agrid[axis];
    if (!g) {
        set_error("rs3d_axis_grid_stats: no axis grid accepted for axis %d", axis);
        return RS3D_E_ARG;
    }
    memset(stats, 0, sizeof *stats);
    stats->rays = n;
    if (n == 0) return RS3D_OK;
    RS3D_TRY(ctx->small.ensure(256));
    RS3D_TRY(ctx->tmp_a.ensure(n));
    unsigned long long* dstats = ctx->small.as<unsigned long long>() + 16;
    RS3D_HIP(hipMemsetAsync(dstats, 0, 4 * sizeof(unsigned long long), ctx->stream));
    RS3D_TRY(launch_grid<true>(ctx, g, points, n, nullptr, RS3D_OUT_FLAG_U8, ctx->tmp_a.ptr, dstats));
    unsigned long long h[4];
    RS3D_HIP(hipMemcpyAsync(h, dstats, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    stats->nodes_visited = h[0];
    stats->tris_tested = h[1];
    stats->hits = h[2];
    stats->restarts = h[3];
    return RS3D_OK;
}

}  // extern "C"
