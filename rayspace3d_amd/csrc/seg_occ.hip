// seg_occ.hip -- segment occupancy bits of a BVH, the first test of every edge ray.
//
// Edge rays are short segments (a mesh edge each) and most of them pass far from every
// triangle of the target: on two overlapping fields of 50,000 spheres each about 93 % of
// one side's edges touch no cube cell of edge twice the mean triangle extent that the other
// side's triangles reach. The bits (one per cell of a uniform grid over the BVH's padded
// bounds, set where a padded triangle box -- the BVH's leaf box, from the shared
// prim_padded_box -- overlaps the cell) answer those rays with a few word reads instead of
// a walk down the tree; the test itself (occ_segment_may_hit, traverse.h) states why no
// hit is lost. Built once per BVH on the first edge launch against it (or ahead with
// rs3d_seg_occupancy_prepare), freed with the BVH.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "rs3d_internal.h"
#include "traverse.h"

namespace rs3d {

struct SegOcc {
    OccArgs args{};
    uint32_t* bits = nullptr;
    uint32_t* snode = nullptr;  // start node per block of OCC_BLOCK^3 cells
    size_t bytes = 0;
    double cell = 0.0;
    uint64_t occupied = 0;
    double build_ms = 0.0;
};

void seg_occ_free(SegOcc* g) {
    if (!g) return;
    if (g->bits) (void)hipFree(g->bits);
    if (g->snode) (void)hipFree(g->snode);
    delete g;
}

constexpr uint64_t SO_MAX_CELLS = 1ull << 30;  // 128 MB of bits

// Sum of the padded boxes' largest extents (the cell size).
__global__ __launch_bounds__(256) void k_so_size(const TriRec* __restrict__ tris, uint64_t n,
                                                 unsigned long long* __restrict__ sum) {
    double v = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const TriRec r = tris[i];
        float plo[3], phi[3];
        prim_padded_box(make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                        make_float3(r.v2.x, r.v2.y, r.v2.z), plo, phi);
        const float e = fmaxf(fmaxf(phi[0] - plo[0], phi[1] - plo[1]), phi[2] - plo[2]);
        if (e < 3.0e38f) v += (double)e;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __shared__ double part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        atomicAdd(reinterpret_cast<double*>(acc_stripe(sum)), part[0] + part[1] + part[2] + part[3]);
}

constexpr int SO_MAX_SPAN = 4096;  // cells of one triangle box; beyond, the bits are rejected

// Whether cell (x, y, z) can hold a point that an fp32 test reports on the triangle: the
// separating-axis test (box axes, triangle normal, the nine edge-axis products) of the
// triangle against the cell widened by the triangle's box padding (a reported hit lies
// within it of the triangle: box_pad bounds the fp32 tests' error, as for the padded box
// itself) plus 2^-20 (nx + ny + nz + 1024) of a cell (occ_cell, which places a segment's
// points, rounds (x - org) inv with a relative error under 2^-23, at most 2^-23 n cells on
// an axis of n cells), all in double. Non-finite triangles meet every cell.
__device__ __forceinline__ bool so_tri_meets_cell(const TriRec& r, int x, int y, int z, const OccArgs& g) {
    const double v[3][3] = {{r.v0.x, r.v0.y, r.v0.z}, {r.v1.x, r.v1.y, r.v1.z}, {r.v2.x, r.v2.y, r.v2.z}};
    float mag = 0.0f, ext = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo = fminf(fminf((float)v[0][k], (float)v[1][k]), (float)v[2][k]);
        const float hi = fmaxf(fmaxf((float)v[0][k], (float)v[1][k]), (float)v[2][k]);
        ext = fmaxf(ext, hi - lo);
        mag = fmaxf(mag, fmaxf(fabsf(lo), fabsf(hi)));
    }
    if (!(mag < 1e30f)) return true;
    const double h = 1.0 / (double)g.inv;
    const double half = 0.5 * h + h * 0x1p-20 * (double)(g.nx + g.ny + g.nz + 1024) + (double)box_pad(mag, ext);
    const double c[3] = {(double)g.org_x + (x + 0.5) * h, (double)g.org_y + (y + 0.5) * h,
                         (double)g.org_z + (z + 0.5) * h};
    double p[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) p[i][k] = v[i][k] - c[k];
    // box axes
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (fmin(fmin(p[0][k], p[1][k]), p[2][k]) > half || fmax(fmax(p[0][k], p[1][k]), p[2][k]) < -half)
            return false;
    }
    const double e[3][3] = {{p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]},
                            {p[2][0] - p[1][0], p[2][1] - p[1][1], p[2][2] - p[1][2]},
                            {p[0][0] - p[2][0], p[0][1] - p[2][1], p[0][2] - p[2][2]}};
    // triangle normal
    {
        const double n0 = e[0][1] * e[1][2] - e[0][2] * e[1][1];
        const double n1 = e[0][2] * e[1][0] - e[0][0] * e[1][2];
        const double n2 = e[0][0] * e[1][1] - e[0][1] * e[1][0];
        const double d = n0 * p[0][0] + n1 * p[0][1] + n2 * p[0][2];
        if (fabs(d) > half * (fabs(n0) + fabs(n1) + fabs(n2))) return false;
    }
    // the nine products of a box axis with a triangle edge
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            // axis = unit k x e[j]
            double a[3] = {0.0, 0.0, 0.0};
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            a[k1] = -e[j][k2];
            a[k2] = e[j][k1];
            const double q0 = a[0] * p[0][0] + a[1] * p[0][1] + a[2] * p[0][2];
            const double q1 = a[0] * p[1][0] + a[1] * p[1][1] + a[2] * p[1][2];
            const double q2 = a[0] * p[2][0] + a[1] * p[2][1] + a[2] * p[2][2];
            const double rad = half * (fabs(a[0]) + fabs(a[1]) + fabs(a[2]));
            if (fmin(fmin(q0, q1), q2) > rad || fmax(fmax(q0, q1), q2) < -rad) return false;
        }
    }
    return true;
}

// One triangle per thread: set the bits of every cell its padded box overlaps and
// so_tri_meets_cell accepts. A box over more than SO_MAX_SPAN cells counts in *big and marks
// nothing (the caller drops the bits).
__global__ __launch_bounds__(256) void k_so_mark(const TriRec* __restrict__ tris, uint64_t n, OccArgs g,
                                                 uint32_t* __restrict__ bits, unsigned long long* __restrict__ big) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1, z0 = 0, z1 = -1;  // empty range for lanes past the end
    if (valid) {
        const TriRec r = tris[i];
        float plo[3], phi[3];
        prim_padded_box(make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                        make_float3(r.v2.x, r.v2.y, r.v2.z), plo, phi);
        x0 = occ_cell(plo[0], g.org_x, g.inv, g.nx), x1 = occ_cell(phi[0], g.org_x, g.inv, g.nx);
        y0 = occ_cell(plo[1], g.org_y, g.inv, g.ny), y1 = occ_cell(phi[1], g.org_y, g.inv, g.ny);
        z0 = occ_cell(plo[2], g.org_z, g.inv, g.nz), z1 = occ_cell(phi[2], g.org_z, g.inv, g.nz);
    }
    if (!valid) return;
    if ((uint64_t)(x1 - x0 + 1) * (uint64_t)(y1 - y0 + 1) * (uint64_t)(z1 - z0 + 1) > (uint64_t)SO_MAX_SPAN) {
        atomicAdd(acc_stripe(big), 1ull);
        return;
    }
    const TriRec r = tris[i];
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                if (!so_tri_meets_cell(r, x, y, z, g)) continue;
                const uint64_t c = ((uint64_t)z * g.ny + y) * g.nx + x;
                const uint32_t m = 1u << (c & 31);
                if (!(bits[c >> 5] & m)) atomicOr(&bits[c >> 5], m);  // a read first: most bits are set already
            }
}

// Leaf range per block: for every block holding a cell the triangle's padded box overlaps,
// the smallest and largest triangle position (Morton order = BVH leaf order). Consecutive
// lanes with the same block range leave the minimum to the first of them and the maximum
// to the last (in Morton order most neighbours share their block).
__global__ __launch_bounds__(256) void k_so_blocks(const TriRec* __restrict__ tris, uint64_t n, OccArgs g,
                                                   int bnz, uint32_t* __restrict__ bmin,
                                                   uint32_t* __restrict__ bmax) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int x0 = 0, x1 = -1, y0 = 0, y1 = -1, z0 = 0, z1 = -1;
    if (valid) {
        const TriRec r = tris[i];
        float plo[3], phi[3];
        prim_padded_box(make_float3(r.v0.x, r.v0.y, r.v0.z), make_float3(r.v1.x, r.v1.y, r.v1.z),
                        make_float3(r.v2.x, r.v2.y, r.v2.z), plo, phi);
        x0 = occ_cell(plo[0], g.org_x, g.inv, g.nx) / OCC_BLOCK, x1 = occ_cell(phi[0], g.org_x, g.inv, g.nx) / OCC_BLOCK;
        y0 = occ_cell(plo[1], g.org_y, g.inv, g.ny) / OCC_BLOCK, y1 = occ_cell(phi[1], g.org_y, g.inv, g.ny) / OCC_BLOCK;
        z0 = occ_cell(plo[2], g.org_z, g.inv, g.nz) / OCC_BLOCK, z1 = occ_cell(phi[2], g.org_z, g.inv, g.nz) / OCC_BLOCK;
    }
    (void)bnz;
    const int lane = threadIdx.x & 63;
    auto same_as = [&](int src) {
        return __shfl(x0, src, 64) == x0 && __shfl(x1, src, 64) == x1 && __shfl(y0, src, 64) == y0 &&
               __shfl(y1, src, 64) == y1 && __shfl(z0, src, 64) == z0 && __shfl(z1, src, 64) == z1;
    };
    const bool same_prev = lane > 0 && same_as(lane - 1);
    const int nxt = lane < 63 ? lane + 1 : lane;
    const bool next_valid = __shfl((int)valid, nxt, 64) != 0;
    const bool same_next = lane < 63 && next_valid && same_as(nxt);
    if (!valid || (same_prev && same_next)) return;
    for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
                const uint64_t b = ((uint64_t)z * g.bny + y) * g.bnx + x;
                if (!same_prev) atomicMin(&bmin[b], (uint32_t)i);
                if (!same_next) atomicMax(&bmax[b], (uint32_t)i);
            }
}

// Start node per block: over the block's 2x2x2 neighbourhood (blocks b .. b+1 per axis)
// the leaf range [a, b] of the triangles there, and the deepest node whose leaf range holds
// it, found from the root down (a node covers [F, L]; its left child covers [F, gamma] and
// is node gamma when internal, its right child [gamma + 1, L]; a leaf child ends the
// descent at its parent). Blocks with no triangle nearby keep the root.
__global__ __launch_bounds__(256) void k_so_start(const BNode* __restrict__ nodes, uint64_t n,
                                                  const uint32_t* __restrict__ bmin,
                                                  const uint32_t* __restrict__ bmax, int bnx, int bny, int bnz,
                                                  uint32_t* __restrict__ snode) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nblocks = (uint64_t)bnx * bny * bnz;
    if (k >= nblocks) return;
    const int bx = (int)(k % bnx), by = (int)((k / bnx) % bny), bz = (int)(k / ((uint64_t)bnx * bny));
    uint32_t a = 0xffffffffu, b = 0u;
    for (int dz = 0; dz < 2 && bz + dz < bnz; ++dz)
        for (int dy = 0; dy < 2 && by + dy < bny; ++dy)
            for (int dx = 0; dx < 2 && bx + dx < bnx; ++dx) {
                const uint64_t q = ((uint64_t)(bz + dz) * bny + (by + dy)) * bnx + (bx + dx);
                a = min(a, bmin[q]);
                b = max(b, bmax[q]);
            }
    int p = 0;
    if (n > 1 && a <= b) {
        uint32_t F = 0;
        for (int depth = 0; depth < 128; ++depth) {
            const int4 lk = nodes[p].link;
            const uint32_t gamma = lk.z > 0 ? F + (uint32_t)lk.z - 1 : (uint32_t)lk.x;
            if (b <= gamma) {
                if (lk.z > 0) break;
                p = lk.x;
            } else if (a > gamma) {
                if (lk.w > 0) break;
                p = lk.y;
                F = gamma + 1;
            } else {
                break;
            }
        }
    }
    snode[k] = (uint32_t)p;
}

__global__ __launch_bounds__(256) void k_so_count(const uint32_t* __restrict__ bits, uint64_t nw,
                                                  unsigned long long* __restrict__ acc) {
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (uint64_t)gridDim.x * blockDim.x)
        c += (unsigned long long)__popc(bits[i]);
    c = block_sum_256(c);
    if (threadIdx.x == 0 && c) atomicAdd(acc_stripe(acc), c);
}

namespace {

int build(rs3d_ctx* ctx, rs3d_bvh* bvh, SegOcc** out) {
    *out = nullptr;
    const uint64_t n = bvh->num_tris;
    hipStream_t s = ctx->stream;
    double size_sum = 0.0;
    {
        unsigned long long* acc;
        RS3D_TRY(acc_slot(ctx, 0, 0, &acc));  // 0 bits = +0.0
        unsigned nb;
        RS3D_TRY(grid_blocks(n, 256, &nb));
        k_so_size<<<nb < 1024 ? nb : 1024, 256, 0, s>>>(bvh->tris, n, acc);
        RS3D_HIP(hipGetLastError());
        uint64_t st[ACC_STRIPES];
        RS3D_TRY(acc_read(ctx, acc, st));
        for (int i = 0; i < ACC_STRIPES; ++i) {
            double d;
            memcpy(&d, &st[i], sizeof d);
            size_sum += d;
        }
    }
    const double ext[3] = {(double)bvh->bmax[0] - bvh->bmin[0], (double)bvh->bmax[1] - bvh->bmin[1],
                           (double)bvh->bmax[2] - bvh->bmin[2]};
    if (!(std::isfinite(ext[0]) && std::isfinite(ext[1]) && std::isfinite(ext[2]))) return RS3D_OK;  // rejected
    double h = size_sum / (double)n * (double)ctx->opt_seg_cell_pct / 100.0;
    const double vol = std::max(ext[0], 1e-30) * std::max(ext[1], 1e-30) * std::max(ext[2], 1e-30);
    h = std::max(h, std::cbrt(vol / (double)SO_MAX_CELLS));
    if (!(h > 0.0) || !std::isfinite(h)) return RS3D_OK;  // rejected: degenerate
    int nc[3];
    uint64_t cells = 1;
    for (int k = 0; k < 3; ++k) {
        nc[k] = (int)std::min<double>(std::max<double>(std::ceil(ext[k] / h), 1.0), 1 << 20);
        cells *= (uint64_t)nc[k];
    }
    if (cells > 2 * SO_MAX_CELLS) return RS3D_OK;  // rejected (flat bounds pushing one axis)
    SegOcc* g = new SegOcc();
    g->cell = h;
    const uint64_t nwords = (cells + 31) / 32;
    hipError_t e = hipMalloc(&g->bits, nwords * sizeof(uint32_t));
    if (e != hipSuccess) {
        delete g;
        set_error("segment occupancy: hipMalloc(%llu) failed: %s", (unsigned long long)(nwords * 4),
                  hipGetErrorString(e));
        return RS3D_E_NOMEM;
    }
    g->bytes = nwords * sizeof(uint32_t);
    OccArgs& a = g->args;
    a.bits = g->bits;
    a.org_x = bvh->bmin[0];
    a.org_y = bvh->bmin[1];
    a.org_z = bvh->bmin[2];
    a.inv = (float)(1.0 / h);
    a.nx = nc[0];
    a.ny = nc[1];
    a.nz = nc[2];
    const int rc = [&]() -> int {
        RS3D_TRY(fill_u32(ctx, g->bits, nwords, 0u));
        unsigned long long* big;
        RS3D_TRY(acc_slot(ctx, 2, 0, &big));
        unsigned nb;
        RS3D_TRY(grid_blocks(n, 256, &nb));
        k_so_mark<<<nb, 256, 0, s>>>(bvh->tris, n, a, g->bits, big);
        RS3D_HIP(hipGetLastError());
        uint64_t nbig = 0;
        RS3D_TRY(acc_read_sum(ctx, big, &nbig));
        if (nbig) return RS3D_E_CAPACITY;  // a triangle box over too many cells: no bits
        unsigned long long* acc;
        RS3D_TRY(acc_slot(ctx, 1, 0, &acc));
        unsigned nbw;
        RS3D_TRY(grid_blocks(nwords, 256, &nbw));
        k_so_count<<<nbw < 1024 ? nbw : 1024, 256, 0, s>>>(g->bits, nwords, acc);
        RS3D_HIP(hipGetLastError());
        uint64_t occ = 0;
        RS3D_TRY(acc_read_sum(ctx, acc, &occ));
        g->occupied = occ;
        // start nodes of the blocks
        const int bnz = (a.nz + OCC_BLOCK - 1) / OCC_BLOCK;
        a.bnx = (a.nx + OCC_BLOCK - 1) / OCC_BLOCK;
        a.bny = (a.ny + OCC_BLOCK - 1) / OCC_BLOCK;
        const uint64_t nblocks = (uint64_t)a.bnx * a.bny * bnz;
        uint32_t *bmin = nullptr, *bmax = nullptr;
        hipError_t e2;
        if ((e2 = hipMalloc(&g->snode, nblocks * sizeof(uint32_t))) != hipSuccess ||
            (e2 = hipMalloc(&bmin, nblocks * sizeof(uint32_t))) != hipSuccess ||
            (e2 = hipMalloc(&bmax, nblocks * sizeof(uint32_t))) != hipSuccess) {
            if (bmin) (void)hipFree(bmin);
            set_error("segment occupancy: hipMalloc of the start nodes failed: %s", hipGetErrorString(e2));
            return RS3D_E_NOMEM;
        }
        const int rb = [&]() -> int {
            RS3D_TRY(fill_u32(ctx, bmin, nblocks, 0xffffffffu));
            RS3D_TRY(fill_u32(ctx, bmax, nblocks, 0u));
            k_so_blocks<<<nb, 256, 0, s>>>(bvh->tris, n, a, bnz, bmin, bmax);
            RS3D_HIP(hipGetLastError());
            unsigned nbb;
            RS3D_TRY(grid_blocks(nblocks, 256, &nbb));
            k_so_start<<<nbb, 256, 0, s>>>(bvh->nodes, n, bmin, bmax, a.bnx, a.bny, bnz, g->snode);
            RS3D_HIP(hipGetLastError());
            RS3D_HIP(hipStreamSynchronize(s));
            return RS3D_OK;
        }();
        (void)hipFree(bmin);
        (void)hipFree(bmax);
        RS3D_TRY(rb);
        a.snode = g->snode;
        g->bytes += nblocks * sizeof(uint32_t);
        return RS3D_OK;
    }();
    if (rc == RS3D_E_CAPACITY) {  // rejected: edge rays walk the tree
        seg_occ_free(g);
        return RS3D_OK;
    }
    if (rc != RS3D_OK) {
        seg_occ_free(g);
        return rc;
    }
    *out = g;
    return RS3D_OK;
}

int ensure(rs3d_ctx* ctx, rs3d_bvh* bvh) {
    if (bvh->socc_state != 0) return RS3D_OK;
    if (bvh->num_tris == 0) {
        bvh->socc_state = 2;
        return RS3D_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    SegOcc* g = nullptr;
    const int rc = build(ctx, bvh, &g);
    if (rc == RS3D_E_NOMEM) {
        // the bits only let edge rays skip the walk: without memory for them every edge
        // walks the tree (same results); state 3 is retried only by the prepare call
        (void)hipGetLastError();
        bvh->socc_state = 3;
        return RS3D_OK;
    }
    if (rc != RS3D_OK) return rc;
    if (g) g->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    bvh->socc = g;
    bvh->socc_state = g ? 1 : 2;
    return RS3D_OK;
}

}  // namespace

int seg_occ_args(rs3d_ctx* ctx, rs3d_bvh* bvh, OccArgs* args, bool build) {
    *args = OccArgs{};
    if (!ctx->opt_seg_occupancy) return RS3D_OK;
    if (build) RS3D_TRY(ensure(ctx, bvh));
    if (bvh->socc_state == 1 && bvh->socc) *args = bvh->socc->args;
    return RS3D_OK;
}

uint64_t seg_occ_bytes(const rs3d_bvh* bvh) { return bvh->socc ? bvh->socc->bytes : 0; }

uint64_t seg_occ_bytes_bound(const rs3d_bvh* bvh) {
    (void)bvh;
    // bits of at most 2^31 cells, 12 B per block of 64 cells while the start nodes are
    // built (4 B kept), and slack
    return (1ull << 31) / 8 + (1ull << 31) / 64 * 12 + (16ull << 20);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_seg_occupancy_prepare(rs3d_ctx* ctx, rs3d_bvh* bvh, rs3d_seg_occupancy_info* info) {
    if (!ctx || !bvh) {
        set_error("rs3d_seg_occupancy_prepare: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(ctx->device));
    if (bvh->socc_state == 3) bvh->socc_state = 0;  // refused for memory: try again
    RS3D_TRY(ensure(ctx, bvh));
    if (info) {
        memset(info, 0, sizeof *info);
        info->state = bvh->socc_state;
        const SegOcc* g = bvh->socc;
        if (g) {
            info->cells_x = g->args.nx;
            info->cells_y = g->args.ny;
            info->cells_z = g->args.nz;
            info->cell_edge = g->cell;
            info->occupied_cells = g->occupied;
            info->device_bytes = g->bytes;
            info->build_ms = g->build_ms;
        }
    }
    return RS3D_OK;
}

}  // extern "C"
