// edge_order.hip -- spatial order of an edge payload for the edge-ray joins.
//
// The .pre payload lists unique edges in lexicographic order of their quantised endpoints
// (preprocess/src/EdgePreprocessor.cpp:110-130), so the 64 lanes of a wave hold edges from
// one thin slab in x spread over the whole y-z range and walk unrelated parts of the
// target BVH. rs3d_edges_spatial_order rewrites the edge records (start, end, source
// offset, source count) along a 30-bit Morton curve of the segment midpoints (bounds of
// all midpoints, 10 bits per axis, stable radix sort, so equal codes keep their order);
// the source list is shared unchanged. Join results are sorted pair sets and do not depend
// on the order in which edges are traced.
#include "rs3d_internal.h"

namespace rs3d {

// Float <-> int mapping whose signed order is the float order (for atomic min/max).
__device__ __forceinline__ int eo_f2ord(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float eo_ord2f(int i) { return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff)); }

__device__ __forceinline__ float3 eo_mid(const float* __restrict__ s, const float* __restrict__ e, uint64_t i) {
    return make_float3(0.5f * s[3 * i] + 0.5f * e[3 * i], 0.5f * s[3 * i + 1] + 0.5f * e[3 * i + 1],
                       0.5f * s[3 * i + 2] + 0.5f * e[3 * i + 2]);
}

__global__ void k_eo_init(int* b) {
    const int t = threadIdx.x;
    if (t < 6) b[t] = t < 3 ? 0x7fffffff : (int)0x80000000;
}

// Midpoint bounds: per-lane min/max, wave reduction, one atomic per wave and axis.
__global__ __launch_bounds__(256) void k_eo_bounds(const float* __restrict__ s, const float* __restrict__ e,
                                                   uint64_t n, int* __restrict__ b) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const float3 m = eo_mid(s, e, i);
        lo[0] = fminf(lo[0], m.x), lo[1] = fminf(lo[1], m.y), lo[2] = fminf(lo[2], m.z);
        hi[0] = fmaxf(hi[0], m.x), hi[1] = fmaxf(hi[1], m.y), hi[2] = fmaxf(hi[2], m.z);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        for (int off = 32; off > 0; off >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], off, 64));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], off, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            atomicMin(&b[k], eo_f2ord(lo[k]));
            atomicMax(&b[3 + k], eo_f2ord(hi[k]));
        }
    }
}

__device__ __forceinline__ uint32_t eo_spread10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(256) void k_eo_keys(const float* __restrict__ s, const float* __restrict__ e,
                                                 uint64_t n, const int* __restrict__ b, uint32_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 m = eo_mid(s, e, i);
    const float c[3] = {m.x, m.y, m.z};
    uint32_t code = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo = eo_ord2f(b[k]), ext = eo_ord2f(b[3 + k]) - lo;
        const float sc = ext > 0.0f ? 1024.0f / ext : 0.0f;
        const float u = fminf(fmaxf((c[k] - lo) * sc, 0.0f), 1023.0f);
        code |= eo_spread10((uint32_t)u) << (2 - k);
    }
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_eo_gather(const uint32_t* __restrict__ perm, uint64_t n,
                                                   const float* __restrict__ s, const float* __restrict__ e,
                                                   const int32_t* __restrict__ off, const int32_t* __restrict__ cnt,
                                                   float* __restrict__ s_out, float* __restrict__ e_out,
                                                   int32_t* __restrict__ off_out, int32_t* __restrict__ cnt_out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = perm[i];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        s_out[3 * i + k] = s[3 * j + k];
        e_out[3 * i + k] = e[3 * j + k];
    }
    off_out[i] = off[j];
    cnt_out[i] = cnt[j];
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_edges_spatial_order(rs3d_ctx* ctx, const rs3d_edges* edges, float* starts_out, float* ends_out,
                             int32_t* src_offsets_out, int32_t* src_counts_out) {
    if (!ctx || !edges ||
        (edges->num_edges &&
         (!edges->starts || !edges->ends || !edges->src_offsets || !edges->src_counts || !starts_out ||
          !ends_out || !src_offsets_out || !src_counts_out))) {
        set_error("rs3d_edges_spatial_order: bad arguments");
        return RS3D_E_ARG;
    }
    const uint64_t n = edges->num_edges;
    if (n == 0) return RS3D_OK;
    if (n > 0xffffffffull) {
        set_error("rs3d_edges_spatial_order: more than 2^32-1 edges");
        return RS3D_E_ARG;
    }
    const uint64_t out_lo = (uint64_t)(uintptr_t)starts_out, out_hi = out_lo + n * 12;
    const uint64_t in_lo = (uint64_t)(uintptr_t)edges->starts, in_hi = in_lo + n * 12;
    if (out_lo < in_hi && in_lo < out_hi) {
        set_error("rs3d_edges_spatial_order: outputs may not alias the inputs");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->ray_keys.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->ray_vals.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->small.ensure(256));
    int* b = ctx->small.as<int>();
    uint32_t* keys = ctx->ray_keys.as<uint32_t>();
    uint32_t* vals = ctx->ray_vals.as<uint32_t>();
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_eo_init<<<1, 64, 0, ctx->stream>>>(b);
    k_eo_bounds<<<nb < 1024 ? nb : 1024, 256, 0, ctx->stream>>>(edges->starts, edges->ends, n, b);
    k_eo_keys<<<nb, 256, 0, ctx->stream>>>(edges->starts, edges->ends, n, b, keys, vals);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(radix_sort_u32_pairs(ctx, keys, vals, n, 30));
    k_eo_gather<<<nb, 256, 0, ctx->stream>>>(vals, n, edges->starts, edges->ends, edges->src_offsets,
                                             edges->src_counts, starts_out, ends_out, src_offsets_out,
                                             src_counts_out);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

}  // extern "C"
