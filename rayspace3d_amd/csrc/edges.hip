// edges.hip -- unique-edge extraction on the GPU (replaces the host std::map pass of
// preprocess/src/EdgePreprocessor.cpp:52-130 that writes the .pre edge payload).
//
// Same records as the host extractor: edges are keyed by their two endpoints quantised
// with llround(x * 1e5), ordered lexicographically by (smaller endpoint, larger endpoint);
// an edge's start/end are the endpoints of its first occurrence in triangle order (edges
// (v0,v1), (v1,v2), (v2,v0) of each triangle); its sources are the distinct objects of its
// occurrences in order of first appearance. Triangles with an index out of range are
// skipped.
//
// Method: (1) quantise every vertex and sort the vertices lexicographically by their
// quantised position (three stable LSD sorts, z then y then x), so each distinct position
// gets a rank that preserves the lexicographic order; (2) every edge occurrence becomes the
// 64-bit key (min rank << 32 | max rank), whose order is the lexicographic edge order, and
// one stable sort groups the occurrences of each edge in triangle order; (3) one thread per
// edge run writes the endpoints of its first occurrence and its distinct sources.
#include <algorithm>
#include <cmath>

#include "rs3d_internal.h"

namespace rs3d {

constexpr uint64_t SIGN64 = 0x8000000000000000ull;

// Quantised coordinate k of every vertex, as an unsigned key (sign bit flipped, so the
// unsigned order is the signed order).
__global__ __launch_bounds__(256) void k_quant_coord(const float* __restrict__ V, uint64_t nV, int k,
                                                     uint64_t* __restrict__ key,
                                                     uint32_t* __restrict__ idx, int init_idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    const uint32_t v = init_idx ? (uint32_t)i : idx[i];
    const long long q = llround((double)V[3 * (uint64_t)v + k] * 1e5);
    key[i] = (uint64_t)q ^ SIGN64;
    if (init_idx) idx[i] = v;
}

__device__ __forceinline__ bool same_point(const float* __restrict__ V, uint32_t a, uint32_t b) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (llround((double)V[3 * (uint64_t)a + k] * 1e5) != llround((double)V[3 * (uint64_t)b + k] * 1e5))
            return false;
    return true;
}

__global__ __launch_bounds__(256) void k_point_flags(const float* __restrict__ V,
                                                     const uint32_t* __restrict__ order, uint64_t nV,
                                                     uint8_t* __restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    flags[i] = (i == 0 || !same_point(V, order[i], order[i - 1])) ? 1 : 0;
}

// rank of vertex order[i] = number of distinct positions before it
__global__ __launch_bounds__(256) void k_point_rank(const uint32_t* __restrict__ order,
                                                    const uint8_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ excl, uint64_t nV,
                                                    uint32_t* __restrict__ rank) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nV) return;
    rank[order[i]] = excl[i] + flags[i] - 1u;
}

// Occurrence o = 3 t + e of edge e of triangle t; invalid triangles get the largest key.
__global__ __launch_bounds__(256) void k_occ_keys(const uint32_t* __restrict__ I, uint64_t nT,
                                                  uint64_t nV, const uint32_t* __restrict__ rank,
                                                  uint64_t* __restrict__ key,
                                                  uint32_t* __restrict__ occ) {
    const uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= 3 * nT) return;
    const uint64_t t = o / 3;
    const int e = (int)(o - 3 * t);
    const uint32_t v0 = I[3 * t], v1 = I[3 * t + 1], v2 = I[3 * t + 2];
    occ[o] = (uint32_t)o;
    if (v0 >= nV || v1 >= nV || v2 >= nV) {
        key[o] = ~0ull;
        return;
    }
    const uint32_t a = e == 0 ? v0 : (e == 1 ? v1 : v2);
    const uint32_t b = e == 0 ? v1 : (e == 1 ? v2 : v0);
    const uint64_t ra = rank[a], rb = rank[b];
    key[o] = ra < rb ? (ra << 32) | rb : (rb << 32) | ra;
}

__global__ __launch_bounds__(256) void k_run_flags(const uint64_t* __restrict__ key, uint64_t n,
                                                   uint8_t* __restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || key[i] != key[i - 1]) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_run_starts(const uint8_t* __restrict__ flags,
                                                    const uint32_t* __restrict__ excl, uint64_t n,
                                                    uint32_t* __restrict__ start) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) start[excl[i]] = (uint32_t)i;
}

__device__ __forceinline__ int occ_obj(const int32_t* __restrict__ tri_obj, uint32_t o) {
    return tri_obj ? tri_obj[o / 3] : (int)(o / 3);
}

// Distinct sources of run r (occurrences [start[r], start[r+1])) in first-appearance
// order; count only when out is null.
__device__ __forceinline__ int32_t run_sources(const uint32_t* __restrict__ occ,
                                               const int32_t* __restrict__ tri_obj, uint32_t s,
                                               uint32_t e, int32_t* out) {
    int32_t cnt = 0;
    for (uint32_t k = s; k < e; ++k) {
        const int src = occ_obj(tri_obj, occ[k]);
        bool dup = false;
        for (uint32_t m = s; m < k && !dup; ++m) dup = occ_obj(tri_obj, occ[m]) == src;
        if (!dup) {
            if (out) out[cnt] = src;
            ++cnt;
        }
    }
    return cnt;
}

__global__ __launch_bounds__(256) void k_run_count(const uint32_t* __restrict__ start,
                                                   uint64_t nruns, uint32_t nvalid,
                                                   const uint32_t* __restrict__ occ,
                                                   const int32_t* __restrict__ tri_obj,
                                                   int32_t* __restrict__ counts) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    const uint32_t s = start[r], e = r + 1 < nruns ? start[r + 1] : nvalid;
    counts[r] = run_sources(occ, tri_obj, s, e, nullptr);
}

__global__ __launch_bounds__(256) void k_run_write(const uint32_t* __restrict__ start,
                                                   uint64_t nruns, uint32_t nvalid,
                                                   const uint32_t* __restrict__ occ,
                                                   const int32_t* __restrict__ tri_obj,
                                                   const float* __restrict__ V,
                                                   const uint32_t* __restrict__ I,
                                                   const int64_t* __restrict__ soff,
                                                   float* __restrict__ starts, float* __restrict__ ends,
                                                   int32_t* __restrict__ offsets,
                                                   int32_t* __restrict__ sources) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nruns) return;
    const uint32_t s = start[r], e = r + 1 < nruns ? start[r + 1] : nvalid;
    const uint32_t o = occ[s];
    const uint64_t t = o / 3;
    const int k = (int)(o - 3 * t);
    const uint32_t a = I[3 * t + k], b = I[3 * t + (k + 1) % 3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        starts[3 * r + c] = V[3 * (uint64_t)a + c];
        ends[3 * r + c] = V[3 * (uint64_t)b + c];
    }
    offsets[r] = (int32_t)soff[r];
    run_sources(occ, tri_obj, s, e, sources + soff[r]);
}

__global__ void k_count_below(const uint64_t* __restrict__ key, uint64_t n,
                              unsigned long long* __restrict__ out) {
    // keys are sorted; the invalid occurrences (~0) sit at the end: binary search
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (key[mid] != ~0ull) lo = mid + 1;
        else hi = mid;
    }
    *out = lo;
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_extract_edges(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                       const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                       float* starts, float* ends, int32_t* src_offsets, int32_t* src_counts,
                       int32_t* src_objects, uint64_t max_edges, uint64_t max_sources,
                       uint64_t* num_edges, uint64_t* num_sources) {
    if (!ctx || !num_edges || !num_sources || (num_triangles && (!vertices || !indices)) ||
        (max_edges && (!starts || !ends || !src_offsets || !src_counts)) ||
        (max_sources && !src_objects)) {
        set_error("rs3d_extract_edges: bad arguments");
        return RS3D_E_ARG;
    }
    *num_edges = 0;
    *num_sources = 0;
    if (num_triangles == 0) return RS3D_OK;
    if (num_vertices >= 0xffffffffull || 3 * num_triangles >= 0xffffffffull) {
        set_error("rs3d_extract_edges: at most 2^32-2 vertices and edge occurrences");
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const uint64_t nV = num_vertices, nO = 3 * num_triangles;
    // scratch: vertex keys / order / ranks, occurrence keys / ids, flags and scans
    DevBuf vkey, vord, rank, okey, occ, flags, excl, start, cnts;
    struct Release {
        DevBuf* b[9];
        ~Release() {
            for (DevBuf* d : b) d->release();
        }
    } rel{{&vkey, &vord, &rank, &okey, &occ, &flags, &excl, &start, &cnts}};
    const uint64_t nmax = std::max<uint64_t>(nV, nO);
    RS3D_TRY(vkey.ensure(std::max<uint64_t>(nV, 1) * 8));
    RS3D_TRY(vord.ensure(std::max<uint64_t>(nV, 1) * 4));
    RS3D_TRY(rank.ensure(std::max<uint64_t>(nV, 1) * 4));
    RS3D_TRY(okey.ensure(nO * 8));
    RS3D_TRY(occ.ensure(nO * 4));
    RS3D_TRY(flags.ensure(nmax + 16));
    RS3D_TRY(excl.ensure((nmax + 16) * 4));
    RS3D_TRY(start.ensure(nO * 4));
    RS3D_TRY(cnts.ensure(nO * 12 + 512));
    unsigned nbv = 1, nbo;
    if (nV) RS3D_TRY(grid_blocks(nV, 256, &nbv));
    RS3D_TRY(grid_blocks(nO, 256, &nbo));
    RS3D_TRY(ctx->small.ensure(256));
    uint32_t* tot = reinterpret_cast<uint32_t*>(ctx->small.as<uint8_t>() + 128);
    // (1) vertices in lexicographic order of their quantised positions, and ranks
    if (nV) {
        for (int k = 2; k >= 0; --k) {
            k_quant_coord<<<nbv, 256, 0, s>>>(vertices, nV, k, vkey.as<uint64_t>(), vord.as<uint32_t>(),
                                               k == 2 ? 1 : 0);
            RS3D_HIP(hipGetLastError());
            RS3D_TRY(radix_sort_u64_pairs(ctx, vkey.as<uint64_t>(), vord.as<uint32_t>(), nV));
        }
        k_point_flags<<<nbv, 256, 0, s>>>(vertices, vord.as<uint32_t>(), nV, flags.as<uint8_t>());
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags.as<uint8_t>(), excl.as<uint32_t>(), nV, tot));
        k_point_rank<<<nbv, 256, 0, s>>>(vord.as<uint32_t>(), flags.as<uint8_t>(), excl.as<uint32_t>(),
                                          nV, rank.as<uint32_t>());
        RS3D_HIP(hipGetLastError());
    }
    // (2) edge occurrences grouped by edge, in lexicographic edge order, triangle order inside
    k_occ_keys<<<nbo, 256, 0, s>>>(indices, num_triangles, nV, rank.as<uint32_t>(), okey.as<uint64_t>(),
                                    occ.as<uint32_t>());
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(radix_sort_u64_pairs(ctx, okey.as<uint64_t>(), occ.as<uint32_t>(), nO));
    unsigned long long* nvalid_d = reinterpret_cast<unsigned long long*>(ctx->small.as<uint8_t>() + 192);
    k_count_below<<<1, 64, 0, s>>>(okey.as<uint64_t>(), nO, nvalid_d);
    RS3D_HIP(hipGetLastError());
    uint64_t nvalid = 0;
    RS3D_TRY(read_u64(ctx, nvalid_d, &nvalid));
    if (nvalid == 0) return RS3D_OK;
    unsigned nbn;
    RS3D_TRY(grid_blocks(nvalid, 256, &nbn));
    k_run_flags<<<nbn, 256, 0, s>>>(okey.as<uint64_t>(), nvalid, flags.as<uint8_t>());
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags.as<uint8_t>(), excl.as<uint32_t>(), nvalid, tot));
    uint32_t nruns32 = 0;
    RS3D_HIP(hipMemcpyAsync(&nruns32, tot, sizeof nruns32, hipMemcpyDeviceToHost, s));
    RS3D_HIP(hipStreamSynchronize(s));
    const uint64_t nruns = nruns32;
    k_run_starts<<<nbn, 256, 0, s>>>(flags.as<uint8_t>(), excl.as<uint32_t>(), nvalid, start.as<uint32_t>());
    RS3D_HIP(hipGetLastError());
    // (3) per edge: distinct sources, their offsets, then the records
    unsigned nbr;
    RS3D_TRY(grid_blocks(nruns, 256, &nbr));
    int32_t* counts_tmp = reinterpret_cast<int32_t*>(cnts.as<uint8_t>());
    int64_t* soff = reinterpret_cast<int64_t*>(cnts.as<uint8_t>() + ((nruns * 4 + 255) & ~255ull));
    k_run_count<<<nbr, 256, 0, s>>>(start.as<uint32_t>(), nruns, (uint32_t)nvalid, occ.as<uint32_t>(),
                                     tri_to_obj, counts_tmp);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(ctx->small.ensure(256));
    int64_t* stot = reinterpret_cast<int64_t*>(ctx->small.as<uint8_t>() + 64);
    RS3D_TRY(scan_i32_to_i64_exclusive(ctx, counts_tmp, soff, nruns, stot));
    int64_t nsrc = 0;
    RS3D_HIP(hipMemcpyAsync(&nsrc, stot, sizeof nsrc, hipMemcpyDeviceToHost, s));
    RS3D_HIP(hipStreamSynchronize(s));
    *num_edges = nruns;
    *num_sources = (uint64_t)nsrc;
    if (nruns > max_edges || (uint64_t)nsrc > max_sources) {
        set_error("rs3d_extract_edges: %llu edges / %lld sources exceed the capacity %llu / %llu",
                  (unsigned long long)nruns, (long long)nsrc, (unsigned long long)max_edges,
                  (unsigned long long)max_sources);
        return RS3D_E_CAPACITY;
    }
    RS3D_HIP(hipMemcpyAsync(src_counts, counts_tmp, nruns * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    k_run_write<<<nbr, 256, 0, s>>>(start.as<uint32_t>(), nruns, (uint32_t)nvalid, occ.as<uint32_t>(),
                                     tri_to_obj, vertices, indices, soff, starts, ends, src_offsets,
                                     src_objects);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipStreamSynchronize(s));
    return RS3D_OK;
}

}  // extern "C"
