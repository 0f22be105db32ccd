// lbvh.hip -- linear BVH builder (replaces OptiX's GAS build,
// query/src/optix/OptixAccelerationStructure.cpp:22-68).
//
// 1. per-triangle padded AABB + centroid; centroid bounds by block reduction + ordered-int atomics
// 2. 30-bit Morton codes of centroids, stable LSD radix sort (ties keep triangle order)
// 3. triangles gathered into 48-byte records in Morton order
// 4. Karras (HPG 2012) radix-tree emission over the unique 62-bit keys (code << 32 | rank)
// 5. bottom-up refit: the second thread to reach a node merges its children's boxes
//    (leaf windows per workgroup, then one workgroup for the nodes crossing windows)
// 6. 64-byte node emission; subtrees of at most leaf_size triangles become leaves
#include <cstring>

#include "rs3d_internal.h"

namespace rs3d {

constexpr int BVH_BLOCK = 256;

__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__host__ __device__ inline float ord2f(int i) {
    int b = i >= 0 ? i : (i ^ 0x7fffffff);
    float f;
    memcpy(&f, &b, sizeof f);
    return f;
}

__device__ __forceinline__ float3 ld3(const float* p, uint64_t i) {
    return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

// bounds[0..2] = centroid min, [3..5] = centroid max, [6..8] = box min, [9..11] = box max
// bounds[16..18] = vertex min, [19..21] = vertex max; bounds[12] = bad-index flag,
// [13] = phase-3 refit records, [14] = refit record overflow
__global__ void k_init_bounds(int* bounds) {
    int t = threadIdx.x;
    if (t < 12) bounds[t] = ((t / 3) % 2 == 0) ? 0x7fffffff : (int)0x80000000;
    if (t >= 12 && t <= 14) bounds[t] = 0;
    if (t >= 16 && t < 22) bounds[t] = t < 19 ? 0x7fffffff : (int)0x80000000;
}

// Block-wide min/max of v[] (entries k with (k / 3) % 2 == 0 are minima, the others
// maxima), merged into dst[] with one ordered-int atomic per entry per block. Every
// thread of the block must call it.
template <int NV>
__device__ __forceinline__ void commit_bounds(float (&v)[NV], int* __restrict__ dst) {
    __shared__ float red[BVH_BLOCK / 64][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const bool is_min = (k / 3) % 2 == 0;
        for (int off = 32; off > 0; off >>= 1) {
            const float o = __shfl_xor(v[k], off, 64);
            v[k] = is_min ? fminf(v[k], o) : fmaxf(v[k], o);
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        const bool is_min = (k / 3) % 2 == 0;
        float r = red[0][k];
        for (int q = 1; q < BVH_BLOCK / 64; ++q) r = is_min ? fminf(r, red[q][k]) : fmaxf(r, red[q][k]);
        if (is_min)
            atomicMin(&dst[k], f2ord(r));
        else
            atomicMax(&dst[k], f2ord(r));
    }
}

// Grid-stride over triangles: padded box per triangle into pbox, and the centroid and
// padded-box bounds of the whole set into bounds[0..11].
__global__ __launch_bounds__(BVH_BLOCK) void k_prim_bounds(const float* __restrict__ V,
                                                           const uint32_t* __restrict__ I,
                                                           uint64_t n, uint64_t nV,
                                                           float4* __restrict__ pbox,
                                                           int* __restrict__ bounds,
                                                           int* __restrict__ bad) {
    // acc[0..2] centroid min, [3..5] centroid max, [6..8] box min, [9..11] box max
    float acc[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) acc[k] = (k / 3) % 2 == 0 ? INFINITY : -INFINITY;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t a = I[3 * i], b = I[3 * i + 1], c = I[3 * i + 2];
        if (a >= nV || b >= nV || c >= nV) {
            atomicOr(bad, 1);
            a = b = c = 0;
        }
        const float3 p0 = ld3(V, a), p1 = ld3(V, b), p2 = ld3(V, c);
        const float lo[3] = {fminf(fminf(p0.x, p1.x), p2.x), fminf(fminf(p0.y, p1.y), p2.y),
                             fminf(fminf(p0.z, p1.z), p2.z)};
        const float hi[3] = {fmaxf(fmaxf(p0.x, p1.x), p2.x), fmaxf(fmaxf(p0.y, p1.y), p2.y),
                             fmaxf(fmaxf(p0.z, p1.z), p2.z)};
        float plo[3], phi[3];
        prim_padded_box(p0, p1, p2, plo, phi);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float cc = 0.5f * lo[k] + 0.5f * hi[k];
            acc[k] = fminf(acc[k], cc);
            acc[3 + k] = fmaxf(acc[3 + k], cc);
            acc[6 + k] = fminf(acc[6 + k], plo[k]);
            acc[9 + k] = fmaxf(acc[9 + k], phi[k]);
        }
        pbox[2 * i] = make_float4(plo[0], plo[1], plo[2], 0.0f);
        pbox[2 * i + 1] = make_float4(phi[0], phi[1], phi[2], 0.0f);
    }
    commit_bounds<12>(acc, bounds);
}

// Exact min/max over all vertices (the dataset box of query/src/geometry/BoundingBox.cpp:4-10).
__global__ __launch_bounds__(BVH_BLOCK) void k_vertex_bounds(const float* __restrict__ V, uint64_t nV,
                                                             int* __restrict__ vb) {
    float acc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nV;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float3 p = ld3(V, i);
        acc[0] = fminf(acc[0], p.x);
        acc[1] = fminf(acc[1], p.y);
        acc[2] = fminf(acc[2], p.z);
        acc[3] = fmaxf(acc[3], p.x);
        acc[4] = fmaxf(acc[4], p.y);
        acc[5] = fmaxf(acc[5], p.z);
    }
    commit_bounds<6>(acc, vb);
}

__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t quant10(float c, float lo, float scale) {
    float u = (c - lo) * scale;
    u = fminf(fmaxf(u, 0.0f), 1023.0f);
    return (uint32_t)u;
}

__global__ __launch_bounds__(BVH_BLOCK) void k_morton(const float4* __restrict__ pbox, uint64_t n,
                                                      const int* __restrict__ bounds,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lo[3], sc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = ord2f(bounds[k]);
        const float ext = ord2f(bounds[3 + k]) - lo[k];
        sc[k] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    const float4 a = pbox[2 * i], b = pbox[2 * i + 1];
    // centroid of the padded box equals the centroid of the raw box up to rounding;
    // any deterministic choice gives a valid ordering.
    const float cx = 0.5f * a.x + 0.5f * b.x, cy = 0.5f * a.y + 0.5f * b.y,
                cz = 0.5f * a.z + 0.5f * b.z;
    const uint32_t code = (expand_bits10(quant10(cx, lo[0], sc[0])) << 2) |
                          (expand_bits10(quant10(cy, lo[1], sc[1])) << 1) |
                          expand_bits10(quant10(cz, lo[2], sc[2]));
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(BVH_BLOCK) void k_gather_tris(const float* __restrict__ V,
                                                           const uint32_t* __restrict__ I,
                                                           const int32_t* __restrict__ tri_obj,
                                                           uint64_t nV,
                                                           const uint32_t* __restrict__ perm,
                                                           const float4* __restrict__ pbox,
                                                           uint64_t n, TriRec* __restrict__ tris,
                                                           float4* __restrict__ sbox) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t t = perm[j];
    uint32_t a = I[3 * (uint64_t)t], b = I[3 * (uint64_t)t + 1], c = I[3 * (uint64_t)t + 2];
    if (a >= nV || b >= nV || c >= nV) a = b = c = 0;
    const float3 p0 = ld3(V, a), p1 = ld3(V, b), p2 = ld3(V, c);
    const int32_t obj = tri_obj ? tri_obj[t] : (int32_t)t;
    TriRec r;
    r.v0 = make_float4(p0.x, p0.y, p0.z, __int_as_float((int)t));
    r.v1 = make_float4(p1.x, p1.y, p1.z, __int_as_float(obj));
    r.v2 = make_float4(p2.x, p2.y, p2.z, 0.0f);
    tris[j] = r;
    sbox[2 * j] = pbox[2 * (uint64_t)t];
    sbox[2 * j + 1] = pbox[2 * (uint64_t)t + 1];
}

__device__ __forceinline__ int kdelta(const uint32_t* __restrict__ codes, int64_t n, int64_t i,
                                      int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = ((uint64_t)codes[i] << 32) | (uint64_t)i;
    const uint64_t b = ((uint64_t)codes[j] << 32) | (uint64_t)j;
    return __clzll(a ^ b);
}

// child encoding: >= 0 internal node index, < 0 leaf primitive ~p
__global__ __launch_bounds__(BVH_BLOCK) void k_karras(const uint32_t* __restrict__ codes, int64_t n,
                                                      int2* __restrict__ child,
                                                      int2* __restrict__ range,
                                                      int* __restrict__ parent_node,
                                                      int* __restrict__ parent_leaf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (kdelta(codes, n, i, i + 1) - kdelta(codes, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(codes, n, i, i - d);
    int64_t lmax = 2;
    while (kdelta(codes, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(codes, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = kdelta(codes, n, i, j);
    int64_t s = 0;
    for (int64_t div = 2;; div <<= 1) {
        const int64_t t = (l + div - 1) / div;
        if (kdelta(codes, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t first = i < j ? i : j, last = i < j ? j : i;
    int c0, c1;
    if (first == gamma) {
        c0 = ~(int)gamma;
        parent_leaf[gamma] = (int)i;
    } else {
        c0 = (int)gamma;
        parent_node[gamma] = (int)i;
    }
    if (last == gamma + 1) {
        c1 = ~(int)(gamma + 1);
        parent_leaf[gamma + 1] = (int)i;
    } else {
        c1 = (int)(gamma + 1);
        parent_node[gamma + 1] = (int)i;
    }
    child[i] = make_int2(c0, c1);
    range[i] = make_int2((int)first, (int)last);
    if (i == 0) parent_node[0] = -1;
}

// Bottom-up refit in three phases, so that threads only ever synchronise with threads of
// their own workgroup (workgroup-scope fences; no device-wide fence per tree level).
// A Karras node covers a contiguous leaf range [first, last] and its index is one end of
// that range. Phase 1: a workgroup owns a window of REFIT_WIN consecutive leaves and refits
// every node whose range lies inside the window; a thread that reaches a node crossing
// the window's edge records it and stops. Phase 2: a workgroup owns MID_WINS consecutive
// windows and continues from their records, refitting every node inside its span; nodes
// crossing the span are recorded again. Phase 3: one workgroup finishes from those records
// (the top of the tree). A record means "one child of this node is complete"; each node's
// second arrival refits it, whichever phase that happens in.
constexpr int REFIT_WIN = 1024;
constexpr int MID_WINS = 256;   // phase-2 span: 256 windows = 262,144 leaves
constexpr int SLOT_CAP = 64;    // phase-2 records per window (more go to phase 3)

__device__ __forceinline__ void refit_node(const int2* __restrict__ child,
                                           const float4* __restrict__ sbox, float4* nbox,
                                           int* height, int p) {
    const int2 c = child[p];
    float4 lo0, hi0, lo1, hi1;
    int h0 = 0, h1 = 0;
    if (c.x < 0) {
        lo0 = sbox[2 * (int64_t)(~c.x)];
        hi0 = sbox[2 * (int64_t)(~c.x) + 1];
    } else {
        lo0 = nbox[2 * (int64_t)c.x];
        hi0 = nbox[2 * (int64_t)c.x + 1];
        h0 = height[c.x];
    }
    if (c.y < 0) {
        lo1 = sbox[2 * (int64_t)(~c.y)];
        hi1 = sbox[2 * (int64_t)(~c.y) + 1];
    } else {
        lo1 = nbox[2 * (int64_t)c.y];
        hi1 = nbox[2 * (int64_t)c.y + 1];
        h1 = height[c.y];
    }
    nbox[2 * (int64_t)p] = make_float4(fminf(lo0.x, lo1.x), fminf(lo0.y, lo1.y),
                                       fminf(lo0.z, lo1.z), 0.0f);
    nbox[2 * (int64_t)p + 1] = make_float4(fmaxf(hi0.x, hi1.x), fmaxf(hi0.y, hi1.y),
                                           fmaxf(hi0.z, hi1.z), 0.0f);
    height[p] = (h0 > h1 ? h0 : h1) + 1;
}

// Record for phase 3 (capacity cap; overflow raises *bad, which fails the build).
__device__ __forceinline__ void push_top(int p, int* __restrict__ top, int* __restrict__ num_top, int64_t cap,
                                         int* __restrict__ bad) {
    const int k = atomicAdd(num_top, 1);
    if (k < cap) top[k] = p;
    else atomicExch(bad, 1);
}

__global__ __launch_bounds__(REFIT_WIN) void k_refit_local(const int2* __restrict__ child,
                                                           const int2* __restrict__ range,
                                                           const int* __restrict__ parent_node,
                                                           const int* __restrict__ parent_leaf,
                                                           const float4* __restrict__ sbox,
                                                           int64_t n, float4* nbox, int* height,
                                                           int* __restrict__ slots,
                                                           int* __restrict__ win_count,
                                                           int* __restrict__ top, int* __restrict__ num_top,
                                                           int* __restrict__ bad) {
    __shared__ int arrived[REFIT_WIN];
    __shared__ int nslot;
    const int64_t base = (int64_t)blockIdx.x * REFIT_WIN;
    const int64_t base2 = (int64_t)(blockIdx.x / MID_WINS) * ((int64_t)REFIT_WIN * MID_WINS);
    const int64_t end2 = base2 + (int64_t)REFIT_WIN * MID_WINS;
    arrived[threadIdx.x] = 0;
    if (threadIdx.x == 0) nslot = 0;
    __syncthreads();
    const int64_t j = base + threadIdx.x;
    if (j < n) {
        int p = parent_leaf[j];
        while (p >= 0) {
            const int2 r = range[p];
            if (r.x < base || r.y >= base + REFIT_WIN) {
                // crossing node: phase 2 when inside this span, else phase 3
                bool mid = r.x >= base2 && r.y < end2;
                if (mid) {
                    const int k = atomicAdd(&nslot, 1);
                    if (k < SLOT_CAP) slots[(int64_t)blockIdx.x * SLOT_CAP + k] = p;
                    else mid = false;
                }
                if (!mid) push_top(p, top, num_top, n, bad);
                break;
            }
            __threadfence_block();
            if (atomicAdd(&arrived[p - base], 1) == 0) break;  // sibling still in progress
            __threadfence_block();
            refit_node(child, sbox, nbox, height, p);
            p = parent_node[p];
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && win_count) win_count[blockIdx.x] = nslot < SLOT_CAP ? nslot : SLOT_CAP;
}

__global__ __launch_bounds__(MID_WINS) void k_refit_mid(const int2* __restrict__ child,
                                                        const int2* __restrict__ range,
                                                        const int* __restrict__ parent_node,
                                                        const float4* __restrict__ sbox, float4* nbox,
                                                        int* height, int* flags, const int* __restrict__ slots,
                                                        const int* __restrict__ win_count, int64_t num_win,
                                                        int64_t n, int* __restrict__ top,
                                                        int* __restrict__ num_top, int* __restrict__ bad) {
    const int64_t w = (int64_t)blockIdx.x * MID_WINS + threadIdx.x;
    const int64_t base2 = (int64_t)blockIdx.x * ((int64_t)REFIT_WIN * MID_WINS);
    const int64_t end2 = base2 + (int64_t)REFIT_WIN * MID_WINS;
    if (w >= num_win) return;
    const int cnt = win_count[w];
    for (int k = 0; k < cnt; ++k) {
        int p = slots[w * SLOT_CAP + k];
        while (p >= 0) {
            const int2 r = range[p];
            if (r.x < base2 || r.y >= end2) {
                push_top(p, top, num_top, n, bad);
                break;
            }
            __threadfence_block();
            if (__hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
            __threadfence_block();
            refit_node(child, sbox, nbox, height, p);
            p = parent_node[p];
        }
    }
}

__global__ __launch_bounds__(REFIT_WIN) void k_refit_top(const int2* __restrict__ child,
                                                         const int* __restrict__ parent_node,
                                                         const float4* __restrict__ sbox,
                                                         float4* nbox, int* height, int* flags,
                                                         const int* __restrict__ pending,
                                                         const int* __restrict__ num_pending, int64_t cap) {
    const int m = (int)min((int64_t)*num_pending, cap);
    for (int it = threadIdx.x; it < m; it += REFIT_WIN) {
        int p = pending[it];
        while (p >= 0) {
            __threadfence_block();
            if (__hip_atomic_fetch_add(&flags[p], 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
            __threadfence_block();
            refit_node(child, sbox, nbox, height, p);
            p = parent_node[p];
        }
    }
}

__device__ __forceinline__ void child_box(const int2* __restrict__ range,
                                          const float4* __restrict__ sbox,
                                          const float4* __restrict__ nbox, int c, int leaf_size,
                                          float4& lo, float4& hi, int& link, int& count) {
    if (c < 0) {
        const int p = ~c;
        lo = sbox[2 * (int64_t)p];
        hi = sbox[2 * (int64_t)p + 1];
        link = p;
        count = 1;
        return;
    }
    lo = nbox[2 * (int64_t)c];
    hi = nbox[2 * (int64_t)c + 1];
    const int2 r = range[c];
    const int sz = r.y - r.x + 1;
    if (sz <= leaf_size) {
        link = r.x;
        count = sz;
    } else {
        link = c;
        count = 0;
    }
}

__device__ __forceinline__ BNode make_node(float4 lo0, float4 hi0, float4 lo1, float4 hi1, int l0,
                                           int l1, int n0, int n1) {
    BNode nd;
    nd.a = make_float4(lo0.x, hi0.x, lo0.y, hi0.y);
    nd.b = make_float4(lo0.z, hi0.z, lo1.x, hi1.x);
    nd.c = make_float4(lo1.y, hi1.y, lo1.z, hi1.z);
    nd.link = make_int4(l0, l1, n0, n1);
    return nd;
}

__global__ __launch_bounds__(BVH_BLOCK) void k_emit_nodes(const int2* __restrict__ child,
                                                          const int2* __restrict__ range,
                                                          const float4* __restrict__ sbox,
                                                          const float4* __restrict__ nbox,
                                                          int64_t n, int leaf_size,
                                                          BNode* __restrict__ nodes) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 1) {
        if (i == 0) {
            const float4 e_lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
            const float4 e_hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
            nodes[0] = make_node(sbox[0], sbox[1], e_lo, e_hi, 0, -1, 1, 0);
        }
        return;
    }
    if (i >= n - 1) return;
    const int2 c = child[i];
    float4 lo0, hi0, lo1, hi1;
    int l0, l1, n0, n1;
    child_box(range, sbox, nbox, c.x, leaf_size, lo0, hi0, l0, n0);
    child_box(range, sbox, nbox, c.y, leaf_size, lo1, hi1, l1, n1);
    nodes[i] = make_node(lo0, hi0, lo1, hi1, l0, l1, n0, n1);
}

}  // namespace rs3d

using namespace rs3d;

namespace {
struct TmpAlloc {
    std::vector<void*> ptrs;
    ~TmpAlloc() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <typename T>
    int get(T** out, size_t count) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, count * sizeof(T) > 0 ? count * sizeof(T) : 16);
        if (e != hipSuccess) {
            set_error("BVH build: hipMalloc(%zu) failed: %s", count * sizeof(T), hipGetErrorString(e));
            return RS3D_E_NOMEM;
        }
        ptrs.push_back(p);
        *out = static_cast<T*>(p);
        return RS3D_OK;
    }
};
}  // namespace

extern "C" {

int rs3d_bvh_build(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                   const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                   const rs3d_bvh_opts* opts, rs3d_bvh** out) {
    if (!ctx || !out || (num_triangles && (!vertices || !indices))) {
        set_error("rs3d_bvh_build: bad arguments");
        return RS3D_E_ARG;
    }
    *out = nullptr;
    if (num_triangles >= 0x7fffffffull) {
        set_error("rs3d_bvh_build: %llu triangles exceed the 2^31-1 limit",
                  (unsigned long long)num_triangles);
        return RS3D_E_ARG;
    }
    int leaf = opts && opts->leaf_size > 0 ? opts->leaf_size : 1;
    if (leaf > 8) {
        set_error("rs3d_bvh_build: leaf_size %d > 8", leaf);
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(ctx->device));
    const int64_t n = (int64_t)num_triangles;
    rs3d_bvh* bvh = new rs3d_bvh();
    bvh->ctx = ctx;
    bvh->num_tris = num_triangles;
    bvh->leaf_size = leaf;
    if (n == 0) {
        *out = bvh;
        return RS3D_OK;
    }
    hipStream_t s = ctx->stream;
    TmpAlloc tmp;
    float4 *pbox, *sbox, *nbox;
    uint32_t *keys, *vals;
    int2 *child, *range;
    int *parent_node, *parent_leaf, *flags, *height, *bounds;
    const int64_t ni = n > 1 ? n - 1 : 1;
    int rc = RS3D_OK;
    auto fail = [&](int code) {
        delete bvh;
        return code;
    };
    if ((rc = tmp.get(&pbox, 2 * n)) || (rc = tmp.get(&sbox, 2 * n)) || (rc = tmp.get(&nbox, 2 * ni)) ||
        (rc = tmp.get(&keys, n)) || (rc = tmp.get(&vals, n)) || (rc = tmp.get(&child, ni)) ||
        (rc = tmp.get(&range, ni)) || (rc = tmp.get(&parent_node, ni)) ||
        (rc = tmp.get(&parent_leaf, n)) || (rc = tmp.get(&flags, ni)) ||
        (rc = tmp.get(&height, ni)) || (rc = tmp.get(&bounds, 32)))
        return fail(rc);
    hipError_t e;
    if ((e = hipMalloc(&bvh->nodes, ni * sizeof(BNode))) != hipSuccess ||
        (e = hipMalloc(&bvh->tris, n * sizeof(TriRec))) != hipSuccess) {
        set_error("rs3d_bvh_build: hipMalloc failed: %s", hipGetErrorString(e));
        if (bvh->nodes) (void)hipFree(bvh->nodes);
        bvh->nodes = nullptr;
        return fail(RS3D_E_NOMEM);
    }
    bvh->num_nodes = ni;
    bvh->device_bytes = ni * sizeof(BNode) + n * sizeof(TriRec);
    unsigned nb;
    if ((rc = grid_blocks(n, BVH_BLOCK, &nb))) return fail(rc);
    int* bad = bounds + 12;
    k_init_bounds<<<1, 64, 0, s>>>(bounds);
    {
        unsigned nbv = (unsigned)std::min<uint64_t>((num_vertices + BVH_BLOCK - 1) / BVH_BLOCK, 1024);
        k_vertex_bounds<<<nbv ? nbv : 1, BVH_BLOCK, 0, s>>>(vertices, num_vertices, bounds + 16);
    }
    k_prim_bounds<<<std::min(nb, 1024u), BVH_BLOCK, 0, s>>>(vertices, indices, n, num_vertices, pbox,
                                                            bounds, bad);
    k_morton<<<nb, BVH_BLOCK, 0, s>>>(pbox, n, bounds, keys, vals);
    if ((e = hipGetLastError()) != hipSuccess) {
        set_error("BVH kernels: %s", hipGetErrorString(e));
        return fail(RS3D_E_HIP);
    }
    if ((rc = radix_sort_u32_pairs(ctx, keys, vals, n, 30))) return fail(rc);
    k_gather_tris<<<nb, BVH_BLOCK, 0, s>>>(vertices, indices, tri_to_obj, num_vertices, vals, pbox, n,
                                           bvh->tris, sbox);
    if (n > 1) {
        unsigned nbi;
        if ((rc = grid_blocks(n - 1, BVH_BLOCK, &nbi))) return fail(rc);
        k_karras<<<nbi, BVH_BLOCK, 0, s>>>(keys, n, child, range, parent_node, parent_leaf);
        if ((e = hipMemsetAsync(flags, 0, (n - 1) * sizeof(int), s)) != hipSuccess) {
            set_error("hipMemsetAsync: %s", hipGetErrorString(e));
            return fail(RS3D_E_HIP);
        }
        int* num_top = bounds + 13;
        int* refit_bad = bounds + 14;
        int* top = reinterpret_cast<int*>(vals);  // the sort permutation is dead here
        const int64_t nw = (n + REFIT_WIN - 1) / REFIT_WIN;
        // phase-2 records in the Morton keys, dead after k_karras (nw * (SLOT_CAP + 1) <= n
        // whenever there is more than one window)
        int* slots = nw > 1 ? reinterpret_cast<int*>(keys) : nullptr;
        int* win_count = nw > 1 ? slots + nw * SLOT_CAP : nullptr;
        k_refit_local<<<(unsigned)nw, REFIT_WIN, 0, s>>>(child, range, parent_node, parent_leaf, sbox, n,
                                                         nbox, height, slots, win_count, top, num_top,
                                                         refit_bad);
        if (nw > 1) {
            const unsigned nm = (unsigned)((nw + MID_WINS - 1) / MID_WINS);
            k_refit_mid<<<nm, MID_WINS, 0, s>>>(child, range, parent_node, sbox, nbox, height, flags, slots,
                                                win_count, nw, n, top, num_top, refit_bad);
            k_refit_top<<<1, REFIT_WIN, 0, s>>>(child, parent_node, sbox, nbox, height, flags, top, num_top,
                                                n);
        }
        k_emit_nodes<<<nbi, BVH_BLOCK, 0, s>>>(child, range, sbox, nbox, n, leaf, bvh->nodes);
    } else {
        k_emit_nodes<<<1, 64, 0, s>>>(child, range, sbox, nbox, n, leaf, bvh->nodes);
    }
    if ((e = hipGetLastError()) != hipSuccess) {
        set_error("BVH kernels: %s", hipGetErrorString(e));
        return fail(RS3D_E_HIP);
    }
    int hb[22];
    int root_h = 0;
    if ((e = hipMemcpyAsync(hb, bounds, sizeof hb, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (n > 1 && (e = hipMemcpyAsync(&root_h, height, sizeof(int), hipMemcpyDeviceToHost, s)) != hipSuccess) ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
        set_error("BVH build sync: %s", hipGetErrorString(e));
        return fail(RS3D_E_HIP);
    }
    if (hb[12]) {
        set_error("rs3d_bvh_build: triangle index out of range (>= %llu vertices)",
                  (unsigned long long)num_vertices);
        return fail(RS3D_E_ARG);
    }
    if (hb[14]) {
        set_error("rs3d_bvh_build: refit records overflowed");
        return fail(RS3D_E_HIP);
    }
    for (int k = 0; k < 3; ++k) {
        bvh->bmin[k] = ord2f(hb[6 + k]);
        bvh->bmax[k] = ord2f(hb[9 + k]);
        bvh->vmin[k] = ord2f(hb[16 + k]);
        bvh->vmax[k] = ord2f(hb[19 + k]);
    }
    bvh->max_depth = n > 1 ? root_h : 1;
    *out = bvh;
    return RS3D_OK;
}

int rs3d_bvh_free(rs3d_bvh* bvh) {
    if (!bvh) return RS3D_OK;
    if (bvh->ctx) (void)hipSetDevice(bvh->ctx->device);
    if (bvh->nodes) (void)hipFree(bvh->nodes);
    if (bvh->tris) (void)hipFree(bvh->tris);
    for (int a = 0; a < 3; ++a) axis_grid_free(bvh->agrid[a]);
    seg_occ_free(bvh->socc);
    delete bvh;
    return RS3D_OK;
}

int rs3d_bvh_export(const rs3d_bvh* bvh, void* nodes, void* tris) {
    if (!bvh || !bvh->ctx) {
        set_error("rs3d_bvh_export: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(bvh->ctx->device));
    hipStream_t s = bvh->ctx->stream;
    if (nodes && bvh->num_nodes && bvh->nodes)
        RS3D_HIP(hipMemcpyAsync(nodes, bvh->nodes, bvh->num_nodes * sizeof(BNode), hipMemcpyDeviceToDevice, s));
    if (tris && bvh->num_tris && bvh->tris)
        RS3D_HIP(hipMemcpyAsync(tris, bvh->tris, bvh->num_tris * sizeof(TriRec), hipMemcpyDeviceToDevice, s));
    RS3D_HIP(hipStreamSynchronize(s));
    return RS3D_OK;
}

int rs3d_bvh_info_get(const rs3d_bvh* bvh, rs3d_bvh_info* info) {
    if (!bvh || !info) {
        set_error("rs3d_bvh_info_get: bad arguments");
        return RS3D_E_ARG;
    }
    info->num_triangles = bvh->num_tris;
    info->num_nodes = bvh->num_nodes;
    info->max_depth = bvh->max_depth;
    info->leaf_size = bvh->leaf_size;
    for (int k = 0; k < 3; ++k) {
        info->bounds_min[k] = bvh->bmin[k];
        info->bounds_max[k] = bvh->bmax[k];
    }
    info->device_bytes = bvh->device_bytes;
    return RS3D_OK;
}

}  // extern "C"
