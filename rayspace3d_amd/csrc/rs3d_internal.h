// rs3d_internal.h -- shared host/device definitions of librs3d (MI355X, gfx950).
//
// HBM layouts:
//   BNode (64 B, one per internal node of the binary LBVH): the boxes of both children
//   and their links, so one node fetch decides both children.
//   TriRec (48 B, one per triangle in Morton order): three vertices; v0.w carries the
//   original triangle index (tie-break key and reported primitive id), v1.w the object id.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rs3d.h"

namespace rs3d {

// ------------------------------------------------------------------------------------
// Device data layouts
// ------------------------------------------------------------------------------------
struct alignas(16) BNode {
    float4 a;  // lo0.x hi0.x lo0.y hi0.y
    float4 b;  // lo0.z hi0.z lo1.x hi1.x
    float4 c;  // lo1.y hi1.y lo1.z hi1.z
    int4 link; // child0, child1 (node index, or first triangle), count0, count1
               // count > 0: leaf of `count` triangles starting at the link;
               // count == 0 and link >= 0: internal node; link < 0: empty child.
};
static_assert(sizeof(BNode) == 64, "BNode must be 64 bytes");

struct alignas(16) TriRec {
    float4 v0;  // .w = original triangle index (bits)
    float4 v1;  // .w = object id (bits)
    float4 v2;  // .w unused
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

struct GridCellRec {
    uint32_t center;
    uint32_t touch;
    float avg;
    float ratio;
};

// ------------------------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), two-sided.
// Every fp32 operation is written out in a fixed order and the library is compiled with
// -ffp-contract=off, so results match the CPU oracle bit for bit.
// ------------------------------------------------------------------------------------
struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

__host__ __device__ inline float get3(const float3& v, int k) {
    return k == 0 ? v.x : (k == 1 ? v.y : v.z);
}

__host__ __device__ inline RayPre ray_pre(float3 d) {
    RayPre r;
    int kz = 0;
    float m = fabsf(d.x);
    if (fabsf(d.y) > m) { kz = 1; m = fabsf(d.y); }
    if (fabsf(d.z) > m) { kz = 2; }
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    const float dz = get3(d, kz);
    if (dz < 0.0f) {
        int t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.Sx = get3(d, kx) / dz;
    r.Sy = get3(d, ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}

// Returns true and t (unclipped) when the triangle is hit.
__device__ __forceinline__ bool tri_hit(const float3 o, const RayPre& r, const float3 v0,
                                        const float3 v1, const float3 v2, float& tout) {
    const float3 A = make_float3(v0.x - o.x, v0.y - o.y, v0.z - o.z);
    const float3 B = make_float3(v1.x - o.x, v1.y - o.y, v1.z - o.z);
    const float3 C = make_float3(v2.x - o.x, v2.y - o.y, v2.z - o.z);
    const float Akz = get3(A, r.kz), Bkz = get3(B, r.kz), Ckz = get3(C, r.kz);
    const float Ax = get3(A, r.kx) - r.Sx * Akz;
    const float Ay = get3(A, r.ky) - r.Sy * Akz;
    const float Bx = get3(B, r.kx) - r.Sx * Bkz;
    const float By = get3(B, r.ky) - r.Sy * Bkz;
    const float Cx = get3(C, r.kx) - r.Sx * Ckz;
    const float Cy = get3(C, r.ky) - r.Sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * Akz;
    const float Bz = r.Sz * Bkz;
    const float Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    tout = T / det;
    return true;
}

// tri_hit for the ray +AX (the only direction of point and object rays), without the shear.
// For d = +e_AX, ray_pre gives kz = AX, kx = AX+1, ky = AX+2 (mod 3), Sx = Sy = +0, Sz = 1,
// so the sheared coordinates are A[kx] - (+0 * A[kz]) and A[kz] * 1. With A[kz] finite the
// first is A[kx] up to the sign of a zero, and a zero's sign changes none of the decisions
// (U, V, W and det are compared with 0; a zero term leaves a nonzero sum unchanged; a zero
// T gives t = +-0, outside every window). With A[kz] infinite tri_hit takes NaN coordinates
// and returns t = NaN or a miss, and this form returns t = +-inf, NaN or a miss: no window
// [thr, tmax) with tmax <= 1e16 admits any of them. The hits that count are identical.
template <int AX>
__host__ __device__ __forceinline__ bool tri_hit_axis(const float3 o, const float3 v0, const float3 v1, const float3 v2,
                                             float& tout) {
    constexpr int KX = (AX + 1) % 3, KY = (AX + 2) % 3;
    const float Ax = get3(v0, KX) - get3(o, KX), Ay = get3(v0, KY) - get3(o, KY);
    const float Bx = get3(v1, KX) - get3(o, KX), By = get3(v1, KY) - get3(o, KY);
    const float Cx = get3(v2, KX) - get3(o, KX), Cy = get3(v2, KY) - get3(o, KY);
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = get3(v0, AX) - get3(o, AX);
    const float Bz = get3(v1, AX) - get3(o, AX);
    const float Cz = get3(v2, AX) - get3(o, AX);
    const float T = U * Az + V * Bz + W * Cz;
    tout = T / det;
    return true;
}

// Outward padding of primitive boxes, relative to coordinate magnitude and extent, so
// box culling can never remove a triangle the fp32 test would report (pruning only).
__host__ __device__ inline float box_pad(float mag, float ext) {
    return (mag + ext) * 3.814697265625e-6f /* 2^-18 */ + 1e-35f;
}

// Padded box of one triangle (the BVH's leaf boxes; every structure that culls triangles
// for the same rays uses this one function, so they all cull alike).
__device__ __forceinline__ void prim_padded_box(float3 p0, float3 p1, float3 p2, float (&plo)[3],
                                                float (&phi)[3]) {
    const float lo[3] = {fminf(fminf(p0.x, p1.x), p2.x), fminf(fminf(p0.y, p1.y), p2.y),
                         fminf(fminf(p0.z, p1.z), p2.z)};
    const float hi[3] = {fmaxf(fmaxf(p0.x, p1.x), p2.x), fmaxf(fmaxf(p0.y, p1.y), p2.y),
                         fmaxf(fmaxf(p0.z, p1.z), p2.z)};
    float ext = 0.0f, mag = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ext = fmaxf(ext, hi[k] - lo[k]);
        mag = fmaxf(mag, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
    }
    const float pad = box_pad(mag, ext);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        plo[k] = lo[k] - pad;
        phi[k] = hi[k] + pad;
    }
}

// ------------------------------------------------------------------------------------
// Host-side objects
// ------------------------------------------------------------------------------------
void set_error(const char* fmt, ...);
const char* last_error();

#define RS3D_HIP(call)                                                                  \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::rs3d::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                              __FILE__, __LINE__);                                      \
            return e_ == hipErrorOutOfMemory ? RS3D_E_NOMEM : RS3D_E_HIP;               \
        }                                                                               \
    } while (0)

#define RS3D_TRY(call)                       \
    do {                                     \
        int r_ = (call);                     \
        if (r_ != RS3D_OK) return r_;        \
    } while (0)

// Block-wide sum of a 64-bit value over a 256-thread block (every thread passes its
// value; thread 0 gets the total). Uses 32 bytes of LDS.
__device__ __forceinline__ unsigned long long block_sum_256(unsigned long long v) {
    __shared__ unsigned long long part[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return threadIdx.x == 0 ? part[0] + part[1] + part[2] + part[3] : 0ull;
}

// The stripe of a striped accumulator this block adds into.
__device__ __forceinline__ unsigned long long* acc_stripe(unsigned long long* acc) {
    return acc + (blockIdx.x % 64u) * 16u;
}

// Grow-only device buffer owned by a context (scratch that persists across calls).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

}  // namespace rs3d

struct rs3d_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // stream-ordered pool for build temporaries (axis grids), keeping up to
    // CTX_TMP_POOL_KEEP bytes cached between builds (null: plain hipMalloc / hipFree)
    hipMemPool_t tmp_pool = nullptr;
    int num_cus = 0;
    // Scratch, reused across calls on this context (calls on one context are serial).
    rs3d::DevBuf sort_keys, sort_vals, sort_hist, sort_csum, scan_tmp, tmp_a, tmp_b, tmp_c, small;
    rs3d::DevBuf ray_keys, ray_vals;  // ray reordering (coherence) keys and permutation
    rs3d::DevBuf accs;                // striped accumulators (rs3d::acc_slot)
    rs3d::DevBuf edge_sel;            // edge launches: selection count or flags, and index list
    int64_t opt_ray_sort_min = 1 << 16;  // reorder rays for launches of at least this many
    int64_t opt_rays_per_wave = 64;      // lane-refill run length of point launches (0: off)
    int64_t opt_axis_grid = 1;           // point launches may use an accepted axis grid
    int64_t opt_grid_rays_per_wave = 256;   // run length of grid point launches
    int64_t opt_grid_cell_pct = 35;         // grid cell edge, % of the mean triangle extent
    bool opt_grid_cell_pct_set = false;     // set by the caller (small grids keep it as given)
    int64_t opt_seg_occupancy = 1;          // edge rays test the target's occupancy bits first
    int64_t opt_seg_cell_pct = 200;         // occupancy cell edge, % of the mean triangle extent
    uint64_t* pinned = nullptr;  // small pinned host buffer for counts
    // rs3d_filter_refine on a grid: stripes its launches leave zeroed (device), and the word
    // their counts are written to (host-coherent pinned memory)
    rs3d::DevBuf fr_acc;
    unsigned long long* fr_total = nullptr;  // host-coherent: the counts (all ones until they land)
    bool fr_dirty = false;  // a launch may have left the stripes non-zero
    // traversal-kernel timing (rs3d_ctx_profile)
    int profile = 0;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool prof_pending = false;
    double prof_ms = 0.0;
    uint64_t prof_launches = 0;
};

namespace rs3d {
struct AxisGrid;
void axis_grid_free(AxisGrid* g);
struct SegOcc;
void seg_occ_free(SegOcc* g);
}  // namespace rs3d

struct rs3d_bvh {
    rs3d_ctx* ctx = nullptr;
    rs3d::AxisGrid* agrid[3] = {nullptr, nullptr, nullptr};  // per ray axis, built lazily
    int agrid_state[3] = {0, 0, 0};  // 0 untried, 1 built, 2 rejected, 3 refused for memory
    rs3d::SegOcc* socc = nullptr;  // segment occupancy bits, built lazily
    int socc_state = 0;            // 0 untried, 1 built, 2 rejected, 3 refused for memory
    uint64_t num_tris = 0;
    uint64_t num_nodes = 0;
    int leaf_size = 4;
    int max_depth = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};  // padded box bounds
    float vmin[3] = {0, 0, 0}, vmax[3] = {0, 0, 0};  // exact bounds of all vertices
    rs3d::BNode* nodes = nullptr;  // device
    rs3d::TriRec* tris = nullptr;  // device, Morton order
    size_t device_bytes = 0;
};

namespace rs3d {
// Internal primitives used across translation units (all asynchronous on ctx->stream
// unless they return a host value).
int scan_u32_exclusive(rs3d_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n);
int scan_i32_to_i64_exclusive(rs3d_ctx* ctx, const int32_t* in, int64_t* out, uint64_t n,
                              int64_t* total_dev);
int scan_u8_to_u32_exclusive(rs3d_ctx* ctx, const uint8_t* in, uint32_t* out, uint64_t n,
                             uint32_t* total_dev);
int radix_sort_u32_pairs(rs3d_ctx* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, int end_bit);
// Same sort, result left in place or in the context's alternate buffers (no copy back);
// *out_keys / *out_vals point at the sorted arrays. With seg (one entry per chunk of
// RS_SEG_KEYS keys: {first chunk of its segment, chunks in the segment}) every segment of
// whole chunks is sorted on its own.
constexpr uint64_t RS_SEG_KEYS = 16 * 2048;
int radix_sort_u32_pairs_to(rs3d_ctx* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, int end_bit,
                            uint32_t** out_keys, uint32_t** out_vals, const uint2* seg = nullptr);
int radix_sort_u64(rs3d_ctx* ctx, uint64_t* keys, uint64_t n);
int radix_sort_u64_pairs(rs3d_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n);  // stable
// Whether point launches along +axis go to the axis grid of the BVH: the context allows it
// and a grid is accepted (built here on first use unless build is false).
int axis_grid_ready(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, bool* ready, bool build = true);
// Upper bound of the device bytes an axis grid build of this BVH takes at its peak.
uint64_t axis_grid_bytes_bound(const rs3d_bvh* bvh);
// Filter-refine on the grid in one launch (axis_grid_ready said yes): box = dataset box
// (host float[6]); out_obj gets every point's answer; counts (a zeroed striped
// accumulator) accumulates candidates << 32 | inside.
int axis_grid_filter_refine(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, const float* pts, uint64_t n,
                            const float* box, int32_t* out_obj, unsigned long long* counts);
// One wave after such launches: the stripes of counts added into *total (host-coherent
// memory) and zeroed again, so that the next launch finds them ready.
int axis_grid_counts_finish(rs3d_ctx* ctx, unsigned long long* counts, unsigned long long* total);
// Segment occupancy bits of a BVH for edge rays (built here on first use when the context
// allows them); *args gets bits = null when there are none.
struct OccArgs;
int seg_occ_args(rs3d_ctx* ctx, rs3d_bvh* bvh, OccArgs* args, bool build = true);
// Upper bound of the device bytes of a BVH's segment occupancy bits (2^30 cells at most).
uint64_t seg_occ_bytes_bound(const rs3d_bvh* bvh);
uint64_t seg_occ_bytes(const rs3d_bvh* bvh);  // bytes of the built bits (0 when none)
// Point launch on the grid (axis_grid_ready said yes).
int axis_grid_launch(rs3d_ctx* ctx, rs3d_bvh* bvh, int axis, const float* pts, uint64_t n,
                     const uint32_t* imap, int out_kind, void* out);
int unique_u64(rs3d_ctx* ctx, const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* count_host);
// Scan and compaction pieces on caller-provided scratch (no allocation): the exclusive scan
// of u8 flags needs scan_scratch_elems(n) u32 words of tmp; scatter_index writes i at
// out[pos[i]] for every flagged i.
uint64_t scan_scratch_elems(uint64_t n);
int scan_u8_to_u32_exclusive_in(rs3d_ctx* ctx, const uint8_t* in, uint32_t* out, uint64_t n,
                                uint32_t* total_dev, uint32_t* tmp, uint64_t tmp_elems);
int scatter_index(rs3d_ctx* ctx, const uint8_t* flags, const uint32_t* pos, uint64_t n, uint32_t* out);
// The BVH walk of point rays as given (no reordering, no axis grid; allocates nothing).
int trace_points_walk(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* pts, uint64_t n,
                      const uint32_t* imap, int axis, int out_kind, void* out);
int fill_i32(rs3d_ctx* ctx, int32_t* p, uint64_t n, int32_t v);
// Adds the number of p[i] >= 0 into the striped accumulator acc.
int count_nonneg(rs3d_ctx* ctx, const int32_t* p, uint64_t n, unsigned long long* acc);
// Looks up, once per device (with ctx's device current), the device address of the fixed
// scratch array of rs3d_filter_refine_noalloc; called by rs3d_ctx_create.
int noalloc_init(rs3d_ctx* ctx);
// Striped accumulators. Device-scope atomics on one address serialise (about 10 ns each on
// this GPU, so ten thousand blocks adding into one counter cost ~0.1 ms), so reduction
// kernels combine within the block and add into one of ACC_STRIPES words 128 B apart
// (acc_stripe picks it from the block index); the reader combines the stripes.
constexpr int ACC_STRIPES = 64;
constexpr int ACC_PITCH = 16;  // 8-byte words from one stripe to the next
constexpr int ACC_SLOTS = 8;   // accumulators per context (6 and 7: the u64 radix sort)
// Size of a context's pinned host buffer (`pinned`): one accumulator's stripes.
constexpr size_t CTX_PINNED_BYTES = (size_t)ACC_STRIPES * ACC_PITCH * sizeof(uint64_t);
// Bytes of freed build temporaries a context's pool keeps for later builds (the rest goes
// back to the device); the pool is destroyed with the context.
constexpr uint64_t CTX_TMP_POOL_KEEP = 4ull << 30;
// Slot k (< ACC_SLOTS) of the context, every stripe set to init (fill kernel, graph-safe).
int acc_slot(rs3d_ctx* ctx, int k, uint64_t init, unsigned long long** acc);
// The stripes of an accumulator, copied to host (synchronises the stream).
int acc_read(rs3d_ctx* ctx, const unsigned long long* acc, uint64_t (&stripes)[ACC_STRIPES]);
int acc_read_sum(rs3d_ctx* ctx, const unsigned long long* acc, uint64_t* sum);
int read_u64(rs3d_ctx* ctx, const void* dev, uint64_t* host);
int grid_blocks(uint64_t n, int block, unsigned* out);
int fill_u32(rs3d_ctx* ctx, void* p, uint64_t n, uint32_t v);  // kernel fill (graph-safe)
int prof_collect(rs3d_ctx* ctx);  // folds a finished timed launch into the totals
int pick_depth(const rs3d_bvh* bvh, int* depth);  // LDS stack size for a BVH
int ht_clear_internal(rs3d_ctx* ctx, uint64_t* slots, uint64_t size, uint64_t* failed);
int ht_compact_internal(rs3d_ctx* ctx, const uint64_t* slots, uint64_t size, uint64_t* out,
                        uint64_t max_out, uint64_t* count);
int edge_hash_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e, int swap,
                       const rs3d_hash_table* ht, int chain_rule, int max_iter);
}  // namespace rs3d
