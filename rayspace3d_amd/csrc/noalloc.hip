// noalloc.hip -- filter-refine containment that allocates nothing (rs3d_filter_refine_noalloc).
//
// Callers that cannot allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid hipMalloc) get the answers of rs3d_filter_refine from this entry
// point. Its scratch -- the count accumulator and, without an axis grid, the filter flags,
// their scan and the candidate list -- lives in one file-scope device array, `scratch`,
// which every GPU the library runs on holds from module load until the process ends. Host
// code never passes the array's name to a kernel: the array's address on each device is
// looked up once per device (hipGetSymbolAddress, with that device current) and kept in
// g_scratch[device]; kernels take that address as an argument.
//
// The lookup happens in rs3d_ctx_create (noalloc_init): the runtime materialises the array
// on a device at its first symbol lookup there, so the entry point itself never triggers it.
//
// Every call on one GPU shares that GPU's array and leaves its contents behind; each call
// clears the accumulator it adds into (a fill kernel, no memset) and overwrites the flags,
// positions and candidates it reads. Calls that overlap on one GPU (two host threads, or two
// streams in flight) race on the array and must be serialised by the caller.
#include <mutex>

#include "rs3d_internal.h"

namespace rs3d {
namespace {

// Scratch layout (bytes): [0, 8192) the 64-stripe count accumulator; [8192, 8448) the scan
// total; [8448, 41216) the scan's tile partials for up to NOALLOC_MAX_WALK points; then
// u8 flags, u32 scan positions and u32 candidates, each of NOALLOC_MAX_WALK entries.
constexpr uint64_t NOALLOC_MAX_WALK = RS3D_NOALLOC_MAX_POINTS;
constexpr size_t SC_ACC = 0;
constexpr size_t SC_TOTAL = (size_t)ACC_STRIPES * ACC_PITCH * sizeof(unsigned long long);  // 8192
constexpr size_t SC_SCAN = SC_TOTAL + 256;
constexpr size_t SC_SCAN_BYTES = 32768;  // >= scan partials of NOALLOC_MAX_WALK points (16,456 B), checked at set-up
constexpr size_t SC_FLAGS = SC_SCAN + SC_SCAN_BYTES;
constexpr size_t SC_POS = SC_FLAGS + NOALLOC_MAX_WALK;
constexpr size_t SC_CAND = SC_POS + NOALLOC_MAX_WALK * sizeof(uint32_t);
constexpr size_t SC_BYTES = SC_CAND + NOALLOC_MAX_WALK * sizeof(uint32_t);
static_assert(SC_FLAGS % 256 == 0 && SC_POS % 256 == 0 && SC_CAND % 256 == 0, "aligned sections");
static_assert(SC_BYTES == RS3D_NOALLOC_SCRATCH_BYTES, "scratch size as declared in rs3d.h");

__device__ __align__(256) unsigned char scratch[SC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[MAX_DEVICES] = {nullptr};

// The device address of `scratch` on ctx's device, looked up on the first call there.
int scratch_ptr(rs3d_ctx* ctx, unsigned char** out) {
    const int dev = ctx->device;
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("rs3d_filter_refine_noalloc: device %d beyond the %d supported", dev, MAX_DEVICES);
        return RS3D_E_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        if (scan_scratch_elems(NOALLOC_MAX_WALK) * sizeof(uint32_t) > SC_SCAN_BYTES) {
            set_error("rs3d_filter_refine_noalloc: scan partials exceed their scratch section");
            return RS3D_E_ARG;
        }
        int cur = -1;
        RS3D_HIP(hipGetDevice(&cur));
        if (cur != dev) RS3D_HIP(hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) {
            set_error("rs3d_filter_refine_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s",
                      dev, hipGetErrorString(e));
            return RS3D_E_HIP;
        }
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return RS3D_OK;
}

__global__ void k_zero_stripes(unsigned long long* __restrict__ acc) {
    const int i = threadIdx.x;
    if (i < ACC_STRIPES) acc[i * ACC_PITCH] = 0ull;
}

int read_stripes_sum(rs3d_ctx* ctx, const unsigned long long* acc, uint64_t* sum) {
    // ctx->pinned (created with the context) receives the stripes: no allocation here
    return acc_read_sum(ctx, acc, sum);
}

}  // namespace

int noalloc_init(rs3d_ctx* ctx) {
    unsigned char* p = nullptr;
    return scratch_ptr(ctx, &p);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_filter_refine_noalloc(rs3d_ctx* ctx, const rs3d_bvh* bvh, const float* points, uint64_t n,
                               int axis, int32_t* out_obj, uint64_t* candidate_count,
                               uint64_t* inside_count) {
    if (!ctx || !bvh || (n && (!points || !out_obj)) || axis < 0 || axis > 2 ||
        n > 0xffffffffull) {
        set_error("rs3d_filter_refine_noalloc: bad arguments");
        return RS3D_E_ARG;
    }
    if (candidate_count) *candidate_count = 0;
    if (inside_count) *inside_count = 0;
    if (n == 0) return RS3D_OK;
    // a grid is used only when it already exists: building one allocates
    const bool grid = ctx->opt_axis_grid && bvh->num_tris && bvh->agrid_state[axis] == 1 &&
                      bvh->agrid[axis] != nullptr;
    if (!grid && n > NOALLOC_MAX_WALK) {
        set_error("rs3d_filter_refine_noalloc: %llu points exceed the %llu its fixed scratch holds "
                  "without a prepared axis grid (rs3d_axis_grid_prepare lifts the limit)",
                  (unsigned long long)n, (unsigned long long)NOALLOC_MAX_WALK);
        return RS3D_E_CAPACITY;
    }
    unsigned char* sc = nullptr;
    RS3D_TRY(scratch_ptr(ctx, &sc));
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(sc + SC_ACC);
    k_zero_stripes<<<1, 64, 0, ctx->stream>>>(acc);
    RS3D_HIP(hipGetLastError());
    const float box[6] = {bvh->vmin[0], bvh->vmin[1], bvh->vmin[2],
                          bvh->vmax[0], bvh->vmax[1], bvh->vmax[2]};
    if (grid) {
        // one pass: filter, trace and count per point (same answers as the three phases)
        RS3D_TRY(axis_grid_filter_refine(ctx, const_cast<rs3d_bvh*>(bvh), axis, points, n, box, out_obj, acc));
        uint64_t h = 0;
        RS3D_TRY(read_stripes_sum(ctx, acc, &h));
        if (candidate_count) *candidate_count = h >> 32;
        if (inside_count) *inside_count = h & 0xffffffffull;
        return RS3D_OK;
    }
    RS3D_TRY(fill_i32(ctx, out_obj, n, -1));
    if (bvh->num_tris == 0) return RS3D_OK;
    uint8_t* flags = sc + SC_FLAGS;
    uint32_t* pos = reinterpret_cast<uint32_t*>(sc + SC_POS);
    uint32_t* cand = reinterpret_cast<uint32_t*>(sc + SC_CAND);
    uint32_t* total = reinterpret_cast<uint32_t*>(sc + SC_TOTAL);
    RS3D_TRY(rs3d_filter_box(ctx, box, points, n, axis, flags));
    RS3D_TRY(scan_u8_to_u32_exclusive_in(ctx, flags, pos, n, total,
                                         reinterpret_cast<uint32_t*>(sc + SC_SCAN),
                                         SC_SCAN_BYTES / sizeof(uint32_t)));
    RS3D_TRY(scatter_index(ctx, flags, pos, n, cand));
    uint64_t nc = 0;
    RS3D_TRY(read_u64(ctx, total, &nc));
    nc &= 0xffffffffull;  // the total is a u32 (the word after it is not part of it)
    if (candidate_count) *candidate_count = nc;
    if (nc == 0) return RS3D_OK;
    RS3D_TRY(trace_points_walk(ctx, bvh, points, nc, cand, axis, RS3D_OUT_OBJ_I32, out_obj));
    if (inside_count) {
        RS3D_TRY(count_nonneg(ctx, out_obj, n, acc));
        uint64_t h = 0;
        RS3D_TRY(read_stripes_sum(ctx, acc, &h));
        *inside_count = h;
    }
    return RS3D_OK;
}

}  // extern "C"
