// join.hip -- pair hash table, overlap join driver, selectivity estimator and the
// reference-named C entry points (query/src/cuda/mesh_query_deduplication.h,
// result_compaction.h, scan_utils.h).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "rs3d_internal.h"

namespace rs3d {
int edge_count_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e, int32_t* counts);
int edge_write_keys_internal(rs3d_ctx* ctx, const rs3d_bvh* t, const rs3d_edges* e,
                             const int64_t* offsets, int64_t bias, int swap, uint64_t* out);

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* __restrict__ p, uint64_t n,
                                                  unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(256) void k_slot_flags(const uint64_t* __restrict__ slots, uint64_t n,
                                                    uint8_t* __restrict__ flags) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = slots[i] != ~0ull ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_scatter_slots(const uint64_t* __restrict__ slots,
                                                       const uint8_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ pos, uint64_t n,
                                                       uint64_t max_out, uint64_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i] && pos[i] < max_out) out[pos[i]] = slots[i];
}

__global__ __launch_bounds__(256) void k_keys_to_pairs(const uint64_t* __restrict__ keys, uint64_t n,
                                                       rs3d_pair* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        rs3d_pair p;
        p.object_id_mesh1 = (int32_t)(keys[i] >> 32);
        p.object_id_mesh2 = (int32_t)(keys[i] & 0xffffffffull);
        out[i] = p;
    }
}

static unsigned fill_blocks(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return (unsigned)(b ? b : 1);
}

int ht_clear_internal(rs3d_ctx* ctx, uint64_t* slots, uint64_t size, uint64_t* failed) {
    k_fill_u64<<<fill_blocks(size), 256, 0, ctx->stream>>>(
        reinterpret_cast<unsigned long long*>(slots), size, ~0ull);
    if (failed)
        k_fill_u64<<<1, 64, 0, ctx->stream>>>(reinterpret_cast<unsigned long long*>(failed), 1, 0ull);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

// Scan-based compaction of occupied slots in slot order. Returns the occupied count in
// *count (may exceed max_out; only max_out keys are written).
int ht_compact_internal(rs3d_ctx* ctx, const uint64_t* slots, uint64_t size, uint64_t* out,
                        uint64_t max_out, uint64_t* count) {
    if (size > 0xffffffffull) {
        set_error("hash table compaction supports at most 2^32-1 slots");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->tmp_a.ensure(size));
    RS3D_TRY(ctx->tmp_b.ensure(size * sizeof(uint32_t)));
    RS3D_TRY(ctx->small.ensure(256));
    uint8_t* flags = ctx->tmp_a.as<uint8_t>();
    uint32_t* pos = ctx->tmp_b.as<uint32_t>();
    uint32_t* tot = ctx->small.as<uint32_t>() + 8;
    unsigned nb;
    RS3D_TRY(grid_blocks(size, 256, &nb));
    k_slot_flags<<<nb, 256, 0, ctx->stream>>>(slots, size, flags);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags, pos, size, tot));
    k_scatter_slots<<<nb, 256, 0, ctx->stream>>>(slots, flags, pos, size, max_out, out);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t h;
    memcpy(&h, ctx->pinned, sizeof h);
    *count = h;
    return RS3D_OK;
}

// ------------------------------------------------------------------------------------
// Selectivity estimate (query/src/cuda/estimated_overlap.cu:8-41): per-cell value in fp32
// exactly as the reference computes it; deterministic fixed-order double reduction.
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_estimate(const GridCellRec* __restrict__ A,
                                                  const GridCellRec* __restrict__ B, uint64_t n,
                                                  float cell_volume, float eps, float gamma,
                                                  double* __restrict__ partial) {
    __shared__ double red[256];
    double s = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const GridCellRec a = A[i], b = B[i];
        if (a.touch == 0 || b.touch == 0) continue;
        const float cs = a.avg + b.avg + eps;
        const float mink = cs * cs * cs;
        float prob = mink / cell_volume;
        const float ratio = sqrtf(a.ratio * b.ratio);
        prob *= powf(ratio, gamma);
        prob = fminf(prob, 1.0f);
        s += (double)((float)a.touch * (float)b.touch * prob);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += partial[i];
        *out = s;
    }
}

// Default context for the reference-named entry points (synchronous, current device).
static std::mutex g_default_mu;
static rs3d_ctx* g_default_ctx[64] = {nullptr};

rs3d_ctx* default_ctx() {
    int ndev = 0, dev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e == hipSuccess && ndev > 0) e = hipGetDevice(&dev);
    if (e != hipSuccess || ndev <= 0 || dev < 0 || dev >= 64) {
        set_error("no HIP device is available (%s, %d devices)", hipGetErrorString(e), ndev);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default_ctx[dev]) {
        rs3d_ctx* c = nullptr;
        if (rs3d_ctx_create(dev, &c) != RS3D_OK) return nullptr;
        g_default_ctx[dev] = c;
    }
    return g_default_ctx[dev];
}

// Prints the calling thread's last error and ends the process with EXIT_FAILURE, as the
// reference's CUDA_CHECK does (query/src/optix/OptixHelpers.h:15-21).
[[noreturn]] void fail_fast(const char* fn) {
    std::fprintf(stderr, "librs3d: %s failed: %s\n", fn, last_error());
    std::fflush(stderr);
    std::exit(EXIT_FAILURE);
}

rs3d_ctx* default_ctx_or_die(const char* fn) {
    rs3d_ctx* c = default_ctx();
    if (!c) {
        if (!*last_error()) set_error("no HIP device is available");
        fail_fast(fn);
    }
    return c;
}

// Moves count keys from src down to dst (dst < src, ranges may overlap) with one
// workgroup walking tiles upward: each tile is loaded whole before any lane stores it, and
// every later tile reads above everything stored so far.
__global__ __launch_bounds__(1024) void k_move_down(uint64_t* __restrict__ base, uint64_t dst,
                                                    uint64_t src, uint64_t count) {
    constexpr int PER = 4;
    constexpr uint64_t TILE = 1024 * PER;
    for (uint64_t t = 0; t < count; t += TILE) {
        uint64_t v[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint64_t i = t + (uint64_t)j * 1024 + threadIdx.x;
            if (i < count) v[j] = base[src + i];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const uint64_t i = t + (uint64_t)j * 1024 + threadIdx.x;
            if (i < count) base[dst + i] = v[j];
        }
        __syncthreads();
    }
}

// Device bytes the dedup may count on: RS3D_DEDUP_FREE_BYTES when set (a cap callers use to
// leave memory to others, and tests use to force the halving path), else the device's free
// memory, plus the scratch the context already holds.
static uint64_t dedup_budget(rs3d_ctx* ctx) {
    uint64_t free_b = 0;
    const char* env = std::getenv("RS3D_DEDUP_FREE_BYTES");
    if (env && *env) {
        free_b = std::strtoull(env, nullptr, 10);
    } else {
        size_t f = 0, t = 0;
        if (hipMemGetInfo(&f, &t) == hipSuccess) free_b = f;
    }
    return free_b + ctx->sort_keys.bytes + ctx->sort_hist.bytes + ctx->tmp_a.bytes + ctx->tmp_b.bytes +
           ctx->scan_tmp.bytes;
}

// Scratch of one sort + unique of n keys (the alternate key buffer, flags, positions, tile
// histograms and scan partials, with the buffers' growth margin), plus 512 MB of headroom
// as the reference keeps.
static uint64_t dedup_need(uint64_t n) { return n * 15 + n / 16 + (512ull << 20); }

static int dedup_inplace(rs3d_ctx* ctx, uint64_t* keys, uint64_t n, uint64_t* unique_count) {
    if (n <= 1) {
        *unique_count = n;
        return RS3D_OK;
    }
    if (n <= 0xffffffffull && dedup_need(n) <= dedup_budget(ctx)) {
        const int rc = rs3d_sort_unique_u64(ctx, keys, n, unique_count);
        if (rc != RS3D_E_NOMEM) return rc;
        (void)hipGetLastError();  // an allocation that failed: halve instead
    }
    const uint64_t half = n / 2;
    uint64_t u1 = 0, u2 = 0;
    RS3D_TRY(dedup_inplace(ctx, keys, half, &u1));
    RS3D_TRY(dedup_inplace(ctx, keys + half, n - half, &u2));
    if (u2 && u1 < half) {
        k_move_down<<<1, 1024, 0, ctx->stream>>>(keys, u1, half, u2);
        RS3D_HIP(hipGetLastError());
    }
    const uint64_t m = u1 + u2;
    if (m < n) return dedup_inplace(ctx, keys, m, unique_count);
    // no duplicates within either half: the merged sort must run at this size
    if (n > 0xffffffffull) {
        set_error("dedup of %llu distinct keys exceeds the sort's 2^32-1", (unsigned long long)n);
        return RS3D_E_CAPACITY;
    }
    return rs3d_sort_unique_u64(ctx, keys, n, unique_count);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_ht_clear(rs3d_ctx* ctx, rs3d_hash_table* ht) {
    if (!ctx || !ht || !ht->slots || ht->size == 0) {
        set_error("rs3d_ht_clear: bad arguments");
        return RS3D_E_ARG;
    }
    return ht_clear_internal(ctx, ht->slots, ht->size, ht->failed);
}

int rs3d_ht_compact(rs3d_ctx* ctx, const rs3d_hash_table* ht, uint64_t* keys_out, uint64_t max_out,
                    uint64_t* count) {
    if (!ctx || !ht || !ht->slots || ht->size == 0 || (max_out && !keys_out)) {
        set_error("rs3d_ht_compact: bad arguments");
        return RS3D_E_ARG;
    }
    uint64_t c = 0;
    RS3D_TRY(ht_compact_internal(ctx, ht->slots, ht->size, keys_out, max_out, &c));
    if (count) *count = c;
    if (c > max_out) {
        set_error("rs3d_ht_compact: %llu keys exceed the output capacity %llu",
                  (unsigned long long)c, (unsigned long long)max_out);
        return RS3D_E_CAPACITY;
    }
    return RS3D_OK;
}

int rs3d_overlap_join(rs3d_ctx* ctx, const rs3d_bvh* bvh1, const rs3d_edges* edges1,
                      const rs3d_bvh* bvh2, const rs3d_edges* edges2, int mode, uint64_t* keys_out,
                      uint64_t max_out, uint64_t* count) {
    if (!ctx || !bvh1 || !bvh2 || !edges1 || !edges2 || (max_out && !keys_out) || mode < 0 ||
        mode > 1) {
        set_error("rs3d_overlap_join: bad arguments");
        return RS3D_E_ARG;
    }
    if (count) *count = 0;
    const uint64_t n1 = edges1->num_edges, n2 = edges2->num_edges;
    uint64_t uniq = 0;
    uint64_t* result = nullptr;
    if (mode == 0) {
        // two-pass: count, scan, write raw keys, radix sort + unique
        RS3D_TRY(ctx->tmp_c.ensure((n1 + n2) * (sizeof(int32_t) + sizeof(int64_t)) + 1024));
        int32_t* c1 = ctx->tmp_c.as<int32_t>();
        int32_t* c2 = c1 + n1;
        int64_t* o1 = reinterpret_cast<int64_t*>(ctx->tmp_c.as<uint8_t>() +
                                                 (((n1 + n2) * sizeof(int32_t) + 255) & ~255ull));
        int64_t* o2 = o1 + n1;
        RS3D_TRY(edge_count_internal(ctx, bvh2, edges1, c1));
        RS3D_TRY(edge_count_internal(ctx, bvh1, edges2, c2));
        RS3D_TRY(ctx->small.ensure(256));
        int64_t* tots = ctx->small.as<int64_t>() + 4;
        RS3D_TRY(scan_i32_to_i64_exclusive(ctx, c1, o1, n1, tots));
        RS3D_TRY(scan_i32_to_i64_exclusive(ctx, c2, o2, n2, tots + 1));
        int64_t ht[2];
        RS3D_HIP(hipMemcpyAsync(ht, tots, sizeof ht, hipMemcpyDeviceToHost, ctx->stream));
        RS3D_HIP(hipStreamSynchronize(ctx->stream));
        const uint64_t total = (uint64_t)(ht[0] + ht[1]);
        if (total == 0) return RS3D_OK;
        static thread_local DevBuf raw;  // raw keys of this thread's joins
        RS3D_TRY(raw.ensure(total * sizeof(uint64_t)));
        uint64_t* keys = raw.as<uint64_t>();
        RS3D_TRY(edge_write_keys_internal(ctx, bvh2, edges1, o1, 0, 0, keys));
        RS3D_TRY(edge_write_keys_internal(ctx, bvh1, edges2, o2, ht[0], 1, keys));
        RS3D_TRY(rs3d_sort_unique_u64(ctx, keys, total, &uniq));
        result = keys;
    } else {
        // hash: on-the-fly dedup in an open-addressing table, compaction, radix sort
        static thread_local DevBuf slots_buf, keys_buf;
        uint64_t size = 1ull << 20;
        while (true) {
            RS3D_TRY(slots_buf.ensure(size * sizeof(uint64_t) + 256));
            rs3d_hash_table t;
            t.slots = slots_buf.as<uint64_t>();
            t.size = size;
            t.bitwise = 1;
            t.max_probes = 0;
            t.failed = reinterpret_cast<uint64_t*>(slots_buf.as<uint8_t>() + size * sizeof(uint64_t));
            RS3D_TRY(ht_clear_internal(ctx, t.slots, t.size, t.failed));
            RS3D_TRY(edge_hash_internal(ctx, bvh2, edges1, 0, &t, 0, 100));
            RS3D_TRY(edge_hash_internal(ctx, bvh1, edges2, 1, &t, 0, 100));
            uint64_t failed = 0;
            RS3D_TRY(read_u64(ctx, t.failed, &failed));
            uint64_t occ = 0;
            RS3D_TRY(keys_buf.ensure(size * sizeof(uint64_t)));
            if (failed == 0) {
                RS3D_TRY(ht_compact_internal(ctx, t.slots, t.size, keys_buf.as<uint64_t>(), size, &occ));
                if (occ * 2 <= size) {
                    RS3D_TRY(rs3d_sort_unique_u64(ctx, keys_buf.as<uint64_t>(), occ, &uniq));
                    result = keys_buf.as<uint64_t>();
                    break;
                }
            }
            // table too full (or an insert failed): grow and redo both directions
            size <<= 2;
            if (size > (1ull << 31)) {
                set_error("rs3d_overlap_join: hash table would exceed 2^31 slots");
                return RS3D_E_CAPACITY;
            }
        }
    }
    if (count) *count = uniq;
    if (uniq > max_out) {
        set_error("rs3d_overlap_join: %llu pairs exceed the output capacity %llu",
                  (unsigned long long)uniq, (unsigned long long)max_out);
        return RS3D_E_CAPACITY;
    }
    if (uniq)
        RS3D_HIP(hipMemcpyAsync(keys_out, result, uniq * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

int rs3d_estimate_raw(rs3d_ctx* ctx, const void* cells_a, const void* cells_b, uint64_t num_cells,
                      float cell_volume, float epsilon, float gamma, double* result) {
    if (!ctx || !result || (num_cells && (!cells_a || !cells_b))) {
        set_error("rs3d_estimate_raw: bad arguments");
        return RS3D_E_ARG;
    }
    const int nb = 1024;
    RS3D_TRY(ctx->tmp_a.ensure((nb + 1) * sizeof(double)));
    double* partial = ctx->tmp_a.as<double>();
    k_estimate<<<nb, 256, 0, ctx->stream>>>(static_cast<const GridCellRec*>(cells_a),
                                             static_cast<const GridCellRec*>(cells_b), num_cells,
                                             cell_volume, epsilon, gamma, partial);
    k_sum_partials<<<1, 64, 0, ctx->stream>>>(partial, nb, partial + nb);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, partial + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(result, ctx->pinned, sizeof(double));
    return RS3D_OK;
}

// ---- reference-named entry points (synchronous, like the reference) ----------------
// Failures are not turned into counts: like the reference's CUDA_CHECK
// (query/src/optix/OptixHelpers.h:15-29) they print the error and end the process with
// EXIT_FAILURE. The dedup first halves its input when the sort's scratch would not fit
// (query/src/cuda/mesh_query_deduplication.cu:56-90).

int count_hits_gpu(const rs3d_ray_result* d_results, int num_results) {
    if (num_results == 0) return 0;
    rs3d_ctx* ctx = default_ctx_or_die("count_hits_gpu");
    if (num_results < 0 || !d_results) {
        set_error("count_hits_gpu: bad arguments (num_results %d)", num_results);
        fail_fast("count_hits_gpu");
    }
    static thread_local DevBuf out;
    if (out.ensure((size_t)num_results * sizeof(rs3d_ray_result)) != RS3D_OK) fail_fast("count_hits_gpu");
    uint64_t c = 0;
    if (rs3d_compact_hits(ctx, d_results, (uint64_t)num_results, out.as<rs3d_ray_result>(), &c) != RS3D_OK)
        fail_fast("count_hits_gpu");
    return (int)c;
}

void compact_hits_gpu(const rs3d_ray_result* d_input, rs3d_ray_result* d_output, int num_results,
                      int* h_num_hits) {
    if (num_results == 0) {
        if (h_num_hits) *h_num_hits = 0;
        return;
    }
    rs3d_ctx* ctx = default_ctx_or_die("compact_hits_gpu");
    if (num_results < 0 || !d_input || !d_output) {
        set_error("compact_hits_gpu: bad arguments (num_results %d)", num_results);
        fail_fast("compact_hits_gpu");
    }
    uint64_t c = 0;
    if (rs3d_compact_hits(ctx, d_input, (uint64_t)num_results, d_output, &c) != RS3D_OK)
        fail_fast("compact_hits_gpu");
    if (h_num_hits) *h_num_hits = (int)c;
}

long long exclusive_scan_gpu(const int* d_input, long long* d_output, int num_elements) {
    if (num_elements == 0) return 0;
    rs3d_ctx* ctx = default_ctx_or_die("exclusive_scan_gpu");
    if (num_elements < 0 || !d_input || !d_output) {
        set_error("exclusive_scan_gpu: bad arguments (num_elements %d)", num_elements);
        fail_fast("exclusive_scan_gpu");
    }
    int64_t total = 0;
    if (rs3d_scan_i32_i64(ctx, d_input, reinterpret_cast<int64_t*>(d_output), (uint64_t)num_elements,
                          &total) != RS3D_OK)
        fail_fast("exclusive_scan_gpu");
    return (long long)total;
}

long long merge_and_deduplicate_pairs_gpu(const rs3d_pair* d_results1, long long num_results1,
                                          const rs3d_pair* d_results2, long long num_results2,
                                          rs3d_pair* d_merged_output) {
    const long long total = num_results1 + num_results2;
    if (num_results1 == 0 && num_results2 == 0) return 0;
    rs3d_ctx* ctx = default_ctx_or_die("merge_and_deduplicate_pairs_gpu");
    // a null input array means its results already lie in place in d_merged_output
    if (num_results1 < 0 || num_results2 < 0 || !d_merged_output) {
        set_error("merge_and_deduplicate_pairs_gpu: bad arguments (%lld, %lld)", num_results1, num_results2);
        fail_fast("merge_and_deduplicate_pairs_gpu");
    }
    if (d_results1 && num_results1 > 0 && d_results1 != d_merged_output &&
        hipMemcpyAsync(d_merged_output, d_results1, num_results1 * sizeof(rs3d_pair), hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess) {
        set_error("merge_and_deduplicate_pairs_gpu: copy of the first result array failed");
        fail_fast("merge_and_deduplicate_pairs_gpu");
    }
    if (d_results2 && num_results2 > 0 && d_results2 != d_merged_output + num_results1 &&
        hipMemcpyAsync(d_merged_output + num_results1, d_results2, num_results2 * sizeof(rs3d_pair),
                       hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess) {
        set_error("merge_and_deduplicate_pairs_gpu: copy of the second result array failed");
        fail_fast("merge_and_deduplicate_pairs_gpu");
    }
    // The pair struct read as a little-endian u64 is (mesh2 << 32 | mesh1): sorting it
    // reproduces the reference's (b, a) order.
    uint64_t u = 0;
    if (dedup_inplace(ctx, reinterpret_cast<uint64_t*>(d_merged_output), (uint64_t)total, &u) != RS3D_OK)
        fail_fast("merge_and_deduplicate_pairs_gpu");
    return (long long)u;
}

int compact_hash_table_pairs(const unsigned long long* d_hash_table, unsigned long long table_size,
                             rs3d_pair* d_output, int max_output_size) {
    if (table_size == 0 || max_output_size == 0) return 0;
    rs3d_ctx* ctx = default_ctx_or_die("compact_hash_table_pairs");
    if (!d_hash_table || max_output_size < 0 || !d_output) {
        set_error("compact_hash_table_pairs: bad arguments (max_output_size %d)", max_output_size);
        fail_fast("compact_hash_table_pairs");
    }
    static thread_local DevBuf keys;
    if (keys.ensure((size_t)max_output_size * sizeof(uint64_t)) != RS3D_OK) fail_fast("compact_hash_table_pairs");
    uint64_t c = 0;
    if (ht_compact_internal(ctx, reinterpret_cast<const uint64_t*>(d_hash_table), table_size,
                            keys.as<uint64_t>(), (uint64_t)max_output_size, &c) != RS3D_OK)
        fail_fast("compact_hash_table_pairs");
    const uint64_t w = c < (uint64_t)max_output_size ? c : (uint64_t)max_output_size;
    if (w) {
        unsigned nb;
        if (grid_blocks(w, 256, &nb) != RS3D_OK) fail_fast("compact_hash_table_pairs");
        k_keys_to_pairs<<<nb, 256, 0, ctx->stream>>>(keys.as<uint64_t>(), w, d_output);
        const hipError_t e = hipGetLastError();
        const hipError_t e2 = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess || e2 != hipSuccess) {
            set_error("compact_hash_table_pairs: %s", hipGetErrorString(e != hipSuccess ? e : e2));
            fail_fast("compact_hash_table_pairs");
        }
    }
    return (int)w;
}

}  // extern "C"
