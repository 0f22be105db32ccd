// prims.hip -- device-wide primitives: exclusive scans, stable LSD radix sort, unique,
// stream compaction. All are HBM-bound; tiles are staged through LDS so global loads
// and stores stay coalesced.
#include <cstdarg>
#include <cstring>

#include "rs3d_internal.h"

namespace rs3d {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

const char* last_error() { return g_last_error.c_str(); }

int DevBuf::ensure(size_t need) {
    if (need <= bytes) return RS3D_OK;
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    size_t alloc = need < 4096 ? 4096 : need + need / 8;
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(%zu) failed: %s", alloc, hipGetErrorString(e));
        return RS3D_E_NOMEM;
    }
    bytes = alloc;
    return RS3D_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int grid_blocks(uint64_t n, int block, unsigned* out) {
    uint64_t b = (n + block - 1) / block;
    if (b > 0x7fffffffull) {
        set_error("launch of %llu blocks exceeds the grid limit", (unsigned long long)b);
        return RS3D_E_ARG;
    }
    *out = (unsigned)(b ? b : 1);
    return RS3D_OK;
}

int prof_collect(rs3d_ctx* ctx) {
    if (!ctx->prof_pending) return RS3D_OK;
    RS3D_HIP(hipEventSynchronize(ctx->ev_stop));
    float ms = 0.0f;
    RS3D_HIP(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    ctx->prof_ms += ms;
    ctx->prof_launches++;
    ctx->prof_pending = false;
    return RS3D_OK;
}

__global__ void k_acc_init(unsigned long long* __restrict__ acc, unsigned long long init) {
    const int i = threadIdx.x;
    if (i < ACC_STRIPES) acc[i * ACC_PITCH] = init;
}

int acc_slot(rs3d_ctx* ctx, int k, uint64_t init, unsigned long long** acc) {
    static_assert(ACC_STRIPES == 64 && ACC_PITCH == 16, "acc_stripe assumes 64 stripes 128 B apart");
    if (k < 0 || k >= ACC_SLOTS) {
        set_error("accumulator slot %d out of range", k);
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->accs.ensure((size_t)ACC_SLOTS * ACC_STRIPES * ACC_PITCH * sizeof(unsigned long long)));
    *acc = ctx->accs.as<unsigned long long>() + (size_t)k * ACC_STRIPES * ACC_PITCH;
    k_acc_init<<<1, 64, 0, ctx->stream>>>(*acc, init);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int acc_read(rs3d_ctx* ctx, const unsigned long long* acc, uint64_t (&stripes)[ACC_STRIPES]) {
    // one contiguous copy of the 8 KB block (a strided copy is slower than the extra bytes)
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, acc, ACC_STRIPES * ACC_PITCH * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < ACC_STRIPES; ++i) stripes[i] = ctx->pinned[i * ACC_PITCH];
    return RS3D_OK;
}

int acc_read_sum(rs3d_ctx* ctx, const unsigned long long* acc, uint64_t* sum) {
    uint64_t v[ACC_STRIPES];
    RS3D_TRY(acc_read(ctx, acc, v));
    uint64_t t = 0;
    for (int i = 0; i < ACC_STRIPES; ++i) t += v[i];
    *sum = t;
    return RS3D_OK;
}

int read_u64(rs3d_ctx* ctx, const void* dev, uint64_t* host) {
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, dev, sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    *host = ctx->pinned[0];
    return RS3D_OK;
}

// ------------------------------------------------------------------------------------
// Exclusive scan: reduce-then-scan over tiles of SCAN_TILE elements.
// ------------------------------------------------------------------------------------
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix and
// the block total in `total`.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T& total) {
    __shared__ T wsum[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = wave_inclusive_scan(v, lane);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    T wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        T s = wsum[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + inc - v;
}

template <typename In, typename Out>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const In* __restrict__ in, uint64_t n,
                                                            Out* __restrict__ partial) {
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    Out s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += (Out)in[i];
    }
    Out total;
    (void)block_exclusive_scan<Out>(s, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

template <typename In, typename Out>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const In* in, uint64_t n, Out* out,
                                                           const Out* __restrict__ offsets,
                                                           Out* __restrict__ total_out) {
    __shared__ Out tile[SCAN_TILE];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        tile[j * SCAN_BLOCK + threadIdx.x] = i < n ? (Out)in[i] : (Out)0;
    }
    __syncthreads();
    Out v[SCAN_ITEMS];
    Out s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = tile[threadIdx.x * SCAN_ITEMS + j];
        s += v[j];
    }
    Out total;
    Out pre = block_exclusive_scan<Out>(s, total);
    const Out off = offsets ? offsets[blockIdx.x] : (Out)0;
    pre += off;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        tile[threadIdx.x * SCAN_ITEMS + j] = pre;
        pre += v[j];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        uint64_t i = base + (uint64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) out[i] = tile[j * SCAN_BLOCK + threadIdx.x];
    }
    if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = off + total;
}

// Recursive exclusive scan; `tmp` holds partials of every level. total_dev (optional)
// receives the sum of all inputs.
template <typename In, typename Out>
static int scan_impl(rs3d_ctx* ctx, const In* in, Out* out, uint64_t n, Out* total_dev, Out* tmp,
                     uint64_t tmp_elems) {
    if (n == 0) {
        if (total_dev) RS3D_HIP(hipMemsetAsync(total_dev, 0, sizeof(Out), ctx->stream));
        return RS3D_OK;
    }
    unsigned nb;
    RS3D_TRY(grid_blocks(n, SCAN_TILE, &nb));
    if (nb == 1) {
        k_scan_apply<In, Out><<<1, SCAN_BLOCK, 0, ctx->stream>>>(in, n, out, nullptr, total_dev);
        RS3D_HIP(hipGetLastError());
        return RS3D_OK;
    }
    if (tmp_elems < nb) {
        set_error("scan scratch too small");
        return RS3D_E_ARG;
    }
    Out* partial = tmp;
    k_scan_reduce<In, Out><<<nb, SCAN_BLOCK, 0, ctx->stream>>>(in, n, partial);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY((scan_impl<Out, Out>(ctx, partial, partial, nb, (Out*)nullptr, tmp + nb, tmp_elems - nb)));
    k_scan_apply<In, Out><<<nb, SCAN_BLOCK, 0, ctx->stream>>>(in, n, out, partial, total_dev);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

static uint64_t scan_tmp_elems(uint64_t n) {
    uint64_t tot = 0;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        tot += n;
    }
    return tot + 16;
}

// In-place use (in == out) is safe: each tile is read fully into LDS before it is written.
int scan_u32_exclusive(rs3d_ctx* ctx, const uint32_t* in, uint32_t* out, uint64_t n) {
    uint64_t te = scan_tmp_elems(n);
    RS3D_TRY(ctx->scan_tmp.ensure(te * sizeof(uint32_t)));
    return scan_impl<uint32_t, uint32_t>(ctx, in, out, n, nullptr, ctx->scan_tmp.as<uint32_t>(), te);
}

int scan_i32_to_i64_exclusive(rs3d_ctx* ctx, const int32_t* in, int64_t* out, uint64_t n,
                              int64_t* total_dev) {
    uint64_t te = scan_tmp_elems(n);
    RS3D_TRY(ctx->scan_tmp.ensure(te * sizeof(int64_t)));
    return scan_impl<int32_t, int64_t>(ctx, in, out, n, total_dev, ctx->scan_tmp.as<int64_t>(), te);
}

int scan_u8_to_u32_exclusive(rs3d_ctx* ctx, const uint8_t* in, uint32_t* out, uint64_t n,
                             uint32_t* total_dev) {
    uint64_t te = scan_tmp_elems(n);
    RS3D_TRY(ctx->scan_tmp.ensure(te * sizeof(uint32_t)));
    return scan_impl<uint8_t, uint32_t>(ctx, in, out, n, total_dev, ctx->scan_tmp.as<uint32_t>(), te);
}

uint64_t scan_scratch_elems(uint64_t n) { return scan_tmp_elems(n); }

int scan_u8_to_u32_exclusive_in(rs3d_ctx* ctx, const uint8_t* in, uint32_t* out, uint64_t n,
                                uint32_t* total_dev, uint32_t* tmp, uint64_t tmp_elems) {
    return scan_impl<uint8_t, uint32_t>(ctx, in, out, n, total_dev, tmp, tmp_elems);
}

// ------------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits. Per pass: (1) per-tile digit histograms, counted
// with wave64 ballot matching (one LDS add per distinct digit per wave, so runs of equal
// digits do not serialise on one counter) and stored tile-major ([tile][digit], coalesced);
// (2) the tile-major table becomes each tile's global start per digit (digit-major order:
// all of digit d before digit d + 1, tiles in order inside a digit) by a column scan;
// (3) stable scatter: each tile ranks its keys with ballot matching, reorders them in LDS
// and stores each digit's run with consecutive lanes.
// ------------------------------------------------------------------------------------
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_WAVES = RS_BLOCK / 64;

template <typename K>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const K* __restrict__ keys, uint64_t n,
                                                      int shift, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    K kv[RS_ITEMS];  // all loads issued before the ranking
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * RS_BLOCK + threadIdx.x;
        kv[j] = i < n ? keys[i] : (K)0;
    }
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint64_t i = base + (uint64_t)j * RS_BLOCK + threadIdx.x;
        const bool valid = i < n;
        const uint32_t d = (uint32_t)(kv[j] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        if (valid && (m & lt_mask) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    hist[(uint64_t)blockIdx.x * 256 + threadIdx.x] = h[threadIdx.x];
}

// Column scan of the tile-major histogram (nb rows of 256 digit counts) into global
// starts: per chunk of RS_CHUNK tiles the digit sums are written digit-major
// (csum[d * nchunks + c]); one flat exclusive scan of that array gives, for every (digit,
// chunk), all counts of smaller digits plus this digit's counts in earlier chunks; each
// chunk then rewrites its rows as running starts. Rows are read with independent loads.
constexpr int RS_CHUNK = 16;
static_assert(RS_SEG_KEYS == (uint64_t)RS_CHUNK * RS_TILE, "segment unit = one column-scan chunk");

// Position of (digit d, chunk c) in the flat column-sum array: digit-major over all
// chunks, or, for a segmented sort (seg[c] = {first chunk of c's segment, chunks in it}),
// segment-major, then digit, then chunk, so each segment sorts on its own.
__device__ __forceinline__ uint64_t csum_index(uint32_t d, uint32_t c, uint32_t nchunks,
                                               const uint2* __restrict__ seg) {
    if (!seg) return (uint64_t)d * nchunks + c;
    const uint2 g = seg[c];
    return (uint64_t)g.x * 256 + (uint64_t)d * g.y + (c - g.x);
}

__global__ __launch_bounds__(256) void k_rs_colsum(const uint32_t* __restrict__ hist, uint32_t nb,
                                                   uint32_t nchunks, const uint2* __restrict__ seg,
                                                   uint32_t* __restrict__ csum) {
    const uint32_t r0 = blockIdx.x * RS_CHUNK;
    uint32_t v[RS_CHUNK];
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k)
        v[k] = r0 + k < nb ? hist[(uint64_t)(r0 + k) * 256 + threadIdx.x] : 0u;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k) s += v[k];
    csum[csum_index(threadIdx.x, blockIdx.x, nchunks, seg)] = s;
}

__global__ __launch_bounds__(256) void k_rs_colapply(uint32_t* __restrict__ hist, uint32_t nb,
                                                     uint32_t nchunks, const uint2* __restrict__ seg,
                                                     const uint32_t* __restrict__ csum) {
    const uint32_t r0 = blockIdx.x * RS_CHUNK;
    uint32_t v[RS_CHUNK];
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k)
        v[k] = r0 + k < nb ? hist[(uint64_t)(r0 + k) * 256 + threadIdx.x] : 0u;
    uint32_t run = csum[csum_index(threadIdx.x, blockIdx.x, nchunks, seg)];
#pragma unroll
    for (int k = 0; k < RS_CHUNK; ++k) {
        if (r0 + k < nb) hist[(uint64_t)(r0 + k) * 256 + threadIdx.x] = run;
        run += v[k];
    }
}

// Tile-major counts -> tile-major global starts (in place). seg: see csum_index.
static int rs_offsets(rs3d_ctx* ctx, uint32_t* hist, uint32_t nb, const uint2* seg = nullptr) {
    const uint32_t nchunks = (nb + RS_CHUNK - 1) / RS_CHUNK;
    RS3D_TRY(ctx->sort_csum.ensure((size_t)nchunks * 256 * sizeof(uint32_t)));
    uint32_t* csum = ctx->sort_csum.as<uint32_t>();
    k_rs_colsum<<<nchunks, 256, 0, ctx->stream>>>(hist, nb, nchunks, seg, csum);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(scan_u32_exclusive(ctx, csum, csum, (uint64_t)nchunks * 256));
    k_rs_colapply<<<nchunks, 256, 0, ctx->stream>>>(hist, nb, nchunks, seg, csum);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

// Scatter of one LSD pass. The tile is first reordered by digit inside LDS (stable:
// groups of 64 keys in input order, ballot rank inside a group), then written out so that
// consecutive threads store consecutive addresses of one digit's run.
template <typename K, bool HAS_VAL>
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const K* __restrict__ kin,
                                                         const uint32_t* __restrict__ vin,
                                                         K* __restrict__ kout,
                                                         uint32_t* __restrict__ vout, uint64_t n,
                                                         int shift,
                                                         const uint32_t* __restrict__ hist) {
    constexpr int G = RS_ITEMS * RS_WAVES;  // groups of 64 keys per tile, in input order
    __shared__ uint16_t gcnt[G][256];
    __shared__ uint32_t dstart[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t wsum[RS_WAVES];
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[HAS_VAL ? RS_TILE : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    const uint32_t count = (uint32_t)((n - base) < (uint64_t)RS_TILE ? (n - base) : (uint64_t)RS_TILE);
    for (int i = tid; i < G * 256; i += RS_BLOCK) (&gcnt[0][0])[i] = 0;
    gbase[tid] = hist[(uint64_t)blockIdx.x * 256 + tid];
    __syncthreads();
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    K key[RS_ITEMS];
    uint32_t val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint32_t li = (uint32_t)(j * RS_BLOCK + tid);
        const bool valid = li < count;
        key[j] = valid ? kin[base + li] : (K)0;
        val[j] = (HAS_VAL && valid) ? vin[base + li] : 0u;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        rank[j] = (uint32_t)__popcll(m & lt_mask);
        if (valid && rank[j] == 0) gcnt[j * RS_WAVES + wave][d] = (uint16_t)__popcll(m);
    }
    __syncthreads();
    // thread tid owns digit tid: exclusive prefix over the groups, then over the digits
    uint32_t tot = 0;
    for (int g = 0; g < G; ++g) {
        const uint32_t c = gcnt[g][tid];
        gcnt[g][tid] = (uint16_t)tot;
        tot += c;
    }
    uint32_t inc = tot;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t woff = 0;
    for (int w = 0; w < wave; ++w) woff += wsum[w];
    dstart[tid] = woff + inc - tot;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const uint32_t li = (uint32_t)(j * RS_BLOCK + tid);
        if (li < count) {
            const uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
            const uint32_t pos = dstart[d] + gcnt[j * RS_WAVES + wave][d] + rank[j];
            sk[pos] = key[j];
            if (HAS_VAL) sv[pos] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < count; i += RS_BLOCK) {
        const K k = sk[i];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = gbase[d] + i - dstart[d];
        kout[pos] = k;
        if (HAS_VAL) vout[pos] = sv[i];
    }
}

// Sorts (keys, vals) over bits [begin_bit, end_bit). With out_k/out_v the result is left in
// whichever of the two buffer pairs the last pass wrote and returned there; without them
// it is copied back into keys/vals when that is the alternate pair.
template <typename K, bool HAS_VAL>
static int radix_sort_impl(rs3d_ctx* ctx, K* keys, uint32_t* vals, uint64_t n, int begin_bit,
                           int end_bit, K* alt_keys, uint32_t* alt_vals, K** out_k = nullptr,
                           uint32_t** out_v = nullptr, const uint2* seg = nullptr) {
    if (out_k) *out_k = keys;
    if (out_v) *out_v = vals;
    if (n <= 1) return RS3D_OK;
    if (n > 0xffffffffull) {
        set_error("radix sort supports at most 2^32-1 keys");
        return RS3D_E_ARG;
    }
    unsigned nb;
    RS3D_TRY(grid_blocks(n, RS_TILE, &nb));
    RS3D_TRY(ctx->sort_hist.ensure((size_t)nb * 256 * sizeof(uint32_t)));
    uint32_t* hist = ctx->sort_hist.as<uint32_t>();
    K* src_k = keys;
    K* dst_k = alt_keys;
    uint32_t* src_v = vals;
    uint32_t* dst_v = alt_vals;
    int passes = 0;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        k_rs_hist<K><<<nb, RS_BLOCK, 0, ctx->stream>>>(src_k, n, shift, hist);
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(rs_offsets(ctx, hist, nb, seg));
        k_rs_scatter<K, HAS_VAL><<<nb, RS_BLOCK, 0, ctx->stream>>>(src_k, src_v, dst_k, dst_v, n,
                                                                   shift, hist);
        RS3D_HIP(hipGetLastError());
        std::swap(src_k, dst_k);
        std::swap(src_v, dst_v);
        ++passes;
    }
    if (out_k) {
        *out_k = src_k;
        if (out_v) *out_v = src_v;
        return RS3D_OK;
    }
    if (passes & 1) {
        RS3D_HIP(hipMemcpyAsync(keys, src_k, n * sizeof(K), hipMemcpyDeviceToDevice, ctx->stream));
        if (HAS_VAL)
            RS3D_HIP(hipMemcpyAsync(vals, src_v, n * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                    ctx->stream));
    }
    return RS3D_OK;
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Fill with a kernel rather than hipMemsetAsync, so the call can sit in a captured graph.
int fill_u32(rs3d_ctx* ctx, void* p, uint64_t n, uint32_t v) {
    if (n == 0) return RS3D_OK;
    const unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
    k_fill_u32<<<nb, 256, 0, ctx->stream>>>(static_cast<uint32_t*>(p), n, v);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int radix_sort_u32_pairs(rs3d_ctx* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, int end_bit) {
    if (n <= 1) return RS3D_OK;
    RS3D_TRY(ctx->sort_keys.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->sort_vals.ensure(n * sizeof(uint32_t)));
    if (end_bit < 1) end_bit = 1;
    if (end_bit > 32) end_bit = 32;
    return radix_sort_impl<uint32_t, true>(ctx, keys, vals, n, 0, end_bit,
                                           ctx->sort_keys.as<uint32_t>(),
                                           ctx->sort_vals.as<uint32_t>());
}

int radix_sort_u32_pairs_to(rs3d_ctx* ctx, uint32_t* keys, uint32_t* vals, uint64_t n, int end_bit,
                            uint32_t** out_keys, uint32_t** out_vals, const uint2* seg) {
    *out_keys = keys;
    *out_vals = vals;
    if (n <= 1) return RS3D_OK;
    RS3D_TRY(ctx->sort_keys.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->sort_vals.ensure(n * sizeof(uint32_t)));
    if (end_bit < 1) end_bit = 1;
    if (end_bit > 32) end_bit = 32;
    return radix_sort_impl<uint32_t, true>(ctx, keys, vals, n, 0, end_bit,
                                           ctx->sort_keys.as<uint32_t>(),
                                           ctx->sort_vals.as<uint32_t>(), out_keys, out_vals, seg);
}

// Bitwise OR / AND of all keys, to skip digit positions where every key agrees.
__global__ __launch_bounds__(256) void k_or_and_u64(const uint64_t* __restrict__ keys, uint64_t n,
                                                    unsigned long long* __restrict__ acc_or,
                                                    unsigned long long* __restrict__ acc_and) {
    uint64_t o = 0, a = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t k = keys[i];
        o |= k;
        a &= k;
    }
    for (int off = 32; off > 0; off >>= 1) {
        o |= __shfl_xor(o, off, 64);
        a &= __shfl_xor(a, off, 64);
    }
    __shared__ uint64_t po[4], pa[4];
    if ((threadIdx.x & 63) == 0) {
        po[threadIdx.x >> 6] = o;
        pa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicOr(acc_stripe(acc_or), (unsigned long long)(po[0] | po[1] | po[2] | po[3]));
        atomicAnd(acc_stripe(acc_and), (unsigned long long)(pa[0] & pa[1] & pa[2] & pa[3]));
    }
}

// Stable sort of u64 keys (with u32 values when vals is given), in place. Byte positions
// where every key agrees are skipped.
template <bool HAS_VAL>
static int radix_sort_u64_impl(rs3d_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n) {
    if (n <= 1) return RS3D_OK;
    if (n > 0xffffffffull) {
        set_error("radix sort supports at most 2^32-1 keys");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->sort_keys.ensure(n * sizeof(uint64_t)));
    if (HAS_VAL) RS3D_TRY(ctx->sort_vals.ensure(n * sizeof(uint32_t)));
    // accumulator slots 6 and 7 belong to this sort (callers may hold the others)
    unsigned long long *acc_or, *acc_and;
    RS3D_TRY(acc_slot(ctx, 6, 0ull, &acc_or));
    RS3D_TRY(acc_slot(ctx, 7, ~0ull, &acc_and));
    unsigned nb = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    k_or_and_u64<<<nb, 256, 0, ctx->stream>>>(keys, n, acc_or, acc_and);
    RS3D_HIP(hipGetLastError());
    uint64_t so[ACC_STRIPES], sa[ACC_STRIPES];
    RS3D_TRY(acc_read(ctx, acc_or, so));
    RS3D_TRY(acc_read(ctx, acc_and, sa));
    uint64_t h[2] = {0ull, ~0ull};
    for (int i = 0; i < ACC_STRIPES; ++i) {
        h[0] |= so[i];
        h[1] &= sa[i];
    }
    const uint64_t varying = h[0] & ~h[1];
    // Sort each varying byte, least significant first. Passes over constant bytes are
    // skipped; the result is identical because those digits are equal for all keys.
    uint64_t* src = keys;
    uint64_t* dst = ctx->sort_keys.as<uint64_t>();
    uint32_t* vsrc = vals;
    uint32_t* vdst = HAS_VAL ? ctx->sort_vals.as<uint32_t>() : nullptr;
    unsigned nbt;
    RS3D_TRY(grid_blocks(n, RS_TILE, &nbt));
    RS3D_TRY(ctx->sort_hist.ensure((size_t)nbt * 256 * sizeof(uint32_t)));
    uint32_t* hist = ctx->sort_hist.as<uint32_t>();
    for (int shift = 0; shift < 64; shift += 8) {
        if (((varying >> shift) & 0xffull) == 0) continue;
        k_rs_hist<uint64_t><<<nbt, RS_BLOCK, 0, ctx->stream>>>(src, n, shift, hist);
        RS3D_HIP(hipGetLastError());
        RS3D_TRY(rs_offsets(ctx, hist, nbt));
        k_rs_scatter<uint64_t, HAS_VAL><<<nbt, RS_BLOCK, 0, ctx->stream>>>(src, vsrc, dst, vdst, n,
                                                                           shift, hist);
        RS3D_HIP(hipGetLastError());
        std::swap(src, dst);
        std::swap(vsrc, vdst);
    }
    if (src != keys) {
        RS3D_HIP(hipMemcpyAsync(keys, src, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, ctx->stream));
        if (HAS_VAL)
            RS3D_HIP(hipMemcpyAsync(vals, vsrc, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    }
    return RS3D_OK;
}

int radix_sort_u64(rs3d_ctx* ctx, uint64_t* keys, uint64_t n) {
    return radix_sort_u64_impl<false>(ctx, keys, nullptr, n);
}

int radix_sort_u64_pairs(rs3d_ctx* ctx, uint64_t* keys, uint32_t* vals, uint64_t n) {
    return radix_sort_u64_impl<true>(ctx, keys, vals, n);
}

// ------------------------------------------------------------------------------------
// Unique of a sorted u64 array and flag-driven compactions (scan-based, stable).
// ------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_unique_flags(const uint64_t* __restrict__ in, uint64_t n,
                                                      uint8_t* __restrict__ flags) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || in[i] != in[i - 1]) ? 1 : 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scatter_flagged(const T* __restrict__ in,
                                                         const uint8_t* __restrict__ flags,
                                                         const uint32_t* __restrict__ pos,
                                                         uint64_t n, T* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = in[i];
}

__global__ __launch_bounds__(256) void k_scatter_index(const uint8_t* __restrict__ flags,
                                                       const uint32_t* __restrict__ pos, uint64_t n,
                                                       uint32_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flags[i]) out[pos[i]] = (uint32_t)i;
}

int scatter_index(rs3d_ctx* ctx, const uint8_t* flags, const uint32_t* pos, uint64_t n, uint32_t* out) {
    if (n == 0) return RS3D_OK;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_scatter_index<<<nb, 256, 0, ctx->stream>>>(flags, pos, n, out);
    RS3D_HIP(hipGetLastError());
    return RS3D_OK;
}

int unique_u64(rs3d_ctx* ctx, const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* count_host) {
    if (n == 0) {
        *count_host = 0;
        return RS3D_OK;
    }
    if (n > 0xffffffffull) {
        set_error("unique supports at most 2^32-1 keys");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->tmp_a.ensure(n));
    RS3D_TRY(ctx->tmp_b.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->small.ensure(64));
    uint8_t* flags = ctx->tmp_a.as<uint8_t>();
    uint32_t* pos = ctx->tmp_b.as<uint32_t>();
    uint32_t* tot = ctx->small.as<uint32_t>() + 4;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_unique_flags<<<nb, 256, 0, ctx->stream>>>(in, n, flags);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags, pos, n, tot));
    k_scatter_flagged<uint64_t><<<nb, 256, 0, ctx->stream>>>(in, flags, pos, n, out);
    RS3D_HIP(hipGetLastError());
    uint32_t h = 0;
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(&h, ctx->pinned, sizeof h);
    *count_host = h;
    return RS3D_OK;
}

__global__ __launch_bounds__(256) void k_hit_flags(const rs3d_ray_result* __restrict__ in, uint64_t n,
                                                   uint8_t* __restrict__ flags) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = in[i].polygon_index != -1 ? 1 : 0;
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

const char* rs3d_last_error(void) { return rs3d::last_error(); }

int rs3d_ctx_create(int device, rs3d_ctx** out) {
    if (!out) {
        set_error("rs3d_ctx_create: out is null");
        return RS3D_E_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    RS3D_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("rs3d_ctx_create: device %d not present (%d devices)", device, ndev);
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipSetDevice(device));
    rs3d_ctx* c = new rs3d_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        set_error("hipStreamCreate failed: %s", hipGetErrorString(e));
        return RS3D_E_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    {
        // a pool of its own for build temporaries: freeing them returns memory to the pool
        // (kept up to CTX_TMP_POOL_KEEP) instead of unmapping it, which took ~2 ms per grid
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = device;
        if (hipMemPoolCreate(&c->tmp_pool, &props) == hipSuccess) {
            uint64_t keep = CTX_TMP_POOL_KEEP;
            if (hipMemPoolSetAttribute(c->tmp_pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
                (void)hipMemPoolDestroy(c->tmp_pool);
                c->tmp_pool = nullptr;
            }
        } else {
            c->tmp_pool = nullptr;
        }
        (void)hipGetLastError();
    }
    e = hipHostMalloc((void**)&c->pinned, CTX_PINNED_BYTES, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        set_error("hipHostMalloc failed: %s", hipGetErrorString(e));
        return RS3D_E_NOMEM;
    }
    // the device address of rs3d_filter_refine_noalloc's fixed scratch, looked up once per
    // device here so that the entry point itself never makes the runtime materialise it
    const int rc = noalloc_init(c);
    if (rc != RS3D_OK) {
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return RS3D_OK;
}

int rs3d_ctx_destroy(rs3d_ctx* ctx) {
    if (!ctx) return RS3D_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->sort_keys.release();
    ctx->sort_vals.release();
    ctx->sort_hist.release();
    ctx->sort_csum.release();
    ctx->scan_tmp.release();
    ctx->tmp_a.release();
    ctx->tmp_b.release();
    ctx->tmp_c.release();
    ctx->small.release();
    ctx->accs.release();
    ctx->edge_sel.release();
    ctx->fr_acc.release();
    if (ctx->fr_total) (void)hipHostFree(ctx->fr_total);
    ctx->ray_keys.release();
    ctx->ray_vals.release();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->tmp_pool) (void)hipMemPoolDestroy(ctx->tmp_pool);
    delete ctx;
    return RS3D_OK;
}

int rs3d_ctx_sync(rs3d_ctx* ctx) {
    if (!ctx) {
        set_error("null context");
        return RS3D_E_ARG;
    }
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    return RS3D_OK;
}

void* rs3d_ctx_stream(rs3d_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int rs3d_ctx_set_option(rs3d_ctx* ctx, int option, int64_t value) {
    if (!ctx) {
        set_error("null context");
        return RS3D_E_ARG;
    }
    if (option == RS3D_OPT_RAY_SORT_MIN) {
        ctx->opt_ray_sort_min = value;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_RAYS_PER_WAVE) {
        if (value != 0 && (value < 64 || value > (1 << 20))) {
            set_error("RS3D_OPT_RAYS_PER_WAVE must be 0 or in [64, 2^20], got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_rays_per_wave = value;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_GRID_RAYS_PER_WAVE) {
        if (value < 64 || value > (1 << 20)) {
            set_error("RS3D_OPT_GRID_RAYS_PER_WAVE must be in [64, 2^20], got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_grid_rays_per_wave = value;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_SEG_OCCUPANCY) {
        if (value != 0 && value != 1) {
            set_error("RS3D_OPT_SEG_OCCUPANCY must be 0 or 1, got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_seg_occupancy = value;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_SEG_CELL_PCT) {
        if (value < 25 || value > 10000) {
            set_error("RS3D_OPT_SEG_CELL_PCT must be in [25, 10000], got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_seg_cell_pct = value;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_GRID_CELL_PCT) {
        if (value < 10 || value > 1000) {
            set_error("RS3D_OPT_GRID_CELL_PCT must be in [10, 1000], got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_grid_cell_pct = value;
        ctx->opt_grid_cell_pct_set = true;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_AXIS_GRID) {
        if (value != 0 && value != 1) {
            set_error("RS3D_OPT_AXIS_GRID must be 0 or 1, got %lld", (long long)value);
            return RS3D_E_ARG;
        }
        ctx->opt_axis_grid = value;
        return RS3D_OK;
    }
    set_error("rs3d_ctx_set_option: unknown option %d", option);
    return RS3D_E_ARG;
}

int rs3d_ctx_get_option(const rs3d_ctx* ctx, int option, int64_t* value) {
    if (!ctx || !value) {
        set_error("rs3d_ctx_get_option: bad arguments");
        return RS3D_E_ARG;
    }
    if (option == RS3D_OPT_RAY_SORT_MIN) {
        *value = ctx->opt_ray_sort_min;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_RAYS_PER_WAVE) {
        *value = ctx->opt_rays_per_wave;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_AXIS_GRID) {
        *value = ctx->opt_axis_grid;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_GRID_RAYS_PER_WAVE) {
        *value = ctx->opt_grid_rays_per_wave;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_GRID_CELL_PCT) {
        *value = ctx->opt_grid_cell_pct;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_SEG_OCCUPANCY) {
        *value = ctx->opt_seg_occupancy;
        return RS3D_OK;
    }
    if (option == RS3D_OPT_SEG_CELL_PCT) {
        *value = ctx->opt_seg_cell_pct;
        return RS3D_OK;
    }
    set_error("rs3d_ctx_get_option: unknown option %d", option);
    return RS3D_E_ARG;
}

int rs3d_ctx_profile(rs3d_ctx* ctx, int enable) {
    if (!ctx) {
        set_error("null context");
        return RS3D_E_ARG;
    }
    if (enable && !ctx->ev_start) {
        RS3D_HIP(hipEventCreate(&ctx->ev_start));
        RS3D_HIP(hipEventCreate(&ctx->ev_stop));
    }
    RS3D_TRY(prof_collect(ctx));
    ctx->profile = enable ? 1 : 0;
    return RS3D_OK;
}

int rs3d_ctx_profile_read(rs3d_ctx* ctx, double* total_ms, uint64_t* launches) {
    if (!ctx) {
        set_error("null context");
        return RS3D_E_ARG;
    }
    RS3D_TRY(prof_collect(ctx));
    if (total_ms) *total_ms = ctx->prof_ms;
    if (launches) *launches = ctx->prof_launches;
    ctx->prof_ms = 0.0;
    ctx->prof_launches = 0;
    return RS3D_OK;
}

int rs3d_scan_i32_i64(rs3d_ctx* ctx, const int32_t* in, int64_t* out, uint64_t n, int64_t* total) {
    if (!ctx || (n && (!in || !out))) {
        set_error("rs3d_scan_i32_i64: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_TRY(ctx->small.ensure(64));
    int64_t* tot = ctx->small.as<int64_t>() + 2;
    RS3D_TRY(scan_i32_to_i64_exclusive(ctx, in, out, n, tot));
    uint64_t h = 0;
    RS3D_TRY(read_u64(ctx, tot, &h));
    if (total) *total = (int64_t)h;
    return RS3D_OK;
}

int rs3d_compact_flags(rs3d_ctx* ctx, const uint8_t* flags, uint64_t n, uint32_t* idx_out,
                       uint64_t* count) {
    if (!ctx || (n && (!flags || !idx_out)) || n > 0xffffffffull) {
        set_error("rs3d_compact_flags: bad arguments");
        return RS3D_E_ARG;
    }
    if (n == 0) {
        if (count) *count = 0;
        return RS3D_OK;
    }
    RS3D_TRY(ctx->tmp_b.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->small.ensure(64));
    uint32_t* pos = ctx->tmp_b.as<uint32_t>();
    uint32_t* tot = ctx->small.as<uint32_t>() + 4;
    RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags, pos, n, tot));
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_scatter_index<<<nb, 256, 0, ctx->stream>>>(flags, pos, n, idx_out);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t h;
    memcpy(&h, ctx->pinned, sizeof h);
    if (count) *count = h;
    return RS3D_OK;
}

int rs3d_compact_hits(rs3d_ctx* ctx, const rs3d_ray_result* in, uint64_t n, rs3d_ray_result* out,
                      uint64_t* count) {
    if (!ctx || (n && (!in || !out)) || n > 0xffffffffull) {
        set_error("rs3d_compact_hits: bad arguments");
        return RS3D_E_ARG;
    }
    if (n == 0) {
        if (count) *count = 0;
        return RS3D_OK;
    }
    RS3D_TRY(ctx->tmp_a.ensure(n));
    RS3D_TRY(ctx->tmp_b.ensure(n * sizeof(uint32_t)));
    RS3D_TRY(ctx->small.ensure(64));
    uint8_t* flags = ctx->tmp_a.as<uint8_t>();
    uint32_t* pos = ctx->tmp_b.as<uint32_t>();
    uint32_t* tot = ctx->small.as<uint32_t>() + 4;
    unsigned nb;
    RS3D_TRY(grid_blocks(n, 256, &nb));
    k_hit_flags<<<nb, 256, 0, ctx->stream>>>(in, n, flags);
    RS3D_HIP(hipGetLastError());
    RS3D_TRY(scan_u8_to_u32_exclusive(ctx, flags, pos, n, tot));
    k_scatter_flagged<rs3d_ray_result><<<nb, 256, 0, ctx->stream>>>(in, flags, pos, n, out);
    RS3D_HIP(hipGetLastError());
    RS3D_HIP(hipMemcpyAsync(ctx->pinned, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    uint32_t h;
    memcpy(&h, ctx->pinned, sizeof h);
    if (count) *count = h;
    return RS3D_OK;
}

int rs3d_sort_unique_u64(rs3d_ctx* ctx, uint64_t* keys, uint64_t n, uint64_t* unique_count) {
    if (!ctx || (n && !keys)) {
        set_error("rs3d_sort_unique_u64: bad arguments");
        return RS3D_E_ARG;
    }
    RS3D_TRY(radix_sort_u64(ctx, keys, n));
    if (n == 0) {
        if (unique_count) *unique_count = 0;
        return RS3D_OK;
    }
    // unique into the sort's alternate buffer (n keys, free again after the sort), then back
    RS3D_TRY(ctx->sort_keys.ensure(n * sizeof(uint64_t)));
    uint64_t cnt = 0;
    RS3D_TRY(unique_u64(ctx, keys, n, ctx->sort_keys.as<uint64_t>(), &cnt));
    if (cnt)
        RS3D_HIP(hipMemcpyAsync(keys, ctx->sort_keys.ptr, cnt * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                                ctx->stream));
    RS3D_HIP(hipStreamSynchronize(ctx->stream));
    if (unique_count) *unique_count = cnt;
    return RS3D_OK;
}

int rs3d_sort_pairs_u32(rs3d_ctx* ctx, uint32_t* keys, uint32_t* values, uint64_t n, int end_bit) {
    if (!ctx || (n && (!keys || !values))) {
        set_error("rs3d_sort_pairs_u32: bad arguments");
        return RS3D_E_ARG;
    }
    return radix_sort_u32_pairs(ctx, keys, values, n, end_bit);
}

}  // extern "C"
