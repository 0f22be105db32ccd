// grid.hip -- grid statistics of a mesh on the GPU (the estimator's histogram) and the
// full pair estimate with the replication correction.
//
// Grid statistics follow generateGridStats (preprocess/src/preprocess_dataset.cpp:37-179):
// per object AABB, mesh volume (signed tetrahedra about the AABB centre, summed in
// double), volume ratio, mean extent; the object's centre cell gets CenterCount += 1 and
// every cell its AABB touches gets TouchCount += 1 and the object's size and ratio,
// averaged per cell at the end. Counts are exact; the float sums are order-dependent
// like the reference's unordered_map walk.
#include <cstring>

#include "rs3d_internal.h"

namespace rs3d {

__device__ __forceinline__ int g_f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __host__ __forceinline__ float g_ord2f(int i) {
    int b = i >= 0 ? i : (i ^ 0x7fffffff);
    float f;
    memcpy(&f, &b, sizeof f);
    return f;
}

__global__ __launch_bounds__(256) void k_gs_init(int* __restrict__ obox, double* __restrict__ vol,
                                                 uint64_t n_obj, int* __restrict__ vb) {
    const uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o < n_obj) {
        for (int k = 0; k < 3; ++k) {
            obox[6 * o + k] = g_f2ord(1e30f);
            obox[6 * o + 3 + k] = g_f2ord(-1e30f);
        }
        vol[o] = 0.0;
    }
    if (o == 0) {
        for (int k = 0; k < 3; ++k) {
            vb[k] = g_f2ord(1e30f);
            vb[3 + k] = g_f2ord(-1e30f);
        }
    }
}

__global__ __launch_bounds__(256) void k_gs_vertex_bounds(const float* __restrict__ V, uint64_t nV,
                                                          int* __restrict__ vb) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nV;
         i += (uint64_t)gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], V[3 * i + k]);
            mx[k] = fmaxf(mx[k], V[3 * i + k]);
        }
    for (int k = 0; k < 3; ++k)
        for (int off = 32; off > 0; off >>= 1) {
            mn[k] = fminf(mn[k], __shfl_xor(mn[k], off, 64));
            mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], off, 64));
        }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 3; ++k) {
            atomicMin(&vb[k], g_f2ord(mn[k]));
            atomicMax(&vb[3 + k], g_f2ord(mx[k]));
        }
}

__device__ __forceinline__ bool tri_ok(const uint32_t* I, uint64_t t, uint64_t nV, const int32_t* obj,
                                       uint64_t n_obj) {
    return I[3 * t] < nV && I[3 * t + 1] < nV && I[3 * t + 2] < nV && obj[t] >= 0 &&
           (uint64_t)obj[t] < n_obj;
}

__global__ __launch_bounds__(256) void k_gs_obj_bounds(const float* __restrict__ V, uint64_t nV,
                                                       const uint32_t* __restrict__ I, uint64_t nT,
                                                       const int32_t* __restrict__ obj, uint64_t n_obj,
                                                       int* __restrict__ obox) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT || !tri_ok(I, t, nV, obj, n_obj)) return;
    const uint64_t o = obj[t];
    for (int k = 0; k < 3; ++k) {
        const float a = V[3 * (uint64_t)I[3 * t] + k], b = V[3 * (uint64_t)I[3 * t + 1] + k],
                    c = V[3 * (uint64_t)I[3 * t + 2] + k];
        atomicMin(&obox[6 * o + k], g_f2ord(fminf(fminf(a, b), c)));
        atomicMax(&obox[6 * o + 3 + k], g_f2ord(fmaxf(fmaxf(a, b), c)));
    }
}

__global__ __launch_bounds__(256) void k_gs_volume(const float* __restrict__ V, uint64_t nV,
                                                   const uint32_t* __restrict__ I, uint64_t nT,
                                                   const int32_t* __restrict__ obj, uint64_t n_obj,
                                                   const int* __restrict__ obox, double* __restrict__ vol) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT || !tri_ok(I, t, nV, obj, n_obj)) return;
    const uint64_t o = obj[t];
    float ce[3], t0[3], t1[3], t2[3];
    for (int k = 0; k < 3; ++k) {
        ce[k] = (g_ord2f(obox[6 * o + k]) + g_ord2f(obox[6 * o + 3 + k])) * 0.5f;
        t0[k] = V[3 * (uint64_t)I[3 * t] + k] - ce[k];
        t1[k] = V[3 * (uint64_t)I[3 * t + 1] + k] - ce[k];
        t2[k] = V[3 * (uint64_t)I[3 * t + 2] + k] - ce[k];
    }
    const float cx = t0[1] * t1[2] - t0[2] * t1[1];
    const float cy = t0[2] * t1[0] - t0[0] * t1[2];
    const float cz = t0[0] * t1[1] - t0[1] * t1[0];
    atomicAdd(&vol[o], (double)(cx * t2[0] + cy * t2[1] + cz * t2[2]));
}

__global__ __launch_bounds__(256) void k_gs_cells(const int* __restrict__ obox,
                                                  const double* __restrict__ vol, uint64_t n_obj,
                                                  int res, const float* __restrict__ gb,
                                                  GridCellRec* __restrict__ cells) {
    const uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n_obj) return;
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = g_ord2f(obox[6 * o + k]);
        hi[k] = g_ord2f(obox[6 * o + 3 + k]);
    }
    if (lo[0] > hi[0]) return;  // object without triangles
    const float gmin[3] = {gb[0], gb[1], gb[2]};
    float cs[3];
    for (int k = 0; k < 3; ++k) cs[k] = (gb[3 + k] - gb[k]) / (float)res;
    const float meshVol = fabsf((float)vol[o]) / 6.0f;
    const float w = hi[0] - lo[0], h = hi[1] - lo[1], d = hi[2] - lo[2];
    float aabbVol = w * h * d;
    if (aabbVol < 1e-9f) aabbVol = 1e-9f;
    float ratio = meshVol / aabbVol;
    if (ratio > 1.0f) ratio = 1.0f;
    const float avgSize = (w + h + d) / 3.0f;
    int ac[3];
    for (int k = 0; k < 3; ++k) ac[k] = (int)(((lo[k] + hi[k]) * 0.5f - gmin[k]) / cs[k]);
    if (ac[0] >= 0 && ac[0] < res && ac[1] >= 0 && ac[1] < res && ac[2] >= 0 && ac[2] < res)
        atomicAdd(&cells[((int64_t)ac[2] * res + ac[1]) * res + ac[0]].center, 1u);
    int c0[3], c1[3];
    for (int k = 0; k < 3; ++k) {
        const int q = (int)((lo[k] - gmin[k]) / cs[k]);
        c0[k] = q > 0 ? q : 0;
        const int r = (int)((hi[k] - gmin[k] - 1e-5f) / cs[k]);
        c1[k] = r < res - 1 ? r : res - 1;
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
            for (int x = c0[0]; x <= c1[0]; ++x) {
                GridCellRec* c = &cells[((int64_t)z * res + y) * res + x];
                atomicAdd(&c->touch, 1u);
                atomicAdd(&c->avg, avgSize);
                atomicAdd(&c->ratio, ratio);
            }
}

__global__ __launch_bounds__(256) void k_gs_normalize(GridCellRec* __restrict__ cells, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && cells[i].touch > 0) {
        cells[i].avg /= (float)cells[i].touch;
        cells[i].ratio /= (float)cells[i].touch;
    }
}

__global__ void k_gs_bounds(const int* __restrict__ vb, float world, float* __restrict__ gb) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            gb[k] = world <= 0.0f ? g_ord2f(vb[k]) : 0.0f;
            gb[3 + k] = world <= 0.0f ? g_ord2f(vb[3 + k]) : world;
        }
    }
}

// TouchCount-weighted global means of AvgSizeMean and VolRatio (double accumulation).
__global__ __launch_bounds__(256) void k_gs_global(const GridCellRec* __restrict__ cells, uint64_t n,
                                                   double* __restrict__ acc) {
    __shared__ double red[3][256];
    double s = 0.0, r = 0.0, c = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const GridCellRec g = cells[i];
        if (g.touch > 0) {
            s += (double)g.avg * (double)g.touch;
            r += (double)g.ratio * (double)g.touch;
            c += (double)g.touch;
        }
    }
    red[0][threadIdx.x] = s;
    red[1][threadIdx.x] = r;
    red[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) atomicAdd(&acc[k], red[k][0]);
}

}  // namespace rs3d

using namespace rs3d;

extern "C" {

int rs3d_grid_stats(rs3d_ctx* ctx, const float* vertices, uint64_t num_vertices,
                    const uint32_t* indices, uint64_t num_triangles, const int32_t* tri_to_obj,
                    uint64_t num_objects, int resolution, float world_size, void* cells,
                    float* bounds_out) {
    if (!ctx || resolution <= 0 || resolution > 1024 || !cells ||
        (num_triangles && (!vertices || !indices || !tri_to_obj))) {
        set_error("rs3d_grid_stats: bad arguments");
        return RS3D_E_ARG;
    }
    const uint64_t ncell = (uint64_t)resolution * resolution * resolution;
    const uint64_t nob = num_objects ? num_objects : 1;
    DevBuf tmp;
    RS3D_TRY(tmp.ensure(nob * (6 * sizeof(int) + sizeof(double)) + 1024));
    uint8_t* base = tmp.as<uint8_t>();
    double* vol = reinterpret_cast<double*>(base);
    int* obox = reinterpret_cast<int*>(base + nob * sizeof(double));
    int* vb = reinterpret_cast<int*>(base + nob * (6 * sizeof(int) + sizeof(double)) + 64);
    float* gb = reinterpret_cast<float*>(vb + 8);
    hipStream_t s = ctx->stream;
    unsigned nbo, nbt, nbc;
    RS3D_TRY(grid_blocks(nob, 256, &nbo));
    RS3D_TRY(grid_blocks(num_triangles ? num_triangles : 1, 256, &nbt));
    RS3D_TRY(grid_blocks(ncell, 256, &nbc));
    RS3D_HIP(hipMemsetAsync(cells, 0, ncell * sizeof(GridCellRec), s));
    k_gs_init<<<nbo, 256, 0, s>>>(obox, vol, num_objects, vb);
    const unsigned nbv = (unsigned)std::min<uint64_t>((num_vertices + 255) / 256 + 1, 4096);
    k_gs_vertex_bounds<<<nbv, 256, 0, s>>>(vertices, num_vertices, vb);
    k_gs_bounds<<<1, 64, 0, s>>>(vb, world_size, gb);
    if (num_triangles && num_objects) {
        k_gs_obj_bounds<<<nbt, 256, 0, s>>>(vertices, num_vertices, indices, num_triangles, tri_to_obj,
                                            num_objects, obox);
        k_gs_volume<<<nbt, 256, 0, s>>>(vertices, num_vertices, indices, num_triangles, tri_to_obj,
                                        num_objects, obox, vol);
        k_gs_cells<<<nbo, 256, 0, s>>>(obox, vol, num_objects, resolution, gb,
                                       static_cast<GridCellRec*>(cells));
    }
    k_gs_normalize<<<nbc, 256, 0, s>>>(static_cast<GridCellRec*>(cells), ncell);
    RS3D_HIP(hipGetLastError());
    float hb[6];
    RS3D_HIP(hipMemcpyAsync(hb, gb, sizeof hb, hipMemcpyDeviceToHost, s));
    RS3D_HIP(hipStreamSynchronize(s));
    if (bounds_out) memcpy(bounds_out, hb, sizeof hb);
    tmp.release();
    return RS3D_OK;
}

// Estimated pair count of raytracer_overlap_estimated
// (query/src/applications/raytracer_overlap_estimated_main.cpp:133-195): raw sum over
// cells divided by alpha = max((sA * cbrt(rA) + sB * cbrt(rB))^3 / Vcell, 1), with the
// TouchCount-weighted means sX, rX; alpha_correction = 0 returns the raw sum.
int rs3d_estimate_pairs(rs3d_ctx* ctx, const void* cells_a, const void* cells_b, uint64_t num_cells,
                        const float* grid_min, const float* grid_max, const uint32_t* resolution,
                        float epsilon, float gamma, int alpha_correction, int64_t* estimate) {
    if (!ctx || !grid_min || !grid_max || !resolution || !estimate ||
        (num_cells && (!cells_a || !cells_b))) {
        set_error("rs3d_estimate_pairs: bad arguments");
        return RS3D_E_ARG;
    }
    const float w = grid_max[0] - grid_min[0], h = grid_max[1] - grid_min[1],
                d = grid_max[2] - grid_min[2];
    const float csx = w / resolution[0], csy = h / resolution[1], csz = d / resolution[2];
    float cell_volume = csx * csy * csz;
    double raw = 0.0;
    RS3D_TRY(rs3d_estimate_raw(ctx, cells_a, cells_b, num_cells, cell_volume, epsilon, gamma, &raw));
    const float rawf = (float)raw;
    if (!alpha_correction) {
        *estimate = (int64_t)rawf;
        return RS3D_OK;
    }
    RS3D_TRY(ctx->tmp_a.ensure(8 * sizeof(double)));
    double* acc = ctx->tmp_a.as<double>();
    float g[4];
    for (int side = 0; side < 2; ++side) {
        RS3D_HIP(hipMemsetAsync(acc, 0, 3 * sizeof(double), ctx->stream));
        k_gs_global<<<256, 256, 0, ctx->stream>>>(
            static_cast<const GridCellRec*>(side ? cells_b : cells_a), num_cells, acc);
        double ha[3];
        RS3D_HIP(hipMemcpyAsync(ha, acc, sizeof ha, hipMemcpyDeviceToHost, ctx->stream));
        RS3D_HIP(hipStreamSynchronize(ctx->stream));
        g[2 * side] = ha[2] > 0 ? (float)(ha[0] / ha[2]) : 0.0f;      // mean size
        g[2 * side + 1] = ha[2] > 0 ? (float)(ha[1] / ha[2]) : 1.0f;  // mean ratio
    }
    const float e1 = g[0] * std::cbrt(g[1]);
    const float e2 = g[2] * std::cbrt(g[3]);
    const float comb = e1 + e2;
    const float mink = comb * comb * comb;
    if (cell_volume < 1e-9f) cell_volume = 1e-9f;
    float alpha = mink / cell_volume;
    if (alpha < 1.0f) alpha = 1.0f;
    *estimate = (int64_t)(rawf / alpha);
    return RS3D_OK;
}

}  // extern "C"
