/*
 * oracle.c -- CPU restatement of RaySpace3D's ray-cast spatial join (test checker only).
 *
 * TEST INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg. The product (librs3d.so) never links or calls this file.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so every fp32
 * operation below is one IEEE single-precision operation in the order written.
 *
 * Canonical semantics (shared, operation for operation, with the HIP kernels):
 *  - Triangle test: watertight ray/triangle test of Woop, Benthin and Wald (JCGT 2(1),
 *    2013) with culling off (the reference builds its GAS with face culling disabled,
 *    query/src/optix/OptixAccelerationStructure.cpp:41) and t = T / det in one
 *    correctly rounded division.
 *  - A.1 point chain (query/src/cuda/raytracing.cu:22-65): next hit is the smallest
 *    (t, triangle index) with t >= fl(t_prev + 1e-6); when that add is absorbed
 *    (t >= 32) the next hit must satisfy t > t_prev. Hits must have t < 1e16. Cap 10000.
 *  - A.2 edge chain (query/src/cuda/mesh_overlap_edges.cu:106-181): segment ray over
 *    [1e-6, len + 1e-6], step fl(t + 1e-6) or t + 2e-4 when absorbed, cap 100.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* Triangle test                                                                         */
/* ------------------------------------------------------------------------------------ */

void orc_ray_pre(const float d[3], orc_raypre* r) {
    int kz = 0;
    if (fabsf(d[1]) > fabsf(d[kz])) kz = 1;
    if (fabsf(d[2]) > fabsf(d[kz])) kz = 2;
    int kx = kz + 1;
    if (kx == 3) kx = 0;
    int ky = kx + 1;
    if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) {
        int tmp = kx;
        kx = ky;
        ky = tmp;
    }
    r->kx = kx;
    r->ky = ky;
    r->kz = kz;
    r->Sx = d[kx] / d[kz];
    r->Sy = d[ky] / d[kz];
    r->Sz = 1.0f / d[kz];
}

int orc_tri_hit(const float o[3], const orc_raypre* r, const float v0[3], const float v1[3],
                const float v2[3], float* tout) {
    const int kx = r->kx, ky = r->ky, kz = r->kz;
    float A[3], B[3], C[3];
    for (int i = 0; i < 3; ++i) {
        A[i] = v0[i] - o[i];
        B[i] = v1[i] - o[i];
        C[i] = v2[i] - o[i];
    }
    const float Ax = A[kx] - r->Sx * A[kz];
    const float Ay = A[ky] - r->Sy * A[kz];
    const float Bx = B[kx] - r->Sx * B[kz];
    const float By = B[ky] - r->Sy * B[kz];
    const float Cx = C[kx] - r->Sx * C[kz];
    const float Cy = C[ky] - r->Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        /* Products of two floats are exact in double; one rounding to double, one to float. */
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return 0;
    const float det = U + V + W;
    if (det == 0.0f) return 0;
    const float Az = r->Sz * A[kz];
    const float Bz = r->Sz * B[kz];
    const float Cz = r->Sz * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    *tout = T / det;
    return 1;
}

/* ------------------------------------------------------------------------------------ */
/* Hit lists                                                                             */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    float t;
    uint32_t tri;
} hit_t;

typedef struct {
    hit_t* h;
    int64_t n, cap;
} hitvec;

static void hv_push(hitvec* v, float t, uint32_t tri) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->h = (hit_t*)realloc(v->h, (size_t)v->cap * sizeof(hit_t));
    }
    v->h[v->n].t = t;
    v->h[v->n].tri = tri;
    v->n++;
}

static int hit_cmp(const void* a, const void* b) {
    const hit_t* x = (const hit_t*)a;
    const hit_t* y = (const hit_t*)b;
    if (x->t < y->t) return -1;
    if (x->t > y->t) return 1;
    if (x->tri < y->tri) return -1;
    if (x->tri > y->tri) return 1;
    return 0;
}

/* Conservative padded box of a triangle (pruning only; never decides a hit). */
static void tri_box(const float* V, const uint32_t* I, int64_t t, double lo[3], double hi[3]) {
    const float* a = V + 3 * (int64_t)I[3 * t + 0];
    const float* b = V + 3 * (int64_t)I[3 * t + 1];
    const float* c = V + 3 * (int64_t)I[3 * t + 2];
    double ext = 0.0, mag = 0.0;
    for (int k = 0; k < 3; ++k) {
        double l = a[k], h = a[k];
        if (b[k] < l) l = b[k];
        if (c[k] < l) l = c[k];
        if (b[k] > h) h = b[k];
        if (c[k] > h) h = c[k];
        lo[k] = l;
        hi[k] = h;
        if (h - l > ext) ext = h - l;
        if (fabs(l) > mag) mag = fabs(l);
        if (fabs(h) > mag) mag = fabs(h);
    }
    const double pad = (mag + ext) * 1e-5 + 1e-30;
    for (int k = 0; k < 3; ++k) {
        lo[k] -= pad;
        hi[k] += pad;
    }
}

/* ------------------------------------------------------------------------------------ */
/* Uniform grids (candidate generation only)                                             */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int dims;            /* 2 or 3 */
    int ax[3];           /* axes covered */
    int g[3];            /* cells per axis */
    double lo[3], cell[3];
    int64_t* start;      /* CSR offsets, ncells + 1 */
    uint32_t* items;
} grid_t;

static int64_t grid_ncells(const grid_t* G) {
    int64_t n = 1;
    for (int i = 0; i < G->dims; ++i) n *= G->g[i];
    return n;
}

static int cell_coord(const grid_t* G, int i, double x) {
    double f = floor((x - G->lo[i]) / G->cell[i]);
    if (f < 0) return -1;
    if (f >= G->g[i]) return G->g[i];
    return (int)f;
}

static void grid_build(grid_t* G, int dims, const int* axes, const float* V, const uint32_t* I,
                       int64_t nT, double cells_per_tri) {
    G->dims = dims;
    double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
    for (int64_t t = 0; t < nT; ++t) {
        double lo[3], hi[3];
        tri_box(V, I, t, lo, hi);
        for (int k = 0; k < 3; ++k) {
            if (lo[k] < blo[k]) blo[k] = lo[k];
            if (hi[k] > bhi[k]) bhi[k] = hi[k];
        }
    }
    double target = (double)nT * cells_per_tri;
    if (target < 1) target = 1;
    double per = pow(target, 1.0 / dims);
    for (int i = 0; i < dims; ++i) {
        G->ax[i] = axes[i];
        int g = (int)per;
        if (g < 1) g = 1;
        if (g > 2048) g = 2048;
        G->g[i] = g;
        G->lo[i] = blo[axes[i]];
        double ext = bhi[axes[i]] - blo[axes[i]];
        if (!(ext > 0)) ext = 1.0;
        G->cell[i] = ext / g * (1.0 + 1e-12);
    }
    int64_t nc = grid_ncells(G);
    G->start = (int64_t*)calloc((size_t)nc + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
        int64_t* fill = NULL;
        if (pass == 1) {
            for (int64_t c = 0; c < nc; ++c) G->start[c + 1] += G->start[c];
            G->items = (uint32_t*)malloc((size_t)(G->start[nc] ? G->start[nc] : 1) * sizeof(uint32_t));
            fill = (int64_t*)malloc((size_t)nc * sizeof(int64_t));
            memcpy(fill, G->start, (size_t)nc * sizeof(int64_t));
        }
        for (int64_t t = 0; t < nT; ++t) {
            double lo[3], hi[3];
            tri_box(V, I, t, lo, hi);
            int c0[3] = {0, 0, 0}, c1[3] = {0, 0, 0};
            for (int i = 0; i < dims; ++i) {
                c0[i] = cell_coord(G, i, lo[axes[i]]);
                c1[i] = cell_coord(G, i, hi[axes[i]]);
                if (c0[i] < 0) c0[i] = 0;
                if (c1[i] > G->g[i] - 1) c1[i] = G->g[i] - 1;
            }
            for (int z = c0[2]; z <= c1[2]; ++z)
                for (int y = c0[1]; y <= c1[1]; ++y)
                    for (int x = c0[0]; x <= c1[0]; ++x) {
                        int64_t c = x + (int64_t)G->g[0] * (y + (int64_t)(dims > 1 ? G->g[1] : 1) * z);
                        if (pass == 0)
                            G->start[c + 1]++;
                        else
                            G->items[fill[c]++] = (uint32_t)t;
                    }
        }
        free(fill);
    }
}

static void grid_free(grid_t* G) {
    free(G->start);
    free(G->items);
}

/* ------------------------------------------------------------------------------------ */
/* A.1 point containment                                                                 */
/* ------------------------------------------------------------------------------------ */

static void test_tri_collect(const float* V, const uint32_t* I, int64_t t, const float o[3],
                             const orc_raypre* rp, float tlo, float thi, int thi_inclusive,
                             hitvec* hv) {
    float th;
    const float* a = V + 3 * (int64_t)I[3 * t + 0];
    const float* b = V + 3 * (int64_t)I[3 * t + 1];
    const float* c = V + 3 * (int64_t)I[3 * t + 2];
    if (!orc_tri_hit(o, rp, a, b, c, &th)) return;
    if (!(th >= tlo)) return;
    if (thi_inclusive ? !(th <= thi) : !(th < thi)) return;
    hv_push(hv, th, (uint32_t)t);
}

/* One point of orc_points_contain: the canonical chain over the collected hits. */
static int32_t contain_one(const float* V, const uint32_t* I, int64_t nT, const int32_t* tri_obj,
                           const float* o, const orc_raypre* rp, const grid_t* G,
                           const int* axes, int use_grid, hitvec* hv) {
    const float eps = 1e-6f, tmax = 1e16f;
    hv->n = 0;
    if (use_grid && nT > 0) {
        int cu = cell_coord(G, 0, o[axes[0]]);
        int cv = cell_coord(G, 1, o[axes[1]]);
        if (cu >= 0 && cu < G->g[0] && cv >= 0 && cv < G->g[1]) {
            int64_t c = cu + (int64_t)G->g[0] * cv;
            for (int64_t k = G->start[c]; k < G->start[c + 1]; ++k)
                test_tri_collect(V, I, G->items[k], o, rp, eps, tmax, 0, hv);
        }
    } else {
        for (int64_t t = 0; t < nT; ++t) test_tri_collect(V, I, t, o, rp, eps, tmax, 0, hv);
    }
    qsort(hv->h, (size_t)hv->n, sizeof(hit_t), hit_cmp);
    float thr = eps;
    int count = 0;
    uint32_t first = 0;
    int64_t j = 0;
    while (count < 10000) {
        while (j < hv->n && hv->h[j].t < thr) ++j;
        if (j == hv->n) break;
        if (count == 0) first = hv->h[j].tri;
        ++count;
        const float t = hv->h[j].t;
        float nthr = t + eps;
        if (nthr <= t) nthr = nextafterf(t, INFINITY);
        thr = nthr;
        ++j;
    }
    return (count & 1) ? tri_obj[first] : -1;
}

int64_t orc_points_contain_mt(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                              const int32_t* tri_obj, const float* P, int64_t n, int axis,
                              int use_grid, int32_t* out_obj, int threads) {
    (void)nV;
    float d[3] = {0.0f, 0.0f, 0.0f};
    d[axis] = 1.0f;
    orc_raypre rp;
    orc_ray_pre(d, &rp);
    grid_t G;
    memset(&G, 0, sizeof G);
    int axes[2] = {(axis + 1) % 3, (axis + 2) % 3};
    if (use_grid && nT > 0) grid_build(&G, 2, axes, V, I, nT, 1.0);
    int64_t inside = 0;
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(+ : inside)
    {
        hitvec hv = {0, 0, 0};
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < n; ++i) {
            const float o[3] = {P[3 * i], P[3 * i + 1], P[3 * i + 2]};
            out_obj[i] = contain_one(V, I, nT, tri_obj, o, &rp, &G, axes, use_grid, &hv);
            inside += out_obj[i] >= 0;
        }
        free(hv.h);
    }
    if (use_grid && nT > 0) grid_free(&G);
    return inside;
}

int64_t orc_points_contain(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                           const int32_t* tri_obj, const float* P, int64_t n, int axis,
                           int use_grid, int32_t* out_obj) {
    return orc_points_contain_mt(V, nV, I, nT, tri_obj, P, n, axis, use_grid, out_obj, 1);
}

/* ------------------------------------------------------------------------------------ */
/* A.2 edge rays                                                                         */
/* ------------------------------------------------------------------------------------ */

int64_t orc_edge_pairs(const float* starts, const float* ends, const int32_t* src_off,
                       const int32_t* src_cnt, const int32_t* src_obj, int64_t nE,
                       const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                       const int32_t* tri_obj, int swap, int use_grid, uint64_t* out_keys,
                       int64_t cap, int32_t* per_edge_counts) {
    return orc_edge_pairs_ex(starts, ends, src_off, src_cnt, src_obj, nE, V, nV, I, nT, tri_obj,
                             swap, use_grid, 0, 100, out_keys, cap, per_edge_counts);
}

/* chain_rule 0: fl(t + 1e-6), or t + 2e-4 when absorbed (mesh_overlap_edges.cu:173-177).
 * chain_rule 1: nextafterf(t, tmax) (mesh_containment.cu:150-160). max_iter: chain cap. */
int64_t orc_edge_pairs_ex(const float* starts, const float* ends, const int32_t* src_off,
                          const int32_t* src_cnt, const int32_t* src_obj, int64_t nE,
                          const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                          const int32_t* tri_obj, int swap, int use_grid, int chain_rule,
                          int max_iter, uint64_t* out_keys, int64_t cap, int32_t* per_edge_counts) {
    (void)nV;
    const float eps = 1e-6f;
    grid_t G;
    const int axes[3] = {0, 1, 2};
    uint32_t* stamp = NULL;
    if (use_grid && nT > 0) {
        grid_build(&G, 3, axes, V, I, nT, 1.0);
        stamp = (uint32_t*)malloc((size_t)nT * sizeof(uint32_t));
        memset(stamp, 0xff, (size_t)nT * sizeof(uint32_t));
    }
    hitvec hv = {0, 0, 0};
    int64_t total = 0;
    for (int64_t e = 0; e < nE; ++e) {
        const float* s = starts + 3 * e;
        const float* en = ends + 3 * e;
        const float dx = en[0] - s[0], dy = en[1] - s[1], dz = en[2] - s[2];
        const float len = sqrtf(dx * dx + dy * dy + dz * dz);
        int nres = 0;
        if (len >= eps && nT > 0) {
            const float d[3] = {dx / len, dy / len, dz / len};
            orc_raypre rp;
            orc_ray_pre(d, &rp);
            const float tmax = len + eps;
            hv.n = 0;
            if (use_grid) {
                double lo[3], hi[3];
                for (int k = 0; k < 3; ++k) {
                    double a = s[k], b = (double)s[k] + (double)d[k] * (double)tmax;
                    lo[k] = a < b ? a : b;
                    hi[k] = a < b ? b : a;
                    double pad = (fabs(lo[k]) + fabs(hi[k]) + (double)tmax) * 1e-5 + 1e-30;
                    lo[k] -= pad;
                    hi[k] += pad;
                }
                int c0[3], c1[3];
                for (int i = 0; i < 3; ++i) {
                    c0[i] = cell_coord(&G, i, lo[i]);
                    c1[i] = cell_coord(&G, i, hi[i]);
                    if (c0[i] < 0) c0[i] = 0;
                    if (c1[i] > G.g[i] - 1) c1[i] = G.g[i] - 1;
                }
                for (int z = c0[2]; z <= c1[2]; ++z)
                    for (int y = c0[1]; y <= c1[1]; ++y)
                        for (int x = c0[0]; x <= c1[0]; ++x) {
                            int64_t c = x + (int64_t)G.g[0] * (y + (int64_t)G.g[1] * z);
                            for (int64_t k = G.start[c]; k < G.start[c + 1]; ++k) {
                                uint32_t t = G.items[k];
                                if (stamp[t] == (uint32_t)e) continue;
                                stamp[t] = (uint32_t)e;
                                test_tri_collect(V, I, t, s, &rp, eps, tmax, 1, &hv);
                            }
                        }
            } else {
                for (int64_t t = 0; t < nT; ++t) test_tri_collect(V, I, t, s, &rp, eps, tmax, 1, &hv);
            }
            qsort(hv.h, (size_t)hv.n, sizeof(hit_t), hit_cmp);
            float thr = eps;
            int64_t j = 0;
            for (int iter = 0; iter < max_iter; ++iter) {
                if (thr > tmax) break;
                while (j < hv.n && hv.h[j].t < thr) ++j;
                if (j == hv.n) break;
                const float t = hv.h[j].t;
                const int32_t tgt = tri_obj[hv.h[j].tri];
                for (int q = 0; q < src_cnt[e]; ++q) {
                    const int32_t src = src_obj[src_off[e] + q];
                    const uint32_t a = (uint32_t)(swap ? tgt : src);
                    const uint32_t b = (uint32_t)(swap ? src : tgt);
                    if (total < cap) out_keys[total] = ((uint64_t)a << 32) | (uint64_t)b;
                    ++total;
                    ++nres;
                }
                float nthr;
                if (chain_rule == 0) {
                    nthr = t + eps;
                    if (nthr <= t) nthr = t + 2e-4f;
                } else {
                    nthr = nextafterf(t, tmax);
                    if (nthr <= t) nthr = t + eps;
                }
                thr = nthr;
                ++j;
            }
        }
        if (per_edge_counts) per_edge_counts[e] = nres;
    }
    free(hv.h);
    if (use_grid && nT > 0) {
        grid_free(&G);
        free(stamp);
    }
    return total;
}

/* ------------------------------------------------------------------------------------ */
/* Sort / unique / hash (query/src/cuda/mesh_query_deduplication.cu:56-120,             */
/* query/src/cuda/mesh_overlap_edges.cu:62-104)                                          */
/* ------------------------------------------------------------------------------------ */

static int u64_cmp(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int64_t orc_sort_unique_u64(uint64_t* keys, int64_t n) {
    if (n <= 1) return n;
    qsort(keys, (size_t)n, sizeof(uint64_t), u64_cmp);
    int64_t w = 1;
    for (int64_t i = 1; i < n; ++i)
        if (keys[i] != keys[w - 1]) keys[w++] = keys[i];
    return w;
}

uint64_t orc_fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

int orc_ht_insert(uint64_t* table, uint64_t size, int bitwise, int max_probes, uint64_t key) {
    if (size == 0) return 0;
    const uint64_t k = orc_fmix64(key);
    uint32_t h = bitwise ? (uint32_t)(k & (size - 1)) : (uint32_t)(k % size);
    for (int i = 0; i < max_probes; ++i) {
        if (table[h] == ~0ULL) {
            table[h] = key;
            return 1;
        }
        if (table[h] == key) return 1;
        h = bitwise ? ((h + 1) & (uint32_t)(size - 1)) : ((h + 1) % (uint32_t)size);
    }
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* A.5 estimator (query/src/cuda/estimated_overlap.cu:8-41;                              */
/* query/src/applications/raytracer_overlap_estimated_main.cpp:129-195)                  */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    uint32_t center, touch;
    float avg, ratio;
} cell_t;

double orc_estimate_raw(const void* cellsA, const void* cellsB, int64_t ncells, float cell_volume,
                        float epsilon, float gamma) {
    const cell_t* A = (const cell_t*)cellsA;
    const cell_t* B = (const cell_t*)cellsB;
    double sum = 0.0;
    for (int64_t i = 0; i < ncells; ++i) {
        if (A[i].touch == 0 || B[i].touch == 0) continue;
        const float cs = A[i].avg + B[i].avg + epsilon;
        const float mink = cs * cs * cs;
        float prob = mink / cell_volume;
        const float ratio = sqrtf(A[i].ratio * B[i].ratio);
        prob *= powf(ratio, gamma);
        prob = fminf(prob, 1.0f);
        sum += (double)((float)A[i].touch * (float)B[i].touch * prob);
    }
    return sum;
}

static float global_avg(const cell_t* c, int64_t n, int ratio) {
    double tot = 0.0;
    long long cnt = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (c[i].touch > 0) {
            tot += (double)(ratio ? c[i].ratio : c[i].avg) * (double)c[i].touch;
            cnt += c[i].touch;
        }
    }
    if (cnt == 0) return ratio ? 1.0f : 0.0f;
    return (float)(tot / cnt);
}

int64_t orc_estimate_pairs(const void* cellsA, const void* cellsB, int64_t ncells,
                           const float minb[3], const float maxb[3], const uint32_t res[3],
                           float epsilon, float gamma, int alpha_correction) {
    const float w = maxb[0] - minb[0], h = maxb[1] - minb[1], dd = maxb[2] - minb[2];
    const float csx = w / res[0], csy = h / res[1], csz = dd / res[2];
    float cell_volume = csx * csy * csz;
    const double raw = orc_estimate_raw(cellsA, cellsB, ncells, cell_volume, epsilon, gamma);
    const float rawf = (float)raw;
    if (!alpha_correction) return (int64_t)rawf;
    const cell_t* A = (const cell_t*)cellsA;
    const cell_t* B = (const cell_t*)cellsB;
    const float e1 = global_avg(A, ncells, 0) * cbrtf(global_avg(A, ncells, 1));
    const float e2 = global_avg(B, ncells, 0) * cbrtf(global_avg(B, ncells, 1));
    const float comb = e1 + e2;
    const float mink = comb * comb * comb;
    if (cell_volume < 1e-9f) cell_volume = 1e-9f;
    float alpha = mink / cell_volume;
    if (alpha < 1.0f) alpha = 1.0f;
    return (int64_t)(rawf / alpha);
}

/* ------------------------------------------------------------------------------------ */
/* Grid statistics (preprocess/src/preprocess_dataset.cpp:37-179)                        */
/* ------------------------------------------------------------------------------------ */

void orc_grid_stats(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                    const int32_t* tri_obj, int res, float world_size, void* cells_out,
                    float* bounds_out) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t v = 0; v < nV; ++v)
        for (int k = 0; k < 3; ++k) {
            mn[k] = fminf(mn[k], V[3 * v + k]);
            mx[k] = fmaxf(mx[k], V[3 * v + k]);
        }
    float gmin[3], gmax[3];
    for (int k = 0; k < 3; ++k) {
        gmin[k] = world_size <= 0 ? mn[k] : 0.0f;
        gmax[k] = world_size <= 0 ? mx[k] : world_size;
    }
    float cs[3];
    for (int k = 0; k < 3; ++k) cs[k] = (gmax[k] - gmin[k]) / (float)res;
    cell_t* cells = (cell_t*)cells_out;
    const int64_t ncell = (int64_t)res * res * res;
    memset(cells, 0, (size_t)ncell * sizeof(cell_t));
    int32_t maxobj = -1;
    for (int64_t t = 0; t < nT; ++t)
        if (tri_obj[t] > maxobj) maxobj = tri_obj[t];
    const int64_t nobj = (int64_t)maxobj + 1;
    float* omin = (float*)malloc(sizeof(float) * 3 * (size_t)(nobj > 0 ? nobj : 1));
    float* omax = (float*)malloc(sizeof(float) * 3 * (size_t)(nobj > 0 ? nobj : 1));
    double* vol = (double*)calloc((size_t)(nobj > 0 ? nobj : 1), sizeof(double));
    char* seen = (char*)calloc((size_t)(nobj > 0 ? nobj : 1), 1);
    for (int64_t o = 0; o < nobj; ++o)
        for (int k = 0; k < 3; ++k) {
            omin[3 * o + k] = 1e30f;
            omax[3 * o + k] = -1e30f;
        }
    for (int64_t t = 0; t < nT; ++t) {
        const uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
        if (a >= nV || b >= nV || c >= nV || tri_obj[t] < 0) continue;
        const int64_t o = tri_obj[t];
        seen[o] = 1;
        for (int k = 0; k < 3; ++k) {
            omin[3 * o + k] = fminf(fminf(fminf(omin[3 * o + k], V[3 * a + k]), V[3 * b + k]), V[3 * c + k]);
            omax[3 * o + k] = fmaxf(fmaxf(fmaxf(omax[3 * o + k], V[3 * a + k]), V[3 * b + k]), V[3 * c + k]);
        }
    }
    for (int64_t t = 0; t < nT; ++t) {
        const uint32_t a = I[3 * t], b = I[3 * t + 1], c = I[3 * t + 2];
        if (a >= nV || b >= nV || c >= nV || tri_obj[t] < 0) continue;
        const int64_t o = tri_obj[t];
        float ce[3], t0[3], t1[3], t2[3];
        for (int k = 0; k < 3; ++k) {
            ce[k] = (omin[3 * o + k] + omax[3 * o + k]) * 0.5f;
            t0[k] = V[3 * a + k] - ce[k];
            t1[k] = V[3 * b + k] - ce[k];
            t2[k] = V[3 * c + k] - ce[k];
        }
        const float cx = t0[1] * t1[2] - t0[2] * t1[1];
        const float cy = t0[2] * t1[0] - t0[0] * t1[2];
        const float cz = t0[0] * t1[1] - t0[1] * t1[0];
        vol[o] += (double)(cx * t2[0] + cy * t2[1] + cz * t2[2]);
    }
    for (int64_t o = 0; o < nobj; ++o) {
        if (!seen[o]) continue;
        const float* lo = omin + 3 * o;
        const float* hi = omax + 3 * o;
        const float meshVol = fabsf((float)vol[o]) / 6.0f;
        const float w = hi[0] - lo[0], h = hi[1] - lo[1], d = hi[2] - lo[2];
        float aabbVol = w * h * d;
        if (aabbVol < 1e-9f) aabbVol = 1e-9f;
        float ratio = meshVol / aabbVol;
        if (ratio > 1.0f) ratio = 1.0f;
        float ce[3];
        for (int k = 0; k < 3; ++k) ce[k] = (lo[k] + hi[k]) * 0.5f;
        const float avgSize = (w + h + d) / 3.0f;
        int ac[3];
        for (int k = 0; k < 3; ++k) ac[k] = (int)((ce[k] - gmin[k]) / cs[k]);
        if (ac[0] >= 0 && ac[0] < res && ac[1] >= 0 && ac[1] < res && ac[2] >= 0 && ac[2] < res)
            cells[((int64_t)ac[2] * res + ac[1]) * res + ac[0]].center++;
        int c0[3], c1[3];
        for (int k = 0; k < 3; ++k) {
            int q = (int)((lo[k] - gmin[k]) / cs[k]);
            c0[k] = q > 0 ? q : 0;
            int r = (int)((hi[k] - gmin[k] - 1e-5f) / cs[k]);
            c1[k] = r < res - 1 ? r : res - 1;
        }
        for (int z = c0[2]; z <= c1[2]; ++z)
            for (int y = c0[1]; y <= c1[1]; ++y)
                for (int x = c0[0]; x <= c1[0]; ++x) {
                    cell_t* cc = &cells[((int64_t)z * res + y) * res + x];
                    cc->touch++;
                    cc->avg += avgSize;
                    cc->ratio += ratio;
                }
    }
    for (int64_t i = 0; i < ncell; ++i)
        if (cells[i].touch > 0) {
            cells[i].avg /= (float)cells[i].touch;
            cells[i].ratio /= (float)cells[i].touch;
        }
    for (int k = 0; k < 3; ++k) {
        bounds_out[k] = gmin[k];
        bounds_out[3 + k] = gmax[k];
    }
    free(omin);
    free(omax);
    free(vol);
    free(seen);
}

/* ------------------------------------------------------------------------------------ */
/* Unique edges (preprocess/src/EdgePreprocessor.cpp:41-119)                             */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int64_t q[6];
    int64_t order; /* 3 * tri + edge slot */
} ekey_t;

static int qp_less(const int64_t* a, const int64_t* b) {
    if (a[0] != b[0]) return a[0] < b[0];
    if (a[1] != b[1]) return a[1] < b[1];
    return a[2] < b[2];
}

static int ekey_cmp(const void* pa, const void* pb) {
    const ekey_t* a = (const ekey_t*)pa;
    const ekey_t* b = (const ekey_t*)pb;
    for (int i = 0; i < 6; ++i)
        if (a->q[i] != b->q[i]) return a->q[i] < b->q[i] ? -1 : 1;
    return a->order < b->order ? -1 : (a->order > b->order ? 1 : 0);
}

int64_t orc_extract_edges(const float* V, int64_t nV, const uint32_t* I, int64_t nT,
                          const int32_t* tri_obj, float* starts, float* ends, int32_t* offsets,
                          int32_t* counts, int32_t* sources, int64_t* n_sources) {
    ekey_t* recs = (ekey_t*)malloc(sizeof(ekey_t) * (size_t)(3 * nT + 1));
    int64_t nr = 0;
    for (int64_t t = 0; t < nT; ++t) {
        const uint32_t vv[3] = {I[3 * t], I[3 * t + 1], I[3 * t + 2]};
        if (vv[0] >= nV || vv[1] >= nV || vv[2] >= nV) continue;
        for (int e = 0; e < 3; ++e) {
            const float* p0 = V + 3 * (int64_t)vv[e];
            const float* p1 = V + 3 * (int64_t)vv[(e + 1) % 3];
            int64_t q0[3], q1[3];
            for (int k = 0; k < 3; ++k) {
                q0[k] = (int64_t)llround((double)p0[k] * 1e5);
                q1[k] = (int64_t)llround((double)p1[k] * 1e5);
            }
            ekey_t* r = &recs[nr++];
            if (qp_less(q1, q0)) {
                memcpy(r->q, q1, sizeof q1);
                memcpy(r->q + 3, q0, sizeof q0);
            } else {
                memcpy(r->q, q0, sizeof q0);
                memcpy(r->q + 3, q1, sizeof q1);
            }
            r->order = 3 * t + e;
        }
    }
    qsort(recs, (size_t)nr, sizeof(ekey_t), ekey_cmp);
    int64_t ne = 0, ns = 0;
    for (int64_t i = 0; i < nr;) {
        int64_t j = i;
        while (j < nr && memcmp(recs[j].q, recs[i].q, sizeof recs[i].q) == 0) ++j;
        const int64_t t = recs[i].order / 3;
        const int e = (int)(recs[i].order % 3);
        const uint32_t vv[3] = {I[3 * t], I[3 * t + 1], I[3 * t + 2]};
        if (starts) {
            for (int k = 0; k < 3; ++k) {
                starts[3 * ne + k] = V[3 * (int64_t)vv[e] + k];
                ends[3 * ne + k] = V[3 * (int64_t)vv[(e + 1) % 3] + k];
            }
            offsets[ne] = (int32_t)ns;
        }
        int32_t cnt = 0;
        for (int64_t k = i; k < j; ++k) {
            const int32_t src = tri_obj[recs[k].order / 3];
            int dup = 0;
            if (sources) {
                for (int32_t q = 0; q < cnt; ++q)
                    if (sources[ns + q] == src) dup = 1;
            } else {
                for (int64_t m = i; m < k; ++m)
                    if (tri_obj[recs[m].order / 3] == src) dup = 1;
            }
            if (!dup) {
                if (sources) sources[ns + cnt] = src;
                ++cnt;
            }
        }
        if (counts) counts[ne] = cnt;
        ns += cnt;
        ++ne;
        i = j;
    }
    free(recs);
    if (n_sources) *n_sources = ns;
    return ne;
}

/* ------------------------------------------------------------------------------------ */
/* A.3 / A.4 per-object containment rays (query/src/cuda/mesh_intersection_anyhit.cu:      */
/* 46-240; query/src/cuda/mesh_containment.cu:196-377)                                   */
/* ------------------------------------------------------------------------------------ */

/* Per-candidate parity list of one object ray: the first max_targets distinct targets
 * in hit order, with parity and hit counts (toggles + the addition). */
typedef struct {
    int32_t* ids;
    unsigned char* par;
    uint32_t* hits;
    int count, max;
    uint64_t additions, toggles, overflow;
} cands_t;

static void cands_hit(cands_t* c, int obj) {
    for (int q = 0; q < c->count; ++q)
        if (c->ids[q] == obj) {
            c->par[q] ^= 1;
            c->hits[q]++;
            c->toggles++;
            return;
        }
    if (c->count < c->max) {
        c->ids[c->count] = obj;
        c->par[c->count] = 1;
        c->hits[c->count] = 1;
        c->count++;
        c->additions++;
    } else {
        c->overflow++;
    }
}

/* Index of the first hit (in (t, tri) order) with t >= thr, or n. */
static int64_t first_at_or_above(const hitvec* hv, float thr) {
    int64_t lo = 0, hi = hv->n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (hv->h[mid].t < thr) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

int64_t orc_object_parity_ex(const float* srcV, const uint32_t* srcI, const int32_t* src_tri_obj,
                             int64_t src_nT, int64_t n_src_obj, const float* V, int64_t nV,
                             const uint32_t* I, int64_t nT, const int32_t* tri_obj, int max_targets,
                             int swap, orc_objray_opts* opt, uint64_t* out_keys, int64_t cap,
                             int64_t* overflow) {
    (void)nV;
    const float d[3] = {0.0f, 0.0f, 1.0f};
    orc_raypre rp;
    orc_ray_pre(d, &rp);
    const int mode = opt ? opt->chain_mode : 0;
    const int max_iter = opt ? opt->max_iter : 0;
    int32_t* first = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_src_obj > 0 ? n_src_obj : 1));
    for (int64_t o = 0; o < n_src_obj; ++o) first[o] = -1;
    for (int64_t t = 0; t < src_nT; ++t) {
        const int32_t o = src_tri_obj ? src_tri_obj[t] : -1;
        if (o >= 0 && o < n_src_obj && first[o] < 0) first[o] = (int32_t)t;
    }
    const int mt = max_targets > 0 ? max_targets : 1;
    cands_t c;
    c.ids = (int32_t*)malloc(sizeof(int32_t) * (size_t)mt);
    c.par = (unsigned char*)malloc((size_t)mt);
    c.hits = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)mt);
    c.max = max_targets;
    grid_t G;
    const int axes[2] = {0, 1};
    if (nT > 0) grid_build(&G, 2, axes, V, I, nT, 1.0);
    hitvec hv = {0, 0, 0};
    int64_t total = 0, over = 0;
    uint64_t* st = opt ? opt->stats : NULL;
    for (int64_t s = 0; s < n_src_obj; ++s) {
        const int has_origin = opt && opt->origins ? 1 : first[s] >= 0;
        if (!has_origin) continue;
        const float* o = opt && opt->origins ? opt->origins + 3 * s : srcV + 3 * (int64_t)srcI[3 * (int64_t)first[s]];
        hv.n = 0;
        if (nT > 0) {
            const int cu = cell_coord(&G, 0, o[0]), cv = cell_coord(&G, 1, o[1]);
            if (cu >= 0 && cu < G.g[0] && cv >= 0 && cv < G.g[1]) {
                const int64_t cc = cu + (int64_t)G.g[0] * cv;
                for (int64_t k = G.start[cc]; k < G.start[cc + 1]; ++k)
                    test_tri_collect(V, I, G.items[k], o, &rp, 1e-4f, 1e10f, 1, &hv);
            }
        }
        qsort(hv.h, (size_t)hv.n, sizeof(hit_t), hit_cmp);
        c.count = 0;
        c.additions = c.toggles = c.overflow = 0;
        uint64_t iters = 0, hits = 0, suppressed = 0;
        if (mode == 0) {
            /* all hits in (t, tri) order, a same-object hit within 1e-5 of the previous merged */
            int has_last = 0, last_obj = -1;
            float last_t = 0.0f;
            iters = 1;
            for (int64_t j = 0; j < hv.n; ++j) {
                const float t = hv.h[j].t;
                const int obj = tri_obj[hv.h[j].tri];
                const int same = has_last && obj == last_obj && fabsf(t - last_t) <= 1e-5f;
                has_last = 1;
                last_obj = obj;
                last_t = t;
                ++hits;
                if (same) {
                    ++suppressed;
                    continue;
                }
                cands_hit(&c, obj);
            }
        } else if (mode == 1) {
            /* closest-hit loop of mesh_intersection.cu:134-230: one trace per iteration over
             * [thr, 1e10], next threshold nextafterf(t, 1e10) */
            const float tmax = 1e10f;
            float thr = 1e-4f, last_t = -1.0f;
            uint32_t last_tri = 0xffffffffu;
            const int cap_it = max_iter > 0 ? max_iter : 2048;
            for (int it = 0; it < cap_it; ++it) {
                ++iters;
                if (thr > tmax) break;
                const int64_t j = first_at_or_above(&hv, thr);
                if (j == hv.n) break;
                const float t = hv.h[j].t;
                const uint32_t tri = hv.h[j].tri;
                ++hits;
                const int same = tri == last_tri && last_t >= 0.0f && fabsf(t - last_t) <= 1e-5f;
                if (same) ++suppressed;
                else cands_hit(&c, tri_obj[tri]);
                last_t = t;
                last_tri = tri;
                float nx = nextafterf(t, tmax);
                if (nx <= t) nx = t + 1e-6f;
                if (nx <= thr) nx = nextafterf(thr, tmax);
                if (nx <= thr) break;
                thr = nx;
            }
        } else {
            /* closest-hit loop of mesh_containment.cu:336-366: one trace per iteration over
             * [thr, 1e10], a hit at 1e10 ends it, next threshold fl(t + 1e-4) (the same hit
             * again when that add is absorbed) */
            const float tmax = 1e10f;
            float thr = 1e-4f;
            const int cap_it = max_iter > 0 ? max_iter : 2000;
            for (int it = 0; it < cap_it; ++it) {
                ++iters;
                const int64_t j = first_at_or_above(&hv, thr);
                if (j == hv.n) break;
                const float t = hv.h[j].t;
                if (t >= tmax) break;
                ++hits;
                cands_hit(&c, tri_obj[hv.h[j].tri]);
                thr = t + 1e-4f;
            }
        }
        int odd = 0;
        const int max_out = opt && opt->max_out > 0 ? opt->max_out : c.count;
        for (int q = 0; q < c.count; ++q) {
            if (!(c.par[q] & 1)) continue;
            if (odd >= max_out) break;
            if (opt && opt->pair_ids && odd < opt->max_pair) {
                opt->pair_ids[s * (int64_t)opt->max_pair + odd] = c.ids[q];
                opt->pair_hits[s * (int64_t)opt->max_pair + odd] = c.hits[q];
            }
            ++odd;
            const uint32_t a = (uint32_t)(swap ? c.ids[q] : (int32_t)s);
            const uint32_t b = (uint32_t)(swap ? (int32_t)s : c.ids[q]);
            if (total < cap) out_keys[total] = ((uint64_t)a << 32) | b;
            ++total;
        }
        if (opt && opt->pair_ids)
            for (int q = odd; q < opt->max_pair; ++q) {
                opt->pair_ids[s * (int64_t)opt->max_pair + q] = -1;
                opt->pair_hits[s * (int64_t)opt->max_pair + q] = 0;
            }
        if (opt && opt->iterations) {
            opt->iterations[s] = (uint32_t)iters;
            opt->cand_count[s] = (uint32_t)c.count;
            opt->cand_overflow[s] = (uint32_t)c.overflow;
        }
        if (st) {
            st[0] += 1;                  /* rays */
            st[1] += iters;              /* iterations */
            st[2] += hits;               /* hits */
            if (iters > st[3]) st[3] = iters;  /* max iterations per ray */
            st[4] += suppressed;
            st[5] += c.additions;
            st[6] += c.toggles;
            st[7] += c.overflow;
            st[8] += (uint64_t)(odd < 256 ? odd : 256);  /* targets (bounded as the reference) */
        }
        over += (int64_t)c.overflow;
    }
    free(hv.h);
    free(first);
    free(c.ids);
    free(c.par);
    free(c.hits);
    if (nT > 0) grid_free(&G);
    if (overflow) *overflow = over;
    return total;
}

int64_t orc_object_parity(const float* srcV, const uint32_t* srcI, const int32_t* src_tri_obj,
                          int64_t src_nT, int64_t n_src_obj, const float* V, int64_t nV,
                          const uint32_t* I, int64_t nT, const int32_t* tri_obj, int max_targets,
                          int swap, uint64_t* out_keys, int64_t cap, int64_t* overflow) {
    return orc_object_parity_ex(srcV, srcI, src_tri_obj, src_nT, n_src_obj, V, nV, I, nT, tri_obj,
                                max_targets, swap, NULL, out_keys, cap, overflow);
}
