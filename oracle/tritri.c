/*
 * tritri.c -- CPU restatement of the exact triangle-triangle predicate and the exact
 * intersection join built on it (test checker only; part of oracle/liboracle.so).
 *
 * TEST INFRASTRUCTURE: loaded only by tests/ and bench.py's cpu_baseline leg.
 *
 * Predicate: Guigue & Devillers, "Fast and Robust Triangle-Triangle Overlap Test Using
 * Orientation Predicates", JGT 8(1), 2003, in the decision form of the reference's
 * TriInt (preprocess/src/tdbase_lib/geometry/TriInt.cpp:400-492): the segment-construction
 * variant, whose result depends only on the orientation signs it tests. Closed triangles:
 * touching and coplanar overlap count. All arithmetic is fp32, one IEEE operation each,
 * in the order of the published formulas (-ffp-contract=off).
 *
 * Join: pairs (object of a triangle of A, object of a triangle of B) whose triangles'
 * closed bounding boxes overlap and which the predicate reports as intersecting.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    float x, y, z;
} v3;

static v3 ld(const float* p) {
    v3 r = {p[0], p[1], p[2]};
    return r;
}
static v3 sub(v3 a, v3 b) {
    v3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}
static v3 cross(v3 a, v3 b) {
    v3 r;
    r.x = a.y * b.z - a.z * b.y;
    r.y = a.z * b.x - a.x * b.z;
    r.z = a.x * b.y - a.y * b.x;
    return r;
}
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

typedef struct {
    float u, v;
} v2;

/* orientation of (a, b) about c */
static float orient(v2 a, v2 b, v2 c) { return (a.u - c.u) * (b.v - c.v) - (a.v - c.v) * (b.u - c.u); }

/* vertex region of T2's P2 (counter-clockwise triangles) */
static int test_vertex(v2 P1, v2 Q1, v2 R1, v2 P2, v2 Q2, v2 R2) {
    if (orient(R2, P2, Q1) >= 0.0f) {
        if (orient(R2, Q2, Q1) <= 0.0f) {
            if (orient(P1, P2, Q1) > 0.0f) return orient(P1, Q2, Q1) <= 0.0f;
            if (orient(P1, P2, R1) >= 0.0f) return orient(Q1, R1, P2) >= 0.0f;
            return 0;
        }
        if (orient(P1, Q2, Q1) <= 0.0f) {
            if (orient(R2, Q2, R1) <= 0.0f) return orient(Q1, R1, Q2) >= 0.0f;
            return 0;
        }
        return 0;
    }
    if (orient(R2, P2, R1) >= 0.0f) {
        if (orient(Q1, R1, R2) >= 0.0f) return orient(P1, P2, R1) >= 0.0f;
        if (orient(Q1, R1, Q2) >= 0.0f) return orient(R2, R1, Q2) >= 0.0f;
        return 0;
    }
    return 0;
}

/* edge region of T2's P2 */
static int test_edge(v2 P1, v2 Q1, v2 R1, v2 P2, v2 Q2, v2 R2) {
    (void)Q2;
    if (orient(R2, P2, Q1) >= 0.0f) {
        if (orient(P1, P2, Q1) >= 0.0f) return orient(P1, Q1, R2) >= 0.0f;
        if (orient(Q1, R1, P2) >= 0.0f) return orient(R1, P1, P2) >= 0.0f;
        return 0;
    }
    if (orient(R2, P2, R1) >= 0.0f) {
        if (orient(P1, P2, R1) >= 0.0f) {
            if (orient(P1, R1, R2) >= 0.0f) return 1;
            return orient(Q1, R1, R2) >= 0.0f;
        }
        return 0;
    }
    return 0;
}

static int ccw_overlap_2d(v2 p1, v2 q1, v2 r1, v2 p2, v2 q2, v2 r2) {
    if (orient(p2, q2, p1) >= 0.0f) {
        if (orient(q2, r2, p1) >= 0.0f) {
            if (orient(r2, p2, p1) >= 0.0f) return 1;
            return test_edge(p1, q1, r1, p2, q2, r2);
        }
        if (orient(r2, p2, p1) >= 0.0f) return test_edge(p1, q1, r1, r2, p2, q2);
        return test_vertex(p1, q1, r1, p2, q2, r2);
    }
    if (orient(q2, r2, p1) >= 0.0f) {
        if (orient(r2, p2, p1) >= 0.0f) return test_edge(p1, q1, r1, q2, r2, p2);
        return test_vertex(p1, q1, r1, q2, r2, p2);
    }
    return test_vertex(p1, q1, r1, r2, p2, q2);
}

static int overlap_2d(v2 p1, v2 q1, v2 r1, v2 p2, v2 q2, v2 r2) {
    const int cw1 = orient(p1, q1, r1) < 0.0f;
    const int cw2 = orient(p2, q2, r2) < 0.0f;
    if (cw1) return cw2 ? ccw_overlap_2d(p1, r1, q1, p2, r2, q2) : ccw_overlap_2d(p1, r1, q1, p2, q2, r2);
    return cw2 ? ccw_overlap_2d(p1, q1, r1, p2, r2, q2) : ccw_overlap_2d(p1, q1, r1, p2, q2, r2);
}

/* Coplanar case: project on the axis plane where T1's normal n1 is largest. On the YZ and
 * XZ planes the first two vertices of each triangle trade places. */
static int coplanar(v3 p1, v3 q1, v3 r1, v3 p2, v3 q2, v3 r2, v3 n1) {
    const float nx = n1.x < 0 ? -n1.x : n1.x;
    const float ny = n1.y < 0 ? -n1.y : n1.y;
    const float nz = n1.z < 0 ? -n1.z : n1.z;
    v2 P1, Q1, R1, P2, Q2, R2;
    if (nx > nz && nx >= ny) {
        P1 = (v2){q1.z, q1.y}; Q1 = (v2){p1.z, p1.y}; R1 = (v2){r1.z, r1.y};
        P2 = (v2){q2.z, q2.y}; Q2 = (v2){p2.z, p2.y}; R2 = (v2){r2.z, r2.y};
    } else if (ny > nz && ny >= nx) {
        P1 = (v2){q1.x, q1.z}; Q1 = (v2){p1.x, p1.z}; R1 = (v2){r1.x, r1.z};
        P2 = (v2){q2.x, q2.z}; Q2 = (v2){p2.x, p2.z}; R2 = (v2){r2.x, r2.z};
    } else {
        P1 = (v2){p1.x, p1.y}; Q1 = (v2){q1.x, q1.y}; R1 = (v2){r1.x, r1.y};
        P2 = (v2){p2.x, p2.y}; Q2 = (v2){q2.x, q2.y}; R2 = (v2){r2.x, r2.y};
    }
    return overlap_2d(P1, Q1, R1, P2, Q2, R2);
}

/* Decision of the segment construction: p1 is T1's vertex alone on its side of T2's
 * plane, p2 T2's vertex alone on its side of T1's plane (canonical form). */
static int segment_decision(v3 p1, v3 q1, v3 r1, v3 p2, v3 q2, v3 r2) {
    const v3 v = sub(p2, p1);
    const v3 a = sub(q1, p1);
    const v3 b = sub(r2, p1);
    if (dot(v, cross(a, b)) > 0.0f) return dot(v, cross(sub(r1, p1), b)) <= 0.0f;
    return !(dot(v, cross(a, sub(q2, p1))) < 0.0f);
}

/* T2's canonical form for a T1 already in canonical form (d* = signed distances of T2's
 * vertices to T1's plane, in the order given). */
static int second_form(v3 p1, v3 q1, v3 r1, v3 p2, v3 q2, v3 r2, float dp2, float dq2, float dr2,
                       v3 n1) {
    if (dp2 > 0.0f) {
        if (dq2 > 0.0f) return segment_decision(p1, r1, q1, r2, p2, q2);
        if (dr2 > 0.0f) return segment_decision(p1, r1, q1, q2, r2, p2);
        return segment_decision(p1, q1, r1, p2, q2, r2);
    }
    if (dp2 < 0.0f) {
        if (dq2 < 0.0f) return segment_decision(p1, q1, r1, r2, p2, q2);
        if (dr2 < 0.0f) return segment_decision(p1, q1, r1, q2, r2, p2);
        return segment_decision(p1, r1, q1, p2, q2, r2);
    }
    if (dq2 < 0.0f) {
        if (dr2 >= 0.0f) return segment_decision(p1, r1, q1, q2, r2, p2);
        return segment_decision(p1, q1, r1, p2, q2, r2);
    }
    if (dq2 > 0.0f) {
        if (dr2 > 0.0f) return segment_decision(p1, r1, q1, p2, q2, r2);
        return segment_decision(p1, q1, r1, q2, r2, p2);
    }
    if (dr2 > 0.0f) return segment_decision(p1, q1, r1, r2, p2, q2);
    if (dr2 < 0.0f) return segment_decision(p1, r1, q1, r2, p2, q2);
    return coplanar(p1, q1, r1, p2, q2, r2, n1);
}

int orc_tri_tri(const float* t1, const float* t2) {
    const v3 p1 = ld(t1), q1 = ld(t1 + 3), r1 = ld(t1 + 6);
    const v3 p2 = ld(t2), q2 = ld(t2 + 3), r2 = ld(t2 + 6);
    const v3 n2 = cross(sub(p2, r2), sub(q2, r2));
    const float dp1 = dot(sub(p1, r2), n2), dq1 = dot(sub(q1, r2), n2), dr1 = dot(sub(r1, r2), n2);
    if (dp1 * dq1 > 0.0f && dp1 * dr1 > 0.0f) return 0;
    const v3 n1 = cross(sub(q1, p1), sub(r1, p1));
    const float dp2 = dot(sub(p2, r1), n1), dq2 = dot(sub(q2, r1), n1), dr2 = dot(sub(r2, r1), n1);
    if (dp2 * dq2 > 0.0f && dp2 * dr2 > 0.0f) return 0;
    /* T1 to canonical form; T2 follows with its second and third vertices exchanged when
     * T1's lone vertex lies on the positive side. */
    if (dp1 > 0.0f) {
        if (dq1 > 0.0f) return second_form(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, n1);
        if (dr1 > 0.0f) return second_form(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, n1);
        return second_form(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, n1);
    }
    if (dp1 < 0.0f) {
        if (dq1 < 0.0f) return second_form(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, n1);
        if (dr1 < 0.0f) return second_form(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, n1);
        return second_form(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, n1);
    }
    if (dq1 < 0.0f) {
        if (dr1 >= 0.0f) return second_form(q1, r1, p1, p2, r2, q2, dp2, dr2, dq2, n1);
        return second_form(p1, q1, r1, p2, q2, r2, dp2, dq2, dr2, n1);
    }
    if (dq1 > 0.0f) {
        if (dr1 > 0.0f) return second_form(p1, q1, r1, p2, r2, q2, dp2, dr2, dq2, n1);
        return second_form(q1, r1, p1, p2, q2, r2, dp2, dq2, dr2, n1);
    }
    if (dr1 > 0.0f) return second_form(r1, p1, q1, p2, q2, r2, dp2, dq2, dr2, n1);
    if (dr1 < 0.0f) return second_form(r1, p1, q1, p2, r2, q2, dp2, dr2, dq2, n1);
    return coplanar(p1, q1, r1, p2, q2, r2, n1);
}

/* ---- join ---------------------------------------------------------------------------- */

typedef struct {
    float lo[3], hi[3];
    float v[9];
    int obj;
    int valid;
} tri_box;

static void make_boxes(const float* V, int64_t nV, const uint32_t* I, const int32_t* O, int64_t nT,
                       tri_box* out) {
    for (int64_t t = 0; t < nT; ++t) {
        tri_box* b = &out[t];
        b->valid = I[3 * t] < (uint64_t)nV && I[3 * t + 1] < (uint64_t)nV && I[3 * t + 2] < (uint64_t)nV;
        b->obj = O ? O[t] : (int32_t)t;
        if (!b->valid) continue;
        for (int k = 0; k < 3; ++k) memcpy(b->v + 3 * k, V + 3 * (uint64_t)I[3 * t + k], 12);
        for (int a = 0; a < 3; ++a) {
            float lo = b->v[a], hi = b->v[a];
            for (int k = 1; k < 3; ++k) {
                const float c = b->v[3 * k + a];
                lo = c < lo ? c : lo;
                hi = c > hi ? c : hi;
            }
            b->lo[a] = lo;
            b->hi[a] = hi;
        }
    }
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* Uniform grid over B's triangle boxes. Every A triangle visits the cells its box touches;
 * a candidate pair is tested only in the cell that holds the lower corner of the two
 * boxes' intersection, so each pair is tested once. */
typedef struct {
    double lo[3], inv;
    int64_t n[3];
    int64_t* start; /* cell -> first entry (n cells + 1) */
    int64_t* items; /* B triangle indices */
} tgrid;

static int64_t cell_of(const tgrid* g, int a, float x) {
    int64_t c = (int64_t)(((double)x - g->lo[a]) * g->inv);
    return c < 0 ? 0 : (c >= g->n[a] ? g->n[a] - 1 : c);
}

int64_t orc_tri_join(const float* Va, int64_t nVa, const uint32_t* Ia, const int32_t* Oa, int64_t nTa,
                     const float* Vb, int64_t nVb, const uint32_t* Ib, const int32_t* Ob, int64_t nTb,
                     uint64_t* out_keys, int64_t cap, int threads) {
    tri_box* A = (tri_box*)malloc(sizeof(tri_box) * (size_t)(nTa > 0 ? nTa : 1));
    tri_box* B = (tri_box*)malloc(sizeof(tri_box) * (size_t)(nTb > 0 ? nTb : 1));
    make_boxes(Va, nVa, Ia, Oa, nTa, A);
    make_boxes(Vb, nVb, Ib, Ob, nTb, B);
    tgrid g;
    double hi[3], ext = 0.0;
    int64_t nb = 0;
    for (int a = 0; a < 3; ++a) g.lo[a] = 1e300, hi[a] = -1e300;
    for (int64_t t = 0; t < nTb; ++t) {
        if (!B[t].valid) continue;
        ++nb;
        for (int a = 0; a < 3; ++a) {
            g.lo[a] = B[t].lo[a] < g.lo[a] ? B[t].lo[a] : g.lo[a];
            hi[a] = B[t].hi[a] > hi[a] ? B[t].hi[a] : hi[a];
            ext += (double)B[t].hi[a] - (double)B[t].lo[a];
        }
    }
    int64_t raw_cap = 1 << 16, raw_n = 0;
    uint64_t* raw = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)raw_cap);
    if (nb > 0) {
        /* cell edge: twice the mean box extent, and no more cells than 2 per triangle */
        double h = 2.0 * ext / (3.0 * (double)nb);
        const double vol = (hi[0] - g.lo[0] + 1e-30) * (hi[1] - g.lo[1] + 1e-30) * (hi[2] - g.lo[2] + 1e-30);
        const double hmin = cbrt(vol / (2.0 * (double)nb));
        h = h > hmin ? h : hmin;
        h = h > 0 ? h : 1.0;
        g.inv = 1.0 / h;
        int64_t ncell = 1;
        for (int a = 0; a < 3; ++a) {
            g.n[a] = (int64_t)((hi[a] - g.lo[a]) * g.inv) + 1;
            ncell *= g.n[a];
        }
        g.start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) {
                for (int64_t c = 0; c < ncell; ++c) g.start[c + 1] += g.start[c];
                g.items = (int64_t*)malloc(sizeof(int64_t) * (size_t)(g.start[ncell] > 0 ? g.start[ncell] : 1));
                for (int64_t c = ncell; c > 0; --c) g.start[c] = g.start[c - 1];
                g.start[0] = 0;
            }
            for (int64_t t = 0; t < nTb; ++t) {
                if (!B[t].valid) continue;
                int64_t c0[3], c1[3];
                for (int a = 0; a < 3; ++a) c0[a] = cell_of(&g, a, B[t].lo[a]), c1[a] = cell_of(&g, a, B[t].hi[a]);
                for (int64_t z = c0[2]; z <= c1[2]; ++z)
                    for (int64_t y = c0[1]; y <= c1[1]; ++y)
                        for (int64_t x = c0[0]; x <= c1[0]; ++x) {
                            const int64_t c = (z * g.n[1] + y) * g.n[0] + x;
                            if (pass == 0) g.start[c + 1]++;
                            else g.items[g.start[c + 1]++] = t;
                        }
            }
        }
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 0 ? threads : 1)
        for (int64_t i = 0; i < nTa; ++i) {
            const tri_box* a = &A[i];
            if (!a->valid) continue;
            int64_t c0[3], c1[3];
            for (int k = 0; k < 3; ++k) c0[k] = cell_of(&g, k, a->lo[k]), c1[k] = cell_of(&g, k, a->hi[k]);
            for (int64_t z = c0[2]; z <= c1[2]; ++z)
                for (int64_t y = c0[1]; y <= c1[1]; ++y)
                    for (int64_t x = c0[0]; x <= c1[0]; ++x) {
                        const int64_t c = (z * g.n[1] + y) * g.n[0] + x;
                        for (int64_t q = g.start[c]; q < g.start[c + 1]; ++q) {
                            const tri_box* b = &B[g.items[q]];
                            int ov = 1;
                            for (int k = 0; k < 3; ++k) ov &= a->lo[k] <= b->hi[k] && b->lo[k] <= a->hi[k];
                            if (!ov) continue;
                            /* test the pair only in the cell of its boxes' common lower corner */
                            const int64_t rx = cell_of(&g, 0, a->lo[0] > b->lo[0] ? a->lo[0] : b->lo[0]);
                            const int64_t ry = cell_of(&g, 1, a->lo[1] > b->lo[1] ? a->lo[1] : b->lo[1]);
                            const int64_t rz = cell_of(&g, 2, a->lo[2] > b->lo[2] ? a->lo[2] : b->lo[2]);
                            if (rx != x || ry != y || rz != z) continue;
                            if (!orc_tri_tri(a->v, b->v)) continue;
                            const uint64_t key = ((uint64_t)(uint32_t)a->obj << 32) | (uint32_t)b->obj;
#pragma omp critical(orc_tri_join_raw)
                            {
                                if (raw_n == raw_cap) {
                                    raw_cap *= 2;
                                    raw = (uint64_t*)realloc(raw, sizeof(uint64_t) * (size_t)raw_cap);
                                }
                                raw[raw_n++] = key;
                            }
                        }
                    }
        }
        free(g.start);
        free(g.items);
    }
    qsort(raw, (size_t)raw_n, sizeof(uint64_t), cmp_u64);
    int64_t u = 0;
    for (int64_t k = 0; k < raw_n; ++k)
        if (k == 0 || raw[k] != raw[k - 1]) {
            if (u < cap) out_keys[u] = raw[k];
            ++u;
        }
    free(raw);
    free(A);
    free(B);
    return u;
}

void orc_tri_tri_many(const float* A, const float* B, int64_t n, uint8_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = (uint8_t)orc_tri_tri(A + 9 * i, B + 9 * i);
}
