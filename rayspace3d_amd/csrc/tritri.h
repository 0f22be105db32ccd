// tritri.h -- exact triangle-triangle intersection predicate on the device.
//
// Guigue & Devillers, "Fast and Robust Triangle-Triangle Overlap Test Using Orientation
// Predicates" (JGT 8(1), 2003), in the decision form of the reference's TriInt
// (preprocess/src/tdbase_lib/geometry/TriInt.cpp:400-492, the segment-construction
// variant): the answer depends only on the signs it tests, so the segment itself is not
// built. Closed triangles: touching and coplanar overlap count. Every operation is one
// fp32 IEEE operation in the published order (the library is built with
// -ffp-contract=off), so the result equals TriInt's bit for bit, degenerate inputs
// included. Comparisons are written as in the published code, which fixes where NaN
// inputs go.
#pragma once

#include <hip/hip_runtime.h>

namespace rs3d {

struct TV2 {
    float u, v;
};

__device__ __forceinline__ float3 tsub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 tcross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float tdot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Orientation of (a, b) about c.
__device__ __forceinline__ float torient(TV2 a, TV2 b, TV2 c) {
    return (a.u - c.u) * (b.v - c.v) - (a.v - c.v) * (b.u - c.u);
}

// 2D test, T2's first vertex in the region of one of T1's vertices.
__device__ __forceinline__ bool tt_vertex(TV2 P1, TV2 Q1, TV2 R1, TV2 P2, TV2 Q2, TV2 R2) {
    if (torient(R2, P2, Q1) >= 0.0f) {
        if (torient(R2, Q2, Q1) <= 0.0f) {
            if (torient(P1, P2, Q1) > 0.0f) return torient(P1, Q2, Q1) <= 0.0f;
            return torient(P1, P2, R1) >= 0.0f && torient(Q1, R1, P2) >= 0.0f;
        }
        return torient(P1, Q2, Q1) <= 0.0f && torient(R2, Q2, R1) <= 0.0f && torient(Q1, R1, Q2) >= 0.0f;
    }
    if (torient(R2, P2, R1) >= 0.0f) {
        if (torient(Q1, R1, R2) >= 0.0f) return torient(P1, P2, R1) >= 0.0f;
        return torient(Q1, R1, Q2) >= 0.0f && torient(R2, R1, Q2) >= 0.0f;
    }
    return false;
}

// 2D test, T2's first vertex in the region of one of T1's edges (R2 = T2's last vertex).
__device__ __forceinline__ bool tt_edge(TV2 P1, TV2 Q1, TV2 R1, TV2 P2, TV2 R2) {
    if (torient(R2, P2, Q1) >= 0.0f) {
        if (torient(P1, P2, Q1) >= 0.0f) return torient(P1, Q1, R2) >= 0.0f;
        return torient(Q1, R1, P2) >= 0.0f && torient(R1, P1, P2) >= 0.0f;
    }
    if (torient(R2, P2, R1) >= 0.0f)
        return torient(P1, P2, R1) >= 0.0f && (torient(P1, R1, R2) >= 0.0f || torient(Q1, R1, R2) >= 0.0f);
    return false;
}

// Both triangles counter-clockwise. The three orientations of p1 against T2's edges pick
// the region of p1; the region test runs once on T2's vertices rotated to match (the
// orientations are pure, so evaluating all three first changes no result).
__device__ __forceinline__ bool tt_ccw_2d(TV2 p1, TV2 q1, TV2 r1, TV2 p2, TV2 q2, TV2 r2) {
    const bool o1 = torient(p2, q2, p1) >= 0.0f;
    const bool o2 = torient(q2, r2, p1) >= 0.0f;
    const bool o3 = torient(r2, p2, p1) >= 0.0f;
    if (o1 && o2 && o3) return true;
    if (o1 + o2 + o3 == 2) {
        // edge region: (P2, R2) = (p2, r2), (r2, q2) or (q2, p2)
        const TV2 P2 = !o3 ? p2 : (!o2 ? r2 : q2);
        const TV2 R2 = !o3 ? r2 : (!o2 ? q2 : p2);
        return tt_edge(p1, q1, r1, P2, R2);
    }
    // vertex region: T2 as (p2, q2, r2), (q2, r2, p2) or (r2, p2, q2)
    const int rot = o1 ? 0 : (o2 ? 1 : 2);
    const TV2 A = rot == 0 ? p2 : (rot == 1 ? q2 : r2);
    const TV2 B = rot == 0 ? q2 : (rot == 1 ? r2 : p2);
    const TV2 C = rot == 0 ? r2 : (rot == 1 ? p2 : q2);
    return tt_vertex(p1, q1, r1, A, B, C);
}

__device__ __forceinline__ bool tt_overlap_2d(TV2 p1, TV2 q1, TV2 r1, TV2 p2, TV2 q2, TV2 r2) {
    // Clockwise triangles are turned counter-clockwise by exchanging their last two vertices.
    const bool cw1 = torient(p1, q1, r1) < 0.0f;
    const bool cw2 = torient(p2, q2, r2) < 0.0f;
    const TV2 b1 = cw1 ? r1 : q1, c1 = cw1 ? q1 : r1;
    const TV2 b2 = cw2 ? r2 : q2, c2 = cw2 ? q2 : r2;
    return tt_ccw_2d(p1, b1, c1, p2, b2, c2);
}

// Coplanar triangles: projection on the axis plane where T1's normal n1 is largest; on the
// YZ and XZ planes the first two vertices of each triangle trade places.
__device__ __forceinline__ TV2 tt_project(float3 p, int plane) {
    return plane == 0 ? TV2{p.z, p.y} : (plane == 1 ? TV2{p.x, p.z} : TV2{p.x, p.y});
}
__device__ __forceinline__ bool tt_coplanar(float3 p1, float3 q1, float3 r1, float3 p2, float3 q2, float3 r2,
                                            float3 n1) {
    const float nx = n1.x < 0 ? -n1.x : n1.x;
    const float ny = n1.y < 0 ? -n1.y : n1.y;
    const float nz = n1.z < 0 ? -n1.z : n1.z;
    const int plane = (nx > nz && nx >= ny) ? 0 : ((ny > nz && ny >= nx) ? 1 : 2);
    const bool trade = plane != 2;
    return tt_overlap_2d(tt_project(trade ? q1 : p1, plane), tt_project(trade ? p1 : q1, plane),
                         tt_project(r1, plane), tt_project(trade ? q2 : p2, plane),
                         tt_project(trade ? p2 : q2, plane), tt_project(r2, plane));
}

// Decision of the segment construction for the canonical form (p1 alone on its side of
// T2's plane, p2 alone on its side of T1's plane).
__device__ __forceinline__ bool tt_segment(float3 p1, float3 q1, float3 r1, float3 p2, float3 q2, float3 r2) {
    const float3 v = tsub(p2, p1);
    const float3 a = tsub(q1, p1);
    const float3 b = tsub(r2, p1);
    if (tdot(v, tcross(a, b)) > 0.0f) return tdot(v, tcross(tsub(r1, p1), b)) <= 0.0f;
    return !(tdot(v, tcross(a, tsub(q2, p1))) < 0.0f);
}

__device__ __forceinline__ float3 tsel(int k, float3 a, float3 b, float3 c) { return k == 0 ? a : (k == 1 ? b : c); }

// The canonical forms are selections of vertices: T1 rotated (rot1: 0 = p q r, 1 = q r p,
// 2 = r p q) with T2's last two vertices exchanged when T1's lone vertex is on the
// positive side (swap2); then T1's last two exchanged (flip1) and T2 rotated (rot2: 0 =
// P Q R, 1 = R P Q, 2 = Q R P). Both steps fall back to the coplanar test when every
// distance is zero. One segment test and one coplanar test run on the selected vertices.
__device__ __forceinline__ bool tri_tri_intersect(float3 p1, float3 q1, float3 r1, float3 p2, float3 q2,
                                                  float3 r2) {
    const float3 n2 = tcross(tsub(p2, r2), tsub(q2, r2));
    const float dp1 = tdot(tsub(p1, r2), n2), dq1 = tdot(tsub(q1, r2), n2), dr1 = tdot(tsub(r1, r2), n2);
    if (dp1 * dq1 > 0.0f && dp1 * dr1 > 0.0f) return false;
    const float3 n1 = tcross(tsub(q1, p1), tsub(r1, p1));
    const float dp2 = tdot(tsub(p2, r1), n1), dq2 = tdot(tsub(q2, r1), n1), dr2 = tdot(tsub(r2, r1), n1);
    if (dp2 * dq2 > 0.0f && dp2 * dr2 > 0.0f) return false;
    int rot1 = 0;
    bool swap2 = false, cop = false;
    if (dp1 > 0.0f) {
        if (dq1 > 0.0f) rot1 = 2, swap2 = true;
        else if (dr1 > 0.0f) rot1 = 1, swap2 = true;
    } else if (dp1 < 0.0f) {
        if (dq1 < 0.0f) rot1 = 2;
        else if (dr1 < 0.0f) rot1 = 1;
        else swap2 = true;
    } else if (dq1 < 0.0f) {
        if (dr1 >= 0.0f) rot1 = 1, swap2 = true;
    } else if (dq1 > 0.0f) {
        if (dr1 > 0.0f) swap2 = true;
        else rot1 = 1;
    } else {
        if (dr1 > 0.0f) rot1 = 2;
        else if (dr1 < 0.0f) rot1 = 2, swap2 = true;
        else cop = true;
    }
    const float3 P1 = tsel(rot1, p1, q1, r1), Q1 = tsel(rot1, q1, r1, p1), R1 = tsel(rot1, r1, p1, q1);
    const float3 P2 = p2, Q2 = swap2 ? r2 : q2, R2 = swap2 ? q2 : r2;
    const float DP = dp2, DQ = swap2 ? dr2 : dq2, DR = swap2 ? dq2 : dr2;
    int rot2 = 0;
    bool flip1 = false;
    if (!cop) {
        if (DP > 0.0f) {
            if (DQ > 0.0f) flip1 = true, rot2 = 1;
            else if (DR > 0.0f) flip1 = true, rot2 = 2;
        } else if (DP < 0.0f) {
            if (DQ < 0.0f) rot2 = 1;
            else if (DR < 0.0f) rot2 = 2;
            else flip1 = true;
        } else if (DQ < 0.0f) {
            if (DR >= 0.0f) flip1 = true, rot2 = 2;
        } else if (DQ > 0.0f) {
            if (DR > 0.0f) flip1 = true;
            else rot2 = 2;
        } else {
            if (DR > 0.0f) rot2 = 1;
            else if (DR < 0.0f) flip1 = true, rot2 = 1;
            else cop = true;
        }
    }
    // (when the outer step finds T1 coplanar, rot1 = 0 and swap2 = false: the original vertices)
    if (cop) return tt_coplanar(P1, Q1, R1, P2, Q2, R2, n1);
    return tt_segment(P1, flip1 ? R1 : Q1, flip1 ? Q1 : R1, tsel(rot2, P2, R2, Q2), tsel(rot2, Q2, P2, R2),
                      tsel(rot2, R2, Q2, P2));
}

}  // namespace rs3d
