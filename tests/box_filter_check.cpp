// Host check of the dataset-box filter: box_filter_pass (the margin-guarded classification,
// falling back to the triangle tests) must equal box_filter_full (the triangle tests) on
// every point. Boxes of many scales (flat and degenerate ones too); points uniform around
// the box, on and near its faces (multiples of half the margin, and up to 20 ulps), and on
// and near the faces' diagonal, for all three axes. Built by test_box_filter.py with hipcc
// as host code: usage box_filter_check <boxes>.
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <random>
#include "traverse.h"
using namespace rs3d;
static std::mt19937_64 rng(12345);
static float uf(float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); }
static float nudge(float x, int k) { for (int i = 0; i < std::abs(k); ++i) x = nextafterf(x, k > 0 ? INFINITY : -INFINITY); return x; }
template <int AX> long run(const FBox& b, long& fast, long& mism, long n) {
    constexpr int U = (AX + 1) % 3, W = (AX + 2) % 3;
    const float lo[3] = {b.x0, b.y0, b.z0}, hi[3] = {b.x1, b.y1, b.z1};
    float mag = 0; for (int k = 0; k < 3; ++k) mag = fmaxf(mag, fmaxf(fabsf(lo[k]), fabsf(hi[k])));
    const float m = mag * 1.52587890625e-5f + 1e-5f;
    long cnt = 0;
    for (long i = 0; i < n; ++i) {
        float p[3];
        for (int k = 0; k < 3; ++k) {
            const float e = hi[k] - lo[k];
            int mode = rng() % 10;
            float v;
            if (mode < 3) v = uf(lo[k] - 0.2f * e - 3 * m, hi[k] + 0.2f * e + 3 * m);
            else if (mode == 3) v = lo[k];
            else if (mode == 4) v = hi[k];
            else if (mode == 5) v = lo[k] + (float)((int)(rng() % 9) - 4) * 0.5f * m;
            else if (mode == 6) v = hi[k] + (float)((int)(rng() % 9) - 4) * 0.5f * m;
            else if (mode == 7) v = nudge(rng() % 2 ? lo[k] : hi[k], (int)(rng() % 41) - 20);
            else v = uf(lo[k], hi[k]);
            p[k] = v;
        }
        if (rng() % 4 == 0) {  // on / near the diagonal of the faces
            const float s = uf(-0.2f, 1.2f);
            p[U] = lo[U] + s * (hi[U] - lo[U]);
            p[W] = lo[W] + s * (hi[W] - lo[W]);
            if (rng() % 2) p[W] = nudge(p[W], (int)(rng() % 2001) - 1000);
            if (rng() % 3 == 0) p[U] = p[U] + (float)((int)(rng() % 9) - 4) * (hi[U] - lo[U]) * 1e-5f;
        }
        const float3 o = make_float3(p[0], p[1], p[2]);
        const bool full = box_filter_full<AX>(o, b);
        const bool got = box_filter_pass<AX>(o, b);
        // fast path used? (replica of the guard)
        {
            const float u = p[U], w = p[W], a = p[AX];
            const float mg = fmaxf(mag, fmaxf(fmaxf(fabsf(u), fabsf(w)), fabsf(a)));
            if (mg < 1e12f) {
                const float mm = mg * 1.52587890625e-5f + 1e-5f;
                const float eu = hi[U] - lo[U], ew = hi[W] - lo[W];
                const float dg = (u - lo[U]) * ew - (w - lo[W]) * eu;
                if (fabsf(dg) > mg * mg * 6.103515625e-5f && eu > 2 * mm && ew > 2 * mm && hi[AX] - lo[AX] > 2 * mm) {
                    const bool in_sq = u > lo[U] + mm && u < hi[U] - mm && w > lo[W] + mm && w < hi[W] - mm;
                    const bool out_sq = u < lo[U] - mm || u > hi[U] + mm || w < lo[W] - mm || w > hi[W] + mm;
                    if (out_sq || (in_sq && ((a > lo[AX] + mm && a < hi[AX] - mm) || a > hi[AX] + mm || a < lo[AX] - mm))) ++fast;
                }
            }
        }
        if (got != full) {
            if (mism < 10) printf("MISMATCH ax %d box (%g %g %g)-(%g %g %g) p (%.9g %.9g %.9g) full %d got %d\n", AX,
                                  b.x0, b.y0, b.z0, b.x1, b.y1, b.z1, p[0], p[1], p[2], full, got);
            ++mism;
        }
        ++cnt;
    }
    return cnt;
}
int main(int argc, char** argv) {
    const int nbox = argc > 1 ? atoi(argv[1]) : 6000;
    long total = 0, fast = 0, mism = 0;
    const float scales[] = {1e-3f, 1.0f, 7.5f, 1000.0f, 3e5f, 1e9f, 2e11f, 4.5e11f};
    for (int t = 0; t < nbox; ++t) {
        const float sc = scales[t % 8];
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            const float c = uf(-2.0f, 2.0f) * sc * (t % 3 == 0 ? 10.0f : 1.0f);
            float e = uf(0.0f, 1.0f) * sc;
            if (t % 11 == 0 && k == t % 3) e = e * 1e-6f;   // flat boxes
            if (t % 13 == 0 && k == (t + 1) % 3) e = 0.0f;  // degenerate
            lo[k] = c; hi[k] = c + e;
        }
        if (t % 17 == 0) { lo[0] = 0; lo[1] = 0; lo[2] = 0; }
        const FBox b{lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
        total += run<0>(b, fast, mism, 3000);
        total += run<1>(b, fast, mism, 3000);
        total += run<2>(b, fast, mism, 3000);
    }
    printf("points %ld fast-path decided %ld mismatches %ld\n", total, fast, mism);
    return mism != 0;
}
