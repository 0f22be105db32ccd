// check_launchers -- drives every method of the intersection and containment launcher
// classes and the two selectivity estimators on two .pre datasets, as a caller of the
// reference's host interfaces would (MeshIntersectionLauncher.h:17-22,
// MeshContainmentLauncher.h:26-30, estimated_overlap.h:6, estimated_intersection.h:6), and
// writes what each produced to a text file (one "name count key..." line per pair set, keys
// (a << 32 | b) sorted; "stat name value" lines; "estimate name value" lines) for a test to
// compare with an independent computation.
//
//   check_launchers --mesh1 A.pre --mesh2 B.pre --out results.txt
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "cli_common.hpp"

using namespace rs3d_host;

namespace {

std::vector<uint64_t> tableKeys(DeviceContext& ctx, DeviceHashTable& t) {
    DeviceBuffer out(t.table.size * sizeof(uint64_t));
    uint64_t n = 0;
    check(rs3d_ht_compact(ctx.get(), &t.table, out.as<uint64_t>(), t.table.size, &n), "rs3d_ht_compact");
    std::vector<uint64_t> k(n);
    out.download(k.data(), n * sizeof(uint64_t));
    std::sort(k.begin(), k.end());
    return k;
}

std::vector<uint64_t> pairKeys(const std::vector<MeshQueryResult>& pairs) {
    std::vector<uint64_t> k;
    for (const auto& p : pairs) k.push_back(((uint64_t)(uint32_t)p.object_id_mesh1 << 32) | (uint32_t)p.object_id_mesh2);
    std::sort(k.begin(), k.end());
    k.erase(std::unique(k.begin(), k.end()), k.end());
    return k;
}

void put(std::ofstream& f, const char* name, const std::vector<uint64_t>& k) {
    f << name << ' ' << k.size();
    for (uint64_t x : k) f << ' ' << x;
    f << '\n';
}

// Two-pass form of one launch direction: pass 1 counts, exclusive_scan_gpu, pass 2 writes.
template <typename Launch>
std::vector<MeshQueryResult> twoPass(MeshIntersectionLaunchParams p, int n, Launch launch) {
    std::vector<MeshQueryResult> out;
    if (n <= 0) return out;
    DeviceBuffer counts((size_t)n * sizeof(int32_t)), offsets((size_t)n * sizeof(int64_t));
    p.hash_table = nullptr;
    p.pass = 1;
    p.collision_counts = counts.as<int32_t>();
    launch(p, n);
    const long long total = exclusive_scan_gpu(counts.as<int>(), offsets.as<long long>(), n);
    if (total <= 0) return out;
    DeviceBuffer res((size_t)total * sizeof(MeshQueryResult));
    p.pass = 2;
    p.collision_offsets = offsets.as<int64_t>();
    p.results = res.as<MeshQueryResult>();
    launch(p, n);
    out.resize((size_t)total);
    res.download(out.data(), out.size() * sizeof(MeshQueryResult));
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    std::string m1p, m2p, outp = "launchers.txt";
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string a = argv[i];
        if (a == "--mesh1") m1p = argv[i + 1];
        else if (a == "--mesh2") m2p = argv[i + 1];
        else if (a == "--out") outp = argv[i + 1];
    }
    if (m1p.empty() || m2p.empty()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <A.pre> --mesh2 <B.pre> [--out <path>]" << std::endl;
        return 1;
    }
    GeometryData m1 = loadGeometryFromFile(m1p), m2 = loadGeometryFromFile(m2p);
    if (m1.vertices.empty() || m2.vertices.empty() || !requirePrecomputedEdges(m1, m1p, "Mesh1") ||
        !requirePrecomputedEdges(m2, m2p, "Mesh2"))
        return 1;
    DeviceContext ctx(0);
    GeometryUploader u1, u2;
    u1.upload(m1);
    u2.upload(m2);
    AccelerationStructure as1(ctx, u1), as2(ctx, u2);
    as1.build();
    as2.build();
    EdgeMeshData e1 = PrecomputedEdgeData::uploadFromGeometry(m1, ctx);
    EdgeMeshData e2 = PrecomputedEdgeData::uploadFromGeometry(m2, ctx);
    const rs3d_mesh_view v1 = u1.view(), v2 = u2.view();
    const int n1 = (int)v1.num_objects, n2 = (int)v2.num_objects;
    std::ofstream f(outp);

    // ---- MeshIntersectionLauncher ----
    MeshIntersectionLauncher il(ctx);
    MeshIntersectionLaunchParams p1, p2;
    p1.mesh1 = v1;
    p1.edges = &e1.view;
    p1.mesh2 = as2.getHandle();
    p2.mesh1 = v2;
    p2.edges = &e2.view;
    p2.mesh2 = as1.getHandle();
    for (int anyhit = 1; anyhit >= 0; --anyhit) {
        DeviceHashTable ht;
        ht.create(16777217ull, false, 1000);  // odd size, modulo hashing as the intersection main sizes it
        ht.clear(ctx);
        DeviceBuffer stats(sizeof(rs3d_intersection_stats));
        stats.fill(0);
        for (MeshIntersectionLaunchParams* p : {&p1, &p2}) {
            p->hash_table = &ht.table;
            p->use_anyhit_containment = anyhit;
            p->profiling_stats = anyhit ? stats.as<rs3d_intersection_stats>() : nullptr;
        }
        il.launchMesh1ToMesh2(p1, e1.num_edges);
        il.launchOverlapMesh2ToMesh1(p2, e2.num_edges);
        il.launchContainmentMesh1ToMesh2(p1, n1);
        il.launchContainmentMesh2ToMesh1(p2, n2);
        if (ht.failedInserts(ctx)) {
            std::cerr << "pair table overflow" << std::endl;
            return 1;
        }
        put(f, anyhit ? "intersection_anyhit" : "intersection_closest", tableKeys(ctx, ht));
        if (anyhit) {
            rs3d_intersection_stats h{};
            stats.download(&h, sizeof h);
            f << "stat overlap_trace_calls " << h.overlap_trace_calls << "\nstat overlap_iterations_total "
              << h.overlap_iterations_total << "\nstat overlap_hits_total " << h.overlap_hits_total
              << "\nstat overlap_max_iterations_per_trace " << h.overlap_max_iterations_per_trace
              << "\nstat containment_rays_total " << h.containment_rays_total << "\nstat containment_iterations_total "
              << h.containment_iterations_total << "\nstat containment_hits_total " << h.containment_hits_total
              << "\nstat containment_candidate_additions " << h.containment_candidate_additions
              << "\nstat containment_candidate_toggles " << h.containment_candidate_toggles
              << "\nstat containment_targets_total " << h.containment_targets_total << '\n';
        }
    }
    for (MeshIntersectionLaunchParams* p : {&p1, &p2}) {
        p->use_anyhit_containment = 1;
        p->profiling_stats = nullptr;
    }
    {
        // overlap 1->2 and containment 1->2, each as a hash launch and as a two-pass launch
        DeviceHashTable ho, hc;
        ho.create(1u << 20, true, 0);
        hc.create(1u << 20, true, 0);
        ho.clear(ctx);
        hc.clear(ctx);
        MeshIntersectionLaunchParams q = p1;
        q.hash_table = &ho.table;
        il.launchOverlapMesh1ToMesh2(q, e1.num_edges);
        q.hash_table = &hc.table;
        il.launchContainmentMesh1ToMesh2(q, n1);
        put(f, "overlap12_hash", tableKeys(ctx, ho));
        put(f, "contain12_hash", tableKeys(ctx, hc));
        auto ov = twoPass(p1, e1.num_edges, [&](const MeshIntersectionLaunchParams& p, int n) {
            il.launchOverlapMesh1ToMesh2(p, n);
        });
        // the reference's dedup of raw pairs: sorted unique in the struct's (b, a) order
        const long long nu = (long long)ov.size();
        std::vector<MeshQueryResult> merged(ov.size());
        if (nu) {
            DeviceBuffer raw(ov.size() * sizeof(MeshQueryResult)), dst(ov.size() * sizeof(MeshQueryResult));
            raw.upload(ov.data(), ov.size() * sizeof(MeshQueryResult));
            const long long u = merge_and_deduplicate_pairs_gpu(raw.as<rs3d_pair>(), nu, nullptr, 0, dst.as<rs3d_pair>());
            merged.resize((size_t)u);
            dst.download(merged.data(), merged.size() * sizeof(MeshQueryResult));
        }
        bool ba_order = true;
        for (size_t i = 1; i < merged.size(); ++i) {
            const auto& x = merged[i - 1];
            const auto& y = merged[i];
            if (std::make_pair(x.object_id_mesh2, x.object_id_mesh1) >= std::make_pair(y.object_id_mesh2, y.object_id_mesh1))
                ba_order = false;
        }
        f << "merged_in_ba_order " << (ba_order ? 1 : 0) << '\n';
        put(f, "overlap12_twopass", pairKeys(merged));
        auto ct = twoPass(p1, n1, [&](const MeshIntersectionLaunchParams& p, int n) {
            il.launchContainmentMesh1ToMesh2(p, n);
        });
        put(f, "contain12_twopass", pairKeys(ct));
    }

    // ---- MeshContainmentLauncher ----
    MeshContainmentLauncher cl(ctx);
    const size_t nB = countObjects(m2);
    std::vector<Vec3> first(nB, Vec3{0, 0, 0});
    std::vector<bool> seen(nB, false);
    for (size_t t = 0; t < m2.indices.size(); ++t) {
        const int o = m2.triangleToObject[t];
        if (o >= 0 && (size_t)o < nB && !seen[o]) {
            first[o] = m2.vertices[m2.indices[t].x];
            seen[o] = true;
        }
    }
    DeviceBuffer dFirst(std::max<size_t>(nB, 1) * sizeof(Vec3));
    if (nB) dFirst.upload(first.data(), nB * sizeof(Vec3));
    for (int anyhit = 1; anyhit >= 0; --anyhit) {
        DeviceHashTable iht, cht;
        iht.create(16777216ull, false, 1000);
        cht.create(16777216ull, false, 1000);
        iht.clear(ctx);
        cht.clear(ctx);
        MeshContainmentLaunchParams p;
        p.intersection_hash_table = &iht.table;
        p.containment_hash_table = &cht.table;
        p.b_first_vertices = dFirst.as<float>();
        p.b_num_objects = (int)nB;
        p.use_anyhit_point_in_mesh = anyhit;
        p.src_edges = &e2.view;  // mesh2 = B, mesh1 = A
        p.target = as1.getHandle();
        p.swap_ids = 0;
        cl.launchEdgeCheck(p, e2.num_edges);
        p.src_edges = &e1.view;
        p.target = as2.getHandle();
        p.swap_ids = 1;
        cl.launchEdgeCheck(p, e1.num_edges);
        p.target = as1.getHandle();
        cl.launchPointInMesh(p, (int)nB);
        put(f, anyhit ? "containment_anyhit" : "containment_closest", tableKeys(ctx, cht));
        if (anyhit) put(f, "containment_overlap", tableKeys(ctx, iht));
    }

    // ---- estimators ----
    if (m1.grid.hasGrid && m2.grid.hasGrid) {
        const float w = m1.grid.maxBound.x - m1.grid.minBound.x, h = m1.grid.maxBound.y - m1.grid.minBound.y,
                    d = m1.grid.maxBound.z - m1.grid.minBound.z;
        const float cv = (w / m1.grid.resolution[0]) * (h / m1.grid.resolution[1]) * (d / m1.grid.resolution[2]);
        const int nc = (int)m1.grid.cells.size();
        f << std::setprecision(9) << "estimate overlap_raw "
          << estimateOverlapSelectivity(m1.grid.cells.data(), m2.grid.cells.data(), nc, cv) << '\n'
          << "estimate intersection_raw "
          << estimateIntersectionSelectivity(m1.grid.cells.data(), m2.grid.cells.data(), nc, cv, 0.001f, 0.8f) << '\n'
          << "estimate empty " << estimateOverlapSelectivity(m1.grid.cells.data(), m2.grid.cells.data(), 0, cv) << '\n'
          << "estimate pairs " << estimatePairs(ctx, m1, m2, 0.001f, 0.8f, true, false) << '\n';
    }
    PrecomputedEdgeData::freeEdgeData(e1);
    PrecomputedEdgeData::freeEdgeData(e2);
    std::cout << "check_launchers: results in " << outp << std::endl;
    return 0;
}
