// raytracer_mesh_overlap -- edge-ray mesh overlap join, two-pass count/scan/write with GPU
// sort/unique (query/src/applications/rayspace_mesh_overlap_main.cpp:42-146, 275-495).
// Output order is the reference's: pairs sorted as the little-endian u64 of the struct,
// i.e. by (object_id_mesh2, object_id_mesh1). --exact replaces the edge rays with the exact
// triangle-triangle intersection join (rs3d_tri_join), which also reports touching and
// coplanar contacts and needs no edge payload.
#include <algorithm>
#include <chrono>
#include <fstream>
#include <iostream>
#include <set>

#include "cli_common.hpp"

using namespace rs3d_host;

struct QueryResults {
    DeviceBuffer merged;
    long long numUnique = 0;
};

struct OverlapOptions : rs3d_cli::MeshPairOptions {
    bool exact = false;
    OverlapOptions() : rs3d_cli::MeshPairOptions("mesh_overlap_timing.json") {}
    bool parseApplicationOption(const std::string& a, int&, int, char**) override {
        if (a == "--exact") {
            exact = true;
            return true;
        }
        return false;
    }
};

// Exact join: sorted unique (mesh1, mesh2) keys, as pairs in the output order.
static QueryResults exactJoin(DeviceContext& ctx, AccelerationStructure& as1, AccelerationStructure& as2,
                              bool verbose) {
    uint64_t count = 0;
    int rc = rs3d_tri_join(ctx.get(), as1.getHandle(), as2.getHandle(), nullptr, 0, &count);
    if (rc != RS3D_E_CAPACITY) check(rc, "rs3d_tri_join");
    QueryResults q;
    if (count) {
        DeviceBuffer keys(count * sizeof(uint64_t));
        check(rs3d_tri_join(ctx.get(), as1.getHandle(), as2.getHandle(), keys.as<uint64_t>(), count, &count),
              "rs3d_tri_join");
        std::vector<uint64_t> k(count);
        keys.download(k.data(), count * sizeof(uint64_t));
        std::vector<MeshQueryResult> pairs = rs3d_cli::keysToPairs(k);
        std::sort(pairs.begin(), pairs.end(), [](const MeshQueryResult& x, const MeshQueryResult& y) {
            return x.object_id_mesh2 != y.object_id_mesh2 ? x.object_id_mesh2 < y.object_id_mesh2
                                                          : x.object_id_mesh1 < y.object_id_mesh1;
        });
        q.merged.alloc(count * sizeof(MeshQueryResult));
        q.merged.upload(pairs.data(), count * sizeof(MeshQueryResult));
    }
    q.numUnique = (long long)count;
    if (verbose) std::cout << "Exact triangle join: Found " << q.numUnique << " unique object pairs." << std::endl;
    return q;
}

static long long us_since(std::chrono::high_resolution_clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::high_resolution_clock::now() - t0).count();
}

static QueryResults twoPass(DeviceContext& ctx, MeshOverlapEdgesLauncher& L, MeshOverlapEdgesLaunchParams& p1,
                            MeshOverlapEdgesLaunchParams& p2, int n1, int n2, PerformanceTimer* timer, bool verbose) {
    DeviceBuffer c1(std::max(n1, 1) * sizeof(int32_t)), o1(std::max(n1, 1) * sizeof(int64_t));
    DeviceBuffer c2(std::max(n2, 1) * sizeof(int32_t)), o2(std::max(n2, 1) * sizeof(int64_t));
    p1.collision_counts = c1.as<int32_t>();
    p1.pass = 1;
    auto t0 = std::chrono::high_resolution_clock::now();
    L.launchMesh1ToMesh2(p1, n1);
    if (timer) timer->addMeasurement("Raytrace_Mesh1ToMesh2_Pass1", us_since(t0));
    int64_t total1 = 0, total2 = 0;
    check(rs3d_scan_i32_i64(ctx.get(), c1.as<int32_t>(), o1.as<int64_t>(), n1, &total1), "scan");
    p2.collision_counts = c2.as<int32_t>();
    p2.pass = 1;
    t0 = std::chrono::high_resolution_clock::now();
    L.launchMesh2ToMesh1(p2, n2);
    if (timer) timer->addMeasurement("Raytrace_Mesh2ToMesh1_Pass1", us_since(t0));
    check(rs3d_scan_i32_i64(ctx.get(), c2.as<int32_t>(), o2.as<int64_t>(), n2, &total2), "scan");
    if (verbose)
        std::cout << "Pass 1: Found " << total1 << " (edges) + " << total2 << " (edges mesh2->mesh1) = "
                  << (total1 + total2) << " potential overlaps." << std::endl;
    QueryResults q;
    const long long all = total1 + total2;
    if (all > 0) q.merged.alloc(all * sizeof(MeshQueryResult));
    p1.collision_offsets = o1.as<int64_t>();
    p1.results = q.merged.as<MeshQueryResult>();
    p1.pass = 2;
    t0 = std::chrono::high_resolution_clock::now();
    L.launchMesh1ToMesh2(p1, n1);
    if (timer) timer->addMeasurement("Raytrace_Mesh1ToMesh2_Pass2", us_since(t0));
    p2.collision_offsets = o2.as<int64_t>();
    p2.results = all > 0 ? q.merged.as<MeshQueryResult>() + total1 : nullptr;
    p2.pass = 2;
    t0 = std::chrono::high_resolution_clock::now();
    L.launchMesh2ToMesh1(p2, n2);
    if (timer) timer->addMeasurement("Raytrace_Mesh2ToMesh1_Pass2", us_since(t0));
    q.numUnique = all > 0 ? merge_and_deduplicate_pairs_gpu(nullptr, all, nullptr, 0, q.merged.as<MeshQueryResult>()) : 0;
    if (verbose) std::cout << "Deduplication: Found " << q.numUnique << " unique object pairs." << std::endl;
    return q;
}

int main(int argc, char* argv[]) {
    PerformanceTimer timer;
    timer.start("Data Reading");
    OverlapOptions opt;
    opt.allowNoExport = true;
    opt.parse(argc, argv);
    if (opt.helpRequested) {
        std::vector<rs3d_cli::HelpEntry> o;
        rs3d_cli::MeshPairOptions::meshPairHelp(o);
        rs3d_cli::MeshPairOptions::runHelp(o);
        o.emplace_back("--no-export", "Do not export results to CSV");
        o.emplace_back("--exact", "Exact triangle-triangle intersection join instead of edge rays");
        rs3d_cli::printHelp(argv[0], "--mesh1 <path> --mesh2 <path> [options]",
                            "Mesh overlap join with edge rays in both directions.", o);
        return 0;
    }
    opt.sanitizeRunCounts();
    if (!opt.hasRequiredMeshInputs()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <path> --mesh2 <path> [options]" << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    DeviceContext ctx(0);
    timer.next("Load Mesh1");
    GeometryData m1 = loadGeometryFromFile(opt.mesh1Path);
    if (m1.vertices.empty() || (!opt.exact && !requirePrecomputedEdges(m1, opt.mesh1Path, "Mesh1"))) return 1;
    timer.next("Load Mesh2");
    GeometryData m2 = loadGeometryFromFile(opt.mesh2Path);
    if (m2.vertices.empty() || (!opt.exact && !requirePrecomputedEdges(m2, opt.mesh2Path, "Mesh2"))) return 1;
    timer.next("Upload Mesh1");
    GeometryUploader u1;
    u1.upload(m1);
    timer.next("Upload Mesh2");
    GeometryUploader u2;
    u2.upload(m2);
    timer.next("Build Mesh1 Index");
    AccelerationStructure as1(ctx, u1);
    as1.build();
    timer.next("Build Mesh2 Index");
    AccelerationStructure as2(ctx, u2);
    as2.build();
    EdgeMeshData e1, e2;
    if (!opt.exact) {
        timer.next("Upload Mesh1");
        e1 = PrecomputedEdgeData::uploadFromGeometry(m1, ctx);
        timer.next("Upload Mesh2");
        e2 = PrecomputedEdgeData::uploadFromGeometry(m2, ctx);
    }
    timer.next("Create Launcher");
    MeshOverlapEdgesLauncher L(ctx);
    timer.next("Prepare Kernel Parameters");
    MeshOverlapEdgesLaunchParams p1, p2;
    p1.edges = &e1.view;
    p1.target = as2.getHandle();
    p1.swap_pair_order = 0;
    p2.edges = &e2.view;
    p2.target = as1.getHandle();
    p2.swap_pair_order = 1;
    timer.next("Warmup");
    for (int w = 0; w < opt.warmupRuns; ++w) {
        if (opt.exact) (void)exactJoin(ctx, as1, as2, false);
        else (void)twoPass(ctx, L, p1, p2, e1.num_edges, e2.num_edges, nullptr, false);
    }
    timer.next("Query");
    QueryResults q;
    for (int r = 0; r < opt.numberOfRuns; ++r) {
        if (opt.exact) q = exactJoin(ctx, as1, as2, true);
        else q = twoPass(ctx, L, p1, p2, e1.num_edges, e2.num_edges, r + 1 == opt.numberOfRuns ? &timer : nullptr, true);
    }
    timer.next("Download Results");
    std::vector<MeshQueryResult> uniq(q.numUnique);
    q.merged.download(uniq.data(), q.numUnique * sizeof(MeshQueryResult));
    timer.next("Output");
    std::set<int> o1(m1.triangleToObject.begin(), m1.triangleToObject.end());
    std::set<int> o2(m2.triangleToObject.begin(), m2.triangleToObject.end());
    std::cout << "\n=== Mesh Overlap Join Summary ===\nMesh1 triangles: " << u1.getNumTriangles()
              << "\nMesh1 objects: " << o1.size() << "\nMesh2 triangles: " << u2.getNumTriangles()
              << "\nMesh2 objects: " << o2.size() << "\nUnique object pairs: " << q.numUnique << std::endl;
    if (opt.exportResults) {
        std::ofstream csv("mesh_overlap_results.csv");
        csv << "object_id_mesh1,object_id_mesh2\n";
        for (const auto& r : uniq) csv << r.object_id_mesh1 << "," << r.object_id_mesh2 << "\n";
    }
    timer.next("Cleanup");
    PrecomputedEdgeData::freeEdgeData(e1);
    PrecomputedEdgeData::freeEdgeData(e2);
    timer.finish(opt.outputJsonPath);
    return 0;
}
