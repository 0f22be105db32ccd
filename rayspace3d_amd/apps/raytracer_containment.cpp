// raytracer_containment -- containment join through MeshContainmentLauncher: overlap set
// from edge rays B->A and A->B, then strict containment of B objects (first vertex, built
// on the host) in A objects that do not overlap
// (query/src/applications/raytracer_containment_main.cpp:96-436). CSV and phase names as
// the reference; pairs are sorted by (a, b). A table insert that exhausts its probes is
// detected and the run repeated with twice the slots.
#include <algorithm>
#include <fstream>
#include <iostream>

#include "cli_common.hpp"

using namespace rs3d_host;

struct ContainmentOptions : rs3d_cli::MeshPairOptions {
    bool useAnyhitPointInMesh = true, includeOverlapPairs = false;
    ContainmentOptions() : MeshPairOptions("containment_timing.json") { allowNoExport = true; }
    bool parseApplicationOption(const std::string& a, int&, int, char**) override {
        if (a == "--use-anyhit-point-in-mesh") { useAnyhitPointInMesh = true; return true; }
        if (a == "--include-overlap-pairs") { includeOverlapPairs = true; return true; }
        return false;
    }
};

int main(int argc, char* argv[]) {
    PerformanceTimer timer;
    timer.start("Data Reading");
    ContainmentOptions opt;
    opt.parse(argc, argv);
    if (opt.helpRequested) {
        std::vector<rs3d_cli::HelpEntry> o;
        rs3d_cli::MeshPairOptions::meshPairHelp(o);
        rs3d_cli::MeshPairOptions::runHelp(o);
        o.emplace_back("--no-export", "Do not export results to CSV");
        o.emplace_back("--use-anyhit-point-in-mesh", "Use AnyHit shader accumulation for point-in-mesh parity (default: enabled)");
        o.emplace_back("--include-overlap-pairs", "Include overlap/touch pairs (union of overlap + containment)");
        rs3d_cli::printHelp(argv[0], "--mesh1 <A.pre> --mesh2 <B.pre> [options]",
                            "Containment join: pairs (a, b) where B object b lies strictly inside A object a.", o);
        return 0;
    }
    opt.sanitizeRunCounts();
    if (!opt.hasRequiredMeshInputs()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <path> --mesh2 <path> [options]" << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    DeviceContext ctx(0);
    timer.next("Load Mesh A");
    GeometryData ma = loadGeometryFromFile(opt.mesh1Path);
    if (ma.vertices.empty() || !requirePrecomputedEdges(ma, opt.mesh1Path, "Mesh A")) return 1;
    timer.next("Load Mesh B");
    GeometryData mb = loadGeometryFromFile(opt.mesh2Path);
    if (mb.vertices.empty() || !requirePrecomputedEdges(mb, opt.mesh2Path, "Mesh B")) return 1;
    timer.next("Upload Mesh A");
    GeometryUploader ua;
    ua.upload(ma);
    timer.next("Upload Mesh B");
    GeometryUploader ub;
    ub.upload(mb);
    timer.next("Build A Index");
    AccelerationStructure asa(ctx, ua);
    asa.build();
    timer.next("Build B Index");
    AccelerationStructure asb(ctx, ub);
    asb.build();
    timer.next("Prepare Kernel Parameters");
    EdgeMeshData ea = PrecomputedEdgeData::uploadFromGeometry(ma, ctx);
    EdgeMeshData eb = PrecomputedEdgeData::uploadFromGeometry(mb, ctx);
    // first vertex of each B object, on the host (raytracer_containment_main.cpp:195-211)
    const size_t numBObjects = countObjects(mb);
    std::vector<Vec3> bFirst(numBObjects, Vec3{0, 0, 0});
    {
        std::vector<bool> seen(numBObjects, false);
        for (size_t t = 0; t < mb.indices.size(); ++t) {
            const int o = mb.triangleToObject[t];
            if (o >= 0 && (size_t)o < numBObjects && !seen[o]) {
                bFirst[o] = mb.vertices[mb.indices[t].x];
                seen[o] = true;
            }
        }
    }
    DeviceBuffer dFirst(std::max<size_t>(numBObjects, 1) * sizeof(Vec3));
    if (numBObjects) dFirst.upload(bFirst.data(), numBObjects * sizeof(Vec3));
    MeshContainmentLauncher launcher(ctx);
    uint64_t slots = 16777216ull;  // 16 M slots per table (raytracer_containment_main.cpp:228-229)
    std::vector<uint64_t> contain, overlap;
    auto runOnce = [&]() {
        while (true) {
            DeviceHashTable iht, cht;
            iht.create(slots, false, 1000);
            cht.create(slots, false, 1000);
            iht.clear(ctx);
            cht.clear(ctx);
            MeshContainmentLaunchParams params;
            params.intersection_hash_table = &iht.table;
            params.containment_hash_table = &cht.table;
            params.b_first_vertices = dFirst.as<float>();
            params.b_num_objects = (int)numBObjects;
            params.use_anyhit_point_in_mesh = opt.useAnyhitPointInMesh ? 1 : 0;
            // phase 1a: B edges against A
            params.src_edges = &eb.view;
            params.target = asa.getHandle();
            params.swap_ids = 0;
            launcher.launchEdgeCheck(params, eb.num_edges);
            // phase 1b: A edges against B
            params.src_edges = &ea.view;
            params.target = asb.getHandle();
            params.swap_ids = 1;
            launcher.launchEdgeCheck(params, ea.num_edges);
            // phase 2: point in mesh against A
            params.target = asa.getHandle();
            launcher.launchPointInMesh(params, (int)numBObjects);
            if ((iht.failedInserts(ctx) > 0 || cht.failedInserts(ctx) > 0) && slots < (1ull << 30)) {
                slots *= 2;
                continue;
            }
            if (launcher.lastPointInMeshOverflow())
                std::cerr << "Warning: " << launcher.lastPointInMeshOverflow()
                          << " candidate objects beyond the 512-target cap" << std::endl;
            auto compact = [&](DeviceHashTable& t, std::vector<uint64_t>& keys) {
                DeviceBuffer out(slots * sizeof(uint64_t));
                uint64_t n = 0;
                check(rs3d_ht_compact(ctx.get(), &t.table, out.as<uint64_t>(), slots, &n), "rs3d_ht_compact");
                keys.resize(n);
                out.download(keys.data(), n * sizeof(uint64_t));
                std::sort(keys.begin(), keys.end());
            };
            compact(cht, contain);
            if (opt.includeOverlapPairs) compact(iht, overlap);
            return;
        }
    };
    timer.next("Warmup");
    for (int w = 0; w < opt.warmupRuns; ++w) runOnce();
    timer.next("Query");
    for (int r = 0; r < opt.numberOfRuns; ++r) runOnce();
    timer.next("Download Results");
    timer.next("Output");
    std::cout << "Containment pairs: " << contain.size() << std::endl;
    if (opt.includeOverlapPairs) std::cout << "Overlap pairs: " << overlap.size() << std::endl;
    if (opt.exportResults) {
        std::ofstream csv("mesh_containment_results.csv");
        if (opt.includeOverlapPairs) {
            csv << "a_object_id,b_object_id,pair_type\n";
            for (const auto& p : rs3d_cli::keysToPairs(contain))
                csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << ",containment\n";
            for (const auto& p : rs3d_cli::keysToPairs(overlap))
                csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << ",overlap\n";
        } else {
            csv << "a_object_id,b_object_id\n";
            for (const auto& p : rs3d_cli::keysToPairs(contain)) csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << "\n";
        }
    }
    timer.next("Cleanup");
    PrecomputedEdgeData::freeEdgeData(ea);
    PrecomputedEdgeData::freeEdgeData(eb);
    timer.finish(opt.outputJsonPath);
    return 0;
}
