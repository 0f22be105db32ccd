// raytracer_containment -- containment join: overlap set from edge rays B->A and A->B,
// then strict containment of B objects (first vertex) in A objects that do not overlap
// (query/src/applications/raytracer_containment_main.cpp:96-436). CSV and phase names
// as the reference; pairs are sorted by (a, b).
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
        o.emplace_back("--use-anyhit-point-in-mesh", "All-hits parity for point-in-mesh (the only mode)");
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
    const rs3d_mesh_view va = ua.view(), vb = ub.view();
    std::vector<uint64_t> contain, overlap;
    auto runOnce = [&]() {
        uint64_t cap = 1u << 20;
        while (true) {
            DeviceBuffer c(cap * sizeof(uint64_t)), o(cap * sizeof(uint64_t));
            uint64_t nc = 0, no = 0, over = 0;
            const int rc = rs3d_containment_join(ctx.get(), &va, asa.getHandle(), &ea.view, &vb, asb.getHandle(),
                                                 &eb.view, 0, c.as<uint64_t>(), cap, &nc, o.as<uint64_t>(), cap,
                                                 &no, &over);
            if (rc == RS3D_E_CAPACITY) {
                cap = std::max(nc, no);
                continue;
            }
            check(rc, "rs3d_containment_join");
            contain.resize(nc);
            overlap.resize(no);
            c.download(contain.data(), nc * sizeof(uint64_t));
            o.download(overlap.data(), no * sizeof(uint64_t));
            if (over) std::cerr << "Warning: " << over << " candidate objects beyond the 512-target cap" << std::endl;
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
