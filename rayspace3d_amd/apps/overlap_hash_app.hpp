// overlap_hash_app.hpp -- the estimated overlap executables: grid-histogram selectivity
// estimate -> power-of-two table size -> edge rays in both directions inserted into one
// pair hash table -> scan-based compaction. Follows
// query/src/applications/raytracer_overlap_estimated_main.cpp:38-485 and, with `direct`,
// raytracer_overlap_direct_estimation_main.cpp:223-685 (query direction, pair CSV, table
// size override, alpha switch). An insert that exhausts the table is detected and the query
// is rerun with a table twice as large, so the pair set is always complete.
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <fstream>
#include <iostream>
#include <set>
#include <string>

#include "cli_common.hpp"

namespace rs3d_cli {

struct OverlapHashOptions : MeshPairOptions {
    bool estimateOnly = false, trackHashContention = false, useAlphaCorrection = true, valid = true;
    float gamma = 0.8f, epsilon = 0.001f, hashTableFreeMemFraction = 0.0f;
    int direction = 0;  // 0 both, 1 mesh1_to_mesh2, 2 mesh2_to_mesh1
    std::string pairsOutputPath;
    unsigned long long manualHashTableSize = 0;
    bool direct = false;
    OverlapHashOptions() : MeshPairOptions("estimated_overlap_timing.json") {}
    bool parseApplicationOption(const std::string& a, int& i, int argc, char** argv) override {
        if (a == "--gamma" && i + 1 < argc) { gamma = std::stof(argv[++i]); return true; }
        if (a == "--epsilon" && i + 1 < argc) { epsilon = std::stof(argv[++i]); return true; }
        if (a == "--estimate-only") { estimateOnly = true; return true; }
        if (!direct) return false;
        if (a == "--query-direction" && i + 1 < argc) {
            const std::string d = argv[++i];
            if (d == "both") direction = 0;
            else if (d == "mesh1_to_mesh2") direction = 1;
            else if (d == "mesh2_to_mesh1") direction = 2;
            else {
                std::cerr << "Error: Invalid --query-direction value: " << d
                          << ". Expected one of: both, mesh1_to_mesh2, mesh2_to_mesh1" << std::endl;
                valid = false;
            }
            return true;
        }
        if (a == "--pairs-output" && i + 1 < argc) { pairsOutputPath = argv[++i]; return true; }
        if (a == "--track-hash-contention") { trackHashContention = true; return true; }
        if (a == "--no-alpha-correction") { useAlphaCorrection = false; return true; }
        if (a == "--hash-table-size" && i + 1 < argc) { manualHashTableSize = std::stoull(argv[++i]); return true; }
        if (a == "--hash-table-free-mem-fraction" && i + 1 < argc) {
            hashTableFreeMemFraction = std::stof(argv[++i]);
            return true;
        }
        return false;
    }
};

inline int runOverlapHash(int argc, char** argv, bool direct) {
    using namespace rs3d_host;
    PerformanceTimer timer;
    OverlapHashOptions opt;
    opt.direct = direct;
    opt.parse(argc, argv);
    if (opt.helpRequested) {
        std::vector<HelpEntry> o;
        MeshPairOptions::meshPairHelp(o);
        MeshPairOptions::runHelp(o);
        o.emplace_back("--gamma <float>", "Gamma parameter for estimation (default: 0.8)");
        o.emplace_back("--epsilon <float>", "Epsilon parameter for estimation (default: 0.001)");
        if (direct) {
            o.emplace_back("--query-direction <d>", "both|mesh1_to_mesh2|mesh2_to_mesh1 (default: both)");
            o.emplace_back("--pairs-output <path>", "Optional CSV export path for unique result pairs");
            o.emplace_back("--track-hash-contention", "Accepted; contention counters are not collected");
            o.emplace_back("--no-alpha-correction", "Disable the replication-factor alpha correction");
            o.emplace_back("--hash-table-size <ull>", "Override hash table size (slots); 0 = auto");
            o.emplace_back("--hash-table-free-mem-fraction <f>", "Size the table from a fraction of free HBM");
        }
        o.emplace_back("--estimate-only", "Only run selectivity estimation, skip actual query");
        printHelp(argv[0], "--mesh1 <path> --mesh2 <path> [options]",
                  "Overlap query: selectivity estimation plus hash-based deduplication.", o);
        return 0;
    }
    if (!opt.valid) return 1;
    opt.sanitizeRunCounts();
    if (!opt.hasRequiredMeshInputs()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <path> --mesh2 <path> [options]" << std::endl;
        return 1;
    }
    if (opt.hashTableFreeMemFraction < 0.0f || opt.hashTableFreeMemFraction > 1.0f) {
        std::cerr << "Error: --hash-table-free-mem-fraction must be in [0, 1]." << std::endl;
        return 1;
    }
    timer.start("Load Mesh1");
    GeometryData m1 = loadGeometryFromFile(opt.mesh1Path);
    if (m1.vertices.empty()) { std::cerr << "Error loading mesh1." << std::endl; return 1; }
    if (!requirePrecomputedEdges(m1, opt.mesh1Path, "Mesh1")) return 1;
    timer.next("Load Mesh2");
    GeometryData m2 = loadGeometryFromFile(opt.mesh2Path);
    if (m2.vertices.empty()) { std::cerr << "Error loading mesh2." << std::endl; return 1; }
    if (!requirePrecomputedEdges(m2, opt.mesh2Path, "Mesh2")) return 1;

    DeviceContext ctx(0);
    auto estimate = [&](bool verbose) {
        long long e = estimatePairs(ctx, m1, m2, opt.epsilon, opt.gamma, opt.useAlphaCorrection, verbose);
        return e < 0 ? 0LL : e;
    };
    auto tableSize = [&](long long est) -> unsigned long long {
        if (opt.manualHashTableSize > 0) return opt.manualHashTableSize;
        if (opt.hashTableFreeMemFraction > 0.0f) {
            size_t freeB = 0, totalB = 0;
            (void)hipMemGetInfo(&freeB, &totalB);
            unsigned long long s = (unsigned long long)(freeB * (double)opt.hashTableFreeMemFraction / 8.0);
            unsigned long long p = 1024;
            while (p * 2 <= s) p *= 2;
            return p;
        }
        unsigned long long size = 16777216ull;
        if (est > 0) {
            unsigned long long target = (unsigned long long)(est / 0.5);
            if (target < 1024) target = 1024;
            if (target > 1073741824ull) target = 1073741824ull;
            size = nextPow2(target);
        }
        return size;
    };
    if (opt.estimateOnly) {
        timer.next("Selectivity Estimation");
        (void)estimate(true);
        timer.finish(opt.outputJsonPath);
        return 0;
    }
    timer.next("Init Device");
    MeshOverlapEdgesLauncher L(ctx);
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
    timer.next("Prepare Kernel Parameters");
    EdgeMeshData e1 = PrecomputedEdgeData::uploadFromGeometry(m1, ctx);
    EdgeMeshData e2 = PrecomputedEdgeData::uploadFromGeometry(m2, ctx);

    auto query = [&](unsigned long long size, PerformanceTimer* t, std::vector<uint64_t>& keys) {
        while (true) {
            DeviceHashTable ht;
            ht.create(size, (size & (size - 1)) == 0, 5000);
            ht.clear(ctx);
            MeshOverlapEdgesLaunchParams p1, p2;
            p1.edges = &e1.view; p1.target = as2.getHandle(); p1.hash_table = &ht.table; p1.swap_pair_order = 0;
            p2.edges = &e2.view; p2.target = as1.getHandle(); p2.hash_table = &ht.table; p2.swap_pair_order = 1;
            auto t0 = std::chrono::high_resolution_clock::now();
            if (opt.direction != 2) L.launchMesh1ToMesh2(p1, e1.num_edges);
            auto t1 = std::chrono::high_resolution_clock::now();
            if (opt.direction != 1) L.launchMesh2ToMesh1(p2, e2.num_edges);
            auto t2 = std::chrono::high_resolution_clock::now();
            if (ht.failedInserts(ctx) > 0 && size < (1ull << 31)) {
                std::cerr << "Note: hash table of " << size << " slots overflowed; retrying with "
                          << size * 2 << std::endl;
                size *= 2;
                continue;
            }
            if (t) {
                t->addMeasurement("Raytrace_Hash_Mesh1ToMesh2",
                                  std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
                t->addMeasurement("Raytrace_Hash_Mesh2ToMesh1",
                                  std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count());
            }
            DeviceBuffer out(size * sizeof(uint64_t));
            uint64_t n = 0;
            check(rs3d_ht_compact(ctx.get(), &ht.table, out.as<uint64_t>(), size, &n), "rs3d_ht_compact");
            keys.resize(n);
            out.download(keys.data(), n * sizeof(uint64_t));
            return;
        }
    };
    timer.next("Warmup");
    std::vector<uint64_t> keys;
    for (int w = 0; w < opt.warmupRuns; ++w) query(tableSize(estimate(false)), nullptr, keys);
    for (int r = 0; r < opt.numberOfRuns; ++r) {
        timer.next("Selectivity Estimation");
        const long long est = estimate(r == 0);
        const unsigned long long size = tableSize(est);
        if (r == 0) std::cout << "Using hash table size: " << size << std::endl;
        timer.next("Execute Hash Query");
        query(size, &timer, keys);
        timer.next("Download Results");
    }
    timer.next("Cleanup");
    PrecomputedEdgeData::freeEdgeData(e1);
    PrecomputedEdgeData::freeEdgeData(e2);
    std::set<int> o1(m1.triangleToObject.begin(), m1.triangleToObject.end());
    std::set<int> o2(m2.triangleToObject.begin(), m2.triangleToObject.end());
    std::cout << "\n=== Mesh Overlap Join Summary ===\nMesh1 triangles: " << u1.getNumTriangles()
              << "\nMesh1 objects: " << o1.size() << "\nMesh2 triangles: " << u2.getNumTriangles()
              << "\nMesh2 objects: " << o2.size() << "\nUnique object pairs: " << keys.size() << std::endl;
    if (direct && !opt.pairsOutputPath.empty()) {
        std::ofstream csv(opt.pairsOutputPath);
        csv << "object_id_mesh1,object_id_mesh2\n";
        for (const auto& p : keysToPairs(keys)) csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << "\n";
    }
    timer.finish(opt.outputJsonPath);
    std::cout << "\nQuery completed in " << (double)timer.getTotalDuration() / 1000.0 << " ms." << std::endl;
    return 0;
}

}  // namespace rs3d_cli
