// intersection_app.hpp -- the intersection join executables: overlap edge rays 1->2 and
// 2->1 plus the containment fallback 1->2 and 2->1 into one pair hash table, compaction,
// sorted CSV (query/src/applications/raytracer_intersection_estimated_main.cpp:77-176,
// 308-774). `estimated` adds the grid estimate that sizes the table (E / load factor);
// without it this is raytracer_mesh_intersection. A table that overflows is detected and
// the query rerun with twice the slots.
#pragma once

#include <algorithm>
#include <chrono>
#include <fstream>
#include <iostream>

#include "cli_common.hpp"

namespace rs3d_cli {

struct IntersectionOptions : MeshPairOptions {
    std::string queryDirectionArg = "both", pairsOutputPath;
    bool estimateOnly = false, enableProfilingStats = false, useAnyhitContainment = true;
    float gamma = 0.8f, epsilon = 0.001f, hashLoadFactor = 0.5f;
    int overlapMaxIterations = 100, containmentMaxIterations = 1024;
    IntersectionOptions() : MeshPairOptions("estimated_intersection_timing.json") {}
    bool parseApplicationOption(const std::string& a, int& i, int argc, char** argv) override {
        if (a == "--gamma" && i + 1 < argc) { gamma = std::stof(argv[++i]); return true; }
        if (a == "--epsilon" && i + 1 < argc) { epsilon = std::stof(argv[++i]); return true; }
        if (a == "--estimate-only") { estimateOnly = true; return true; }
        if (a == "--query-direction" && i + 1 < argc) { queryDirectionArg = argv[++i]; return true; }
        if (a == "--overlap-max-iterations" && i + 1 < argc) { overlapMaxIterations = std::stoi(argv[++i]); return true; }
        if (a == "--containment-max-iterations" && i + 1 < argc) { containmentMaxIterations = std::stoi(argv[++i]); return true; }
        if (a == "--use-anyhit-containment") { useAnyhitContainment = true; return true; }
        if (a == "--hash-load-factor" && i + 1 < argc) { hashLoadFactor = std::stof(argv[++i]); return true; }
        if (a == "--enable-profiling-stats") { enableProfilingStats = true; return true; }
        if (a == "--pairs-output" && i + 1 < argc) { pairsOutputPath = argv[++i]; return true; }
        if ((a == "--pair-hits-output" || a == "--containment-tracking-output") && i + 1 < argc) { ++i; return true; }
        return false;
    }
};

inline int runIntersection(int argc, char** argv, bool estimated) {
    using namespace rs3d_host;
    PerformanceTimer timer;
    IntersectionOptions opt;
    if (!estimated) opt.outputJsonPath = "mesh_intersection_timing.json";
    opt.parse(argc, argv);
    if (opt.helpRequested) {
        std::vector<HelpEntry> o;
        MeshPairOptions::meshPairHelp(o);
        if (estimated) {
            o.emplace_back("--gamma <float>", "Estimation gamma (default: 0.8)");
            o.emplace_back("--epsilon <float>", "Estimation epsilon (default: 0.001)");
            o.emplace_back("--estimate-only", "Run only selectivity estimation");
            o.emplace_back("--hash-load-factor <float>", "Hash load factor in (0,1] (default: 0.5)");
        }
        o.emplace_back("--query-direction <both|mesh1_to_mesh2|mesh2_to_mesh1>", "Query direction (default: both)");
        o.emplace_back("--overlap-max-iterations <int>", "Overlap ray iteration cap (default: 100)");
        o.emplace_back("--containment-max-iterations <int>", "Accepted; containment rays walk every hit");
        o.emplace_back("--use-anyhit-containment", "Accepted (all-hits containment is the only mode)");
        o.emplace_back("--enable-profiling-stats", "Report traversal counters");
        o.emplace_back("--pairs-output <path>", "Intersection pairs CSV path (default: intersection_pairs.csv)");
        printHelp(argv[0], "--mesh1 <path> --mesh2 <path> [options]",
                  "Intersection join: boundary overlap or containment.", o);
        return 0;
    }
    int direction = 0;
    if (opt.queryDirectionArg == "mesh1_to_mesh2") direction = 1;
    else if (opt.queryDirectionArg == "mesh2_to_mesh1") direction = 2;
    else if (opt.queryDirectionArg != "both") {
        std::cerr << "Error: Invalid --query-direction: " << opt.queryDirectionArg << std::endl;
        return 1;
    }
    if (!opt.hasRequiredMeshInputs()) {
        std::cerr << "Usage: " << argv[0] << " --mesh1 <path> --mesh2 <path> [options]" << std::endl;
        return 1;
    }
    if (opt.hashLoadFactor <= 0.0f || opt.hashLoadFactor > 1.0f) {
        std::cerr << "Error: --hash-load-factor must be in (0, 1]." << std::endl;
        return 1;
    }
    const std::string pairsOut = opt.pairsOutputPath.empty() ? "intersection_pairs.csv" : opt.pairsOutputPath;
    timer.start("Load Mesh1");
    GeometryData m1 = loadGeometryFromFile(opt.mesh1Path);
    if (m1.vertices.empty() || !requirePrecomputedEdges(m1, opt.mesh1Path, "Mesh1")) return 1;
    timer.next("Load Mesh2");
    GeometryData m2 = loadGeometryFromFile(opt.mesh2Path);
    if (m2.vertices.empty() || !requirePrecomputedEdges(m2, opt.mesh2Path, "Mesh2")) return 1;
    DeviceContext ctx(0);
    long long est = 0;
    if (estimated) {
        timer.next("Selectivity Estimation");
        est = estimatePairs(ctx, m1, m2, opt.epsilon, opt.gamma, true, true);
        if (est < 0) est = 0;
        if (opt.estimateOnly) {
            timer.finish(opt.outputJsonPath);
            return 0;
        }
    }
    unsigned long long slots = 16777216ull;
    if (est > 0) {
        unsigned long long target = (unsigned long long)(est / opt.hashLoadFactor);
        target = std::max(target, 1024ull);
        target = std::min(target, 1073741824ull);
        if (target % 2 == 0) target += 1;  // odd size, modulo hashing
        slots = target;
    }
    timer.next("Init Device");
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
    const rs3d_mesh_view v1 = u1.view(), v2 = u2.view();
    timer.next("Query");
    std::vector<uint64_t> keys;
    uint64_t overflow = 0;
    while (true) {
        DeviceHashTable ht;
        ht.create(slots, (slots & (slots - 1)) == 0, 0);
        ht.clear(ctx);
        using clk = std::chrono::high_resolution_clock;
        auto t0 = clk::now();
        if (direction != 2)
            check(rs3d_edge_hash_ex(ctx.get(), as2.getHandle(), &e1.view, 0, RS3D_CHAIN_EPS,
                                    opt.overlapMaxIterations, &ht.table), "overlap 1->2");
        ctx.sync();
        auto t1 = clk::now();
        if (direction != 1)
            check(rs3d_edge_hash_ex(ctx.get(), as1.getHandle(), &e2.view, 1, RS3D_CHAIN_EPS,
                                    opt.overlapMaxIterations, &ht.table), "overlap 2->1");
        ctx.sync();
        auto t2 = clk::now();
        uint64_t o1 = 0, o2 = 0;
        if (direction != 2)
            check(rs3d_containment_fallback(ctx.get(), &v1, as2.getHandle(), 0, 256, &ht.table, &o1), "containment 1->2");
        auto t3 = clk::now();
        if (direction != 1)
            check(rs3d_containment_fallback(ctx.get(), &v2, as1.getHandle(), 1, 256, &ht.table, &o2), "containment 2->1");
        ctx.sync();
        auto t4 = clk::now();
        if (ht.failedInserts(ctx) > 0 && slots < (1ull << 31)) {
            slots = slots * 2 + 1;
            continue;
        }
        auto us = [](clk::time_point a, clk::time_point b) {
            return (long long)std::chrono::duration_cast<std::chrono::microseconds>(b - a).count();
        };
        timer.addMeasurement("Raytrace_Overlap_Hash_Mesh1To2", us(t0, t1));
        timer.addMeasurement("Raytrace_Overlap_Hash_Mesh2To1", us(t1, t2));
        timer.addMeasurement("Raytrace_Containment_Hash_Mesh1To2", us(t2, t3));
        timer.addMeasurement("Raytrace_Containment_Hash_Mesh2To1", us(t3, t4));
        DeviceBuffer out(slots * sizeof(uint64_t));
        uint64_t n = 0;
        auto t5 = clk::now();
        check(rs3d_ht_compact(ctx.get(), &ht.table, out.as<uint64_t>(), slots, &n), "rs3d_ht_compact");
        check(rs3d_sort_unique_u64(ctx.get(), out.as<uint64_t>(), n, &n), "rs3d_sort_unique_u64");
        timer.addMeasurement("compact_hash_table_pairs", us(t5, clk::now()));
        keys.resize(n);
        out.download(keys.data(), n * sizeof(uint64_t));
        overflow = o1 + o2;
        break;
    }
    timer.addCounter("Profile_Actual_Intersection_Pairs", keys.size());
    if (opt.enableProfilingStats) timer.addCounter("Profile_Containment_Candidate_Overflow", overflow);
    std::cout << "Intersection pairs: " << keys.size() << std::endl;
    {
        std::ofstream csv(pairsOut);
        csv << "a_object_id,b_object_id\n";
        for (const auto& p : keysToPairs(keys)) csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << "\n";
    }
    timer.next("Cleanup");
    PrecomputedEdgeData::freeEdgeData(e1);
    PrecomputedEdgeData::freeEdgeData(e2);
    timer.finish(opt.outputJsonPath);
    return 0;
}

}  // namespace rs3d_cli
