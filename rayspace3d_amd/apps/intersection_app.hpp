// intersection_app.hpp -- the intersection join executables: overlap edge rays 1->2 and
// 2->1 plus the containment fallback 1->2 and 2->1 into one pair hash table through
// MeshIntersectionLauncher, compaction, sorted CSV, the pair-hit and containment tracking
// CSVs and the profiling counters (query/src/applications/raytracer_intersection_estimated_main.cpp:
// 77-176, 308-774). `estimated` adds the grid estimate that sizes the table (E / load
// factor); without it this is raytracer_mesh_intersection. A table that overflows is
// detected and the query rerun with twice the slots.
#pragma once

#include <algorithm>
#include <chrono>
#include <fstream>
#include <iostream>

#include "cli_common.hpp"

namespace rs3d_cli {

struct IntersectionOptions : MeshPairOptions {
    std::string queryDirectionArg = "both", pairsOutputPath, pairHitsOutputPath, containmentTrackingOutputPath;
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
        if (a == "--pair-hits-output" && i + 1 < argc) { pairHitsOutputPath = argv[++i]; return true; }
        if (a == "--containment-tracking-output" && i + 1 < argc) { containmentTrackingOutputPath = argv[++i]; return true; }
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
        o.emplace_back("--containment-max-iterations <int>", "Containment ray iteration cap (default: 1024)");
        o.emplace_back("--use-anyhit-containment", "Use AnyHit accumulation for containment rays (default: enabled)");
        o.emplace_back("--enable-profiling-stats", "Enable device-side profiling counters");
        o.emplace_back("--pairs-output <path>", "Intersection pairs CSV path (default: intersection_pairs.csv)");
        o.emplace_back("--pair-hits-output <path>", "Pair hit tracking CSV path (default: intersection_pair_hits.csv)");
        o.emplace_back("--containment-tracking-output <path>",
                       "Containment tracking CSV path (default: intersection_containment_tracking.csv)");
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
    const std::string pairHitsOut =
        opt.pairHitsOutputPath.empty() ? "intersection_pair_hits.csv" : opt.pairHitsOutputPath;
    const std::string trackingOut = opt.containmentTrackingOutputPath.empty() ? "intersection_containment_tracking.csv"
                                                                               : opt.containmentTrackingOutputPath;
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
        est = estimatePairs(ctx, m1, m2, opt.epsilon, opt.gamma, true, true, true);
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
    // per-source tracking buffers (PairHitTracking.h:28-110, ContainmentTracking.h:20-95)
    constexpr int kMaxPairTargets = 256;
    const uint64_t n1 = v1.num_objects, n2 = v2.num_objects;
    DeviceBuffer ids1(std::max<uint64_t>(n1, 1) * kMaxPairTargets * 4), hits1(std::max<uint64_t>(n1, 1) * kMaxPairTargets * 4);
    DeviceBuffer ids2(std::max<uint64_t>(n2, 1) * kMaxPairTargets * 4), hits2(std::max<uint64_t>(n2, 1) * kMaxPairTargets * 4);
    DeviceBuffer trk1(std::max<uint64_t>(n1, 1) * 12), trk2(std::max<uint64_t>(n2, 1) * 12);
    DeviceBuffer stats(sizeof(rs3d_intersection_stats));
    MeshIntersectionLauncher launcher(ctx);
    MeshIntersectionLaunchParams p1, p2;
    auto fill = [&](MeshIntersectionLaunchParams& p, const rs3d_mesh_view& src, const EdgeMeshData& e,
                    AccelerationStructure& target, DeviceBuffer& ids, DeviceBuffer& hits, DeviceBuffer& trk,
                    uint64_t n) {
        p.mesh1 = src;
        p.edges = &e.view;
        p.mesh2 = target.getHandle();
        p.use_anyhit_containment = opt.useAnyhitContainment ? 1 : 0;
        p.overlap_max_iterations = opt.overlapMaxIterations;
        p.containment_max_iterations = opt.containmentMaxIterations;
        p.profiling_stats = opt.enableProfilingStats ? stats.as<rs3d_intersection_stats>() : nullptr;
        p.pair_target_ids = ids.as<int32_t>();
        p.pair_target_hits = hits.as<uint32_t>();
        p.max_pair_targets_per_source = kMaxPairTargets;
        p.containment_iterations_per_source = trk.as<uint32_t>();
        p.containment_candidate_count_per_source = trk.as<uint32_t>() + n;
        p.containment_candidate_overflow_per_source = trk.as<uint32_t>() + 2 * n;
    };
    fill(p1, v1, e1, as2, ids1, hits1, trk1, n1);
    fill(p2, v2, e2, as1, ids2, hits2, trk2, n2);
    timer.next("Query");
    std::vector<uint64_t> keys;
    uint64_t overflow = 0;
    while (true) {
        DeviceHashTable ht;
        ht.create(slots, (slots & (slots - 1)) == 0, 0);
        ht.clear(ctx);
        p1.hash_table = p2.hash_table = &ht.table;
        stats.fill(0);
        ids1.fill(0xff);
        ids2.fill(0xff);
        hits1.fill(0);
        hits2.fill(0);
        trk1.fill(0);
        trk2.fill(0);
        using clk = std::chrono::high_resolution_clock;
        auto us = [](clk::time_point a, clk::time_point b) {
            return (long long)std::chrono::duration_cast<std::chrono::microseconds>(b - a).count();
        };
        long long tm[4] = {0, 0, 0, 0};
        uint64_t o1 = 0, o2 = 0;
        auto t0 = clk::now();
        if (direction != 2) launcher.launchOverlapMesh1ToMesh2(p1, e1.num_edges);
        auto t1 = clk::now();
        tm[0] = us(t0, t1);
        if (direction != 1) launcher.launchOverlapMesh2ToMesh1(p2, e2.num_edges);
        auto t2 = clk::now();
        tm[1] = us(t1, t2);
        if (direction != 2) {
            launcher.launchContainmentMesh1ToMesh2(p1, (int)n1);
            o1 = launcher.lastContainmentOverflow();
        }
        auto t3 = clk::now();
        tm[2] = us(t2, t3);
        if (direction != 1) {
            launcher.launchContainmentMesh2ToMesh1(p2, (int)n2);
            o2 = launcher.lastContainmentOverflow();
        }
        auto t4 = clk::now();
        tm[3] = us(t3, t4);
        if (ht.failedInserts(ctx) > 0 && slots < (1ull << 31)) {
            slots = slots * 2 + 1;
            continue;
        }
        if (direction != 2) timer.addMeasurement("Raytrace_Overlap_Hash_Mesh1ToMesh2", tm[0]);
        if (direction != 1) timer.addMeasurement("Raytrace_Overlap_Hash_Mesh2ToMesh1", tm[1]);
        if (direction != 2) timer.addMeasurement("Raytrace_Containment_Hash_Mesh1ToMesh2", tm[2]);
        if (direction != 1) timer.addMeasurement("Raytrace_Containment_Hash_Mesh2ToMesh1", tm[3]);
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
    std::cout << "Intersection pairs: " << keys.size() << std::endl;
    {
        std::ofstream csv(pairsOut);
        csv << "a_object_id,b_object_id\n";
        for (const auto& p : keysToPairs(keys)) csv << p.object_id_mesh1 << "," << p.object_id_mesh2 << "\n";
    }
    {
        // direction,source_object_id,target_object_id,target_ray_hits (PairHitTracking.h:127-165)
        std::vector<int32_t> hid1(n1 * kMaxPairTargets), hid2(n2 * kMaxPairTargets);
        std::vector<uint32_t> hh1(n1 * kMaxPairTargets), hh2(n2 * kMaxPairTargets);
        if (n1) ids1.download(hid1.data(), hid1.size() * 4), hits1.download(hh1.data(), hh1.size() * 4);
        if (n2) ids2.download(hid2.data(), hid2.size() * 4), hits2.download(hh2.data(), hh2.size() * 4);
        std::ofstream csv(pairHitsOut);
        if (!csv) {
            std::cerr << "Error: failed to open pair hit tracking output: " << pairHitsOut << std::endl;
            return 1;
        }
        csv << "direction,source_object_id,target_object_id,target_ray_hits\n";
        auto rows = [&](const char* dir, const std::vector<int32_t>& id, const std::vector<uint32_t>& h, uint64_t n) {
            for (uint64_t s = 0; s < n; ++s)
                for (int i = 0; i < kMaxPairTargets; ++i) {
                    const int32_t t = id[s * kMaxPairTargets + i];
                    if (t >= 0) csv << dir << ',' << s << ',' << t << ',' << h[s * kMaxPairTargets + i] << "\n";
                }
        };
        rows("mesh1_to_mesh2", hid1, hh1, n1);
        rows("mesh2_to_mesh1", hid2, hh2, n2);
    }
    {
        // direction,source_object_id,iterations,candidate_count,candidate_overflow_events
        // (ContainmentTracking.h:97-121)
        std::vector<uint32_t> t1h(3 * n1), t2h(3 * n2);
        if (n1) trk1.download(t1h.data(), t1h.size() * 4);
        if (n2) trk2.download(t2h.data(), t2h.size() * 4);
        std::ofstream csv(trackingOut);
        if (!csv) {
            std::cerr << "Error: failed to open containment tracking output: " << trackingOut << std::endl;
            return 1;
        }
        csv << "direction,source_object_id,iterations,candidate_count,candidate_overflow_events\n";
        for (uint64_t s = 0; s < n1; ++s)
            csv << "mesh1_to_mesh2," << s << ',' << t1h[s] << ',' << t1h[n1 + s] << ',' << t1h[2 * n1 + s] << "\n";
        for (uint64_t s = 0; s < n2; ++s)
            csv << "mesh2_to_mesh1," << s << ',' << t2h[s] << ',' << t2h[n2 + s] << ',' << t2h[2 * n2 + s] << "\n";
    }
    std::cout << "Intersection pairs CSV: " << pairsOut << std::endl;
    std::cout << "Pair hit tracking CSV: " << pairHitsOut << std::endl;
    std::cout << "Containment tracking CSV: " << trackingOut << std::endl;
    if (opt.enableProfilingStats) {
        rs3d_intersection_stats h{};
        stats.download(&h, sizeof h);
        timer.addCounter("Profile_Overlap_Trace_Calls", h.overlap_trace_calls);
        timer.addCounter("Profile_Overlap_Iterations_Total", h.overlap_iterations_total);
        timer.addCounter("Profile_Overlap_Hits_Total", h.overlap_hits_total);
        timer.addCounter("Profile_Overlap_Max_Iterations_Per_Trace", h.overlap_max_iterations_per_trace);
        timer.addCounter("Profile_Containment_Rays_Total", h.containment_rays_total);
        timer.addCounter("Profile_Containment_Iterations_Total", h.containment_iterations_total);
        timer.addCounter("Profile_Containment_Hits_Total", h.containment_hits_total);
        timer.addCounter("Profile_Containment_Max_Iterations_Per_Ray", h.containment_max_iterations_per_ray);
        timer.addCounter("Profile_Containment_Same_Hit_Suppressed", h.containment_same_hit_suppressed);
        timer.addCounter("Profile_Containment_Candidate_Additions", h.containment_candidate_additions);
        timer.addCounter("Profile_Containment_Candidate_Toggles", h.containment_candidate_toggles);
        timer.addCounter("Profile_Containment_Candidate_Overflow", h.containment_candidate_overflow);
        timer.addCounter("Profile_Containment_Targets_Total", h.containment_targets_total);
        const unsigned long long calls = h.overlap_trace_calls, rays = h.containment_rays_total;
        timer.addCounter("Profile_Overlap_Avg_Iterations_x1000", calls ? h.overlap_iterations_total * 1000ull / calls : 0ull);
        timer.addCounter("Profile_Containment_Avg_Iterations_x1000",
                         rays ? h.containment_iterations_total * 1000ull / rays : 0ull);
    }
    if (overflow) std::cerr << "Warning: " << overflow << " containment candidates beyond the 256-target cap" << std::endl;
    timer.next("Cleanup");
    PrecomputedEdgeData::freeEdgeData(e1);
    PrecomputedEdgeData::freeEdgeData(e2);
    timer.finish(opt.outputJsonPath);
    return 0;
}

}  // namespace rs3d_cli
