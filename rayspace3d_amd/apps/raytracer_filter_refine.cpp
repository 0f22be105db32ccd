// raytracer_filter_refine -- dataset-box filter then exact containment
// (query/src/applications/rayspace_filter_refine_main.cpp:20-156).
#include <fstream>
#include <iostream>

#include "cli_common.hpp"

using namespace rs3d_host;

int main(int argc, char* argv[]) {
    PerformanceTimer timer;
    timer.start("Data Reading");
    std::string geometryFilePath, pointDatasetPath, outputJsonPath = "performance_timing_filter_refine.json";
    bool exportResults = true;
    int axis = 2;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--geometry" && i + 1 < argc) geometryFilePath = argv[++i];
        else if (a == "--points" && i + 1 < argc) pointDatasetPath = argv[++i];
        else if (a == "--output" && i + 1 < argc) outputJsonPath = argv[++i];
        else if ((a == "--ptx" || a == "--runs" || a == "--warmup-runs") && i + 1 < argc) ++i;  // parsed, unused as in the reference
        else if (a == "--ray-axis" && i + 1 < argc) axis = std::atoi(argv[++i]);
        else if (a == "--no-export") exportResults = false;
        else if (a == "--help" || a == "-h") {
            rs3d_cli::printHelp(argv[0], "--geometry <path> --points <path> [options]",
                                "Filter-refine point containment (dataset AABB filter, exact refine).",
                                {{"--geometry <path>", "Geometry .pre file"},
                                 {"--points <path>", "WKT POINT Z / xyz point file"},
                                 {"--output <path>", "Timing JSON"},
                                 {"--ray-axis <0|1|2>", "Ray direction (default +z)"},
                                 {"--no-export", "Do not write ray_results.csv"}});
            return 0;
        }
    }
    if (geometryFilePath.empty() || pointDatasetPath.empty()) {
        std::cerr << "Error: --geometry and --points are required." << std::endl;
        return 1;
    }
    GeometryData geometry = loadGeometryFromFile(geometryFilePath);
    if (geometry.vertices.empty()) {
        std::cerr << "Error: Failed to load geometry." << std::endl;
        return 1;
    }
    PointData points = loadPointDataset(pointDatasetPath);
    if (points.numPoints == 0) {
        std::cerr << "Error: Failed to load points." << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    DeviceContext ctx(0);
    FilterRefinePipeline pipe(ctx, axis);
    timer.next("Filter-Refine Execution");
    FilterRefineResult r = pipe.execute(geometry, points);
    timer.next("Output");
    std::cout << "\n=== Filter-Refine Summary ===\nTotal points: " << r.totalPoints
              << "\nCandidates after filter: " << r.candidateCount << "\nFiltered out: " << r.filteredOut
              << "\nInside: " << r.insideCount << "\nOutside: " << r.outsideCount << std::endl;
    if (exportResults) {
        std::ofstream csv("ray_results.csv");
        csv << "pointId,polygonId\n";
        for (const auto& h : r.refineHits) csv << h.ray_id << "," << h.polygon_index << "\n";
    }
    timer.next("Cleanup");
    timer.finish(outputJsonPath);
    return 0;
}
