// raytracer -- point-in-mesh containment (query/src/applications/rayspace_main.cpp:39-260):
// same flags, task/caching loop, phase names, ray_results.csv and timing JSON.
#include <fstream>
#include <iostream>
#include <memory>

#include "cli_common.hpp"

using namespace rs3d_host;

int main(int argc, char* argv[]) {
    PerformanceTimer timer;
    timer.start("Data Reading");
    std::string geometryFilePath, pointDatasetPath, outputJsonPath = "performance_timing.json";
    int numberOfRuns = 1, warmupRuns = 2, axis = 2;
    bool exportResults = true;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--geometry" && i + 1 < argc) geometryFilePath = argv[++i];
        else if (a == "--points" && i + 1 < argc) pointDatasetPath = argv[++i];
        else if (a == "--output" && i + 1 < argc) outputJsonPath = argv[++i];
        else if (a == "--ptx" && i + 1 < argc) ++i;
        else if (a == "--runs" && i + 1 < argc) numberOfRuns = std::atoi(argv[++i]);
        else if (a == "--warmup-runs" && i + 1 < argc) warmupRuns = std::atoi(argv[++i]);
        else if (a == "--ray-axis" && i + 1 < argc) axis = std::atoi(argv[++i]);
        else if (a == "--no-export") exportResults = false;
        else if (a == "--help" || a == "-h") {
            rs3d_cli::printHelp(argv[0], "[--geometry <path(s)>] [--points <path(s)>] [options]",
                                "Point-in-mesh containment, one ray per point.",
                                {{"--geometry <path(s)>", "Geometry .pre file(s), comma separated"},
                                 {"--points <path(s)>", "WKT POINT Z / xyz file(s), comma separated"},
                                 {"--output <path>", "Timing JSON (default performance_timing.json)"},
                                 {"--runs <number>", "Number of query launches"},
                                 {"--warmup-runs <number>", "Warmup launches (default 2)"},
                                 {"--ray-axis <0|1|2>", "Ray direction +x/+y/+z (default 2, as the reference)"},
                                 {"--no-export", "Do not write ray_results.csv"},
                                 {"--ptx <path>", "Accepted for compatibility; ignored"}});
            return 0;
        }
    }
    const auto geometryFiles = splitPaths(geometryFilePath);
    const auto pointFiles = splitPaths(pointDatasetPath);
    if (geometryFiles.empty()) {
        std::cerr << "Error: Geometry file path is required. Use --geometry <path_to_geometry_file>" << std::endl;
        return 1;
    }
    if (pointFiles.empty()) {
        std::cerr << "Error: Points file path is required. Use --points <path_to_point_file>" << std::endl;
        return 1;
    }
    const size_t numTasks = std::max(geometryFiles.size(), pointFiles.size());
    if (geometryFiles.size() > 1 && pointFiles.size() > 1 && geometryFiles.size() != pointFiles.size()) {
        std::cerr << "Error: When both --geometry and --points are lists, they must have the same length." << std::endl;
        return 1;
    }
    timer.next("Application Creation");
    DeviceContext ctx(0);
    std::string cachedGeometryPath, cachedPointsPath;
    GeometryData cachedGeometry;
    GeometryUploader uploader;
    std::unique_ptr<AccelerationStructure> as;
    PointData cachedPoints;
    std::unique_ptr<RayLauncher> launcher;
    ResultProcessor proc;
    for (size_t task = 0; task < numTasks; ++task) {
        const std::string g = geometryFiles[geometryFiles.size() == 1 ? 0 : task];
        const std::string p = pointFiles[pointFiles.size() == 1 ? 0 : task];
        std::cout << "\n=== Task " << (task + 1) << "/" << numTasks << " ===" << std::endl;
        if (g != cachedGeometryPath) {
            timer.next("Load Geometry");
            cachedGeometry = loadGeometryFromFile(g);
            if (cachedGeometry.vertices.empty()) {
                std::cerr << "Error: Failed to load geometry from " << g << std::endl;
                continue;
            }
            cachedGeometryPath = g;
            timer.next("Upload Geometry");
            uploader.upload(cachedGeometry);
            timer.next("Build Index");
            as = std::make_unique<AccelerationStructure>(ctx, uploader);
            as->build();
            as->prepareAxis(axis);  // the axis grid belongs to the index build
        }
        if (p != cachedPointsPath) {
            timer.next("Load Points");
            cachedPoints = loadPointDataset(p);
            if (cachedPoints.numPoints == 0) {
                std::cerr << "Error: Failed to load points from " << p << std::endl;
                continue;
            }
            cachedPointsPath = p;
            timer.next("Upload Points");
            launcher = std::make_unique<RayLauncher>(ctx, uploader, axis);
            launcher->uploadRays(cachedPoints.positions.data(), cachedPoints.numPoints);
        }
        const int numRays = (int)cachedPoints.numPoints;
        timer.next("Warmup");
        if (launcher && as) launcher->runWarmup(*as, warmupRuns);
        timer.next("Query");
        for (int r = 0; r < std::max(numberOfRuns, 1); ++r)
            if (launcher && as) launcher->launch(*as, numRays);
        timer.next("Download Results");
        std::vector<RayResult> hits;
        if (launcher) hits = proc.compactAndDownload(launcher->getResultBuffer(), numRays);
        timer.next("Output");
        std::cout << "\n=== Point-in-Polygon Summary ===\nTotal rays: " << numRays
                  << "\nPoints INSIDE polygons:  " << hits.size()
                  << "\nPoints OUTSIDE polygons: " << (numRays - (int)hits.size()) << std::endl;
        if (exportResults) {
            std::ofstream csv("ray_results.csv");
            csv << "pointId,polygonId\n";
            for (const auto& h : hits) csv << h.ray_id << "," << h.polygon_index << "\n";
        }
    }
    timer.next("Cleanup");
    launcher.reset();
    as.reset();
    timer.finish(outputJsonPath);
    std::cout << "\nTotal tasks processed: " << numTasks << std::endl;
    return 0;
}
