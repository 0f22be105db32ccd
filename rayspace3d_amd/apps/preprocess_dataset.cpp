// preprocess_dataset -- dataset -> .pre v2 (preprocess/src/preprocess_dataset.cpp:183-385):
// same flags, phase names and timing JSON. Loading is host work (OBJ, or an existing .pre
// so that v1 files become v2); the grid histogram and the unique-edge payload are computed
// on the GPU. --mode points converts a text point file into the binary point file that
// loadPointDataset reads without parsing. WKT-polygon and .dt inputs need triangulation and
// TDBase decoding, which are not part of this build: those modes exit with an error.
#include <algorithm>
#include <iomanip>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "cli_common.hpp"

using namespace rs3d_host;

namespace {

// Translates every object by a uniform offset in [-15, 15)^3 (the reference's --shuffle).
void shuffleObjects(GeometryData& g, std::mt19937& rng) {
    if (g.indices.size() != g.triangleToObject.size()) {
        std::cerr << "Warning: indices and triangleToObject size mismatch; skipping shuffle." << std::endl;
        return;
    }
    const size_t nObj = countObjects(g);
    std::vector<std::vector<uint32_t>> verts(nObj);
    for (size_t t = 0; t < g.indices.size(); ++t) {
        const int o = g.triangleToObject[t];
        if (o < 0) continue;
        verts[o].insert(verts[o].end(), {g.indices[t].x, g.indices[t].y, g.indices[t].z});
    }
    std::uniform_real_distribution<float> dist(-15.0f, 15.0f);
    size_t shown = 0;
    for (size_t o = 0; o < nObj; ++o) {
        auto& vs = verts[o];
        if (vs.empty()) continue;
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
        const float tx = dist(rng), ty = dist(rng), tz = dist(rng);
        for (uint32_t v : vs) {
            if (v >= g.vertices.size()) continue;
            g.vertices[v].x += tx;
            g.vertices[v].y += ty;
            g.vertices[v].z += tz;
        }
        if (++shown <= 3)
            std::cout << "  Object " << o << " translated by (" << std::fixed << std::setprecision(2) << tx << ", "
                      << ty << ", " << tz << ")" << std::endl;
    }
}

}  // namespace

int main(int argc, char* argv[]) {
    std::string datasetPath, outputGeometryPath = "geometry_data.txt",
                             outputTimingPath = "preprocessing_timing.json", modeStr = "mesh";
    bool shuffle = false, generateGrid = false, haveSeed = false, skipEdges = false;
    int gridResolution = 128;
    float worldSize = -1.0f;
    unsigned long seed = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--dataset" && i + 1 < argc) datasetPath = argv[++i];
        else if (a == "--mode" && i + 1 < argc) modeStr = argv[++i];
        else if (a == "--output-geometry" && i + 1 < argc) outputGeometryPath = argv[++i];
        else if (a == "--output-timing" && i + 1 < argc) outputTimingPath = argv[++i];
        else if (a == "--shuffle" || a == "shuffle") shuffle = true;
        else if (a == "--generate-grid") generateGrid = true;
        else if ((a == "--grid-resolution" || a == "--euler-grid-size") && i + 1 < argc)
            gridResolution = std::stoi(argv[++i]);
        else if (a == "--world-size" && i + 1 < argc) worldSize = std::stof(argv[++i]);
        else if (a == "--gamma" && i + 1 < argc) ++i;
        else if (a == "--skip-edges") skipEdges = true;
        else if (a == "--seed" && i + 1 < argc) {
            seed = std::stoul(argv[++i]);
            haveSeed = true;
        } else if (a == "--help" || a == "-h") {
            rs3d_cli::printHelp(
                argv[0], "[--mode <mesh|pre|points>] [--dataset <path>] [--output-geometry <path>] [options]",
                "Dataset preprocessing: .pre v2 with the unique-edge payload (and grid statistics).",
                {{"--mode <mesh|pre|points>", "mesh: .obj file (default); pre: existing .pre v1/v2;"},
                 {"", "points: text points -> binary point file"},
                 {"--dataset <path>", "Path to dataset"},
                 {"--output-geometry <path>", "Output .pre (or binary point) file"},
                 {"--output-timing <path>", "Timing JSON (default preprocessing_timing.json)"},
                 {"--shuffle", "Randomly translate each loaded object"},
                 {"--seed <n>", "Seed of --shuffle (default: random_device)"},
                 {"--skip-edges", "Write the geometry without the edge payload (no GPU used)"},
                 {"--generate-grid", "Generate grid statistics for selectivity estimation"},
                 {"--grid-resolution <N>", "Grid resolution (default: 128)"},
                 {"--world-size <S>", "World size (default: auto-detected)"}});
            return 0;
        } else {
            std::cout << "Warning: Unknown argument '" << a << "'" << std::endl;
        }
    }
    if (modeStr != "mesh" && modeStr != "pre" && modeStr != "points") {
        if (modeStr == "wkt" || modeStr == "dt")
            std::cerr << "Error: mode '" << modeStr
                      << "' (polygon triangulation / TDBase decoding) is not supported by this build" << std::endl;
        else
            std::cerr << "Error: Unknown mode '" << modeStr << "'. Valid modes: mesh, pre, points" << std::endl;
        return 1;
    }
    if (datasetPath.empty()) {
        std::cerr << "Error: Dataset path is required. Use --dataset <path>" << std::endl;
        return 1;
    }
    if (gridResolution < 1 || gridResolution > 1024) {
        std::cerr << "Error: --grid-resolution must be in [1, 1024]" << std::endl;
        return 1;
    }
    std::cout << "Dataset Preprocessing Tool\nMode: " << modeStr << "\nInput dataset: " << datasetPath
              << "\nOutput geometry: " << outputGeometryPath << "\nOutput timing: " << outputTimingPath << std::endl;

    PerformanceTimer timer;
    if (modeStr == "points") {
        timer.start("Loading Points");
        PointData pts = loadPointDataset(datasetPath);
        if (pts.numPoints == 0) {
            std::cerr << "Error: no points loaded from " << datasetPath << std::endl;
            return 1;
        }
        timer.next("Writing Points");
        if (!writeBinaryPoints(outputGeometryPath, pts)) return 1;
        timer.finish(outputTimingPath);
        std::cout << "Wrote " << pts.numPoints << " points to " << outputGeometryPath << std::endl;
        return 0;
    }

    timer.start("Loading Dataset");
    std::unique_ptr<IDatasetLoader> loader;
    if (modeStr == "pre") loader = std::make_unique<PreDatasetLoader>();
    else loader = DatasetLoaderFactory::createFromPath(datasetPath);
    if (!loader) {
        std::cerr << "Error: Could not determine loader for path: " << datasetPath << std::endl;
        return 1;
    }
    GeometryData geometry = loader->load(datasetPath);
    if (geometry.vertices.empty()) {
        std::cerr << "Error: Failed to load dataset or dataset is empty." << std::endl;
        return 1;
    }
    if (shuffle) {
        timer.next("Shuffling Objects (Random Translation)");
        std::random_device rd;
        std::mt19937 rng(haveSeed ? seed : rd());
        shuffleObjects(geometry, rng);
    }
    std::unique_ptr<DeviceContext> ctx;
    if (generateGrid || !skipEdges) ctx = std::make_unique<DeviceContext>();
    if (generateGrid) {
        timer.next("Generating Grid Statistics");
        generateGridStatsOnDevice(*ctx, geometry, geometry.grid, gridResolution, worldSize);
    }
    if (skipEdges) {
        geometry.edges = EdgeData();
    } else {
        timer.next("Extracting Edges");
        geometry.edges = extractEdgesOnDevice(*ctx, geometry);
        if (!geometry.edges.hasEdges()) {
            std::cerr << "Error: Edge extraction failed. No valid edge payload generated." << std::endl;
            return 1;
        }
        std::cout << "Extracted " << geometry.edges.numEdges() << " unique edges with "
                  << geometry.edges.sourceObjects.size() << " flattened source-object entries" << std::endl;
    }
    timer.next("Writing Geometry Data");
    if (!writeBinaryFile(outputGeometryPath, geometry)) return 1;
    std::cout << "Geometry data written to: " << outputGeometryPath << " (Binary)" << std::endl;
    timer.finish(outputTimingPath);

    const GridData& gd = geometry.grid;
    std::cout << "\n=== Preprocessing Complete ===\nMode: " << modeStr << "\nProcessed dataset: " << datasetPath
              << "\nTotal objects (triangles): " << geometry.totalTriangles
              << "\nTotal vertices: " << geometry.vertices.size() << "\nTotal triangles: " << geometry.indices.size()
              << "\nTotal extracted edges: " << geometry.edges.numEdges() << "\nUniverse Extents: ["
              << (gd.maxBound.x - gd.minBound.x) << ", " << (gd.maxBound.y - gd.minBound.y) << ", "
              << (gd.maxBound.z - gd.minBound.z) << "]" << std::endl;
    return 0;
}
