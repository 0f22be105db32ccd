// rs3d_host.hpp -- C++ host layer of the ray-cast spatial join: the reference's host
// classes and file formats, re-implemented over the librs3d C ABI (include/rs3d.h).
//
// Mirrors (reference paths):
//   GeometryData/EdgeData/GridData/PointData   common/include/Geometry.h:15-56
//   readBinaryFile/writeBinaryFile (.pre)      common/include/BinaryIO.h:35-180
//   loadGeometryFromFile/requirePrecomputedEdges  query/src/runtime/GeometryIO.cpp:6-20
//   loadPointDataset                           query/src/runtime/PointIO.cpp:12-72
//   IDatasetLoader/ObjMeshDatasetLoader/DatasetLoaderFactory  preprocess/src/*.cpp
//   PerformanceTimer (timing JSON)             query/src/timer.h:9-49, timer.cpp:124-192
//   DeviceContext       <- OptixContext        query/src/optix/OptixContext.cpp:4-8
//   GeometryUploader                           query/src/geometry/GeometryUploader.h:8-43
//   AccelerationStructure <- OptixAccelerationStructure (OptixAccelerationStructure.h:9-34)
//   RayLauncher                                query/src/raytracing/RayLauncher.h:9-40
//   ResultProcessor                            query/src/raytracing/ResultProcessor.h:7-20
//   FilterRefinePipeline                       query/src/raytracing/FilterRefine.h:15-48
//   PrecomputedEdgeData                        query/src/geometry/PrecomputedEdgeData.h:22-26
//   MeshOverlapEdgesLauncher                   query/src/raytracing/MeshOverlapEdgesLauncher.h:43-56
//   MeshIntersectionLauncher                   query/src/raytracing/MeshIntersectionLauncher.h:9-60
//   MeshContainmentLauncher                    query/src/raytracing/MeshContainmentLauncher.h:16-71
//   estimateOverlapSelectivity                 query/src/cuda/estimated_overlap.h:6-13
//   estimateIntersectionSelectivity            query/src/cuda/estimated_intersection.h:6-13
// Device errors are fail-fast (message + exit(1)) as with CUDA_CHECK/OPTIX_CHECK
// (query/src/optix/OptixHelpers.h:15-29).
#pragma once

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rs3d.h"

namespace rs3d_host {

struct Vec3 {
    float x, y, z;
};
struct UInt3 {
    uint32_t x, y, z;
};
static_assert(sizeof(Vec3) == 12 && sizeof(UInt3) == 12, "packed float3/uint3");

struct GridCell {
    uint32_t CenterCount;
    uint32_t TouchCount;
    float AvgSizeMean;
    float VolRatio;
};

struct GridData {
    Vec3 minBound{0, 0, 0};
    Vec3 maxBound{0, 0, 0};
    uint32_t resolution[3] = {0, 0, 0};
    std::vector<GridCell> cells;
    bool hasGrid = false;
};

struct EdgeData {
    std::vector<Vec3> edgeStarts, edgeEnds;
    std::vector<int> sourceObjects, sourceObjectOffsets, sourceObjectCounts;
    bool hasEdges() const {
        return !edgeStarts.empty() && edgeStarts.size() == edgeEnds.size() &&
               edgeStarts.size() == sourceObjectOffsets.size() &&
               edgeStarts.size() == sourceObjectCounts.size();
    }
    size_t numEdges() const { return edgeStarts.size(); }
};

struct GeometryData {
    std::vector<Vec3> vertices;
    std::vector<UInt3> indices;
    std::vector<int> triangleToObject;
    size_t totalTriangles = 0;
    EdgeData edges;
    GridData grid;
};

struct PointData {
    std::vector<Vec3> positions;
    size_t numPoints = 0;
};

using RayResult = rs3d_ray_result;
using MeshQueryResult = rs3d_pair;

// ---- files ---------------------------------------------------------------------------
GeometryData readBinaryFile(const std::string& filename);  // v2 and v1; empty on error
bool writeBinaryFile(const std::string& filename, const GeometryData& geometry);
GeometryData loadGeometryFromFile(const std::string& path);
bool requirePrecomputedEdges(const GeometryData& g, const std::string& path, const char* label);
// Text points (WKT "POINT Z (x y z)" or "x y z" per line), or the binary point file below
// (recognised by its magic). An empty path gives the reference's three default points.
PointData loadPointDataset(const std::string& path);
// Binary point file: 32-byte header {u32 magic 'RS3P', u32 version 1, u64 count, u32 12,
// u32[3] 0} followed by count packed float3 (x, y, z).
bool writeBinaryPoints(const std::string& path, const PointData& points);
size_t countObjects(const GeometryData& g);  // max(triangleToObject) + 1

// ---- dataset loaders (preprocess/src/IDatasetLoader.h, DatasetLoaderFactory.cpp:24-38) ----
class IDatasetLoader {
public:
    virtual ~IDatasetLoader() = default;
    virtual GeometryData load(const std::string& path) = 0;
};

// Wavefront OBJ (preprocess/src/ObjMeshDatasetLoader.cpp:14-169): one object per shape
// ('o'/'g' groups), shapes without faces skipped, quads split along the shorter diagonal.
class ObjMeshDatasetLoader : public IDatasetLoader {
public:
    GeometryData load(const std::string& path) override;
};

// Existing .pre file (v1 or v2) as input: re-preprocessing turns v1 files into v2.
class PreDatasetLoader : public IDatasetLoader {
public:
    GeometryData load(const std::string& path) override;
};

struct DatasetLoaderFactory {
    // ".obj" -> OBJ, ".pre" -> .pre; other extensions (WKT polygons, .dt) have no loader here.
    static std::unique_ptr<IDatasetLoader> createFromPath(const std::string& path);
};

// ---- timing JSON ---------------------------------------------------------------------
class PerformanceTimer {
public:
    PerformanceTimer();
    void start(const std::string& phaseName);
    void next(const std::string& phaseName);
    void finish(const std::string& filename);
    void addMeasurement(const std::string& phaseName, long long durationUs);
    void addCounter(const std::string& counterName, unsigned long long value);
    long long getPhaseDuration(const std::string& phaseName) const;
    long long getTotalDuration() const;

private:
    struct Phase {
        std::string name;
        std::chrono::high_resolution_clock::time_point start_time, end_time;
        long long duration_us = 0;
    };
    struct Counter {
        std::string name;
        unsigned long long value;
    };
    std::vector<Phase> phases, manualMeasurements;
    std::vector<Counter> counters;
    std::chrono::high_resolution_clock::time_point total_start, total_end;
    bool is_running = false;
    void endCurrentPhase();
    void writeResultsToFile(const std::string& filename) const;
};

// ---- device objects -----------------------------------------------------------------
void check(int rc, const char* what);  // prints rs3d_last_error and exits on failure

class DeviceContext {
public:
    explicit DeviceContext(int device = 0);
    ~DeviceContext();
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;
    rs3d_ctx* get() const { return ctx_; }
    void sync();

private:
    rs3d_ctx* ctx_ = nullptr;
};

// Owned device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { alloc(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void alloc(size_t bytes);
    void release();
    void upload(const void* src, size_t bytes);
    void download(void* dst, size_t bytes) const;
    void fill(int byteValue);  // every byte set to byteValue (synchronous)
    template <typename T>
    T* as() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

class GeometryUploader {
public:
    void upload(const GeometryData& geometry);
    const float* getVertices() const { return vertices_.as<float>(); }
    const uint32_t* getIndices() const { return indices_.as<uint32_t>(); }
    const int32_t* getTriangleToObject() const { return triObj_.as<int32_t>(); }
    size_t getNumVertices() const { return numVertices_; }
    size_t getNumIndices() const { return numTriangles_; }
    size_t getNumTriangles() const { return numTriangles_; }
    size_t getNumObjects() const { return numObjects_; }
    rs3d_mesh_view view() const;

private:
    DeviceBuffer vertices_, indices_, triObj_;
    size_t numVertices_ = 0, numTriangles_ = 0, numObjects_ = 0;
};

class AccelerationStructure {
public:
    AccelerationStructure(DeviceContext& ctx, GeometryUploader& geometry, int leafSize = 1);
    ~AccelerationStructure();
    AccelerationStructure(const AccelerationStructure&) = delete;
    AccelerationStructure& operator=(const AccelerationStructure&) = delete;
    void build();
    // Builds the axis grid for rays along +axis now (otherwise the first launch does);
    // point launches use it when it is accepted. Returns whether it was accepted.
    bool prepareAxis(int axis);
    rs3d_bvh* getHandle() const { return bvh_; }

private:
    DeviceContext& ctx_;
    GeometryUploader& geometry_;
    int leafSize_;
    rs3d_bvh* bvh_ = nullptr;
};

class RayLauncher {
public:
    RayLauncher(DeviceContext& ctx, GeometryUploader& geometry, int axis = 2);
    ~RayLauncher() { free(); }
    RayLauncher(const RayLauncher&) = delete;
    RayLauncher& operator=(const RayLauncher&) = delete;
    void uploadRays(const Vec3* rayOrigins, size_t numRays);
    void uploadRays(const std::vector<Vec3>& rayOrigins) { uploadRays(rayOrigins.data(), rayOrigins.size()); }
    void launch(AccelerationStructure& as, int numRays);
    void runWarmup(AccelerationStructure& as, int warmupCount);
    RayResult* getResultBuffer() const { return results_.as<RayResult>(); }
    const float* getRayOrigins() const { return origins_.as<float>(); }
    size_t getNumRays() const { return numRays_; }
    void free();
    bool isUploaded() const { return origins_.as<void>() != nullptr; }

private:
    DeviceContext& ctx_;
    GeometryUploader& geometry_;
    int axis_;
    DeviceBuffer origins_, results_;
    size_t numRays_ = 0;
};

class ResultProcessor {
public:
    std::vector<RayResult> compactAndDownload(RayResult* d_results, int numResults);
    int countHits(RayResult* d_results, int numResults);
    std::vector<int> mapToOriginalIndices(const std::vector<RayResult>& hits,
                                          const std::vector<int>& candidateIndices, size_t totalPoints);
};

struct FilterRefineResult {
    std::vector<int> candidateIndices;
    std::vector<RayResult> filterHits;
    std::vector<RayResult> refineHits;
    std::vector<int> finalResults;
    size_t totalPoints = 0, filteredOut = 0, candidateCount = 0, insideCount = 0, outsideCount = 0;
};

class FilterRefinePipeline {
public:
    explicit FilterRefinePipeline(DeviceContext& ctx, int axis = 2) : ctx_(ctx), axis_(axis) {}
    FilterRefineResult executeFilter(const GeometryData& filterGeometry, const PointData& pointData);
    FilterRefineResult executeRefine(const GeometryData& refineGeometry,
                                     const std::vector<Vec3>& candidatePoints,
                                     const std::vector<int>& candidateIndices, size_t totalPoints);
    // Device-resident filter, stable compaction and refine in one pass.
    FilterRefineResult execute(const GeometryData& queryGeometry, const PointData& pointData);

private:
    DeviceContext& ctx_;
    int axis_;
};

// Device copy of an edge payload.
struct EdgeMeshData {
    DeviceBuffer starts, ends, offsets, counts, sources;
    rs3d_edges view{};
    int num_edges = 0;
};

class PrecomputedEdgeData {
public:
    static EdgeMeshData uploadFromGeometry(const GeometryData& geometry);
    // Upload, then the edge records in Morton order of their midpoints
    // (rs3d_edges_spatial_order): faster edge-ray joins, the same join results.
    static EdgeMeshData uploadFromGeometry(const GeometryData& geometry, DeviceContext& ctx);
    static void freeEdgeData(EdgeMeshData& data);
};

// Edge-ray launches (two-pass or hash) against a target structure.
struct MeshOverlapEdgesLaunchParams {
    const rs3d_edges* edges = nullptr;
    rs3d_bvh* target = nullptr;
    rs3d_hash_table* hash_table = nullptr;  // use_hash_table when non-null
    int32_t* collision_counts = nullptr;    // pass 1
    int64_t* collision_offsets = nullptr;   // pass 2
    MeshQueryResult* results = nullptr;     // pass 2
    int pass = 1;
    int swap_pair_order = 0;
};

class MeshOverlapEdgesLauncher {
public:
    explicit MeshOverlapEdgesLauncher(DeviceContext& ctx) : ctx_(ctx) {}
    void launchMesh1ToMesh2(const MeshOverlapEdgesLaunchParams& params, int numEdges);
    void launchMesh2ToMesh1(const MeshOverlapEdgesLaunchParams& params, int numEdges);

private:
    void launchInternal(const MeshOverlapEdgesLaunchParams& params, int numEdges);
    DeviceContext& ctx_;
};

// Intersection-join launches (query/src/raytracing/MeshIntersectionLauncher.h:9-60) over the
// fields of MeshIntersectionLaunchParams (query/src/cuda/mesh_intersection.h:24-88): mesh1 is
// the source of the launch (the 2->1 launches take params whose mesh1 fields hold mesh 2),
// mesh2 the structure it is traced against (mesh2_handle). Every launch is synchronous.
struct MeshIntersectionLaunchParams {
    rs3d_mesh_view mesh1{};                 // vertices, indices, triangle_to_object, objects
    const rs3d_edges* edges = nullptr;      // precomputed source edge rays (overlap phase)
    rs3d_bvh* mesh2 = nullptr;              // target structure
    rs3d_hash_table* hash_table = nullptr;  // use_hash_table when non-null
    int32_t* collision_counts = nullptr;    // two-pass form: pass 1 counts
    int64_t* collision_offsets = nullptr;   // pass 2 write offsets
    MeshQueryResult* results = nullptr;     // pass 2 output
    int pass = 1;
    int use_anyhit_containment = 1;         // 0: the closest-hit nextafter loop
    int overlap_max_iterations = 100;
    int containment_max_iterations = 1024;
    rs3d_intersection_stats* profiling_stats = nullptr;  // device counters; profiling when set
    int32_t* pair_target_ids = nullptr;     // pair-hit tracking (device, sources x max)
    uint32_t* pair_target_hits = nullptr;
    int max_pair_targets_per_source = 0;
    uint32_t* containment_iterations_per_source = nullptr;  // containment tracking (device)
    uint32_t* containment_candidate_count_per_source = nullptr;
    uint32_t* containment_candidate_overflow_per_source = nullptr;
    int anyhit_max_pair_targets_per_source = 256;  // distinct targets tracked per source ray
};

class MeshIntersectionLauncher {
public:
    explicit MeshIntersectionLauncher(DeviceContext& ctx) : ctx_(ctx) {}
    MeshIntersectionLauncher(const MeshIntersectionLauncher&) = delete;
    MeshIntersectionLauncher& operator=(const MeshIntersectionLauncher&) = delete;
    void launchOverlapMesh1ToMesh2(const MeshIntersectionLaunchParams& params, int launchSize);
    void launchOverlapMesh2ToMesh1(const MeshIntersectionLaunchParams& params, int launchSize);
    void launchContainmentMesh1ToMesh2(const MeshIntersectionLaunchParams& params, int launchSize);
    void launchContainmentMesh2ToMesh1(const MeshIntersectionLaunchParams& params, int launchSize);
    void launchMesh1ToMesh2(const MeshIntersectionLaunchParams& params, int launchSize) {
        launchOverlapMesh1ToMesh2(params, launchSize);
    }
    void launchMesh2ToMesh1(const MeshIntersectionLaunchParams& params, int launchSize) {
        launchOverlapMesh2ToMesh1(params, launchSize);
    }
    bool isValid() const { return true; }
    // Targets beyond anyhit_max_pair_targets_per_source in the last containment launch.
    uint64_t lastContainmentOverflow() const { return containmentOverflow_; }

private:
    void overlap(const MeshIntersectionLaunchParams& p, int launchSize, int swap);
    void containment(const MeshIntersectionLaunchParams& p, int launchSize, int swap);
    DeviceContext& ctx_;
    uint64_t containmentOverflow_ = 0;
};

// Containment-join launches (query/src/raytracing/MeshContainmentLauncher.h:16-71) over the
// fields of MeshContainmentLaunchParams (query/src/cuda/mesh_containment.h:7-48).
struct MeshContainmentLaunchParams {
    const rs3d_edges* src_edges = nullptr;              // phase 1 rays
    rs3d_bvh* target = nullptr;                          // traced against (target_handle)
    rs3d_hash_table* intersection_hash_table = nullptr;  // phase 1 output, phase 2 lookup
    int swap_ids = 0;  // 0: B edges vs A, pairs (A hit, B source); 1: A edges vs B, (A source, B hit)
    const float* b_first_vertices = nullptr;             // device float3, one per B object
    int b_num_objects = 0;
    rs3d_hash_table* containment_hash_table = nullptr;   // phase 2 output
    int use_anyhit_point_in_mesh = 1;                    // 0: the closest-hit t + 1e-4 loop
    int anyhit_max_unique_a_objects = 512;
};

class MeshContainmentLauncher {
public:
    explicit MeshContainmentLauncher(DeviceContext& ctx) : ctx_(ctx) {}
    MeshContainmentLauncher(const MeshContainmentLauncher&) = delete;
    MeshContainmentLauncher& operator=(const MeshContainmentLauncher&) = delete;
    // Phase 1: edge rays of src_edges (the first numEdges) against target, nextafter chain
    // of at most 256 hits, pairs into the intersection table.
    void launchEdgeCheck(const MeshContainmentLaunchParams& params, int numEdges);
    // Phase 2: one +z ray per B first vertex against target (A); odd A objects give (a, b)
    // unless the pair is in the intersection table.
    void launchPointInMesh(const MeshContainmentLaunchParams& params, int numBObjects);
    bool isValid() const { return true; }
    uint64_t lastPointInMeshOverflow() const { return overflow_; }

private:
    DeviceContext& ctx_;
    uint64_t overflow_ = 0;
};

// Hash table owned on the device.
struct DeviceHashTable {
    DeviceBuffer buf;
    rs3d_hash_table table{};
    void create(uint64_t slots, bool bitwise, int maxProbes);
    void clear(DeviceContext& ctx);
    uint64_t failedInserts(DeviceContext& ctx) const;
};

// ---- preprocessing on the device (preprocess/src/preprocess_dataset.cpp:37-179, 346-353) ----
// Unique-edge payload of the geometry (rs3d_extract_edges), copied back to the host.
EdgeData extractEdgesOnDevice(DeviceContext& ctx, const GeometryData& geometry);
// Grid histogram (rs3d_grid_stats) at resolution^3 over the vertex bounds, or [0, worldSize]^3
// when worldSize > 0.
void generateGridStatsOnDevice(DeviceContext& ctx, const GeometryData& geometry, GridData& grid,
                               int resolution, float worldSize);

// Selectivity estimate of two .pre grids (raw / alpha), or -1 when grids are missing.
long long estimatePairs(DeviceContext& ctx, const GeometryData& a, const GeometryData& b,
                        float epsilon, float gamma, bool alphaCorrection, bool verbose,
                        bool intersection = false);

// Raw selectivity estimate of two grids of numCells cells (host arrays), the sum over
// cells of TouchA * TouchB * min((AvgA + AvgB + epsilon)^3 / cellVolume * sqrt(RatioA *
// RatioB)^gamma, 1) (query/src/cuda/estimated_overlap.cu:8-77); -1.0f when device memory
// cannot be allocated. Runs on the current device.
float estimateOverlapSelectivity(const GridCell* h_gridA, const GridCell* h_gridB, int numCells,
                                 float cellVolume, float epsilon = 0.001f, float gamma = 0.8f);
// The same estimate for the intersection join (query/src/cuda/estimated_intersection.cu).
float estimateIntersectionSelectivity(const GridCell* h_gridA, const GridCell* h_gridB, int numCells,
                                      float cellVolume, float epsilon = 0.001f, float gamma = 0.8f);
// Touch-weighted means of the cells' AvgSizeMean and VolRatio
// (query/src/applications/raytracer_overlap_estimated_main.cpp:129-158).
float calculateGlobalAvgSize(const std::vector<GridCell>& cells);
float calculateGlobalAvgVolRatio(const std::vector<GridCell>& cells);

// Split a comma list of paths (query/src/applications/rayspace_main.cpp:23-37).
std::vector<std::string> splitPaths(const std::string& input);

}  // namespace rs3d_host
