// rs3d_host.cpp -- implementation of the C++ host layer (see rs3d_host.hpp).
#include "rs3d_host.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace rs3d_host {

// ------------------------------------------------------------------------------------
// errors and device memory
// ------------------------------------------------------------------------------------
void check(int rc, const char* what) {
    if (rc != RS3D_OK) {
        std::cerr << "Error: " << what << " failed (" << rc << "): " << rs3d_last_error() << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::cerr << "HIP error in " << what << ": " << hipGetErrorString(e) << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
        release();
        ptr_ = o.ptr_;
        bytes_ = o.bytes_;
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    return *this;
}

void DeviceBuffer::fill(int byteValue) {
    if (bytes_) hip_check(hipMemset(ptr_, byteValue, bytes_), "hipMemset");
}

void DeviceBuffer::alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    hip_check(hipMalloc(&ptr_, bytes), "hipMalloc");
    bytes_ = bytes;
}

void DeviceBuffer::release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

void DeviceBuffer::upload(const void* src, size_t bytes) {
    if (bytes) hip_check(hipMemcpy(ptr_, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

void DeviceBuffer::download(void* dst, size_t bytes) const {
    if (bytes) hip_check(hipMemcpy(dst, ptr_, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

DeviceContext::DeviceContext(int device) { check(rs3d_ctx_create(device, &ctx_), "rs3d_ctx_create"); }
DeviceContext::~DeviceContext() { rs3d_ctx_destroy(ctx_); }
void DeviceContext::sync() { check(rs3d_ctx_sync(ctx_), "rs3d_ctx_sync"); }

// ------------------------------------------------------------------------------------
// .pre files (common/include/BinaryIO.h:11-180)
// ------------------------------------------------------------------------------------
namespace {
constexpr uint32_t kMagic = 0x52334442;  // "R3DB"
#pragma pack(push, 1)
struct HeaderV2 {
    uint32_t magic, version;
    uint64_t numVertices, numIndices, numMappings, numEdges, numEdgeSourceObjects, totalTriangles;
    uint8_t hasGrid, hasEdges, padding[6];
};
struct HeaderV1 {
    uint32_t magic, version;
    uint64_t numVertices, numIndices, numMappings, totalTriangles;
    uint8_t hasGrid, padding[7];
};
struct GridParams {
    float minBound[3], maxBound[3];
    uint32_t resolution[3], padding;
};
#pragma pack(pop)
static_assert(sizeof(HeaderV2) == 64 && sizeof(HeaderV1) == 48 && sizeof(GridParams) == 40, "layout");

template <typename T>
bool read_vec(std::ifstream& in, std::vector<T>& v, uint64_t n) {
    v.resize(n);
    if (n) in.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
    return bool(in);
}
}  // namespace

GeometryData readBinaryFile(const std::string& filename) {
    GeometryData g;
    std::ifstream in(filename, std::ios::binary);
    if (!in) {
        std::cerr << "Error: Could not open file for reading: " << filename << std::endl;
        return g;
    }
    uint32_t mv[2] = {0, 0};
    in.read(reinterpret_cast<char*>(mv), sizeof mv);
    if (!in || mv[0] != kMagic) {
        std::cerr << "Error: Invalid file format (Magic mismatch). File: " << filename << std::endl;
        return g;
    }
    in.seekg(0);
    uint64_t nv, ni, nm, ne = 0, nes = 0, tt;
    uint8_t hasGrid, hasEdges = 0;
    if (mv[1] == 2) {
        HeaderV2 h;
        in.read(reinterpret_cast<char*>(&h), sizeof h);
        nv = h.numVertices; ni = h.numIndices; nm = h.numMappings; ne = h.numEdges;
        nes = h.numEdgeSourceObjects; tt = h.totalTriangles; hasGrid = h.hasGrid; hasEdges = h.hasEdges;
    } else if (mv[1] == 1) {
        HeaderV1 h;
        in.read(reinterpret_cast<char*>(&h), sizeof h);
        nv = h.numVertices; ni = h.numIndices; nm = h.numMappings; tt = h.totalTriangles;
        hasGrid = h.hasGrid;
    } else {
        std::cerr << "Error: Unsupported binary geometry version " << mv[1] << " in file: " << filename
                  << ". Expected version 2 (or 1)." << std::endl;
        return g;
    }
    GeometryData out;
    out.totalTriangles = tt;
    bool ok = read_vec(in, out.vertices, nv) && read_vec(in, out.indices, ni) &&
              read_vec(in, out.triangleToObject, nm);
    if (ok && hasEdges) {
        ok = read_vec(in, out.edges.edgeStarts, ne) && read_vec(in, out.edges.edgeEnds, ne) &&
             read_vec(in, out.edges.sourceObjectOffsets, ne) &&
             read_vec(in, out.edges.sourceObjectCounts, ne) && read_vec(in, out.edges.sourceObjects, nes);
    }
    if (ok && hasGrid) {
        GridParams gp;
        in.read(reinterpret_cast<char*>(&gp), sizeof gp);
        out.grid.hasGrid = true;
        out.grid.minBound = {gp.minBound[0], gp.minBound[1], gp.minBound[2]};
        out.grid.maxBound = {gp.maxBound[0], gp.maxBound[1], gp.maxBound[2]};
        for (int k = 0; k < 3; ++k) out.grid.resolution[k] = gp.resolution[k];
        ok = bool(in) && read_vec(in, out.grid.cells, (uint64_t)gp.resolution[0] * gp.resolution[1] * gp.resolution[2]);
    }
    if (!ok) {
        std::cerr << "Error: truncated geometry file: " << filename << std::endl;
        return g;
    }
    return out;
}

bool writeBinaryFile(const std::string& filename, const GeometryData& geometry) {
    std::ofstream out(filename, std::ios::binary);
    if (!out) {
        std::cerr << "Error: Could not open file for writing: " << filename << std::endl;
        return false;
    }
    HeaderV2 h;
    std::memset(&h, 0, sizeof h);
    h.magic = kMagic;
    h.version = 2;
    h.numVertices = geometry.vertices.size();
    h.numIndices = geometry.indices.size();
    h.numMappings = geometry.triangleToObject.size();
    h.numEdges = geometry.edges.edgeStarts.size();
    h.numEdgeSourceObjects = geometry.edges.sourceObjects.size();
    h.totalTriangles = geometry.totalTriangles;
    h.hasGrid = geometry.grid.hasGrid ? 1 : 0;
    h.hasEdges = geometry.edges.hasEdges() ? 1 : 0;
    out.write(reinterpret_cast<const char*>(&h), sizeof h);
    auto wr = [&](const void* p, size_t n) { if (n) out.write(static_cast<const char*>(p), n); };
    wr(geometry.vertices.data(), h.numVertices * sizeof(Vec3));
    wr(geometry.indices.data(), h.numIndices * sizeof(UInt3));
    wr(geometry.triangleToObject.data(), h.numMappings * sizeof(int));
    if (h.hasEdges) {
        const EdgeData& e = geometry.edges;
        wr(e.edgeStarts.data(), h.numEdges * sizeof(Vec3));
        wr(e.edgeEnds.data(), h.numEdges * sizeof(Vec3));
        wr(e.sourceObjectOffsets.data(), h.numEdges * sizeof(int));
        wr(e.sourceObjectCounts.data(), h.numEdges * sizeof(int));
        wr(e.sourceObjects.data(), h.numEdgeSourceObjects * sizeof(int));
    }
    if (h.hasGrid) {
        GridParams gp;
        std::memset(&gp, 0, sizeof gp);
        const GridData& gd = geometry.grid;
        gp.minBound[0] = gd.minBound.x; gp.minBound[1] = gd.minBound.y; gp.minBound[2] = gd.minBound.z;
        gp.maxBound[0] = gd.maxBound.x; gp.maxBound[1] = gd.maxBound.y; gp.maxBound[2] = gd.maxBound.z;
        for (int k = 0; k < 3; ++k) gp.resolution[k] = gd.resolution[k];
        out.write(reinterpret_cast<const char*>(&gp), sizeof gp);
        wr(gd.cells.data(), gd.cells.size() * sizeof(GridCell));
    }
    return bool(out);
}

GeometryData loadGeometryFromFile(const std::string& path) { return readBinaryFile(path); }

bool requirePrecomputedEdges(const GeometryData& g, const std::string& path, const char* label) {
    if (g.edges.hasEdges()) return true;
    std::cerr << "Error: " << label << " loaded from '" << path
              << "' does not contain precomputed edges." << std::endl;
    std::cerr << "Re-run preprocessing to regenerate the .pre file with embedded edge payload."
              << std::endl;
    return false;
}

size_t countObjects(const GeometryData& g) {
    int mx = -1;
    for (int o : g.triangleToObject) mx = std::max(mx, o);
    return (size_t)(mx + 1);
}

std::vector<std::string> splitPaths(const std::string& input) {
    std::vector<std::string> out;
    std::stringstream ss(input);
    std::string item;
    while (std::getline(ss, item, ',')) {
        const size_t a = item.find_first_not_of(" \t"), b = item.find_last_not_of(" \t");
        if (a != std::string::npos) out.push_back(item.substr(a, b - a + 1));
    }
    return out;
}

// ------------------------------------------------------------------------------------
// timing JSON (query/src/timer.cpp:124-192 key scheme)
// ------------------------------------------------------------------------------------
PerformanceTimer::PerformanceTimer() { total_start = std::chrono::high_resolution_clock::now(); }

void PerformanceTimer::start(const std::string& name) {
    if (is_running) endCurrentPhase();
    if (phases.empty()) total_start = std::chrono::high_resolution_clock::now();
    Phase p;
    p.name = name;
    p.start_time = std::chrono::high_resolution_clock::now();
    phases.push_back(p);
    is_running = true;
}

void PerformanceTimer::next(const std::string& name) { start(name); }

void PerformanceTimer::endCurrentPhase() {
    if (!is_running || phases.empty()) return;
    Phase& p = phases.back();
    p.end_time = std::chrono::high_resolution_clock::now();
    p.duration_us = std::chrono::duration_cast<std::chrono::microseconds>(p.end_time - p.start_time).count();
    is_running = false;
}

void PerformanceTimer::addMeasurement(const std::string& name, long long us) {
    Phase p;
    p.name = name;
    p.duration_us = us;
    manualMeasurements.push_back(p);
}

void PerformanceTimer::addCounter(const std::string& name, unsigned long long v) { counters.push_back({name, v}); }

long long PerformanceTimer::getPhaseDuration(const std::string& name) const {
    for (const auto& p : phases)
        if (p.name == name) return p.duration_us;
    return -1;
}

long long PerformanceTimer::getTotalDuration() const {
    return std::chrono::duration_cast<std::chrono::microseconds>(total_end - total_start).count();
}

void PerformanceTimer::finish(const std::string& filename) {
    endCurrentPhase();
    total_end = std::chrono::high_resolution_clock::now();
    std::cout << "\n=== Performance Timing ===" << std::endl;
    for (const auto& p : phases)
        std::cout << std::left << std::setw(30) << (p.name + ":") << std::right << std::setw(12)
                  << p.duration_us << " us (" << std::fixed << std::setprecision(2)
                  << p.duration_us / 1000.0 << " ms)" << std::endl;
    writeResultsToFile(filename);
}

static std::string lower(const std::string& s) {
    std::string o;
    o.reserve(s.size());
    for (char c : s) o.push_back((char)std::tolower((unsigned char)c));
    return o;
}

void PerformanceTimer::writeResultsToFile(const std::string& filename) const {
    std::ofstream f(filename);
    if (!f) {
        std::cerr << "Warning: Could not write timing results to file." << std::endl;
        return;
    }
    f << "{\n  \"phases\": {\n";
    std::unordered_map<std::string, int> seen;
    std::vector<Phase> all(phases);
    all.insert(all.end(), manualMeasurements.begin(), manualMeasurements.end());
    for (size_t i = 0; i < all.size(); ++i) {
        const std::string key = lower(all[i].name) + "_" + std::to_string(++seen[lower(all[i].name)]);
        f << "    \"" << key << "\": {\n      \"duration_us\": " << all[i].duration_us
          << ",\n      \"duration_ms\": " << std::fixed << std::setprecision(2)
          << all[i].duration_us / 1000.0 << "\n    }" << (i + 1 < all.size() ? "," : "") << "\n";
    }
    f << "  },\n  \"counters\": {\n";
    for (size_t i = 0; i < counters.size(); ++i)
        f << "    \"" << lower(counters[i].name) << "\": " << counters[i].value
          << (i + 1 < counters.size() ? "," : "") << "\n";
    f << "  },\n";
    const long long total = getTotalDuration();
    if (total > 0)
        f << "  \"total\": {\n    \"duration_us\": " << total << ",\n    \"duration_ms\": " << std::fixed
          << std::setprecision(2) << total / 1000.0 << "\n  }\n";
    else
        f << "  \"total\": null\n";
    f << "}\n";
    std::cout << "Performance timing written to " << filename << std::endl;
}

// ------------------------------------------------------------------------------------
// geometry upload, BVH, ray launches
// ------------------------------------------------------------------------------------
void GeometryUploader::upload(const GeometryData& g) {
    numVertices_ = g.vertices.size();
    numTriangles_ = g.indices.size();
    numObjects_ = countObjects(g);
    vertices_.alloc(numVertices_ * sizeof(Vec3));
    vertices_.upload(g.vertices.data(), numVertices_ * sizeof(Vec3));
    indices_.alloc(numTriangles_ * sizeof(UInt3));
    indices_.upload(g.indices.data(), numTriangles_ * sizeof(UInt3));
    triObj_.alloc(numTriangles_ * sizeof(int));
    triObj_.upload(g.triangleToObject.data(), std::min(g.triangleToObject.size(), numTriangles_) * sizeof(int));
}

rs3d_mesh_view GeometryUploader::view() const {
    rs3d_mesh_view v;
    v.vertices = getVertices();
    v.num_vertices = numVertices_;
    v.indices = getIndices();
    v.num_triangles = numTriangles_;
    v.tri_to_obj = getTriangleToObject();
    v.num_objects = numObjects_;
    return v;
}

AccelerationStructure::AccelerationStructure(DeviceContext& ctx, GeometryUploader& g, int leafSize)
    : ctx_(ctx), geometry_(g), leafSize_(leafSize) {}

AccelerationStructure::~AccelerationStructure() { rs3d_bvh_free(bvh_); }

void AccelerationStructure::build() {
    rs3d_bvh_free(bvh_);
    bvh_ = nullptr;
    rs3d_bvh_opts o;
    std::memset(&o, 0, sizeof o);
    o.leaf_size = leafSize_;
    check(rs3d_bvh_build(ctx_.get(), geometry_.getVertices(), geometry_.getNumVertices(),
                         geometry_.getIndices(), geometry_.getNumTriangles(),
                         geometry_.getTriangleToObject(), &o, &bvh_),
          "rs3d_bvh_build");
}

bool AccelerationStructure::prepareAxis(int axis) {
    if (!bvh_) return false;
    rs3d_axis_grid_info info;
    check(rs3d_axis_grid_prepare(ctx_.get(), bvh_, axis, &info), "rs3d_axis_grid_prepare");
    return info.state == 1;
}

RayLauncher::RayLauncher(DeviceContext& ctx, GeometryUploader& g, int axis)
    : ctx_(ctx), geometry_(g), axis_(axis) {}

void RayLauncher::uploadRays(const Vec3* o, size_t n) {
    free();
    numRays_ = n;
    if (!n) return;
    origins_.alloc(n * sizeof(Vec3));
    origins_.upload(o, n * sizeof(Vec3));
    results_.alloc(n * sizeof(RayResult));
    hip_check(hipMemset(results_.as<void>(), 0, n * sizeof(RayResult)), "hipMemset");
}

void RayLauncher::launch(AccelerationStructure& as, int numRays) {
    if (!isUploaded() || numRays <= 0 || (size_t)numRays > numRays_) return;
    if (geometry_.getNumTriangles() == 0 && as.getHandle() == nullptr) return;
    check(rs3d_trace_points(ctx_.get(), as.getHandle(), origins_.as<float>(), (uint64_t)numRays, nullptr,
                            axis_, RS3D_OUT_RAY_RESULT, results_.as<void>()),
          "rs3d_trace_points");
    ctx_.sync();
}

void RayLauncher::runWarmup(AccelerationStructure& as, int k) {
    for (int i = 0; i < k; ++i) launch(as, (int)numRays_);
}

void RayLauncher::free() {
    origins_.release();
    results_.release();
    numRays_ = 0;
}

std::vector<RayResult> ResultProcessor::compactAndDownload(RayResult* d, int n) {
    std::vector<RayResult> hits;
    if (n <= 0 || !d) return hits;
    DeviceBuffer tmp(n * sizeof(RayResult));
    int h = 0;
    compact_hits_gpu(d, tmp.as<RayResult>(), n, &h);
    hits.resize(h);
    tmp.download(hits.data(), h * sizeof(RayResult));
    return hits;
}

int ResultProcessor::countHits(RayResult* d, int n) { return (n <= 0 || !d) ? 0 : count_hits_gpu(d, n); }

std::vector<int> ResultProcessor::mapToOriginalIndices(const std::vector<RayResult>& hits,
                                                       const std::vector<int>& cand, size_t total) {
    std::vector<int> out(total, -1);
    for (const auto& h : hits)
        if (h.ray_id >= 0 && (size_t)h.ray_id < cand.size()) {
            const int o = cand[h.ray_id];
            if (o >= 0 && (size_t)o < total) out[o] = h.polygon_index;
        }
    return out;
}

FilterRefineResult FilterRefinePipeline::executeFilter(const GeometryData& filterGeometry,
                                                       const PointData& pd) {
    FilterRefineResult r;
    r.totalPoints = pd.numPoints;
    GeometryUploader up;
    up.upload(filterGeometry);
    AccelerationStructure as(ctx_, up);
    as.build();
    RayLauncher l(ctx_, up, axis_);
    l.uploadRays(pd.positions.data(), pd.numPoints);
    l.launch(as, (int)pd.numPoints);
    ResultProcessor proc;
    r.filterHits = proc.compactAndDownload(l.getResultBuffer(), (int)pd.numPoints);
    for (const auto& h : r.filterHits) r.candidateIndices.push_back(h.ray_id);
    r.candidateCount = r.candidateIndices.size();
    r.filteredOut = r.totalPoints - r.candidateCount;
    return r;
}

FilterRefineResult FilterRefinePipeline::executeRefine(const GeometryData& g,
                                                       const std::vector<Vec3>& pts,
                                                       const std::vector<int>& cand, size_t total) {
    FilterRefineResult r;
    r.totalPoints = total;
    r.candidateIndices = cand;
    r.candidateCount = pts.size();
    if (pts.empty()) {
        r.filteredOut = total;
        r.outsideCount = total;
        r.finalResults.assign(total, -1);
        return r;
    }
    GeometryUploader up;
    up.upload(g);
    AccelerationStructure as(ctx_, up);
    as.build();
    RayLauncher l(ctx_, up, axis_);
    l.uploadRays(pts);
    l.launch(as, (int)pts.size());
    ResultProcessor proc;
    r.refineHits = proc.compactAndDownload(l.getResultBuffer(), (int)pts.size());
    r.finalResults = proc.mapToOriginalIndices(r.refineHits, cand, total);
    r.insideCount = r.refineHits.size();
    r.outsideCount = total - r.insideCount;
    return r;
}

FilterRefineResult FilterRefinePipeline::execute(const GeometryData& g, const PointData& pd) {
    FilterRefineResult r;
    r.totalPoints = pd.numPoints;
    if (pd.numPoints == 0) return r;
    GeometryUploader up;
    up.upload(g);
    AccelerationStructure as(ctx_, up);
    as.build();
    DeviceBuffer pts(pd.numPoints * sizeof(Vec3)), out(pd.numPoints * sizeof(int));
    pts.upload(pd.positions.data(), pd.numPoints * sizeof(Vec3));
    uint64_t cand = 0, inside = 0;
    check(rs3d_filter_refine(ctx_.get(), as.getHandle(), pts.as<float>(), pd.numPoints, axis_,
                             out.as<int32_t>(), &cand, &inside),
          "rs3d_filter_refine");
    r.finalResults.resize(pd.numPoints);
    out.download(r.finalResults.data(), pd.numPoints * sizeof(int));
    r.candidateCount = cand;
    r.filteredOut = pd.numPoints - cand;
    r.insideCount = inside;
    r.outsideCount = pd.numPoints - inside;
    for (size_t i = 0; i < pd.numPoints; ++i)
        if (r.finalResults[i] >= 0) r.refineHits.push_back({(int32_t)i, r.finalResults[i]});
    return r;
}

// ------------------------------------------------------------------------------------
// edges, launches, hash tables
// ------------------------------------------------------------------------------------
EdgeMeshData PrecomputedEdgeData::uploadFromGeometry(const GeometryData& g) {
    EdgeMeshData d;
    const EdgeData& e = g.edges;
    const size_t ne = e.numEdges();
    d.num_edges = (int)ne;
    d.starts.alloc(ne * sizeof(Vec3));
    d.starts.upload(e.edgeStarts.data(), ne * sizeof(Vec3));
    d.ends.alloc(ne * sizeof(Vec3));
    d.ends.upload(e.edgeEnds.data(), ne * sizeof(Vec3));
    d.offsets.alloc(ne * sizeof(int));
    d.offsets.upload(e.sourceObjectOffsets.data(), ne * sizeof(int));
    d.counts.alloc(ne * sizeof(int));
    d.counts.upload(e.sourceObjectCounts.data(), ne * sizeof(int));
    d.sources.alloc(std::max<size_t>(e.sourceObjects.size(), 1) * sizeof(int));
    d.sources.upload(e.sourceObjects.data(), e.sourceObjects.size() * sizeof(int));
    d.view.starts = d.starts.as<float>();
    d.view.ends = d.ends.as<float>();
    d.view.src_offsets = d.offsets.as<int32_t>();
    d.view.src_counts = d.counts.as<int32_t>();
    d.view.src_objects = d.sources.as<int32_t>();
    d.view.num_edges = ne;
    return d;
}

EdgeMeshData PrecomputedEdgeData::uploadFromGeometry(const GeometryData& g, DeviceContext& ctx) {
    EdgeMeshData d = uploadFromGeometry(g);
    const size_t ne = (size_t)d.num_edges;
    if (ne < 2) return d;
    DeviceBuffer st(ne * sizeof(Vec3)), en(ne * sizeof(Vec3)), off(ne * sizeof(int)), cnt(ne * sizeof(int));
    check(rs3d_edges_spatial_order(ctx.get(), &d.view, st.as<float>(), en.as<float>(), off.as<int32_t>(),
                                   cnt.as<int32_t>()),
          "rs3d_edges_spatial_order");
    d.starts = std::move(st);
    d.ends = std::move(en);
    d.offsets = std::move(off);
    d.counts = std::move(cnt);
    d.view.starts = d.starts.as<float>();
    d.view.ends = d.ends.as<float>();
    d.view.src_offsets = d.offsets.as<int32_t>();
    d.view.src_counts = d.counts.as<int32_t>();
    return d;
}

void PrecomputedEdgeData::freeEdgeData(EdgeMeshData& d) {
    d.starts.release();
    d.ends.release();
    d.offsets.release();
    d.counts.release();
    d.sources.release();
    d.view = rs3d_edges{};
    d.num_edges = 0;
}

void MeshOverlapEdgesLauncher::launchInternal(const MeshOverlapEdgesLaunchParams& p, int numEdges) {
    rs3d_edges e = *p.edges;
    e.num_edges = (uint64_t)std::max(numEdges, 0);
    if (p.hash_table)
        check(rs3d_edge_hash(ctx_.get(), p.target, &e, p.swap_pair_order, p.hash_table), "rs3d_edge_hash");
    else if (p.pass == 1)
        check(rs3d_edge_count(ctx_.get(), p.target, &e, p.collision_counts), "rs3d_edge_count");
    else
        check(rs3d_edge_write(ctx_.get(), p.target, &e, p.collision_offsets, p.swap_pair_order, p.results),
              "rs3d_edge_write");
    ctx_.sync();
}

void MeshOverlapEdgesLauncher::launchMesh1ToMesh2(const MeshOverlapEdgesLaunchParams& p, int n) {
    launchInternal(p, n);
}
void MeshOverlapEdgesLauncher::launchMesh2ToMesh1(const MeshOverlapEdgesLaunchParams& p, int n) {
    launchInternal(p, n);
}

void MeshIntersectionLauncher::overlap(const MeshIntersectionLaunchParams& p, int launchSize, int swap) {
    rs3d_edges e = *p.edges;
    e.num_edges = std::min<uint64_t>((uint64_t)std::max(launchSize, 0), e.num_edges);
    const int cap = p.overlap_max_iterations > 0 ? p.overlap_max_iterations : 100;
    if (p.profiling_stats)
        check(rs3d_edge_profile(ctx_.get(), p.mesh2, &e, RS3D_CHAIN_EPS, cap, p.profiling_stats),
              "rs3d_edge_profile");
    if (p.hash_table)
        check(rs3d_edge_hash_ex(ctx_.get(), p.mesh2, &e, swap, RS3D_CHAIN_EPS, cap, p.hash_table),
              "rs3d_edge_hash_ex");
    else if (p.pass == 1)
        check(rs3d_edge_count_ex(ctx_.get(), p.mesh2, &e, RS3D_CHAIN_EPS, cap, p.collision_counts),
              "rs3d_edge_count_ex");
    else
        check(rs3d_edge_write_ex(ctx_.get(), p.mesh2, &e, RS3D_CHAIN_EPS, cap, p.collision_offsets, swap, p.results),
              "rs3d_edge_write_ex");
    ctx_.sync();
}

void MeshIntersectionLauncher::containment(const MeshIntersectionLaunchParams& p, int launchSize, int swap) {
    rs3d_mesh_view src = p.mesh1;
    src.num_objects = std::min<uint64_t>((uint64_t)std::max(launchSize, 0), src.num_objects);
    rs3d_object_ray_opts o{};
    o.chain_mode = p.use_anyhit_containment ? RS3D_OBJ_ALL_HITS : RS3D_OBJ_CHAIN_NEXTAFTER;
    o.max_iterations = p.containment_max_iterations;
    o.pair_target_ids = p.pair_target_ids;
    o.pair_target_hits = p.pair_target_hits;
    o.max_pair_targets = p.max_pair_targets_per_source;
    o.iterations = p.containment_iterations_per_source;
    o.candidate_count = p.containment_candidate_count_per_source;
    o.candidate_overflow = p.containment_candidate_overflow_per_source;
    o.stats = p.profiling_stats;
    if (!p.hash_table) {
        if (p.pass == 1) o.counts = p.collision_counts;
        else {
            o.offsets = p.collision_offsets;
            o.results = p.results;
        }
    }
    containmentOverflow_ = 0;
    check(rs3d_object_rays(ctx_.get(), &src, p.mesh2, RS3D_OBJ_FALLBACK, swap, p.anyhit_max_pair_targets_per_source,
                           &o, p.hash_table, nullptr, &containmentOverflow_),
          "rs3d_object_rays");
    ctx_.sync();
}

void MeshIntersectionLauncher::launchOverlapMesh1ToMesh2(const MeshIntersectionLaunchParams& p, int n) { overlap(p, n, 0); }
void MeshIntersectionLauncher::launchOverlapMesh2ToMesh1(const MeshIntersectionLaunchParams& p, int n) { overlap(p, n, 1); }
void MeshIntersectionLauncher::launchContainmentMesh1ToMesh2(const MeshIntersectionLaunchParams& p, int n) {
    containment(p, n, 0);
}
void MeshIntersectionLauncher::launchContainmentMesh2ToMesh1(const MeshIntersectionLaunchParams& p, int n) {
    containment(p, n, 1);
}

void MeshContainmentLauncher::launchEdgeCheck(const MeshContainmentLaunchParams& p, int numEdges) {
    rs3d_edges e = *p.src_edges;
    e.num_edges = std::min<uint64_t>((uint64_t)std::max(numEdges, 0), e.num_edges);
    // swap_ids 0 stores (hit, source), which is this library's swapped order
    check(rs3d_edge_hash_ex(ctx_.get(), p.target, &e, p.swap_ids == 0 ? 1 : 0, RS3D_CHAIN_NEXTAFTER, 256,
                            p.intersection_hash_table),
          "rs3d_edge_hash_ex");
    ctx_.sync();
}

void MeshContainmentLauncher::launchPointInMesh(const MeshContainmentLaunchParams& p, int numBObjects) {
    rs3d_object_ray_opts o{};
    o.chain_mode = p.use_anyhit_point_in_mesh ? RS3D_OBJ_ALL_HITS : RS3D_OBJ_CHAIN_STEP;
    o.max_iterations = 2000;
    o.origins = p.b_first_vertices;
    o.num_sources = (uint64_t)std::max(std::min(numBObjects, p.b_num_objects), 0);
    overflow_ = 0;
    if (o.num_sources == 0) return;
    check(rs3d_object_rays(ctx_.get(), nullptr, p.target, RS3D_OBJ_POINT_IN_MESH, 0, p.anyhit_max_unique_a_objects,
                           &o, p.containment_hash_table, p.intersection_hash_table, &overflow_),
          "rs3d_object_rays");
    ctx_.sync();
}

void DeviceHashTable::create(uint64_t slots, bool bitwise, int maxProbes) {
    buf.alloc(slots * sizeof(uint64_t) + 256);
    table.slots = buf.as<uint64_t>();
    table.size = slots;
    table.bitwise = bitwise ? 1 : 0;
    table.max_probes = maxProbes;
    table.failed = reinterpret_cast<uint64_t*>(buf.as<uint8_t>() + slots * sizeof(uint64_t));
}

void DeviceHashTable::clear(DeviceContext& ctx) { check(rs3d_ht_clear(ctx.get(), &table), "rs3d_ht_clear"); }

uint64_t DeviceHashTable::failedInserts(DeviceContext& ctx) const {
    ctx.sync();
    uint64_t f = 0;
    hip_check(hipMemcpy(&f, table.failed, sizeof f, hipMemcpyDeviceToHost), "hipMemcpy");
    return f;
}

namespace {
// The library context of the current device for the context-free estimator entry points.
rs3d_ctx* current_device_ctx() {
    static std::mutex mu;
    static rs3d_ctx* ctxs[64] = {nullptr};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!ctxs[dev] && rs3d_ctx_create(dev, &ctxs[dev]) != RS3D_OK) ctxs[dev] = nullptr;
    return ctxs[dev];
}

float estimate_selectivity(const GridCell* a, const GridCell* b, int numCells, float cellVolume, float epsilon,
                           float gamma, const char* what) {
    if (numCells <= 0) return 0.0f;
    rs3d_ctx* ctx = current_device_ctx();
    if (!ctx) return -1.0f;
    const size_t bytes = (size_t)numCells * sizeof(GridCell);
    void *da = nullptr, *db = nullptr;
    if (hipMalloc(&da, bytes) != hipSuccess) return -1.0f;
    if (hipMalloc(&db, bytes) != hipSuccess) {
        (void)hipFree(da);
        return -1.0f;
    }
    hip_check(hipMemcpy(da, a, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(db, b, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    double total = 0.0;
    const int rc = rs3d_estimate_raw(ctx, da, db, (uint64_t)numCells, cellVolume, epsilon, gamma, &total);
    (void)hipFree(da);
    (void)hipFree(db);
    check(rc, what);
    return (float)total;
}
}  // namespace

float estimateOverlapSelectivity(const GridCell* h_gridA, const GridCell* h_gridB, int numCells, float cellVolume,
                                 float epsilon, float gamma) {
    return estimate_selectivity(h_gridA, h_gridB, numCells, cellVolume, epsilon, gamma, "estimateOverlapSelectivity");
}

float estimateIntersectionSelectivity(const GridCell* h_gridA, const GridCell* h_gridB, int numCells,
                                      float cellVolume, float epsilon, float gamma) {
    return estimate_selectivity(h_gridA, h_gridB, numCells, cellVolume, epsilon, gamma,
                                "estimateIntersectionSelectivity");
}

float calculateGlobalAvgSize(const std::vector<GridCell>& cells) {
    double total = 0.0;
    long long count = 0;
    for (const GridCell& c : cells)
        if (c.TouchCount > 0) {
            total += (double)c.AvgSizeMean * (double)c.TouchCount;
            count += c.TouchCount;
        }
    return count ? (float)(total / count) : 0.0f;
}

float calculateGlobalAvgVolRatio(const std::vector<GridCell>& cells) {
    double total = 0.0;
    long long count = 0;
    for (const GridCell& c : cells)
        if (c.TouchCount > 0) {
            total += (double)c.VolRatio * (double)c.TouchCount;
            count += c.TouchCount;
        }
    return count ? (float)(total / count) : 1.0f;
}

long long estimatePairs(DeviceContext& ctx, const GeometryData& a, const GeometryData& b, float eps,
                        float gamma, bool alpha, bool verbose, bool intersection) {
    (void)ctx;
    if (!a.grid.hasGrid || !b.grid.hasGrid) {
        if (verbose) std::cout << "Skipping estimation: Grid data not found." << std::endl;
        return -1;
    }
    const size_t n = (size_t)a.grid.resolution[0] * a.grid.resolution[1] * a.grid.resolution[2];
    if (a.grid.cells.size() != n || b.grid.cells.size() != n) {
        if (verbose) std::cerr << "Error: Grid mismatch. Skipping estimation." << std::endl;
        return -1;
    }
    // raw estimate on the device, then the replication correction on the host
    // (query/src/applications/raytracer_overlap_estimated_main.cpp:255-292)
    const float width = a.grid.maxBound.x - a.grid.minBound.x;
    const float height = a.grid.maxBound.y - a.grid.minBound.y;
    const float depth = a.grid.maxBound.z - a.grid.minBound.z;
    const float csX = width / a.grid.resolution[0], csY = height / a.grid.resolution[1],
                csZ = depth / a.grid.resolution[2];
    float cellVolume = csX * csY * csZ;
    const float raw = intersection
                          ? estimateIntersectionSelectivity(a.grid.cells.data(), b.grid.cells.data(), (int)n,
                                                            cellVolume, eps, gamma)
                          : estimateOverlapSelectivity(a.grid.cells.data(), b.grid.cells.data(), (int)n, cellVolume,
                                                       eps, gamma);
    if (raw < 0.0f) {
        std::cerr << "Error: selectivity estimation could not allocate device memory" << std::endl;
        return -1;
    }
    if (!alpha) return (long long)raw;
    const float avgSize1 = calculateGlobalAvgSize(a.grid.cells), avgSize2 = calculateGlobalAvgSize(b.grid.cells);
    const float avgVolRatio1 = calculateGlobalAvgVolRatio(a.grid.cells);
    const float avgVolRatio2 = calculateGlobalAvgVolRatio(b.grid.cells);
    const float effectiveSize1 = avgSize1 * std::cbrt(avgVolRatio1);
    const float effectiveSize2 = avgSize2 * std::cbrt(avgVolRatio2);
    const float combinedSize = effectiveSize1 + effectiveSize2;
    const float minkowskiVol = combinedSize * combinedSize * combinedSize;
    if (cellVolume < 1e-9f) cellVolume = 1e-9f;
    float alphaF = minkowskiVol / cellVolume;
    if (alphaF < 1.0f) alphaF = 1.0f;
    const long long est = (long long)(raw / alphaF);
    if (verbose) {
        std::cout << "\n=== Selectivity Estimation ===" << std::endl;
        std::cout << "Raw Potential Pairs:       " << (long long)raw << std::endl;
        std::cout << "Avg Object Size (Mesh1):   " << avgSize1 << std::endl;
        std::cout << "Avg Object Size (Mesh2):   " << avgSize2 << std::endl;
        std::cout << "Avg VolRatio (Mesh1):      " << avgVolRatio1 << std::endl;
        std::cout << "Avg VolRatio (Mesh2):      " << avgVolRatio2 << std::endl;
        std::cout << "Effective Size (Mesh1):    " << effectiveSize1 << std::endl;
        std::cout << "Effective Size (Mesh2):    " << effectiveSize2 << std::endl;
        std::cout << "Replication Factor (alpha):" << alphaF << std::endl;
        std::cout << "Final Estimated Pairs:     " << est << std::endl;
        std::cout << "==============================\n" << std::endl;
    }
    return est;
}

}  // namespace rs3d_host
