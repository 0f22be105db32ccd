// rs3d_io.cpp -- dataset input of the host layer: point files (text and binary), the OBJ
// mesh loader, the loader factory, and the device-side preprocessing helpers.
//
// Reference behaviour followed:
//   loadPointDataset        query/src/runtime/PointIO.cpp:12-72 (fgets into a 1024-byte
//                           line buffer; WKT "POINT ... (x y z)" or a bare "x y z" line;
//                           three default points for an empty path)
//   ObjMeshDatasetLoader    preprocess/src/ObjMeshDatasetLoader.cpp:14-169, over
//                           tinyobjloader's LoadObj (third-party, not vendored in the
//                           reference; its published v1.0.7+/v2.0 behaviour is restated here:
//                           number parser, face index rules, shape splitting at 'o'/'g',
//                           quad triangulation along the shorter diagonal)
//   DatasetLoaderFactory    preprocess/src/DatasetLoaderFactory.cpp:24-38
//   EdgePreprocessor / generateGridStats -> rs3d_extract_edges / rs3d_grid_stats
#include "rs3d_host.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <thread>

namespace rs3d_host {
namespace {

// Read-only mapping of a whole file (empty files map to nothing).
class MappedFile {
public:
    explicit MappedFile(const std::string& path) {
        fd_ = ::open(path.c_str(), O_RDONLY);
        if (fd_ < 0) return;
        struct stat st;
        if (::fstat(fd_, &st) != 0) return;
        size_ = (size_t)st.st_size;
        ok_ = true;
        if (size_ == 0) return;
        void* p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd_, 0);
        if (p == MAP_FAILED) {
            ok_ = false;
            return;
        }
        ::madvise(p, size_, MADV_SEQUENTIAL | MADV_WILLNEED);
        data_ = static_cast<const char*>(p);
    }
    ~MappedFile() {
        if (data_) ::munmap(const_cast<char*>(data_), size_);
        if (fd_ >= 0) ::close(fd_);
    }
    MappedFile(const MappedFile&) = delete;
    MappedFile& operator=(const MappedFile&) = delete;
    bool ok() const { return ok_; }
    const char* data() const { return data_; }
    size_t size() const { return size_; }

private:
    int fd_ = -1;
    const char* data_ = nullptr;
    size_t size_ = 0;
    bool ok_ = false;
};

int worker_count(size_t bytes, size_t bytes_per_worker) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t want = bytes / bytes_per_worker + 1;
    return (int)std::min<size_t>({want, (size_t)hw, (size_t)16});
}

// ---- points ----------------------------------------------------------------------------

constexpr uint32_t kPointMagic = 0x50335352u;  // "RS3P" little-endian
constexpr size_t kPointHeader = 32;
constexpr size_t kLineBuf = 1024;  // the reference's fgets buffer

// One piece of a line as fgets(buf, 1024, f) hands it over, parsed by the reference's rules.
void parse_point_piece(const char* line, std::vector<Vec3>& out) {
    const char* p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\0' || *p == '\r' || *p == '\n' || *p == '#') return;
    char* end = nullptr;
    if (std::strncmp(p, "POINT", 5) == 0) {
        const char* s = std::strchr(p, '(');
        if (!s) return;
        ++s;
        const float x = std::strtof(s, &end);
        const float y = std::strtof(end, &end);
        const float z = std::strtof(end, &end);
        out.push_back({x, y, z});
    } else if ((*p >= '0' && *p <= '9') || *p == '-') {
        const float x = std::strtof(p, &end);
        const float y = std::strtof(end, &end);
        const float z = std::strtof(end, &end);
        if (end != p) out.push_back({x, y, z});
    }
}

// Lines [begin, end) of a text point file; begin is a line start. Every line is cut into
// pieces of at most 1023 bytes, each NUL-terminated in a private buffer, so that the parse
// sees exactly what the reference's fgets loop sees (strtof never runs past a line).
void parse_point_range(const char* begin, const char* end, std::vector<Vec3>& out) {
    char buf[kLineBuf];
    const char* p = begin;
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* line_end = nl ? nl + 1 : end;
        while (p < line_end) {
            const size_t n = std::min((size_t)(line_end - p), kLineBuf - 1);
            std::memcpy(buf, p, n);
            buf[n] = '\0';
            parse_point_piece(buf, out);
            p += n;
        }
    }
}

bool read_binary_points(const MappedFile& f, const std::string& path, PointData& pd) {
    uint32_t hdr[8];
    if (f.size() < kPointHeader) {
        std::cerr << "Error: truncated point file: " << path << std::endl;
        return false;
    }
    std::memcpy(hdr, f.data(), kPointHeader);
    uint64_t count;
    std::memcpy(&count, hdr + 2, 8);
    if (hdr[1] != 1 || hdr[4] != 12 || f.size() != kPointHeader + count * 12) {
        std::cerr << "Error: malformed binary point file (version " << hdr[1] << ", stride " << hdr[4]
                  << ", " << count << " points, " << f.size() << " bytes): " << path << std::endl;
        return false;
    }
    pd.positions.resize(count);
    const char* src = f.data() + kPointHeader;
    char* dst = reinterpret_cast<char*>(pd.positions.data());
    const size_t bytes = count * 12;
    const int T = worker_count(bytes, 64u << 20);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const size_t a = bytes * t / T, b = bytes * (t + 1) / T;
        th.emplace_back([=] { std::memcpy(dst + a, src + a, b - a); });
    }
    for (auto& x : th) x.join();
    pd.numPoints = count;
    return true;
}

// ---- OBJ ------------------------------------------------------------------------------

bool is_digit(char c) { return c >= '0' && c <= '9'; }
bool is_blank(char c) { return c == ' ' || c == '\t'; }
bool is_line_end(char c) { return c == '\r' || c == '\n' || c == '\0'; }

// tinyobjloader's tryParseDouble: the integer digits accumulate as m = m * 10 + d, fraction
// digit k adds d * 10^-k (a table for k < 8), and an exponent e gives ldexp(m * 5^e, e).
// Its results can differ from strtod's correctly rounded ones in the last bits, and the
// reference's vertex coordinates are these values rounded to float.
bool obj_parse_double(const char* s, const char* e, double* result) {
    static const double kNegPow10[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
    if (s >= e) return false;
    double m = 0.0;
    int ex = 0;
    bool neg = false, exp_neg = false, leading_dot = false;
    const char* c = s;
    if (*c == '+' || *c == '-') {
        neg = *c == '-';
        ++c;
        if (c != e && *c == '.') leading_dot = true;
    } else if (*c == '.') {
        leading_dot = true;
    } else if (!is_digit(*c)) {
        return false;
    }
    if (!leading_dot) {
        int digits = 0;
        while (c != e && is_digit(*c)) {
            m *= 10;
            m += (int)(*c - '0');
            ++c;
            ++digits;
        }
        if (digits == 0) return false;
    }
    if (c != e) {
        bool try_exp = true;
        if (*c == '.') {
            ++c;
            for (int k = 1; c != e && is_digit(*c); ++k, ++c)
                m += (int)(*c - '0') * (k < 8 ? kNegPow10[k] : std::pow(10.0, -k));
            try_exp = c != e;
        } else if (*c != 'e' && *c != 'E') {
            try_exp = false;
        }
        if (try_exp && (*c == 'e' || *c == 'E')) {
            ++c;
            if (c != e && (*c == '+' || *c == '-')) {
                exp_neg = *c == '-';
                ++c;
            } else if (!is_digit(*c)) {
                return false;
            }
            int digits = 0;
            while (c != e && is_digit(*c)) {
                if (ex > 2147483647 / 10) return false;
                ex = ex * 10 + (int)(*c - '0');
                ++c;
                ++digits;
            }
            if (exp_neg) ex = -ex;
            if (digits == 0) return false;
        }
    }
    *result = (neg ? -1 : 1) * (ex ? std::ldexp(m * std::pow(5.0, ex), ex) : m);
    return true;
}

// tinyobjloader's parseReal: skip blanks, the token runs to the next ' ', '\t' or '\r';
// an unparsable token gives 0.
float obj_parse_real(const char** token) {
    *token += std::strspn(*token, " \t");
    const char* end = *token + std::strcspn(*token, " \t\r");
    double v = 0.0;
    obj_parse_double(*token, end, &v);
    *token = end;
    return (float)v;
}

// tinyobjloader's fixIndex: 1-based, negative = relative to the current count; a vertex
// index of 0 or one before the first element fails the whole load, texture/normal 0 passes.
bool obj_fix_index(int idx, int n, bool allow_zero, int* out) {
    if (idx > 0) {
        *out = idx - 1;
        return true;
    }
    if (idx == 0) {
        *out = -1;
        return allow_zero;
    }
    *out = n + idx;
    return *out >= 0;
}

// One "v[/vt[/vn]]" corner of an 'f' line (tinyobjloader's parseTriple); returns the vertex
// index, or false for an index the loader rejects.
bool obj_parse_corner(const char** token, int nv, int nvt, int nvn, int* v) {
    int dummy;
    if (!obj_fix_index(std::atoi(*token), nv, false, v)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    if ((*token)[0] != '/') return true;
    ++*token;
    if ((*token)[0] == '/') {
        ++*token;
        if (!obj_fix_index(std::atoi(*token), nvn, true, &dummy)) return false;
        *token += std::strcspn(*token, "/ \t\r");
        return true;
    }
    if (!obj_fix_index(std::atoi(*token), nvt, true, &dummy)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    if ((*token)[0] != '/') return true;
    ++*token;
    if (!obj_fix_index(std::atoi(*token), nvn, true, &dummy)) return false;
    *token += std::strcspn(*token, "/ \t\r");
    return true;
}

struct ObjShape {
    std::vector<uint32_t> tri;  // 3 vertex indices (file order, 0-based) per triangle
};

// Exports the pending polygons of a group into triangles (tinyobjloader's
// exportGroupsToShape with triangulation on). Quads with a vertex beyond the vertices read
// so far are dropped; faces with more than 4 corners are rejected.
bool obj_export(const std::vector<std::vector<int>>& faces, const std::vector<float>& v, ObjShape& shape,
                std::string& err) {
    for (const auto& f : faces) {
        const size_t n = f.size();
        if (n < 3) continue;
        if (n == 3) {
            shape.tri.insert(shape.tri.end(), {(uint32_t)f[0], (uint32_t)f[1], (uint32_t)f[2]});
            continue;
        }
        if (n > 4) {
            err = "faces with more than 4 vertices are not supported (triangulate the OBJ first)";
            return false;
        }
        bool in_range = true;
        for (int k = 0; k < 4; ++k) in_range = in_range && (size_t)f[k] * 3 + 2 < v.size();
        if (!in_range) continue;
        const float* p0 = &v[(size_t)f[0] * 3];
        const float* p1 = &v[(size_t)f[1] * 3];
        const float* p2 = &v[(size_t)f[2] * 3];
        const float* p3 = &v[(size_t)f[3] * 3];
        const float ax = p2[0] - p0[0], ay = p2[1] - p0[1], az = p2[2] - p0[2];
        const float bx = p3[0] - p1[0], by = p3[1] - p1[1], bz = p3[2] - p1[2];
        float d02 = ax * ax;
        d02 += ay * ay;
        d02 += az * az;
        float d13 = bx * bx;
        d13 += by * by;
        d13 += bz * bz;
        const uint32_t i0 = f[0], i1 = f[1], i2 = f[2], i3 = f[3];
        if (d02 < d13)
            shape.tri.insert(shape.tri.end(), {i0, i1, i2, i0, i2, i3});
        else
            shape.tri.insert(shape.tri.end(), {i0, i1, i3, i1, i2, i3});
    }
    return true;
}

// Parses an OBJ into its vertex array and face shapes. Lines end at "\n", "\r\n" or "\r".
bool obj_parse(const MappedFile& f, std::vector<float>& v, std::vector<ObjShape>& shapes, std::string& err) {
    std::vector<std::vector<int>> faces;
    ObjShape shape;
    int nvt = 0, nvn = 0;
    size_t line_no = 0;
    std::string line;
    auto flush = [&]() -> bool {
        if (!obj_export(faces, v, shape, err)) return false;
        faces.clear();
        if (!shape.tri.empty()) shapes.push_back(std::move(shape));
        shape = ObjShape();
        return true;
    };
    const char* p = f.data();
    const char* end = p + f.size();
    while (p < end) {
        const char* q = p;
        while (q < end && *q != '\n' && *q != '\r') ++q;
        line.assign(p, q);
        p = (q < end && *q == '\r' && q + 1 < end && q[1] == '\n') ? q + 2 : std::min(q + 1, end);
        ++line_no;
        const char* t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') continue;
        if (t[0] == 'v' && is_blank(t[1])) {
            t += 2;
            const float x = obj_parse_real(&t);
            const float y = obj_parse_real(&t);
            const float z = obj_parse_real(&t);
            v.insert(v.end(), {x, y, z});
        } else if (t[0] == 'v' && t[1] == 't' && is_blank(t[2])) {
            ++nvt;
        } else if (t[0] == 'v' && t[1] == 'n' && is_blank(t[2])) {
            ++nvn;
        } else if (t[0] == 'f' && is_blank(t[1])) {
            t += 2;
            t += std::strspn(t, " \t");
            std::vector<int> face;
            while (!is_line_end(t[0])) {
                int vi;
                if (!obj_parse_corner(&t, (int)(v.size() / 3), nvt, nvn, &vi)) {
                    err = "failed to parse 'f' line " + std::to_string(line_no) +
                          " (a zero or out-of-range relative vertex index)";
                    return false;
                }
                face.push_back(vi);
                t += std::strspn(t, " \t\r");
            }
            faces.push_back(std::move(face));
        } else if ((t[0] == 'g' || t[0] == 'o') && is_blank(t[1])) {
            if (!flush()) return false;
        }
    }
    return flush();
}

}  // namespace

PointData loadPointDataset(const std::string& path) {
    PointData pd;
    if (path.empty()) {
        std::cout << "No point dataset provided, using default test points" << std::endl;
        pd.positions = {{0.0f, 0.0f, -1.0f}, {0.5f, 0.5f, -1.0f}, {1.0f, 1.0f, -1.0f}};
        pd.numPoints = 3;
        return pd;
    }
    MappedFile f(path);
    if (!f.ok()) {
        std::cerr << "Error: Could not open point dataset file: " << path << std::endl;
        return pd;
    }
    uint32_t magic = 0;
    if (f.size() >= 4) std::memcpy(&magic, f.data(), 4);
    if (magic == kPointMagic) {
        if (!read_binary_points(f, path, pd)) pd = PointData();
        return pd;
    }
    // Text: line-aligned ranges parsed in parallel, concatenated in file order.
    const char* base = f.data();
    const size_t n = f.size();
    const int T = worker_count(n, 4u << 20);
    std::vector<size_t> cut(T + 1, n);
    cut[0] = 0;
    for (int t = 1; t < T; ++t) {
        const size_t b = n * t / T;
        const char* nl = static_cast<const char*>(std::memchr(base + b - 1, '\n', n - (b - 1)));
        cut[t] = nl ? (size_t)(nl - base) + 1 : n;
    }
    for (int t = 1; t <= T; ++t) cut[t] = std::max(cut[t], cut[t - 1]);
    std::vector<std::vector<Vec3>> parts(T);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] { parse_point_range(base + cut[t], base + cut[t + 1], parts[t]); });
    for (auto& x : th) x.join();
    size_t total = 0;
    for (const auto& q : parts) total += q.size();
    pd.positions.reserve(total);
    for (auto& q : parts) {
        pd.positions.insert(pd.positions.end(), q.begin(), q.end());
        std::vector<Vec3>().swap(q);
    }
    pd.numPoints = pd.positions.size();
    return pd;
}

bool writeBinaryPoints(const std::string& path, const PointData& points) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::cerr << "Error: Could not open file for writing: " << path << std::endl;
        return false;
    }
    uint32_t hdr[8] = {kPointMagic, 1, 0, 0, 12, 0, 0, 0};
    const uint64_t count = points.positions.size();
    std::memcpy(hdr + 2, &count, 8);
    bool ok = std::fwrite(hdr, 1, sizeof hdr, f) == sizeof hdr;
    if (ok && count) ok = std::fwrite(points.positions.data(), 12, count, f) == count;
    ok = (std::fclose(f) == 0) && ok;
    return ok;
}

GeometryData ObjMeshDatasetLoader::load(const std::string& path) {
    GeometryData g;
    struct stat st;
    if (::stat(path.c_str(), &st) != 0) {
        std::cerr << "Error: File does not exist: " << path << std::endl;
        return g;
    }
    if (!S_ISREG(st.st_mode)) {
        std::cerr << "Error: Path is not a regular file: " << path << std::endl;
        return g;
    }
    const size_t dot = path.find_last_of('.');
    const size_t slash = path.find_last_of('/');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash) || path.substr(dot) != ".obj") {
        std::cerr << "Error: File is not a .obj file: " << path << std::endl;
        return g;
    }
    MappedFile f(path);
    std::vector<float> v;
    std::vector<ObjShape> shapes;
    std::string err;
    if (!f.ok() || !obj_parse(f, v, shapes, err)) {
        std::cerr << "Error loading .obj file: " << (f.ok() ? err : "cannot read " + path) << std::endl;
        return g;
    }
    if (shapes.empty()) {
        std::cerr << "Error: No shapes found in .obj file: " << path << std::endl;
        return g;
    }
    const size_t nv = v.size() / 3;
    Vec3 lo{1e30f, 1e30f, 1e30f}, hi{-1e30f, -1e30f, -1e30f};
    int object = 0;
    std::vector<uint32_t> used, local;
    std::vector<uint32_t> remap(nv, UINT32_MAX);
    for (const ObjShape& s : shapes) {
        // Vertices of the object: the distinct file indices it uses, in ascending order
        // (the reference iterates an unordered_set here, whose order is unspecified).
        used.assign(s.tri.begin(), s.tri.end());
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        if (used.empty()) continue;
        const uint32_t base = (uint32_t)g.vertices.size();
        uint32_t k = 0;
        for (uint32_t gv : used) {
            if (gv >= nv) continue;
            const Vec3 p{v[(size_t)gv * 3], v[(size_t)gv * 3 + 1], v[(size_t)gv * 3 + 2]};
            lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
            hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
            g.vertices.push_back(p);
            remap[gv] = base + k++;
        }
        for (size_t t = 0; t + 2 < s.tri.size(); t += 3) {
            const uint32_t a = s.tri[t], b = s.tri[t + 1], c = s.tri[t + 2];
            if (a >= nv || b >= nv || c >= nv) continue;
            g.indices.push_back({remap[a], remap[b], remap[c]});
            g.triangleToObject.push_back(object);
        }
        for (uint32_t gv : used)
            if (gv < nv) remap[gv] = UINT32_MAX;
        ++object;
    }
    g.totalTriangles = g.indices.size();
    if (!g.vertices.empty()) {
        g.grid.minBound = lo;
        g.grid.maxBound = hi;
    }
    std::cout << "Loaded " << object << " object(s), " << g.vertices.size() << " vertices, "
              << g.totalTriangles << " triangles from " << path << std::endl;
    return g;
}

GeometryData PreDatasetLoader::load(const std::string& path) { return readBinaryFile(path); }

std::unique_ptr<IDatasetLoader> DatasetLoaderFactory::createFromPath(const std::string& path) {
    const size_t dot = path.find_last_of('.');
    const size_t slash = path.find_last_of('/');
    std::string ext = (dot == std::string::npos || (slash != std::string::npos && dot < slash)) ? "" : path.substr(dot);
    for (auto& c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext == ".obj") return std::make_unique<ObjMeshDatasetLoader>();
    if (ext == ".pre") return std::make_unique<PreDatasetLoader>();
    return nullptr;
}

EdgeData extractEdgesOnDevice(DeviceContext& ctx, const GeometryData& g) {
    EdgeData e;
    const size_t nV = g.vertices.size(), nT = g.indices.size();
    if (g.triangleToObject.size() < nT) {
        std::cerr << "Error: triangleToObject has " << g.triangleToObject.size() << " entries for " << nT
                  << " triangles" << std::endl;
        std::exit(1);
    }
    DeviceBuffer dV(std::max<size_t>(nV, 1) * sizeof(Vec3)), dI(std::max<size_t>(nT, 1) * sizeof(UInt3)),
        dO(std::max<size_t>(nT, 1) * sizeof(int));
    dV.upload(g.vertices.data(), nV * sizeof(Vec3));
    dI.upload(g.indices.data(), nT * sizeof(UInt3));
    dO.upload(g.triangleToObject.data(), nT * sizeof(int));
    // At most 3 edges and 3 edge sources per triangle: one pass with outputs of that size.
    const uint64_t cap = 3 * (uint64_t)nT;
    uint64_t ne = 0, ns = 0;
    DeviceBuffer st(std::max<uint64_t>(cap, 1) * 12), en(std::max<uint64_t>(cap, 1) * 12),
        off(std::max<uint64_t>(cap, 1) * 4), cnt(std::max<uint64_t>(cap, 1) * 4), src(std::max<uint64_t>(cap, 1) * 4);
    check(rs3d_extract_edges(ctx.get(), dV.as<float>(), nV, dI.as<uint32_t>(), nT, dO.as<int32_t>(), st.as<float>(),
                             en.as<float>(), off.as<int32_t>(), cnt.as<int32_t>(), src.as<int32_t>(), cap, cap, &ne,
                             &ns),
          "rs3d_extract_edges");
    e.edgeStarts.resize(ne);
    e.edgeEnds.resize(ne);
    e.sourceObjectOffsets.resize(ne);
    e.sourceObjectCounts.resize(ne);
    e.sourceObjects.resize(ns);
    st.download(e.edgeStarts.data(), ne * 12);
    en.download(e.edgeEnds.data(), ne * 12);
    off.download(e.sourceObjectOffsets.data(), ne * 4);
    cnt.download(e.sourceObjectCounts.data(), ne * 4);
    src.download(e.sourceObjects.data(), ns * 4);
    return e;
}

void generateGridStatsOnDevice(DeviceContext& ctx, const GeometryData& g, GridData& grid, int resolution,
                               float worldSize) {
    const size_t nV = g.vertices.size(), nT = g.indices.size();
    const size_t cells = (size_t)resolution * resolution * resolution;
    DeviceBuffer dV(std::max<size_t>(nV, 1) * sizeof(Vec3)), dI(std::max<size_t>(nT, 1) * sizeof(UInt3)),
        dO(std::max<size_t>(nT, 1) * sizeof(int)), dC(std::max<size_t>(cells, 1) * sizeof(GridCell));
    dV.upload(g.vertices.data(), nV * sizeof(Vec3));
    dI.upload(g.indices.data(), nT * sizeof(UInt3));
    dO.upload(g.triangleToObject.data(), std::min(nT, g.triangleToObject.size()) * sizeof(int));
    float b[6];
    check(rs3d_grid_stats(ctx.get(), dV.as<float>(), nV, dI.as<uint32_t>(), nT, dO.as<int32_t>(), countObjects(g),
                          resolution, worldSize, dC.as<void>(), b),
          "rs3d_grid_stats");
    grid.cells.resize(cells);
    dC.download(grid.cells.data(), cells * sizeof(GridCell));
    grid.minBound = {b[0], b[1], b[2]};
    grid.maxBound = {b[3], b[4], b[5]};
    for (int k = 0; k < 3; ++k) grid.resolution[k] = (uint32_t)resolution;
    grid.hasGrid = true;
}

}  // namespace rs3d_host
