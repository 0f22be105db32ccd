// cli_common.hpp -- command-line option classes shared by the query executables. Flag
// names, defaults and help text follow query/src/applications/app_cli_options.h:10-163.
#pragma once

#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "../host/rs3d_host.hpp"

namespace rs3d_cli {

using HelpEntry = std::pair<std::string, std::string>;

inline void printHelp(const char* exe, const std::string& usage, const std::string& desc,
                      const std::vector<HelpEntry>& opts) {
    std::cout << "Usage: " << exe << " " << usage << "\n" << desc << "\nOptions:\n";
    for (const auto& o : opts) std::cout << "  " << std::left << std::setw(34) << o.first << o.second << "\n";
}

// --mesh1/--mesh2/--output/--ptx (+ --runs/--warmup-runs/--no-export when benchmarking).
struct MeshPairOptions {
    std::string mesh1Path, mesh2Path, outputJsonPath, ptxPath;
    int numberOfRuns = 1;
    int warmupRuns = 2;
    bool exportResults = true;
    bool helpRequested = false;
    bool allowNoExport = false;

    explicit MeshPairOptions(std::string defaultJson) : outputJsonPath(std::move(defaultJson)) {}
    virtual ~MeshPairOptions() = default;

    bool hasRequiredMeshInputs() const { return !mesh1Path.empty() && !mesh2Path.empty(); }
    void sanitizeRunCounts() {
        if (numberOfRuns < 1) numberOfRuns = 1;
        if (warmupRuns < 0) warmupRuns = 0;
    }
    void parse(int argc, char** argv) {
        for (int i = 1; i < argc; ++i) {
            const std::string a = argv[i];
            if (a == "--help" || a == "-h") helpRequested = true;
            else if (a == "--mesh1" && i + 1 < argc) mesh1Path = argv[++i];
            else if (a == "--mesh2" && i + 1 < argc) mesh2Path = argv[++i];
            else if (a == "--output" && i + 1 < argc) outputJsonPath = argv[++i];
            else if (a == "--ptx" && i + 1 < argc) ptxPath = argv[++i];  // accepted, unused (no PTX)
            else if (a == "--runs" && i + 1 < argc) numberOfRuns = std::atoi(argv[++i]);
            else if (a == "--warmup-runs" && i + 1 < argc) warmupRuns = std::atoi(argv[++i]);
            else if (allowNoExport && a == "--no-export") exportResults = false;
            else parseApplicationOption(a, i, argc, argv);
        }
    }
    virtual bool parseApplicationOption(const std::string&, int&, int, char**) { return false; }
    static void meshPairHelp(std::vector<HelpEntry>& o) {
        o.emplace_back("--mesh1 <path>", "Path to first mesh dataset (geometry file)");
        o.emplace_back("--mesh2 <path>", "Path to second mesh dataset (geometry file)");
        o.emplace_back("--output <path>", "Path to JSON file for performance timing output");
        o.emplace_back("--ptx <path>", "Accepted for compatibility; ignored (HIP kernels are linked in)");
    }
    static void runHelp(std::vector<HelpEntry>& o) {
        o.emplace_back("--runs <number>", "Number of measured query runs (default: 1)");
        o.emplace_back("--warmup-runs <number>", "Number of warmup iterations (default: 2)");
    }
};

inline unsigned long long nextPow2(unsigned long long v) {
    if (v == 0) return 1;
    --v;
    for (int s = 1; s < 64; s <<= 1) v |= v >> s;
    return v + 1;
}

// Keys (a << 32 | b) -> pairs.
inline std::vector<rs3d_pair> keysToPairs(const std::vector<uint64_t>& keys) {
    std::vector<rs3d_pair> out(keys.size());
    for (size_t i = 0; i < keys.size(); ++i)
        out[i] = {(int32_t)(keys[i] >> 32), (int32_t)(keys[i] & 0xffffffffull)};
    return out;
}

}  // namespace rs3d_cli
