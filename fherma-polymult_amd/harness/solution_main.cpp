// Point-directory driver for the FHERMA boundary: binds fherma_init /
// fherma_run / fherma_free through include/fherma.h exactly as a platform
// harness does (reference main.cpp:137-209), so the library can be run on the
// reference's test points unchanged.
//
//     solution <point dir>
//
//   manifest.json     {"N": .., "W": .., "L": .., "cases": ..}   read
//   point/q.bin       L little-endian u32                         read
//   cases/NNNNNN/     a.bin, b.bin: N x L u32                      read
//   out/NNNNNN/c.bin  N x L u32                                    written
//   out/results.json  {"init_s": .., "cases": [{i, seconds, status[, note]}]}
//
// Only fherma_run is timed. results.json is rewritten after every case, so a run
// that is killed keeps the cases it finished. A failure in fherma_init ends the
// process with a non-zero status and no answers (a refused point).
#include "fherma.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

std::string read_file(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + p.string());
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

// The manifest holds four non-negative integers; a key lookup is all it needs.
uint64_t manifest_value(const std::string& text, const char* key) {
    const std::string k = std::string("\"") + key + "\"";
    const size_t at = text.find(k);
    if (at == std::string::npos) throw std::runtime_error(std::string("manifest.json has no ") + key);
    const size_t colon = text.find_first_not_of(" \t\r\n", at + k.size());
    if (colon == std::string::npos || text[colon] != ':')
        throw std::runtime_error(std::string("manifest.json: no value for ") + key);
    char* end = nullptr;
    const double v = std::strtod(text.c_str() + colon + 1, &end);
    if (end == text.c_str() + colon + 1 || !(v >= 0) || v > 4294967295.0 || v != static_cast<double>(static_cast<uint64_t>(v)))
        throw std::runtime_error(std::string("manifest.json: ") + key + " is not an integer in [0, 2^32)");
    return static_cast<uint64_t>(v);
}

fherma::Tensor<uint32_t> read_tensor(const fs::path& p, std::vector<int64_t> shape) {
    fherma::Tensor<uint32_t> t;
    t.shape = std::move(shape);
    const std::string raw = read_file(p);
    const size_t n = static_cast<size_t>(t.count());
    if (raw.size() != n * sizeof(uint32_t)) {
        throw std::runtime_error(p.filename().string() + ": " + std::to_string(raw.size()) + " bytes for " +
                                 std::to_string(n) + " values");
    }
    t.data.resize(n);
    std::memcpy(t.data.data(), raw.data(), raw.size());
    return t;
}

std::string case_dir(size_t i) {
    char b[16];
    std::snprintf(b, sizeof b, "%06zu", i);
    return b;
}

void write_results(const fs::path& out, double init_s, const std::vector<std::string>& cases) {
    std::ofstream f(out / "results.json");
    f.setf(std::ios::fixed);
    f.precision(9);
    f << "{\"init_s\":" << init_s << ",\"cases\":[";
    for (size_t i = 0; i < cases.size(); ++i) f << (i ? "," : "") << cases[i];
    f << "]}";
}

std::string escape(const std::string& s) {
    std::string o;
    for (char c : s.substr(0, 200)) {
        if (c == '"' || c == '\\') o += '\\';
        o += static_cast<unsigned char>(c) < 0x20 ? ' ' : c;   // no control characters in a JSON string
    }
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: solution <point directory>\n";
        return 2;
    }
    const fs::path root = argv[1];
    const fs::path out = root / "out";
    fherma::Point p;
    size_t total = 0;
    try {   // an unreadable point directory: a message and status 2, no answers
        const std::string manifest = read_file(root / "manifest.json");
        p.N = static_cast<uint32_t>(manifest_value(manifest, "N"));
        p.W = static_cast<uint32_t>(manifest_value(manifest, "W"));
        p.L = static_cast<uint32_t>(manifest_value(manifest, "L"));
        p.q = read_tensor(root / "point" / "q.bin", {static_cast<int64_t>(p.L)});
        total = static_cast<size_t>(manifest_value(manifest, "cases"));
        fs::create_directories(out);
    } catch (const std::exception& e) {
        std::cerr << "solution: " << e.what() << "\n";
        return 2;
    }

    const auto t0 = std::chrono::steady_clock::now();
    void* state = nullptr;
    try {
        state = fherma_init(p);
    } catch (const std::exception& e) {   // a refused point: no answers, non-zero exit
        std::cerr << "fherma_init: " << e.what() << "\n";
        return 1;
    }
    const double init_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::vector<std::string> cases;
    write_results(out, init_s, cases);
    for (size_t i = 0; i < total; ++i) {
        const fs::path in_dir = root / "cases" / case_dir(i);
        std::ostringstream rec;
        rec.setf(std::ios::fixed);
        rec.precision(9);
        try {
            fherma::Inputs in;
            in.a = read_tensor(in_dir / "a.bin", {static_cast<int64_t>(p.N), static_cast<int64_t>(p.L)});
            in.b = read_tensor(in_dir / "b.bin", {static_cast<int64_t>(p.N), static_cast<int64_t>(p.L)});
            const auto s = std::chrono::steady_clock::now();
            fherma::Outputs c = fherma_run(state, in);
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - s).count();
            fs::create_directories(out / case_dir(i));
            std::ofstream f(out / case_dir(i) / "c.bin", std::ios::binary);
            f.write(reinterpret_cast<const char*>(c.c.data.data()),
                    static_cast<std::streamsize>(c.c.data.size() * sizeof(uint32_t)));
            rec << "{\"i\":" << i << ",\"seconds\":" << sec << ",\"status\":\"ok\"}";
        } catch (const std::exception& e) {
            rec << "{\"i\":" << i << ",\"seconds\":null,\"status\":\"crashed\",\"note\":\"" << escape(e.what())
                << "\"}";
        }
        cases.push_back(rec.str());
        write_results(out, init_s, cases);
    }
    fherma_free(state);
    write_results(out, init_s, cases);
    return 0;
}
