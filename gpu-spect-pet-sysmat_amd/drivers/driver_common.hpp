// Shared pieces of the command-line drivers: parameter files, output names,
// file writes.  The drivers keep the reference's command lines, file names and
// the raw float32 det-major .sysmat layout; errors exit non-zero with a message.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/sysmat.h"

namespace drv {

[[noreturn]] inline void fail(const std::string& msg, int code = 1) {
    std::fprintf(stderr, "ERROR: %s\n", msg.c_str());
    std::exit(code);
}

inline void check(int rc, const char* what) {
    if (rc < 0) fail(std::string(what) + ": " + sm_last_error(), 4);
}

// Reads a whole Params_*.dat file, zero-padded to at least `min_len` floats
// (the reference reads a fixed 80000 / 100 floats, PEGen_CircularHole.cpp:39-62;
// longer files are read whole so detectors beyond 6666 crystals work).
inline std::vector<float> read_params(const char* name, size_t min_len) {
    FILE* f = std::fopen(name, "rb");
    if (!f) fail(std::string("cannot open ") + name);
    std::vector<float> v;
    float buf[4096];
    size_t got;
    while ((got = std::fread(buf, sizeof(float), 4096, f)) > 0) v.insert(v.end(), buf, buf + got);
    std::fclose(f);
    if (v.size() < min_len) v.resize(min_len, 0.0f);
    return v;
}

struct Params {
    std::vector<float> col, det, img, phy;
};

inline Params read_all() {
    Params p;
    p.col = read_params("Params_Collimator.dat", 80000);
    p.det = read_params("Params_Detector.dat", 80000);
    p.img = read_params("Params_Image.dat", 100);
    p.phy = read_params("Params_Physics.dat", 100);
    const int n = (int)std::floor(p.det[0] + 0.0001f);
    if (n <= 0 || p.det.size() < (size_t)(1 + 12 * n)) fail("Params_Detector.dat holds fewer crystals than its header");
    const int layers = (int)std::floor(p.col[0] + 0.000001f);
    if (layers < 0 || layers > 9) fail("Params_Collimator.dat: layer count outside 0..9");
    long long holes = 0;
    for (int m = 0; m < layers; ++m) holes += std::max(0LL, (long long)std::floor(p.col[(m + 1) * 10]));
    if (p.col.size() < (size_t)(100 + 9 * holes))
        fail("Params_Collimator.dat holds fewer holes than its layer headers");
    return p;
}

inline std::string shift_name(const char* prefix, const std::vector<float>& img) {
    char buf[2048];
    std::snprintf(buf, sizeof(buf), "%s_shift_%f_%f_%f.sysmat", prefix, img[8], img[9], img[10]);
    return buf;
}

inline void print_layers(const Params& p) {
    const int L = (int)std::floor(p.col[0] + 0.001f);
    for (int m = 0; m < L; ++m) {
        const float* c = &p.col[(m + 1) * 10];
        std::printf("############ Collimator %d ############\n", m);
        std::printf("Number of collimator holes = %g\n", c[0]);
        std::printf("Width of collimator layer(X direction) = %gmm\n", c[1]);
        std::printf("Thickness of collimator layer(Y direction) = %gmm\n", c[2]);
        std::printf("Height of collimator layer(Z direction) = %gmm\n", c[3]);
        std::printf("Collimator Layer to 1st Collimator Layer = %gmm\n", c[4]);
        std::printf("Total Coeff of collimator layer = %g\n", c[5]);
    }
    std::printf("FOV center to 1st Collimator = %g\n", p.img[11]);
}

inline void write_all(const std::string& name, const float* data, size_t n) {
    FILE* f = std::fopen(name.c_str(), "wb+");
    if (!f) fail("cannot create " + name);
    if (std::fwrite(data, sizeof(float), n, f) != n) fail("short write to " + name);
    std::fclose(f);
}

inline std::vector<float> read_floats(const std::string& name, size_t n) {
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) fail("cannot open " + name);
    std::vector<float> v(n, 0.0f);
    size_t got = std::fread(v.data(), sizeof(float), n, f);
    std::fclose(f);
    if (got != n) fail("short read of " + name);
    return v;
}

}  // namespace drv
