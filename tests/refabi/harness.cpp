// Test harness for the C++-linkage drop-ins of libsysmat_hip.so (declared in
// include/sysmat_ref.h).  It calls them the way the reference's own drivers do:
//   * PESysMatGen once per (rotation, Z-slab of 4), with the in-band fields
//     Img[30] = full Z, Img[20] = rotation, Img[2] = slab Z, Img[31] = z0, and
//     the slab rows scattered into a det-major file (PEGen_CircularHole.cpp:139-216);
//   * each of the three scatter() overloads once for rotation 0
//     (ScatterGen_CircularHole.cpp:188, ScatterGen_Collimator.cpp:172,
//     ScatterGen_Crystal.cpp:172), first computing the GeometryRelationShip
//     bitmaps (Phys[8] = Phys[9] = 1), then reading them back (= 0).
// Parameter files are read into zero-filled 80000/80000/100/100-float blocks
// like the reference drivers.  Usage: harness DIR (Params_*.dat in DIR; the
// outputs are written there).  Exit status 0 only when every call returned
// its expected value.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sysmat_ref.h"

static std::vector<float> read_block(const std::string& path, size_t n) {
    std::vector<float> v(n, 0.0f);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::fprintf(stderr, "harness: cannot open %s\n", path.c_str());
        std::exit(1);
    }
    size_t got = std::fread(v.data(), sizeof(float), n, f);
    (void)got;
    std::fclose(f);
    return v;
}

static void write_all(const std::string& path, const float* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(float), n, f) != n) {
        std::fprintf(stderr, "harness: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: harness DIR\n");
        return 2;
    }
    const std::string dir = argv[1];
    std::vector<float> col = read_block(dir + "/Params_Collimator.dat", 80000);
    std::vector<float> det = read_block(dir + "/Params_Detector.dat", 80000);
    std::vector<float> img = read_block(dir + "/Params_Image.dat", 100);
    std::vector<float> phy = read_block(dir + "/Params_Physics.dat", 100);
    const int N = (int)(det[0] + 0.001f);
    const int X = (int)(img[0] + 0.001f), Y = (int)(img[1] + 0.001f), Z = (int)(img[2] + 0.001f);
    const int R = (int)(img[6] + 0.001f);
    const long long V = (long long)X * Y * Z;
    const int slabZ = 4;
    int failures = 0;

    // photoelectric matrix, slab by slab
    std::vector<float> pe((size_t)R * N * V, 0.0f);
    std::vector<float> slab((size_t)N * X * Y * slabZ, 0.0f);
    const float saved_z = img[2];
    img[30] = (float)Z;
    for (int rot = 0; rot < R; ++rot) {
        img[20] = (float)rot;
        for (int z0 = 0; z0 < Z; z0 += slabZ) {
            const int sz = z0 + slabZ > Z ? Z - z0 : slabZ;
            const int bins = X * Y * sz;
            img[2] = (float)sz;
            img[31] = (float)z0;
            const int q = PESysMatGen(col.data(), det.data(), img.data(), slab.data(), 0);
            if (q != bins) {
                std::fprintf(stderr, "harness: PESysMatGen returned %d, expected %d\n", q, bins);
                ++failures;
            }
            for (int d = 0; d < N; ++d)
                std::memcpy(&pe[((size_t)rot * N + d) * V + (size_t)z0 * X * Y], &slab[(size_t)d * bins],
                            sizeof(float) * (size_t)bins);
            std::printf("PESysMatGen rotation %d slab z0=%d slabZ=%d -> %d\n", rot, z0, sz, q);
        }
        img[2] = saved_z;
    }
    write_all(dir + "/harness_PE.sysmat", pe.data(), pe.size());

    // scatter overloads, rotation 0; the combined reference variant touches N*V*R floats
    img[20] = 0.0f;
    img[31] = 0.0f;
    std::vector<float> dst((size_t)N * V * (R > 0 ? R : 1), 0.0f);
    auto check = [&](const char* name, int rc) {
        std::printf("%s -> %d\n", name, rc);
        if (rc != (int)V) {
            std::fprintf(stderr, "harness: %s returned %d, expected %lld\n", name, rc, V);
            ++failures;
        }
    };
    const std::string geo_c = dir + "/harness_GeoCrystal", geo_h = dir + "/harness_GeoCollimator";
    for (int pass = 0; pass < 2; ++pass) {
        phy[8] = phy[9] = pass == 0 ? 1.0f : 0.0f;  // 1: compute and write the bitmaps, 0: read them
        const std::string tag = pass == 0 ? "" : "_reread";
        std::fill(dst.begin(), dst.end(), 0.0f);
        check("scatter(combined)", scatter(col.data(), det.data(), img.data(), phy.data(), pe.data(), geo_c.c_str(),
                                           geo_h.c_str(), dst.data(), 0));
        write_all(dir + "/harness_scatter_combined" + tag + ".bin", dst.data(), (size_t)N * V);
        std::fill(dst.begin(), dst.end(), 0.0f);
        check("scatter(collimator)", scatter(col.data(), det.data(), img.data(), phy.data(), pe.data(),
                                             geo_h.c_str(), dst.data(), 0));
        write_all(dir + "/harness_scatter_collimator" + tag + ".bin", dst.data(), (size_t)N * V);
        std::fill(dst.begin(), dst.end(), 0.0f);
        check("scatter(crystal)", scatter(det.data(), img.data(), phy.data(), pe.data(), geo_c.c_str(), dst.data(), 0));
        write_all(dir + "/harness_scatter_crystal" + tag + ".bin", dst.data(), (size_t)N * V);
    }
    std::printf("harness done, %d failure(s)\n", failures);
    return failures ? 1 : 0;
}
