// PEGen_CircularHole -- photoelectric system matrix, command-line compatible
// with PEGen_RayTracing_CircularHole/PEGen_CircularHole.cpp:37-237:
//   ./PEGen_CircularHole [-cuda GPU_ID] [-h|--help]
// reads Params_{Collimator,Detector,Image,Physics}.dat from the working
// directory and writes PE_SysMat_shift_<x>_<y>_<z>.sysmat (raw float32,
// [rotation][detector][z][y][x]).
// Extensions (not in the reference): -sub NV CX CY CZ sets the sampling
// (voxel sub-points per axis, crystal sub-cells; default 1 8 16 8 = reference),
// -block N the detector rows per GPU launch, -tune picks the voxel tile with the
// autotuner (cached in $HOME/.cache/sysmat_tiles.txt), -csr REL writes
// PE_SysMat_shift_<x>_<y>_<z>.csr instead (rows = rotation-major detectors,
// elements >= REL x row max kept; the .csr layout is given in sysmat.h).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "driver_common.hpp"

int main(int argc, char* argv[]) {
    drv::Params P = drv::read_all();
    drv::print_layers(P);
    SmPeOptions o;
    sm_pe_default_options(&o);
    int block = 0;
    double csr_rel = -1.0;
    bool tune = false;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-cuda") == 0 && i + 1 < argc) {
            o.device = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-sub") == 0 && i + 4 < argc) {
            o.voxel_sub = std::atoi(argv[++i]);
            o.sub_x = std::atoi(argv[++i]);
            o.sub_y = std::atoi(argv[++i]);
            o.sub_z = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-block") == 0 && i + 1 < argc) {
            block = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-tune") == 0) {
            tune = true;
        } else if (std::strcmp(argv[i], "-csr") == 0 && i + 1 < argc) {
            csr_rel = std::atof(argv[++i]);
            if (!(csr_rel >= 0.0 && csr_rel <= 1.0)) drv::fail("-csr threshold must be in [0, 1]", 2);
        } else if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
            std::printf("Usage: %s [-cuda GPU_ID] [-sub NV CX CY CZ] [-block ROWS] [-tune] [-csr REL]\n", argv[0]);
            return 0;
        } else {
            std::fprintf(stderr, "Unknown parameter or missing argument: %s\n", argv[i]);
            std::printf("Usage: %s [-cuda GPU_ID] \n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    const int N = (int)std::floor(P.det[0] + 0.001f);
    const int X = (int)std::floor(P.img[0] + 0.001f), Y = (int)std::floor(P.img[1] + 0.001f);
    const int Z = (int)std::floor(P.img[2] + 0.001f);
    const int R = std::max(1, (int)std::floor(P.img[6] + 0.001f));
    const long long V = (long long)X * Y * Z;
    const long long total = (long long)N * V * R;
    std::printf("FOV dimension : %d %d %d\n", X, Y, Z);
    std::printf("FOV Voxel Size(mm) : %f %f %f\n", P.img[3], P.img[4], P.img[5]);
    std::printf("numDet=%d  fullVox=%lld  rotations=%d  sysmatBytes=%lld (%.2f GB)\n", N, V, R, total * 4,
                (double)total * 4 / (1024.0 * 1024.0 * 1024.0));
    if (V <= 0 || N <= 0) drv::fail("empty system matrix", 2);
    if (tune) {
        const char* home = std::getenv("HOME");
        const std::string cache = std::string(home ? home : ".") + "/.cache/sysmat_tiles.txt";
        SmPeOptions best;
        double ns = 0.0;
        const int rc = sm_pe_autotune(P.col.data(), P.det.data(), P.img.data(), &o, 64, cache.c_str(), &best, &ns);
        if (rc < 0) {
            std::fprintf(stderr, "autotune failed (%s); using the default tile\n", sm_last_error());
        } else {
            o = best;
            std::printf("voxel tile %d x %d x %d (%.2f ns/element on the sample%s)\n", o.tile_x, o.tile_y, o.tile_z,
                        ns, rc == 1 ? ", cached" : "");
        }
    }
    if (csr_rel >= 0.0) {
        const std::string dense_name = drv::shift_name("PE_SysMat", P.img);
        const std::string path = dense_name.substr(0, dense_name.size() - 7) + ".csr";  // .sysmat -> .csr
        std::vector<SmCsr*> parts;
        auto t0 = std::chrono::steady_clock::now();
        long long nnz_total = 0;
        for (int rot = 0; rot < R; ++rot) {
            std::printf("########################\nRotation (%d) processing ...\n########################\n", rot);
            o.rotation = rot;
            o.det_begin = 0;
            o.det_count = N;
            SmCsr* c = nullptr;
            drv::check(sm_pe_compute_csr(P.col.data(), P.det.data(), P.img.data(), &o, csr_rel, &c), "sm_pe_compute_csr");
            long long nnz = 0;
            sm_csr_info(c, nullptr, nullptr, &nnz);
            nnz_total += nnz;
            parts.push_back(c);
        }
        drv::check(sm_csr_write(path.c_str(), parts.data(), (int)parts.size()), "sm_csr_write");
        for (SmCsr* c : parts) sm_csr_destroy(c);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("########################\nPhoton Electric 3D Sysmat Written (%s, %lld of %lld elements kept, "
                    "%.3f s, %.3g elements/s).\n########################\n", path.c_str(), nnz_total, total, s,
                    (double)total / s);
        return 0;
    }
    const std::string name = drv::shift_name("PE_SysMat", P.img);
    FILE* f = std::fopen(name.c_str(), "wb+");
    if (!f) drv::fail("cannot create " + name);
    if (block <= 0) block = (int)std::max(1LL, std::min<long long>(N, (1LL << 30) / V));  // ~4 GiB per launch
    std::vector<float> host((size_t)block * V);
    auto t0 = std::chrono::steady_clock::now();
    for (int rot = 0; rot < R; ++rot) {
        std::printf("########################\nRotation (%d) processing ...\n########################\n", rot);
        o.rotation = rot;
        for (int d0 = 0; d0 < N; d0 += block) {
            const int cnt = std::min(block, N - d0);
            o.det_begin = d0;
            o.det_count = cnt;
            drv::check(sm_pe_compute(P.col.data(), P.det.data(), P.img.data(), &o, host.data()), "sm_pe_compute");
            const long long off = ((long long)rot * N + d0) * V;
            if (std::fseek(f, (long)(off * 4), SEEK_SET) != 0) drv::fail("seek failed in " + name);
            if (std::fwrite(host.data(), sizeof(float), (size_t)cnt * V, f) != (size_t)cnt * V)
                drv::fail("short write to " + name);
            std::printf("WROTE detectors %d..%d\n", d0, d0 + cnt - 1);
            std::fflush(f);
        }
    }
    std::fclose(f);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("########################\nPhoton Electric 3D Sysmat Written (%s, %.3f s, %.3g elements/s).\n"
                "########################\n", name.c_str(), s, (double)total / s);
    return 0;
}
