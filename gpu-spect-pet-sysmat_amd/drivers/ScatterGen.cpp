// Primary-Compton scatter drivers, command-line compatible with the reference's
// three programs; one source, built three times with SCATTER_MODE:
//   0  ScatterGen_CircularHole  (ScatterGen_RayTracing_CircularHole/ScatterGen_CircularHole.cpp:38-226)
//        [-PE file] [-GeoCrystal file] [-GeoCollimator file] [-cuda id]
//        collimator + crystal scatter; writes Scatter_SysMat_shift_* if Phys[2] == 1
//        and SysMat_withScatter_shift_* (PE + scatter) if Phys[3] == 1.
//   1  ScatterGen_Collimator    (ScatterGen_Collimator/ScatterGen_Collimator.cpp:34-219)
//        [-PE file] [-GeoCollimator file] [-cuda id]
//        writes Scatter_SysMat_shift_*, and SysMat_shift_* if Phys[3] == 1.
//   2  ScatterGen_Crystal       (ScatterGen_Crystal/ScatterGen_Crystal.cpp:34-215)
//        [-PE file] [-GeoCrystal file] [-cuda id]
//        writes Scatter_Crystal_SysMat_shift_*, and SysMat_shift_* if Phys[3] == 1.
// Naming a bitmap file on the command line clears Phys[8] / Phys[9] so the file
// is read instead of computed (as the reference does).
// Rotations: rotation r's scatter is computed from rotation r's PE block.  The
// reference drivers pass the start of the PE buffer for every rotation
// (ScatterGen_CircularHole.cpp:188), so their rotations r > 0 use the rotation-0
// PE matrix; -pe-rotation0 (an extension) reproduces that.  The reference's
// combined sum also reads N*V*R floats of per-rotation buffers
// (scatter.cu:2354-2357, out of bounds for R > 1); here each call sums N*V.
#include <chrono>
#include <cstring>
#include <string>

#include "driver_common.hpp"

#ifndef SCATTER_MODE
#define SCATTER_MODE 0
#endif

static void usage(const char* argv0, bool with_cuda) {
    std::printf("  (extension: -pe-rotation0 feeds every rotation the rotation-0 PE block, as the reference)\n");
#if SCATTER_MODE != 2
    std::printf("  (extension: -cells SIZE KEEP replaces the hole-site collimator term by plate cells of SIZE mm\n"
                "   kept within KEEP mm of a hole wall, KEEP < 0: every cell)\n");
#endif
#if SCATTER_MODE == 0
    std::printf("Usage: %s [-PE PE_SysMat_path] [-GeoCrystal GeometryRelationShip_Crystal2Crystal_path] "
                "[-GeoCollimator GeometryRelationShip_Collimator2Crystal_path]%s\n", argv0, with_cuda ? " [-cuda GPU_ID]" : " ");
#elif SCATTER_MODE == 1
    std::printf("Usage: %s [-PE PE_SysMat_path] [-GeoCollimator GeometryRelationShip_Collimator2Crystal_path]%s\n",
                argv0, with_cuda ? " [-cuda GPU_ID]" : "");
#else
    std::printf("Usage: %s [-PE PE_SysMat_path] [-GeoCrystal GeometryRelationShip_Crystal2Crystal_path]%s\n", argv0,
                with_cuda ? " [-cuda GPU_ID]" : "");
#endif
}

int main(int argc, char* argv[]) {
    drv::Params P = drv::read_all();
    drv::print_layers(P);
    std::string fname_pe;
    std::string geo_crystal = "GeometryRelationShip_Crystal2Crystal";
    std::string geo_collimator = "GeometryRelationShip_Collimator2Crystal";
    int cuda_id = 0;
    bool pe_rotation0 = false;
    float cell_size = 0.0f, cell_keep = -1.0f;
    for (int i = 1; i < argc; ++i) {
        if (std::strcmp(argv[i], "-PE") == 0 && i + 1 < argc) {
            fname_pe = argv[++i];
#if SCATTER_MODE != 1
        } else if (std::strcmp(argv[i], "-GeoCrystal") == 0 && i + 1 < argc) {
            geo_crystal = argv[++i];
            P.phy[8] = 0;
#endif
#if SCATTER_MODE != 2
        } else if (std::strcmp(argv[i], "-GeoCollimator") == 0 && i + 1 < argc) {
            geo_collimator = argv[++i];
            P.phy[9] = 0;
#endif
        } else if (std::strcmp(argv[i], "-cuda") == 0 && i + 1 < argc) {
            cuda_id = std::atoi(argv[++i]);
        } else if (std::strcmp(argv[i], "-pe-rotation0") == 0) {
            pe_rotation0 = true;
#if SCATTER_MODE != 2
        } else if (std::strcmp(argv[i], "-cells") == 0 && i + 2 < argc) {
            cell_size = (float)std::atof(argv[++i]);
            cell_keep = (float)std::atof(argv[++i]);
            if (!(cell_size > 0.0f)) {
                std::fprintf(stderr, "-cells: cell size must be positive\n");
                return EXIT_FAILURE;
            }
#endif
        } else if (std::strcmp(argv[i], "-h") == 0 || std::strcmp(argv[i], "--help") == 0) {
            usage(argv[0], true);
            return 0;
        } else {
            std::fprintf(stderr, "Unknown parameter or missing argument: %s\n", argv[i]);
            usage(argv[0], false);
            return EXIT_FAILURE;
        }
    }
    const int N = (int)P.det[0];
    const int X = (int)P.img[0], Y = (int)P.img[1], Z = (int)P.img[2];
    const int R = (int)std::floor(P.img[6] + 0.001f);
    const size_t V = (size_t)X * Y * Z;
    const size_t per_rot = (size_t)N * V;
    const size_t total = per_rot * (size_t)(R > 0 ? R : 0);
    if (fname_pe.empty()) fname_pe = drv::shift_name("PE_SysMat", P.img);
    std::printf("Photon Electric SysMat: %s\n", fname_pe.c_str());
    auto t_io = std::chrono::steady_clock::now();
    std::vector<float> pe = drv::read_floats(fname_pe, total);
    std::printf("Time of io PE System Matrix: %lld ms\n",
                (long long)std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t_io).count());
#if SCATTER_MODE != 1
    std::printf("Geometry RelationShip Crystal2Crystal:  %s\n", geo_crystal.c_str());
#endif
#if SCATTER_MODE != 2
    std::printf("Geometry RelationShip Collimator2Crystal:  %s\n", geo_collimator.c_str());
#endif
    std::vector<float> out(total, 0.0f);
    std::printf("FOV dimension : %d %d %d\n", X, Y, Z);
    std::printf("FOV Voxel Size(mm) : %f %f %f\n", P.img[3], P.img[4], P.img[5]);
    auto t0 = std::chrono::steady_clock::now();
    for (int rot = 0; rot < R; ++rot) {
        std::printf("########################\nRotation (%d) processing ...\n########################\n", rot);
        std::printf("Shift FOV in X = %gmm\nShift FOV in Y = %gmm\nShift FOV in Z = %gmm\n", P.img[8], P.img[9], P.img[10]);
        P.img[20] = (float)rot;
        float* dst = out.data() + (size_t)rot * per_rot;
        const float* pe_rot = pe.data() + (pe_rotation0 ? 0 : (size_t)rot * per_rot);
#if SCATTER_MODE == 0
        int q;
        if (cell_size > 0.0f) {  // crystal term (bitmap file semantics) + plate-cell collimator term
            q = sm_scatter_crystal_only(P.det.data(), P.img.data(), P.phy.data(), pe_rot, geo_crystal.c_str(), dst,
                                        cuda_id);
            if (q >= 0) {
                std::vector<float> cells(per_rot);
                const int c = sm_scatter_cells(P.col.data(), P.det.data(), P.img.data(), P.phy.data(), rot, cell_size,
                                               cell_keep, 0, N, cells.data(), cuda_id);
                if (c < 0) q = c;
                for (size_t i = 0; q >= 0 && i < per_rot; ++i) dst[i] = cells[i] + dst[i];
            }
        } else {
            q = sm_scatter(P.col.data(), P.det.data(), P.img.data(), P.phy.data(), pe_rot, geo_crystal.c_str(),
                           geo_collimator.c_str(), dst, cuda_id);
        }
#elif SCATTER_MODE == 1
        int q;
        if (cell_size > 0.0f) {
            q = sm_scatter_cells(P.col.data(), P.det.data(), P.img.data(), P.phy.data(), rot, cell_size, cell_keep, 0,
                                 N, dst, cuda_id);
            if (q >= 0) q = (int)V;
        } else {
            q = sm_scatter_collimator_only(P.col.data(), P.det.data(), P.img.data(), P.phy.data(), pe_rot,
                                           geo_collimator.c_str(), dst, cuda_id);
        }
#else
        int q = sm_scatter_crystal_only(P.det.data(), P.img.data(), P.phy.data(), pe_rot, geo_crystal.c_str(), dst,
                                        cuda_id);
#endif
        drv::check(q, "scatter");
        std::printf("numImagebin = %d\n", q);
    }
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Time of scatter function: %g min (%.3g elements/s)\n", s / 60.0, (double)total / (s > 0 ? s : 1));
#if SCATTER_MODE == 0
    if (P.phy[2] == 1) {
        drv::write_all(drv::shift_name("Scatter_SysMat", P.img), out.data(), total);
        std::printf("########################\nCompton Scatter Sysmat written.\n########################\n");
    }
    const char* full_prefix = "SysMat_withScatter";
#elif SCATTER_MODE == 1
    drv::write_all(drv::shift_name("Scatter_SysMat", P.img), out.data(), total);
    std::printf("########################\nCompton Scatter Sysmat of Collimator written.\n########################\n");
    const char* full_prefix = "SysMat";
#else
    drv::write_all(drv::shift_name("Scatter_Crystal_SysMat", P.img), out.data(), total);
    std::printf("########################\nInter-crystal Compton Scatter Sysmat written.\n########################\n");
    const char* full_prefix = "SysMat";
#endif
    if (P.phy[3] == 1) {
        for (size_t i = 0; i < total; ++i) out[i] += pe[i];
        drv::write_all(drv::shift_name(full_prefix, P.img), out.data(), total);
        std::printf("########################\nFull Sysmat written.\n########################\n");
    }
    return 0;
}
