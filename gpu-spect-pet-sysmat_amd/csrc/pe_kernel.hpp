// Launch parameters of the PE tracer (pe_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "tables.hpp"

namespace sysmat {

constexpr int kCapCrystals = 4096;  // tile candidate crystals (reference per-element cap: 3000, PESysMatGen.cu:481)
constexpr int kCapHoles = 1024;     // tile candidate holes (reference: 1000 per layer, PESysMatGen.cu:485)
constexpr int kCapGroup = 4096;     // tile-group candidate crystals kept in LDS (beyond: retry with smaller tiles)
static_assert(kCapGroup <= kCapCrystals, "a one-tile group writes its crystal list into the tile list");
constexpr int kGroupTiles = 2;      // tiles per group along each axis
// Plans with fewer (detector, tile) pairs run one tile per workgroup (finer work
// units, shorter tail); so do launches that would have fewer workgroups.
constexpr long long kGroupMinTiles = 1LL << 18;
constexpr long long kGroupMinWorkgroups = 2048;
// A workgroup traces up to kGroupsPerWorkgroup consecutive tile groups of its
// detector row, as long as the launch keeps at least kLoopMinWorkgroups
// workgroups (the detector set-up and its spill stores are paid once per run).
constexpr int kGroupsPerWorkgroup = 8;
constexpr long long kLoopMinWorkgroups = 16384;
// Revision of the tracer's schedule; part of the autotuner's cache key, so tile
// choices tuned for an older schedule are not reused.
constexpr int kPeScheduleRevision = 4;

struct PeParams {
    // image slab (in-band Img[2], Img[30], Img[31] of the reference ABI)
    int X, Y, Zslab, Zfull, z0;
    float wx, wy, wz, sx, sy, sz;
    float rc, rs;              // cos / sin of the image rotation Img[20] * Img[7]
    // sampling
    int nv;                    // voxel sub-points per axis
    int cx, cy, cz;            // crystal sub-division
    int S;                     // sub-rays per element = nv^3 * cx * cy * cz
    // tiling
    int tx, ty, tz, ntx, nty, ntz;
    int gx, gy, gz, ngx, ngy, ngz;  // tiles per workgroup group, group counts
    int grp_per_wg;            // consecutive tile groups per workgroup (set by launch_pe)
    // detector rows of this launch
    int n_det, det_begin, det_count;
    int n_layers, n_holes, n_slabs;
    long long row_stride;      // floats between consecutive detector rows of dst
    int use_bundle;
    float inv4pi;
    double rc64, rs64;         // image rotation in double (FP64 end points of the collimator path)
    double rnv64, rcx64, rcy64, rcz64;  // 1 / nv, 1 / cx, 1 / cy, 1 / cz (correctly rounded)
    double hX64, hY64, hZ64;   // X / 2, Y / 2, Zfull / 2 (FP64 voxel centres of the collimator path)
    double wx64, wy64, wz64, sx64, sy64, sz64;  // voxel size and image shift in FP64
    float fov2c;               // Img[11]
};

// Per-detector candidate lists (CSR over detectors, indices in 16 bits): the
// crystals and holes that may attenuate some sub-ray from the image slab to the
// detector; built once per plan by launch_detector_cull.
struct PeLists {
    const long long* cry_ptr;   // [n_det + 1]
    unsigned short* cry_list;
    const long long* hol_ptr;   // [n_det + 1]
    unsigned short* hol_list;
};

hipError_t launch_pe(const PeParams& p, const Crystal* cry, const Hole* hol, const Plate* pla, const Slab* slb,
                     const PeLists& L, float* dst, int* status, hipStream_t stream);
// fill = 0: counts[2 d] / counts[2 d + 1] = list lengths of detector d; fill = 1:
// write the lists at L's offsets.  One workgroup per detector of the table.
hipError_t launch_detector_cull(const PeParams& p, const Crystal* cry, const Hole* hol, const Slab* slb, int fill,
                                int* counts, const PeLists& L, hipStream_t stream);

}  // namespace sysmat
