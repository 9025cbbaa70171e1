// Launch parameters of the PE tracer (pe_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "tables.hpp"

namespace sysmat {

constexpr int kCapCrystals = 4096;  // tile candidate crystals (reference per-element cap: 3000, PESysMatGen.cu:481)
constexpr int kCapHoles = 1024;     // tile candidate holes (reference: 1000 per layer, PESysMatGen.cu:485)
// Revision of the tracer's schedule; part of the autotuner's cache key, so tile
// choices tuned for an older schedule are not reused.
constexpr int kPeScheduleRevision = 2;

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
    // detector rows of this launch
    int n_det, det_begin, det_count;
    int n_layers, n_holes, n_slabs;
    long long row_stride;      // floats between consecutive detector rows of dst
    int use_bundle;
    float inv4pi;
    double rc64, rs64;         // image rotation in double (FP64 end points of the collimator path)
    float fov2c;               // Img[11]
};

hipError_t launch_pe(const PeParams& p, const Crystal* cry, const Hole* hol, const Plate* pla, const Slab* slb,
                     float* dst, int* status, hipStream_t stream);

}  // namespace sysmat
