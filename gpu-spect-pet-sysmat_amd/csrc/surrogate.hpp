// Scatter surrogate: exact near field + learned far field of the crystal
// (inter-crystal) Compton scatter matrix.
//
// For detector r the sites are ranked once by their single-site weight at the
// image centre (surrogate_rank_kernel).  The best n_near sites are summed
// exactly for every voxel (scatter_kernel with a site cap): S_near.  The rest,
// S_far = S - S_near, is smooth in (r, v) and is predicted by the MLP from 12
// features; the next n_ring sites give a fixed-weight prior feature
// P_ring = sum_j PE[site_j, v] w_j.  Members of an ensemble that disagree
// mark their 4 x 4 x 4 voxel brick for the exact kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "scatter.hpp"

namespace sysmat {

constexpr int kSgFeatures = 12;

struct alignas(16) SgRow {
    float cx, cy, cz, use;  // crystal centre (FOV2C added), 1 when the row has scatter
    float f3, f4, f5, pad;  // normalised centre (detector bounding box)
};

struct SgParams {
    int X, Y, Z;
    float wx, wy, wz, sx, sy, sz, rc, rs;
    float icx, icy, icz, irho;  // rotated image centre, image half diagonal
    float d_ref;                // image centre to detector centroid
    long long V;
    int n_det, n_near, n_keep;
    int det_begin, det_count;
    long long row_stride;
    float pe_scale, s_scale, mu_y, sig_y, phi;  // normalisation (set after training-data statistics)
    float tol, floor_abs;
    int n_members;
    long long n_tiles;          // 4 x 4 x 4 voxel bricks per row
    const float* ftab;          // [n_det][kSgTab + 1] angular factor tables
    const float4* ring;         // [n_det][n_keep][2] ranked sites for the ring prior: (x, y, z, weight),
                                // (unit vector to the detector, site index bits); zero past the row's count
    const int* ring_cnt;        // [n_det] 1 + rank of the row's last site with a non-zero ring weight
                                // (n_near when none): the ring prior's loop bound
};

hipError_t launch_surrogate_features(const SgParams& q, const SgRow* rows, const ScPair* keep, const int* keep_cnt,
                                     const float* pe, const float* near, long long near_stride, const int* vox,
                                     int n_vox, float* x, hipStream_t s);
hipError_t launch_surrogate_ring_table(const ScPair* keep, const int* keep_cnt, int n_det, int n_keep, int n_near,
                                      float4* ring, int* ring_cnt, hipStream_t s);
hipError_t launch_surrogate_apply(const SgParams& q, const SgRow* rows, const ScPair* keep, const int* keep_cnt,
                                  const uint8_t* packed, const float* pe, float* dst, uint8_t* flags, hipStream_t s);

}  // namespace sysmat
