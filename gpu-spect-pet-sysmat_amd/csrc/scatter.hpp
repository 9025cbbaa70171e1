// Primary-Compton scatter system matrices: host/device data layout.
//
// Semantics: crystalScatterSysMatCuda (ScatterGen_RayTracing_CircularHole/scatter.cu:1077-1437)
// and collimatorScatterSysMatCuda (scatter.cu:1544-1936; ScatterGen_Collimator
// variant scatter.cu:1056-1436).  Element (detector r, voxel v) = sum over
// scatter sites s (crystals k != r, or layer-0 holes) of
//   w_s(v) * Range_phi/2pi * KN(theta_min..theta_max)/KN(0..pi) * P_window(E')
//   * exp(-att(E')) * (1 - exp(-L_abs mu(E'))) * mu_pe(E')/mu(E').
// Everything that does not depend on the voxel (site position in r's frame,
// |Q - S|, the absorption chord, the attenuator chord sums behind the
// GeometryRelationShip bitmaps) is computed once per (r, s) pair by
// pair_prep; the per-voxel kernel then does O(1) work per (r, v, s) triple:
// the reference's O(N) bitmap scan and O(n_theta) Riemann sum per triple
// become a 2-term attenuation formula and an Euler-Maclaurin closed form.
#pragma once
#include <hip/hip_runtime.h>

#include "cells.hpp"
#include "tables.hpp"

namespace sysmat {

// Crystals of one y-slab (detector layer) binned by centre on a uniform x-z grid,
// for finding the crystals a segment can cross without scanning all N.
struct alignas(16) SlabGrid {
    double y1, y2;            // slab y-range (FOV2C added)
    double x0, z0;            // grid origin
    double cx, cz;            // cell size
    double m;                 // largest x-z half-extent of the slab's crystals, plus a margin
    int nx, nz;               // cells
    int cell_base;            // first entry of this grid in the cell offset table
    int pad;
};

// Riemann steps of the KN weight never exceed pi / h = 314 (h = 0.01f).
constexpr int kKnSteps = 320;

// Pair-table chunk: the preparation kernel handles (row, chunk of kPrepChunk
// sites) per workgroup and compacts survivors within the chunk.
constexpr int kPrepChunk = 256;

// Per detecting crystal r (FP64 where the value feeds a branch).
struct alignas(16) ScRow {
    double cx, cy, cz;        // crystal centre (sub-point, divide 1x1x1), FOV2C added in FP64
    double crn, srn;          // cos/sin(-rot): world -> crystal frame
    double hx, hy, hz;        // half extents
    double lo, hi;            // energy window (same-window flag or +-res/2 around E0)
    double res;               // energy resolution at E0
    double mu_pe, mu_c;       // photoelectric / Compton attenuation of r at E0
    double rdet;              // sqrt(w^2 + h^2 + t^2) / 2 (Range_phi sphere)
    double th_lo, th_hi;      // scatter angles outside which the window test always fails
    float use;                // 1 when r contributes (mu_t > 0.01, Compton on)
    float pad0, pad1, pad2;
};

// Per (r, site) pair, 192 B.  A_pe / A_c also hold the plate path of a cell site.
struct alignas(16) ScPair {
    double sx, sy, sz;        // site (world)
    double ssx, ssy, ssz;     // site in r's frame
    double bx, by, bz;        // unit vector site -> r's sub-point (world)
    double L;                 // |Q - S|
    double ib[8];             // 1 / |corner_i - S_self|, corners in (x, y, z) bit order below
    float A_pe, A_c;          // sum over bitmap attenuators of chord * mu_pe, chord * mu_c (at E0)
    float l_abs;              // absorption chord through r along S -> Q extended 1000 mm
    float c_fwd;              // cos(asin(R/L)) + margin: "through" test needed when |cos theta| >= this
    int site;                 // crystal index (crystal mode) or hole index (collimator mode)
    float hy1, hy2;           // hole faces (collimator mode, as stored); surrogate tables (crystal
                              // mode): hy1 = weight / angular factor at the image centre
    float pad;                // surrogate tables: single-site weight at the image centre
};

struct ScParams {
    int X, Y, Z;              // image voxel counts (Img[0..2])
    double wx, wy, wz, sx, sy, sz;
    double rc, rs;            // image rotation (Img[20] * Img[7])
    double e0;                // photopeak energy (Phys[7])
    double kn_total;          // KN Riemann sum over [0, pi] (reference normalisation)
    double ikn_total;         // 1 / kn_total
    double h;                 // KN step (0.01f)
    float ihf;                // 1 / h in FP32 (step count outside the guard band)
    const double2* kn_tab;    // (cos n h, sin n h), n = 0 .. kKnSteps
    int n_det;
    int det_begin, det_count; // detector rows of this launch
    int max_sites;            // pair table stride per row
    long long row_stride;     // floats between detector rows of dst
    long long V;              // voxels (PE matrix row length)
    int mode;                 // 0 crystal scatter, 1 collimator (combined), 2 collimator (ScatterGen_Collimator),
                              // 3 collimator plate cells (source weights in the `pe` table, [cell][voxel])
    double w_col;             // collimator source factor Col[17] / Col[15]
    double fov2c;
    int n_tiles;              // voxel tiles: 4x4x4 bricks (brick = 1) or runs of 64 voxels
    int brick;                // 1: wave = 4 x 4 x 4 voxel brick; 0: wave = 64 consecutive voxels
    int n_rgroups;            // detector groups of 4
    long long n_blocks;       // tiles * rgroups (before XCD padding)
    // optional modes (zero / null = off)
    const int* row_list;      // prep / accumulate rows: r = row_list[rr], else det_begin + rr
    int dst_by_row;           // with row_list: 1 = dst row r - det_begin, 0 = dst row rr
    int site_cap;             // accumulate over the first site_cap pairs of each row only
    int site_chunk;           // pairs of a row come in chunks of site_chunk (counts per chunk);
                              // 0: one compacted list per row (counts per row)
    const int2* items;        // list mode: wave w of block b handles items[4b + w] = (rr, voxel tile)
    long long n_items;
    int accumulate;           // 1: dst += element (e.g. crystal scatter onto the collimator term)
    // crystal grids per y-slab (n_grids == 0: attenuators found by scanning every crystal)
    const SlabGrid* grids;
    const int* grid_cell;     // cell -> first item; cell c of grid g spans [grid_cell[b + c], grid_cell[b + c + 1]),
                              // b = cell_base of g (one extra entry per grid)
    const int* grid_item;     // crystal indices, ascending within a cell
    int n_grids;
    // collimator cells (mode 3)
    const Cell* cells;        // this launch's cells; pair.site indexes them and the weight table rows
    const CellPlate* cplates;
    const CellHole* choles;
    int n_cplates;
};

// Voxel tiles of the accumulation kernel.
inline void set_tiles(ScParams& p, bool brick) {
    p.brick = brick ? 1 : 0;
    p.n_tiles = brick ? ((p.X + 3) / 4) * ((p.Y + 3) / 4) * ((p.Z + 3) / 4) : (int)((p.V + 63) / 64);
}

// Chunks of a row's pair table and their size.
__host__ __device__ inline int site_chunks(const ScParams& p) {
    return p.site_chunk > 0 ? (p.max_sites + p.site_chunk - 1) / p.site_chunk : 1;
}
__host__ __device__ inline int site_chunk_size(const ScParams& p) {
    return p.site_chunk > 0 ? p.site_chunk : p.max_sites;
}

// Scatter-surrogate near field: per detector the n_near + n_ring sites with the
// largest single-site weight at the image centre, best first; pair.pad holds the
// weight (crystal mode).
struct SurrogateRank {
    const int* cand;          // [n_det][n_cand] candidate sites (nearest crystals), -1 = none
    int n_cand;               // <= 256
    int n_keep;               // n_near + n_ring
    const float* ftab;        // [n_det][kSgTab + 1] angular factor F_r(cos theta), see sg_angular
};

// Angular factor of a far site as seen by detector r: KN shape x energy-window
// probability at the scattered energy, tabulated on cos(theta) in [-1, 1].
constexpr int kSgTab = 64;
__host__ __device__ inline float sg_angular(const float* __restrict__ tab, float c) {
    float x = (c + 1.0f) * (0.5f * kSgTab);
    x = fminf(fmaxf(x, 0.0f), (float)kSgTab);
    const int i = min((int)x, kSgTab - 1);
    const float f = x - (float)i;
    return tab[i] + f * (tab[i + 1] - tab[i]);
}

hipError_t launch_scatter_prep(const ScParams& p, const ScRow* rows, const Crystal* cry, const Hole* hol,
                               const unsigned* bits, ScPair* pairs, int* counts, hipStream_t stream);
hipError_t launch_scatter(const ScParams& p, const ScRow* rows, const ScPair* pairs, const int* counts,
                          const float* pe, float* dst, hipStream_t stream);
hipError_t launch_surrogate_rank(const ScParams& p, const SurrogateRank& k, const ScRow* rows, const Crystal* cry,
                                 const unsigned* bits, ScPair* keep, int* counts, hipStream_t stream);
hipError_t launch_cell_weights(const ScParams& p, int n_cells, float* weights, hipStream_t stream);
hipError_t launch_bitmap_crystal(const Crystal* cry, int n, unsigned* words, hipStream_t stream);
hipError_t launch_bitmap_collimator(const Hole* hol, int n_holes, const Crystal* cry, int n, unsigned* words,
                                    hipStream_t stream);

}  // namespace sysmat
