// Collimator scatter in plate cells: the plates cut into small cells, each a
// Compton scatter site.
//
// The reference models collimator scatter with one site per layer-0 hole, a
// virtual-disc source weight and no attenuation inside the collimator
// (ScatterGen_RayTracing_CircularHole/scatter.cu:1544-1936).  The cell model
// replaces the sites and the source weight and keeps everything after the
// site (Klein-Nishina angular weight, energy window, crystal attenuation,
// absorption) exactly as for the hole sites:
//
//   plates     layer m box x in +-W/2, y in FOV2C + offset +- t/2, z in +-H/2
//              (Col[(m+1)*10 + 1..4]), FP64 from the float parameters
//   cells      n_x = max(1, round(W/h)) x n_y x n_z cells of W/n_x x t/n_y x H/n_z;
//              a cell is material when its centre lies outside every hole of
//              its layer; it is kept when keep < 0, or when its centre lies
//              within `keep` mm (x-z distance) of the wall of a hole of its layer
//   weight     W(I, c) = mu_c(m) vol / (4 pi |c - I|^2) exp(-A_t(I -> c))
//   plate path A(P -> Q) = sum over plates of mu(m) (|PQ ∩ box| - sum_h |PQ ∩ hole_h ∩ box|)
//                          + sum_h mu(h) |PQ ∩ hole_h ∩ box|   (holes of that plate)
//   out-leg    A_pe, A_c of the pair gain the plate path from the cell to the
//              detector centre, scaled to E' like the crystal attenuators
//   attenuating crystals: the reference's angular bounding-sphere criterion
//              (scatter.cu:1960-2000) seen from the cell centre, in FP64
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace sysmat {

struct alignas(16) Cell {
    double x, y, z;   // centre (world, FOV2C added)
    double w;         // mu_c(layer) * volume / (4 pi)
};

struct alignas(16) CellPlate {
    double x1, x2, y1, y2, z1, z2;
    double mu_t, mu_pe, mu_c;
    int hole_begin, hole_end;
};

struct alignas(16) CellHole {
    double xc, zc, r, y1, y2;   // axis, radius, faces (y1 <= y2, FOV2C added)
    double mu_t, mu_pe, mu_c;
};

struct PathAtt {
    double t, pe, c;
};

// Narrow [t0, t1] to the part of p + t d inside [a, b] on one axis.
__host__ __device__ inline void clip_axis(double p, double d, double a, double b, double& t0, double& t1) {
    if (d == 0.0) {
        if (p < a || p > b) { t0 = 1.0; t1 = 0.0; }
        return;
    }
    double ta = (a - p) / d, tb = (b - p) / d;
    if (ta > tb) { const double s = ta; ta = tb; tb = s; }
    t0 = fmax(t0, ta);
    t1 = fmin(t1, tb);
}

// Parameter interval of the segment P + t (Q - P), t in [0, 1], inside the
// y-axis cylinder (xc, zc, r, y1..y2), intersected with [t0, t1].
__host__ __device__ inline void clip_cylinder(double px, double py, double pz, double dx, double dy, double dz,
                                              const CellHole& H, double& t0, double& t1) {
    clip_axis(py, dy, H.y1, H.y2, t0, t1);
    if (!(t1 > t0)) return;
    const double ex = px - H.xc, ez = pz - H.zc;
    const double a = dx * dx + dz * dz;
    const double b = ex * dx + ez * dz;       // half of the linear coefficient
    const double c = ex * ex + ez * ez - H.r * H.r;
    if (a == 0.0) {
        if (c > 0.0) { t0 = 1.0; t1 = 0.0; }
        return;
    }
    const double disc = b * b - a * c;
    if (disc <= 0.0) { t0 = 1.0; t1 = 0.0; return; }
    const double sq = sqrt(disc);
    // roots of a t^2 + 2 b t + c without cancellation
    const double q = (b >= 0.0) ? -(b + sq) : -(b - sq);
    double ra = q / a, rb = (q != 0.0) ? c / q : ra;
    if (ra > rb) { const double s = ra; ra = rb; rb = s; }
    t0 = fmax(t0, ra);
    t1 = fmin(t1, rb);
}

// Attenuation sums (mu_t, mu_pe, mu_c weighted path lengths) of the segment P -> Q
// through the plates and their holes.
__host__ __device__ inline PathAtt plate_path(const CellPlate* __restrict__ pl, int n_pl,
                                              const CellHole* __restrict__ ho, double px, double py, double pz,
                                              double qx, double qy, double qz) {
    PathAtt a{0.0, 0.0, 0.0};
    const double dx = qx - px, dy = qy - py, dz = qz - pz;
    const double len = sqrt(dx * dx + dy * dy + dz * dz);
    for (int m = 0; m < n_pl; ++m) {
        const CellPlate P = pl[m];
        double b0 = 0.0, b1 = 1.0;
        clip_axis(px, dx, P.x1, P.x2, b0, b1);
        clip_axis(py, dy, P.y1, P.y2, b0, b1);
        clip_axis(pz, dz, P.z1, P.z2, b0, b1);
        if (!(b1 > b0)) continue;
        const double lbox = (b1 - b0) * len;
        double lh = 0.0, ht = 0.0, hpe = 0.0, hc = 0.0;
        for (int h = P.hole_begin; h < P.hole_end; ++h) {
            const CellHole H = ho[h];
            double c0 = b0, c1 = b1;
            clip_cylinder(px, py, pz, dx, dy, dz, H, c0, c1);
            if (!(c1 > c0)) continue;
            const double l = (c1 - c0) * len;
            lh += l;
            ht += H.mu_t * l;
            hpe += H.mu_pe * l;
            hc += H.mu_c * l;
        }
        const double lm = fmax(0.0, lbox - lh);
        a.t += P.mu_t * lm + ht;
        a.pe += P.mu_pe * lm + hpe;
        a.c += P.mu_c * lm + hc;
    }
    return a;
}

// Host: plates, holes and the kept cells of the collimator `col` (FOV2C = fov2c)
// for cell size h and keep distance `keep` (< 0: every material cell).
// Returns "" on success.
inline std::string build_cells(const float* col, float fov2c, double h, double keep, std::vector<Cell>& cells,
                               std::vector<CellPlate>& plates, std::vector<CellHole>& holes) {
    cells.clear();
    plates.clear();
    holes.clear();
    if (!(h > 0.0)) return "cell size must be positive";
    const int n_layers = (int)std::floor(col[0] + 0.000001f);
    if (n_layers < 0 || n_layers > 9) return "collimator layer count out of range (0..9)";
    const double f = fov2c;
    int hb = 0;
    for (int m = 0; m < n_layers; ++m) {
        const float* L = col + (m + 1) * 10;
        const int nh = (int)std::floor(L[0]);
        if (nh < 0) return "negative hole count";
        CellPlate P{};
        P.x1 = -(double)L[1] / 2.0;
        P.x2 = (double)L[1] / 2.0;
        P.y1 = f + (double)L[4] - (double)L[2] / 2.0;
        P.y2 = f + (double)L[4] + (double)L[2] / 2.0;
        P.z1 = -(double)L[3] / 2.0;
        P.z2 = (double)L[3] / 2.0;
        P.mu_t = L[5];
        P.mu_pe = L[6];
        P.mu_c = L[7];
        P.hole_begin = hb;
        P.hole_end = hb + nh;
        for (int q = 0; q < nh; ++q) {
            const float* H = col + 100 + 9 * (hb + q);
            CellHole o{};
            o.xc = H[0];
            o.zc = H[3];
            o.r = H[4];
            const double ya = (double)H[1] + f, yb = (double)H[2] + f;
            o.y1 = ya < yb ? ya : yb;
            o.y2 = ya < yb ? yb : ya;
            o.mu_t = H[5];
            o.mu_pe = H[6];
            o.mu_c = H[7];
            holes.push_back(o);
        }
        hb += nh;
        plates.push_back(P);
    }
    for (int m = 0; m < n_layers; ++m) {
        const CellPlate& P = plates[m];
        const double W = P.x2 - P.x1, T = P.y2 - P.y1, Hh = P.z2 - P.z1;
        if (!(W > 0.0 && T > 0.0 && Hh > 0.0) || !(P.mu_c > 0.0)) continue;
        const long long nx = std::max(1LL, std::llround(W / h)), ny = std::max(1LL, std::llround(T / h)),
                        nz = std::max(1LL, std::llround(Hh / h));
        if (nx * ny * nz > (1LL << 30)) return "too many collimator cells";
        const double cx = W / (double)nx, cy = T / (double)ny, cz = Hh / (double)nz;
        const double w = P.mu_c * (cx * cy * cz) / (4.0 * M_PI);
        for (long long iy = 0; iy < ny; ++iy) {
            const double y = P.y1 + ((double)iy + 0.5) * cy;
            for (long long iz = 0; iz < nz; ++iz) {
                const double z = P.z1 + ((double)iz + 0.5) * cz;
                for (long long ix = 0; ix < nx; ++ix) {
                    const double x = P.x1 + ((double)ix + 0.5) * cx;
                    bool material = true;
                    double dmin = INFINITY;
                    for (int q = P.hole_begin; q < P.hole_end; ++q) {
                        const CellHole& H = holes[q];
                        const double d = std::sqrt((x - H.xc) * (x - H.xc) + (z - H.zc) * (z - H.zc));
                        if (d < H.r && y >= H.y1 && y <= H.y2) material = false;
                        dmin = std::min(dmin, d - H.r);
                    }
                    if (!material) continue;
                    if (keep >= 0.0 && !(dmin <= keep)) continue;
                    cells.push_back(Cell{x, y, z, w});
                }
            }
        }
    }
    return "";
}

}  // namespace sysmat
