// CPU oracle for the photoelectric (PE) system matrix.  TEST INFRASTRUCTURE
// ONLY: the shipped library never links or calls this file; the tests, the
// smoke check and the CPU-baseline leg of the benchmark use it as the checker.
//
// Restates photodetectorCudaMe (PEGen_RayTracing_CircularHole/PESysMatGen.cu:375-764)
// for one Z-slab, element by element, in the reference's operation order:
//   voxel / crystal set-up ............ PESysMatGen.cu:382-478
//   per-element culling ............... PESysMatGen.cu:480-590
//   sub-ray loop (NumZ, NumX, NumY) ... PESysMatGen.cu:598-762
// plus the sampling extension (n_v^3 voxel sub-points averaged, crystal
// sub-grid cx x cy x cz summed; n_v = 1 and 8x16x8 is the reference).
//
// Instantiations: float (FP32 restatement), double (FP64 oracle),
// Counted<float> (algorithmic FLOP count; culling prologue excluded).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "geometry.hpp"

namespace oracle {

thread_local uint64_t OpCounter::flops = 0;
thread_local bool OpCounter::enabled = false;

struct PeOpts {
    int nv;                 // voxel sub-points per axis (1 = voxel centre, the reference)
    int cx, cy, cz;         // crystal sub-division (reference: 8, 16, 8)
    int det_begin, det_end; // detector rows to compute
    int threads;            // OpenMP threads (<=0: all)
    int flags;              // FP32 build only: bit 0 = FP64 element accumulator,
                            // bit 1 = FP64 collimator path (hole chords and plate length),
                            // bit 2 = FP64 absorption chord through the sub-cell,
                            // bit 3 = FP64 crystal chords, bit 4 = FP64 self path,
                            // bit 5 = the collimator path's end points in FP64,
                            // bit 6 = count only attenuators that touch the element (non-zero
                            //         chord on some sub-ray); the value is unchanged
    double jitter;          // conditioning probe: displace every sub-ray's end points by up
    int jseed;              // to `jitter` mm (pseudo-random per sub-ray, seeded by jseed)
};

// splitmix64: deterministic per-(element, sub-ray) displacement in [-1, 1)
static inline double jit_unit(uint64_t& s) {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

struct PeStats {            // optional per-row diagnostics
    double flops;           // counted FLOPs (Counted build only)
    double n_cross_det;     // summed culled-crystal list lengths
    double n_cross_hole;    // summed culled-hole list lengths
    double n_touch_det;     // flag 64: summed counts of crystals with a non-zero chord on some sub-ray
    double n_touch_hole;    // flag 64: the same for holes
};

static inline int ifloor(float v) { return (int)std::floor(v); }

template <class R>
static R pe_element(const float* col, const float* det, const float* img, int row, int col0, const PeOpts& o,
                    PeStats* st) {
    using W = typename Wide<R>::type;
    const int n_layers = (int)std::floor(col[0] + 0.000001f);
    const int n_det = (int)std::floor(det[0] + 0.000001f);
    const R fov2c = R(img[11]);
    const int X = ifloor(img[0]), Y = ifloor(img[1]), Zl = ifloor(img[2]), Zf = ifloor(img[30]);
    (void)Zl;
    const int z0 = (int)std::floor(img[31] + 0.001f);
    const R wx = R(img[3]), wy = R(img[4]), wz = R(img[5]);
    const R rot_img = R(img[20]) * R(img[7]);
    const R sx = R(img[8]), sy = R(img[9]), sz = R(img[10]);

    const int izl = col0 / (Y * X);
    const int rem = col0 % (Y * X);
    const int iy = rem / X, ix = rem % X;
    const int iz = z0 + izl;

    // voxel centre before rotation (PESysMatGen.cu:438-444)
    const R xc0 = (R((float)ix) - R((float)X) / R(2.0f) + R(0.5f)) * wx + sx;
    const R yc0 = (R((float)iy) - R((float)Y) / R(2.0f) + R(0.5f)) * wy + sy;
    const R zc0 = (R((float)iz) - R((float)Zf) / R(2.0f) + R(0.5f)) * wz + sz;
    const R ca = rot_cos(rot_img), sa = rot_sin(rot_img);
    auto rotate_image = [&](R x, R y, R& xr, R& yr) {
        xr = x * ca - y * sa;
        yr = x * sa + y * ca;
    };
    R xI, yI;
    rotate_image(xc0, yc0, xI, yI);
    const R zI = zc0;

    const float* dd = det + 12 * row;
    const R xD = R(dd[1]), yD = R(dd[2]) + fov2c, zD = R(dd[3]);
    const R rotD = R(dd[11]);
    const R wD = R(dd[4]), hD = R(dd[6]), tD = R(dd[5]);
    const R mu_t = R(dd[7]), mu_pe = R(dd[8]);
    const R cD = rot_cos(rotD), sD = rot_sin(rotD);
    const R cDn = rot_cos(-rotD), sDn = rot_sin(-rotD);

    if (!(mu_t > R(0.01f))) return R(0.0f);

    // ---- per-element culling from the voxel centre (PESysMatGen.cu:504-590) ----
    const bool counting = OpCounter::enabled;
    OpCounter::enabled = false;
    std::vector<int> cross_det;
    std::vector<std::vector<int>> cross_hole(n_layers);
    std::vector<int> holes_in(n_layers);
    std::vector<R> thick(n_layers);
    for (int m = 0; m < n_layers; ++m) {
        holes_in[m] = (int)std::floor(col[(m + 1) * 10 + 0]);
        thick[m] = R(col[(m + 1) * 10 + 2]);
    }
    {
        R dist = r_sqrt((yD - yI) * (yD - yI) + (xD - xI) * (xD - xI) + (zD - zI) * (zD - zI));
        R crit_self = r_sqrt(wD * wD + hD * hD + tD * tD) / dist / R(2.0f);
        R ux = (xD - xI) / dist, uy = (yD - yI) / dist, uz = (zD - zI) / dist;
        for (int k = 0; k < n_det; ++k) {
            if (k == row) continue;
            const float* dk = det + 12 * k;
            R wk = R(dk[4]), hk = R(dk[6]), tk = R(dk[5]);
            R crit = r_sqrt(wk * wk + hk * hk + tk * tk);
            R xk = R(dk[1]), yk = R(dk[2]) + fov2c, zk = R(dk[3]);
            R dk2 = r_sqrt((xI - xk) * (xI - xk) + (yI - yk) * (yI - yk) + (zI - zk) * (zI - zk));
            crit = crit / dk2 / R(2.0f);
            R vx = (xk - xI) / dk2, vy = (yk - yI) / dk2, vz = (zk - zI) / dk2;
            if (r_sqrt((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <= crit + crit_self)
                cross_det.push_back(k);
        }
        int h = 0;
        for (int m = 0; m < n_layers; ++m) {
            for (int q = 0; q < holes_in[m]; ++q, ++h) {
                const float* hp = col + 100 + 9 * h;
                R xh = R(hp[0]), y1 = R(hp[1]) + fov2c, y2 = R(hp[2]) + fov2c, zh = R(hp[3]), rh = R(hp[4]);
                R yh = (y1 + y2) / R(2.0f);
                R crit = r_sqrt(rh * rh * R(4.0f) + (y1 - y2) * (y1 - y2)) / R(2.0f);
                R dh = r_sqrt((xI - xh) * (xI - xh) + (yI - yh) * (yI - yh) + (zI - zh) * (zI - zh));
                crit = crit / dh;
                R vx = (xh - xI) / dh, vy = (yh - yI) / dh, vz = (zh - zI) / dh;
                if (r_sqrt((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <= crit + crit_self)
                    cross_hole[m].push_back(h);
            }
        }
    }
    OpCounter::enabled = counting;
    if (st) {
        st->n_cross_det += (double)cross_det.size();
        for (auto& v : cross_hole) st->n_cross_hole += (double)v.size();
    }

    // ---- sub-ray loop (PESysMatGen.cu:600-762) ----
    const R divX = R((float)o.cx), divY = R((float)o.cy), divZ = R((float)o.cz);
    const R x1c = -wD / R(2.0f), x2c = wD / R(2.0f);
    const R y1c = -tD / R(2.0f), y2c = tD / R(2.0f);
    const R z1c = -hD / R(2.0f), z2c = hD / R(2.0f);
    const W fourpi = W(4.0 * M_PI);
    const bool acc_wide = (o.flags & 1) != 0, plate_wide = (o.flags & 2) != 0;
    const int nv = o.nv;
    // One pass over every sub-ray of the element with the given attenuator lists;
    // mark_* (optional) flag the attenuators whose chord is non-zero on some sub-ray.
    auto trace = [&](const std::vector<int>& cross_det, const std::vector<std::vector<int>>& cross_hole,
                     std::vector<char>* mark_det, std::vector<std::vector<char>>* mark_hole) -> R {
        R acc = R(0.0f);
        double acc64 = 0.0;
        for (int vz = 0; vz < nv; ++vz)
        for (int vy = 0; vy < nv; ++vy)
        for (int vx = 0; vx < nv; ++vx) {
            // voxel sub-point (offset 0 exactly when nv == 1)
            const R fnv = R((float)nv);
            R ox = ((R((float)vx) + R(0.5f)) / fnv - R(0.5f)) * wx;
            R oy = ((R((float)vy) + R(0.5f)) / fnv - R(0.5f)) * wy;
            R oz = ((R((float)vz) + R(0.5f)) / fnv - R(0.5f)) * wz;
            R xP, yP;
            if (nv == 1) {
                xP = xI;
                yP = yI;
            } else {
                rotate_image(xc0 + ox, yc0 + oy, xP, yP);
            }
            R zP = (nv == 1) ? zI : zc0 + oz;
            uint64_t jstate = ((uint64_t)row * 2654435761ull) ^ ((uint64_t)col0 << 20) ^
                              ((uint64_t)((vz * nv + vy) * nv + vx) << 44) ^ ((uint64_t)o.jseed << 56);
            if (o.jitter != 0.0) {
                xP = xP + R(o.jitter * jit_unit(jstate));
                yP = yP + R(o.jitter * jit_unit(jstate));
                zP = zP + R(o.jitter * jit_unit(jstate));
            }
            // the same sub-point in double (collimator path with FP64 end points, flag 32)
            double xP64 = to_double(xP), yP64 = to_double(yP), zP64 = to_double(zP);
            if (o.flags & 32) {
                const double wx64 = img[3], wy64 = img[4], wz64 = img[5];
                const double ang = (double)img[20] * (double)img[7];
                const double ca64 = std::cos(ang), sa64 = std::sin(ang);
                const double x0 = ((double)ix - (double)X / 2.0 + 0.5) * wx64 + (double)img[8];
                const double y0 = ((double)iy - (double)Y / 2.0 + 0.5) * wy64 + (double)img[9];
                const double z0d = ((double)iz - (double)Zf / 2.0 + 0.5) * wz64 + (double)img[10];
                const double ox64 = (((double)vx + 0.5) / nv - 0.5) * wx64, oy64 = (((double)vy + 0.5) / nv - 0.5) * wy64;
                const double oz64 = (((double)vz + 0.5) / nv - 0.5) * wz64;
                xP64 = (x0 + ox64) * ca64 - (y0 + oy64) * sa64;
                yP64 = (x0 + ox64) * sa64 + (y0 + oy64) * ca64;
                zP64 = z0d + oz64;
            }
            // image point in the detecting crystal's frame (PESysMatGen.cu:460-464)
            const R xPs = (xP - xD) * cDn - (zP - zD) * sDn;
            const R yPs = yP - yD;
            const R zPs = (xP - xD) * sDn + (zP - zD) * cDn;

            for (int nz = 0; nz < o.cz; ++nz)
            for (int nx = 0; nx < o.cx; ++nx)
            for (int ny = 0; ny < o.cy; ++ny) {
                R xDs = -wD / R(2.0f) + (R((float)nx) + R(0.5f)) / divX * wD;
                R zDs = -hD / R(2.0f) + (R((float)nz) + R(0.5f)) / divZ * hD;
                R yDs = -tD / R(2.0f) + (R((float)ny) + R(0.5f)) / divY * tD;
                if (o.jitter != 0.0) {
                    xDs = xDs + R(o.jitter * jit_unit(jstate));
                    yDs = yDs + R(o.jitter * jit_unit(jstate));
                    zDs = zDs + R(o.jitter * jit_unit(jstate));
                }
                R xQ = xD + (xDs * cD - zDs * sD);
                R zQ = zD + (xDs * sD + zDs * cD);
                R yQ = yD + yDs;

                R dsq = (yQ - yP) * (yQ - yP) + (xQ - xP) * (xQ - xP) + (zQ - zP) * (zQ - zP);
                R dist = r_sqrt(dsq);
                R cosang = (yQ - yP) / dist;
                // projected sub-face solid angles; M_PI promotes the quotient to double
                W W4 = fourpi * W(dsq);
                R sa_x = R(W(tD * hD / divZ / divY) / W4 * W(r_fabs((xQ - xP) * cD - (zQ - zP) * sD)) / W(dist));
                R sa_z = R(W(wD * tD / divX / divY) / W4 * W(r_fabs((xQ - xP) * sD + (zQ - zP) * cD)) / W(dist));
                R sa_y = R(W(wD * hD / divX / divZ) / W4 * W(r_fabs(yQ - yP)) / W(dist));
                sa_x = rmax(sa_x, sa_y);
                R solid = rmax(sa_x, sa_z);

                R att = R(0.0f);
                for (int m = 0; m < n_layers; ++m) {  // plate miss => blocked
                    const float* lp = col + (m + 1) * 10;
                    R bx1 = -R(lp[1]) / R(2.0f), bx2 = R(lp[1]) / R(2.0f);
                    R by1 = -R(lp[2]) / R(2.0f) + fov2c + R(lp[4]);
                    R by2 = R(lp[2]) / R(2.0f) + fov2c + R(lp[4]);
                    R bz1 = -R(lp[3]) / R(2.0f), bz2 = R(lp[3]) / R(2.0f);
                    R len = chord_box(xP, yP, zP, xQ, yQ, zQ, bx1, by1, bz1, bx2, by2, bz2);
                    if (len < R(0.1f)) att += R(100000.0f);
                }
                for (int m = 0; m < n_layers; ++m) {  // plate path minus hole chords
                    if (plate_wide) {  // same steps in double from the FP32 (or FP64, flag 32) end points
                        double xp = to_double(xP), yp = to_double(yP), zp = to_double(zP);
                        double xq = to_double(xQ), yq = to_double(yQ), zq = to_double(zQ);
                        if (o.flags & 32) {
                            const double w64 = dd[4], h64 = dd[6], t64 = dd[5], r64 = dd[11];
                            const double xs = -w64 / 2.0 + ((double)nx + 0.5) / o.cx * w64;
                            const double zs = -h64 / 2.0 + ((double)nz + 0.5) / o.cz * h64;
                            const double ys = -t64 / 2.0 + ((double)ny + 0.5) / o.cy * t64;
                            xp = xP64; yp = yP64; zp = zP64;
                            xq = (double)dd[1] + (xs * std::cos(r64) - zs * std::sin(r64));
                            zq = (double)dd[3] + (xs * std::sin(r64) + zs * std::cos(r64));
                            yq = ((double)dd[2] + (double)img[11]) + ys;
                        }
                        const double f2c = to_double(fov2c);
                        double holes_len = 0.0, a64 = 0.0;
                        double dsq64 = (yq - yp) * (yq - yp) + (xq - xp) * (xq - xp) + (zq - zp) * (zq - zp);
                        double in_plate = to_double(thick[m]) / ((yq - yp) / std::sqrt(dsq64));
                        for (size_t q = 0; q < cross_hole[m].size(); ++q) {
                            const int h = cross_hole[m][q];
                            const float* hp = col + 100 + 9 * h;
                            double c = chord_hole<double>(xp, yp, zp, xq, yq, zq, hp[0], (double)hp[1] + f2c,
                                                          (double)hp[2] + f2c, hp[3], hp[4]);
                            if (mark_hole && c != 0.0) (*mark_hole)[m][q] = 1;
                            holes_len += c;
                            a64 += (double)hp[5] * c;
                        }
                        if ((in_plate - holes_len) >= 0.00001f) a64 += (double)col[(m + 1) * 10 + 5] * (in_plate - holes_len);
                        att += R((float)a64);
                        continue;
                    }
                    R holes_len = R(0.0f);
                    R in_plate = thick[m] / cosang;
                    R mu_plate = R(col[(m + 1) * 10 + 5]);
                    for (size_t q = 0; q < cross_hole[m].size(); ++q) {
                        const int h = cross_hole[m][q];
                        const float* hp = col + 100 + 9 * h;
                        R c = chord_hole(xP, yP, zP, xQ, yQ, zQ, R(hp[0]), R(hp[1]) + fov2c, R(hp[2]) + fov2c,
                                         R(hp[3]), R(hp[4]));
                        if (mark_hole && c != R(0.0f)) (*mark_hole)[m][q] = 1;
                        holes_len += c;
                        att += R(hp[5]) * c;
                    }
                    if ((in_plate - holes_len) >= R(0.00001f)) att += mu_plate * (in_plate - holes_len);
                }
                for (size_t q = 0; q < cross_det.size(); ++q) {  // other crystals on the way
                    const int k = cross_det[q];
                    const float* dk = det + 12 * k;
                    R xk = R(dk[1]), yk = R(dk[2]) + fov2c, zk = R(dk[3]);
                    R wk = R(dk[4]), hk = R(dk[6]), tk = R(dk[5]);
                    R mk = R(dk[7]), rk = R(dk[11]);
                    R ck = rot_cos(-rk), sk = rot_sin(-rk);
                    R xPk = (xP - xk) * ck - (zP - zk) * sk, yPk = yP - yk, zPk = (xP - xk) * sk + (zP - zk) * ck;
                    R xQk = (xQ - xk) * ck - (zQ - zk) * sk, yQk = yQ - yk, zQk = (xQ - xk) * sk + (zQ - zk) * ck;
                    R c;
                    if (o.flags & 8) {  // crystal chord in double from the FP32 frame coordinates
                        c = R((float)chord_box<double>(to_double(xPk), to_double(yPk), to_double(zPk), to_double(xQk),
                                                       to_double(yQk), to_double(zQk), to_double(R(-0.5f) * wk),
                                                       to_double(R(-0.5f) * tk), to_double(R(-0.5f) * hk),
                                                       to_double(R(0.5f) * wk), to_double(R(0.5f) * tk),
                                                       to_double(R(0.5f) * hk)));
                    } else {
                        c = chord_box(xPk, yPk, zPk, xQk, yQk, zQk, R(-0.5f) * wk, R(-0.5f) * tk, R(-0.5f) * hk,
                                      R(0.5f) * wk, R(0.5f) * tk, R(0.5f) * hk);
                    }
                    if (mark_det && c != R(0.0f)) (*mark_det)[q] = 1;
                    att += c * mk;
                }
                // sub-cell box in the crystal frame (PESysMatGen.cu:727-734)
                R ux1 = (R((float)nx) / divX - R(0.5f)) * wD, ux2 = ((R((float)nx) + R(1.0f)) / divX - R(0.5f)) * wD;
                R uy1 = (R((float)ny) / divY - R(0.5f)) * tD, uy2 = ((R((float)ny) + R(1.0f)) / divY - R(0.5f)) * tD;
                R uz1 = (R((float)nz) / divZ - R(0.5f)) * hD, uz2 = ((R((float)nz) + R(1.0f)) / divZ - R(0.5f)) * hD;
                R in_crystal = chord_box_open(xPs, yPs, zPs, xDs, yDs, zDs, x1c, y1c, z1c, x2c, y2c, z2c);
                R in_cell = chord_box_open(xPs, yPs, zPs, xDs, yDs, zDs, ux1, uy1, uz1, ux2, uy2, uz2);
                if (o.flags & 16) {  // self path in double
                    double a1 = chord_box_open<double>(to_double(xPs), to_double(yPs), to_double(zPs), to_double(xDs),
                                                       to_double(yDs), to_double(zDs), to_double(x1c), to_double(y1c),
                                                       to_double(z1c), to_double(x2c), to_double(y2c), to_double(z2c));
                    double a2 = chord_box_open<double>(to_double(xPs), to_double(yPs), to_double(zPs), to_double(xDs),
                                                       to_double(yDs), to_double(zDs), to_double(ux1), to_double(uy1),
                                                       to_double(uz1), to_double(ux2), to_double(uy2), to_double(uz2));
                    att += R((float)((a1 - a2) * to_double(mu_t)));
                } else {
                    att += (in_crystal - in_cell) * mu_t;
                }
                R eff = r_exp(-att);
                const R ext = R(1000.0f);
                R xt = xDs + ext * (xDs - xPs) / dist;
                R yt = yDs + ext * (yDs - yPs) / dist;
                R zt = zDs + ext * (zDs - zPs) / dist;
                R l_abs;
                if (o.flags & 4) {  // absorption chord in double from the FP32 end points
                    double d64 = std::sqrt(to_double(dsq));
                    double xds = to_double(xDs), yds = to_double(yDs), zds = to_double(zDs);
                    double xps = to_double(xPs), yps = to_double(yPs), zps = to_double(zPs);
                    l_abs = R((float)chord_box<double>(xps, yps, zps, xds + 1000.0 * (xds - xps) / d64,
                                                       yds + 1000.0 * (yds - yps) / d64, zds + 1000.0 * (zds - zps) / d64,
                                                       to_double(ux1), to_double(uy1), to_double(uz1), to_double(ux2),
                                                       to_double(uy2), to_double(uz2)));
                } else {
                    l_abs = chord_box(xPs, yPs, zPs, xt, yt, zt, ux1, uy1, uz1, ux2, uy2, uz2);
                }
                R absorb = (R(1.0f) - r_exp(-l_abs * mu_t)) * mu_pe / mu_t;
                if (acc_wide)
                    acc64 += to_double(eff * solid * absorb);
                else
                    acc += eff * solid * absorb;
            }
        }
        if (acc_wide) acc = R((float)(acc64 / (double)(nv * nv * nv)));
        else if (nv > 1) acc = acc / R((float)(nv * nv * nv));
        return acc;
    };
    if (o.flags & 64) {
        // Counting basis "attenuators that touch the element": a first, uncounted pass
        // finds the crystals and holes with a non-zero chord on at least one sub-ray; the
        // others add exactly zero, so dropping them leaves the value unchanged and the
        // counted pass measures only the work on attenuators the element needs.
        OpCounter::enabled = false;
        std::vector<char> md(cross_det.size(), 0);
        std::vector<std::vector<char>> mh(n_layers);
        for (int m = 0; m < n_layers; ++m) mh[m].assign(cross_hole[m].size(), 0);
        trace(cross_det, cross_hole, &md, &mh);
        OpCounter::enabled = counting;
        std::vector<int> kd;
        std::vector<std::vector<int>> kh(n_layers);
        for (size_t q = 0; q < cross_det.size(); ++q)
            if (md[q]) kd.push_back(cross_det[q]);
        for (int m = 0; m < n_layers; ++m)
            for (size_t q = 0; q < cross_hole[m].size(); ++q)
                if (mh[m][q]) kh[m].push_back(cross_hole[m][q]);
        if (st) {
            st->n_touch_det += (double)kd.size();
            for (auto& v : kh) st->n_touch_hole += (double)v.size();
        }
        return trace(kd, kh, nullptr, nullptr);
    }
    return trace(cross_det, cross_hole, nullptr, nullptr);
}

template <class R>
static void pe_run(const float* col, const float* det, const float* img, float* dst, const PeOpts& o,
                   double* row_stats) {
    const int X = ifloor(img[0]), Y = ifloor(img[1]), Zl = ifloor(img[2]);
    const long bins = (long)X * Y * Zl;
    const int rows = o.det_end - o.det_begin;
    int nth = o.threads > 0 ? o.threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int r = 0; r < rows; ++r) {
        PeStats st{0, 0, 0, 0, 0};
        OpCounter::flops = 0;
        OpCounter::enabled = true;
        for (long c = 0; c < bins; ++c) {
            R v = pe_element<R>(col, det, img, o.det_begin + r, (int)c, o, row_stats ? &st : nullptr);
            dst[(long)r * bins + c] = (float)to_double(v);
        }
        OpCounter::enabled = false;
        if (row_stats) {
            row_stats[5 * r + 0] = (double)OpCounter::flops;
            row_stats[5 * r + 1] = st.n_cross_det;
            row_stats[5 * r + 2] = st.n_cross_hole;
            row_stats[5 * r + 3] = st.n_touch_det;
            row_stats[5 * r + 4] = st.n_touch_hole;
        }
    }
}

}  // namespace oracle

using oracle::PeOpts;

extern "C" {

// dst: (det_end - det_begin) x (X*Y*img[2]) floats, det-major, for the slab
// described in-band by img[2] (slab Z), img[30] (full Z), img[31] (z0).
int or_pe_f32(const float* col, const float* det, const float* img, float* dst, const PeOpts* o) {
    oracle::pe_run<float>(col, det, img, dst, *o, nullptr);
    return 0;
}

// FP64 oracle; values rounded to float on output.  dst64 (optional) is unused
// here and kept for ABI symmetry with the scatter oracle.
int or_pe_f64(const float* col, const float* det, const float* img, float* dst, const PeOpts* o) {
    oracle::pe_run<double>(col, det, img, dst, *o, nullptr);
    return 0;
}

// Op-counting build: row_stats[5*r + {0..4}] = counted FLOPs of row r (all its
// elements), summed culled-crystal and culled-hole list lengths, and (flag 64)
// the summed counts of culled crystals / holes that touch their element.
int or_pe_count(const float* col, const float* det, const float* img, float* dst, const PeOpts* o,
                double* row_stats) {
    oracle::pe_run<oracle::Counted<float>>(col, det, img, dst, *o, row_stats);
    return 0;
}

float or_chord_box_f32(const float* a) {
    return oracle::chord_box<float>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11]);
}
double or_chord_box_f64(const double* a) {
    return oracle::chord_box<double>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11]);
}
float or_chord_box_open_f32(const float* a) {
    return oracle::chord_box_open<float>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11]);
}
double or_chord_box_open_f64(const double* a) {
    return oracle::chord_box_open<double>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11]);
}
float or_chord_hole_f32(const float* a) {
    return oracle::chord_hole<float>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);
}
double or_chord_hole_f64(const double* a) {
    return oracle::chord_hole<double>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10]);
}

}  // extern "C"
