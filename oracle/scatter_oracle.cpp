// CPU oracle for the primary-Compton scatter system matrices.  TEST
// INFRASTRUCTURE ONLY: the shipped library never links or calls this file.
//
// Restates, in the reference's operation order:
//   geometryRelationShip_Crystal2Crystal    scatter.cu:1440-1540 (bitmap builder)
//   geometryRelationShip_Collimator2Crystal scatter.cu:1939-2009 (bitmap builder)
//   crystalScatterSysMatCuda                scatter.cu:1077-1437 (= ScatterGen_Crystal/scatter.cu:1061-1421)
//   collimatorScatterSysMatCuda             scatter.cu:1544-1936 (virtual_r 0.1, projected disc)
//       and its ScatterGen_Collimator variant (ScatterGen_Collimator/scatter.cu:1056-1436,
//       virtual_r 0.2, solid angle (r^2 + dy r) / (4 L^2), :1212-1215)
//   the plate-cell collimator model (no reference kernel; cells_run below)
// One element (detector r, voxel v) sums the scatter sites (crystals k != r,
// or layer-0 holes) in ascending order; the reference's atomicAdd order is
// nondeterministic, so the FP32 sum order is a free choice.  The reference's
// O(N) bitmap scan per (r, v, site) is done here once per (r, site) into an
// attenuator list: the same set in the same order, independent of v.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "physics.hpp"

namespace oracle {

struct ScOpts {
    int det_begin, det_end;  // detector rows
    int threads;             // <= 0: all
    int variant;             // collimator: 0 combined (ScatterGen_RayTracing_CircularHole), 1 ScatterGen_Collimator
    const int* vox;          // optional voxel sample: rows hold these voxels only, and the crystal
    long long n_vox;         // scatter's PE input holds these columns only (N x n_vox)
};

static inline int nfloor(float v, float eps) { return (int)std::floor(v + eps); }

// ---- bitmap criteria and builders (FP32, as the reference kernels) ----------
// Crystal k attenuates the pair (i, j), i < j: its bounding sphere, seen from i
// or from j, overlaps the other crystal of the pair (scatter.cu:1440-1540).
static bool crystal_criterion(const float* det, int i, int j, int k) {
    const float* di = det + 12 * i;
    const float* dj = det + 12 * j;
    const float* dk = det + 12 * k;
    float Ri = std::sqrt(di[4] * di[4] + di[6] * di[6] + di[5] * di[5]) / 2.0f;
    float Rj = std::sqrt(dj[4] * dj[4] + dj[6] * dj[6] + dj[5] * dj[5]) / 2.0f;
    float Lij = distance<float>(di[1], di[2], di[3], dj[1], dj[2], dj[3]);
    float cij = Ri / Lij, cji = Rj / Lij;
    float uijx = (di[1] - dj[1]) / Lij, uijy = (di[2] - dj[2]) / Lij, uijz = (di[3] - dj[3]) / Lij;
    float ujix = (dj[1] - di[1]) / Lij, ujiy = (dj[2] - di[2]) / Lij, ujiz = (dj[3] - di[3]) / Lij;
    float Rk = std::sqrt(dk[4] * dk[4] + dk[6] * dk[6] + dk[5] * dk[5]) / 2.0f;
    float Lik = distance<float>(dk[1], dk[2], dk[3], di[1], di[2], di[3]);
    float cki = Rk / Lik;
    float Ljk = distance<float>(dk[1], dk[2], dk[3], dj[1], dj[2], dj[3]);
    float ckj = Rk / Ljk;
    float ukix = (dk[1] - di[1]) / Lik, ukiy = (dk[2] - di[2]) / Lik, ukiz = (dk[3] - di[3]) / Lik;
    float ukjx = (dk[1] - dj[1]) / Ljk, ukjy = (dk[2] - dj[2]) / Ljk, ukjz = (dk[3] - dj[3]) / Ljk;
    bool cross = distance<float>(ukix, ukiy, ukiz, ujix, ujiy, ujiz) <= cki + cji;
    if (!cross) cross = distance<float>(ukjx, ukjy, ukjz, uijx, uijy, uijz) <= ckj + cij;
    return cross;
}

// Crystal k attenuates the path hole h -> crystal j (scatter.cu:1939-2009).
static bool collimator_criterion(const float* col, const float* det, int h, int j, int k) {
    const float* hp = col + 100 + 9 * h;
    float xh = hp[0], yh = (hp[1] + hp[2]) / 2.0f, zh = hp[3];
    const float* dj = det + 12 * j;
    float Rj = std::sqrt(dj[4] * dj[4] + dj[6] * dj[6] + dj[5] * dj[5]) / 2.0f;
    float Lij = distance<float>(xh, yh, zh, dj[1], dj[2], dj[3]);
    float cji = Rj / Lij;
    float ujix = (dj[1] - xh) / Lij, ujiy = (dj[2] - yh) / Lij, ujiz = (dj[3] - zh) / Lij;
    const float* dk = det + 12 * k;
    float Rk = std::sqrt(dk[4] * dk[4] + dk[6] * dk[6] + dk[5] * dk[5]) / 2.0f;
    float Lik = distance<float>(dk[1], dk[2], dk[3], xh, yh, zh);
    float cki = Rk / Lik;
    float ukix = (dk[1] - xh) / Lik, ukiy = (dk[2] - yh) / Lik, ukiz = (dk[3] - zh) / Lik;
    return distance<float>(ukix, ukiy, ukiz, ujix, ujiy, ujiz) <= cki + cji;
}

static void build_crystal_bitmap(const float* det, uint32_t* words, long long n_words) {
    const int n = (int)det[0];
    std::memset(words, 0, sizeof(uint32_t) * (size_t)n_words);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            long long pair = (long long)i * n - ((long long)i * (i + 1)) / 2 + j;
            for (int k = 0; k < n; ++k) {
                if (k == i || k == j) continue;
                if (crystal_criterion(det, i, j, k)) {
                    long long bit = pair * n + k;
#pragma omp atomic
                    words[bit >> 5] |= (1u << (bit & 31));
                }
            }
        }
    }
}

static void build_collimator_bitmap(const float* col, const float* det, uint32_t* words, long long n_words) {
    const int n = (int)det[0];
    const int nh = (int)col[10];
    std::memset(words, 0, sizeof(uint32_t) * (size_t)n_words);
#pragma omp parallel for schedule(dynamic, 1)
    for (int h = 0; h < nh; ++h) {
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < n; ++k) {
                if (collimator_criterion(col, det, h, j, k)) {
                    long long bit = collimator_bit(h, j, k, n);
#pragma omp atomic
                    words[bit >> 5] |= (1u << (bit & 31));
                }
            }
        }
    }
}

// ---- shared per-site evaluation (scatter.cu:1232-1430 / 1728-1929) ----------
// Detector r's record and the physics set-up of its row.
template <class R>
struct Row {
    R xD, yD, zD, w, h, t, rot, mu_t, mu_pe, mu_c, res;
    R lo, hi, e0, kn_total;
    R cr, sr, crn, srn;  // cos/sin(rot), cos/sin(-rot)
};

template <class R>
static Row<R> make_row(const float* det, const float* img, const float* phy, int r) {
    Row<R> w;
    const float* d = det + 12 * r;
    const R fov2c = R(img[11]);
    w.xD = R(d[1]);
    w.yD = R(d[2]) + fov2c;
    w.zD = R(d[3]);
    w.w = R(d[4]);
    w.h = R(d[6]);
    w.t = R(d[5]);
    w.rot = R(d[11]);
    w.mu_t = R(d[7]);
    w.mu_pe = R(d[8]);
    w.mu_c = R(d[9]);
    w.res = R(d[10]);
    w.e0 = R(phy[7]);
    if (nfloor(phy[4], 0.5f) > 0) {
        w.lo = R(phy[5]);
        w.hi = R(phy[6]);
    } else {
        w.lo = (R(1.0f) - w.res / R(2.0f)) * w.e0;
        w.hi = (R(1.0f) + w.res / R(2.0f)) * w.e0;
    }
    // normalisation integral over [0, pi] with pi rounded to the argument type
    w.kn_total = kn_riemann(w.e0, R(0.0f), R((float)M_PI), R(0.01f));
    w.cr = rot_cos(w.rot);
    w.sr = rot_sin(w.rot);
    w.crn = rot_cos(-w.rot);
    w.srn = rot_sin(-w.rot);
    return w;
}

// Contribution of one scatter site S to element (r, image point I); `atten`
// lists the attenuating crystals of (site, r); `weight` is the site's source
// term (prob_Compton_*).  Returns the product in the reference's mixed precision.
template <class R>
static R site_term(const Row<R>& D, const float* det, const float* img, R xI, R yI, R zI, R xS, R yS, R zS,
                   R weight, const std::vector<int>& atten, R plate_pe = R(0.0f), R plate_c = R(0.0f)) {
    using W = typename Wide<R>::type;
    const R fov2c = R(img[11]);
    // single sub-cell (divide 1x1x1): the sub-point is the crystal centre
    const R xDs = -D.w / R(2.0f) + (R(0.0f) + R(0.5f)) / R(1.0f) * D.w;
    const R zDs = -D.h / R(2.0f) + (R(0.0f) + R(0.5f)) / R(1.0f) * D.h;
    const R yDs = -D.t / R(2.0f) + (R(0.0f) + R(0.5f)) / R(1.0f) * D.t;
    const R xQ = D.xD + (xDs * D.cr - zDs * D.sr);
    const R zQ = D.zD + (xDs * D.sr + zDs * D.cr);
    const R yQ = D.yD + yDs;

    const R theta = cone_angle(xI, yI, zI, xS, yS, zS, xQ, yQ, zQ);
    const R es = scatter_energy(theta, D.e0);
    const R res_s = D.res * r_sqrt(es / D.e0);
    if (((es * (R(1.0f) + R(2.0f) * res_s / R(2.35482f))) <= D.lo) ||
        (es * (R(1.0f) - R(2.0f) * res_s / R(2.35482f)) >= D.hi))
        return R(0.0f);
    const R p_e = window_probability(es, res_s, D.lo, D.hi);
    const R L = distance(xS, yS, zS, xQ, yQ, zQ);
    const R Rdet = r_sqrt(D.w * D.w / R(1.0f) / R(1.0f) + D.h * D.h / R(1.0f) / R(1.0f) +
                          D.t * D.t / R(1.0f) / R(1.0f)) / R(2.0f);
    R phi;
    if (L * r_sin(theta) * R(2.0f) <= Rdet)
        phi = R(W(R(2.0f)) * W(M_PI));
    else
        phi = R(4.0f) * r_asin(rmin(Rdet / R(2.0f) / L / r_sin(theta), R(1.0f)));
    // theta range over the sub-cell corners, in the detector frame
    const R xSs = (xS - D.xD) * D.crn - (zS - D.zD) * D.srn;
    const R ySs = yS - D.yD;
    const R zSs = (xS - D.xD) * D.srn + (zS - D.zD) * D.crn;
    const R xIs = (xI - D.xD) * D.crn - (zI - D.zD) * D.srn;
    const R yIs = yI - D.yD;
    const R zIs = (xI - D.xD) * D.srn + (zI - D.zD) * D.crn;
    const R ux1 = (R(0.0f) / R(1.0f) - R(0.5f)) * D.w, ux2 = ((R(0.0f) + R(1.0f)) / R(1.0f) - R(0.5f)) * D.w;
    const R uy1 = (R(0.0f) / R(1.0f) - R(0.5f)) * D.t, uy2 = ((R(0.0f) + R(1.0f)) / R(1.0f) - R(0.5f)) * D.t;
    const R uz1 = (R(0.0f) / R(1.0f) - R(0.5f)) * D.h, uz2 = ((R(0.0f) + R(1.0f)) / R(1.0f) - R(0.5f)) * D.h;
    const R ext = R(1000.0f);
    const R dIS = distance(xSs, ySs, zSs, xIs, yIs, zIs);
    R xt = xSs + ext * (xSs - xIs) / dIS;
    R yt = ySs + ext * (ySs - yIs) / dIS;
    R zt = zSs + ext * (zSs - zIs) / dIS;
    const R through = chord_box(xIs, yIs, zIs, xt, yt, zt, ux1, uy1, uz1, ux2, uy2, uz2);
    R th[8];
    th[0] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux1, uy1, uz1);
    th[1] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux2, uy1, uz1);
    th[2] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux1, uy2, uz1);
    th[3] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux1, uy1, uz2);
    th[4] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux2, uy2, uz1);
    th[5] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux2, uy1, uz2);
    th[6] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux1, uy2, uz2);
    th[7] = cone_angle(xIs, yIs, zIs, xSs, ySs, zSs, ux2, uy2, uz2);
    R tmin = th[0], tmax = th[0];
    for (int i = 1; i < 8; ++i) {
        if (th[i] > tmax) tmax = th[i];
        if (th[i] < tmin) tmin = th[i];
    }
    if (through > R(0.001f)) tmin = R(0.0f);
    const R ratio = kn_riemann(D.e0, tmin, tmax, R(0.01f)) / D.kn_total;
    // attenuation site -> detector through the listed crystals, mu scaled to E'
    R att = R(0.0f);
    for (int k : atten) {
        const float* dk = det + 12 * k;
        R xk = R(dk[1]), yk = R(dk[2]) + fov2c, zk = R(dk[3]);
        R wk = R(dk[4]), hk = R(dk[6]), tk = R(dk[5]), rk = R(dk[11]);
        R pe = R(dk[8]), cc = R(dk[9]);
        pe = pe * D.e0 * D.e0 * D.e0 / es / es / es;
        cc = cc * D.e0 / es;
        R tot = pe + cc;
        R ck = rot_cos(-rk), sk = rot_sin(-rk);
        R xSk = (xS - xk) * ck - (zS - zk) * sk, ySk = yS - yk, zSk = (xS - xk) * sk + (zS - zk) * ck;
        R xQk = (xQ - xk) * ck - (zQ - zk) * sk, yQk = yQ - yk, zQk = (xQ - xk) * sk + (zQ - zk) * ck;
        R c = chord_box(xSk, ySk, zSk, xQk, yQk, zQk, R(-0.5f) * wk, R(-0.5f) * tk, R(-0.5f) * hk, R(0.5f) * wk,
                        R(0.5f) * tk, R(0.5f) * hk);
        att = att + c * tot;
    }
    // collimator plate path of a cell site (zero for crystal and hole sites), mu scaled to E'
    att = att + plate_pe * D.e0 * D.e0 * D.e0 / es / es / es + plate_c * D.e0 / es;
    // inside the detecting crystal, before the (single) sub-cell
    const R a1 = chord_box_open(xSs, ySs, zSs, xDs, yDs, zDs, -D.w / R(2.0f), -D.t / R(2.0f), -D.h / R(2.0f),
                                D.w / R(2.0f), D.t / R(2.0f), D.h / R(2.0f));
    const R a2 = chord_box_open(xSs, ySs, zSs, xDs, yDs, zDs, ux1, uy1, uz1, ux2, uy2, uz2);
    const R pe_abs = D.mu_pe * D.e0 * D.e0 * D.e0 / es / es / es;
    const R c_abs = D.mu_c * D.e0 / es;
    const R tot_abs = pe_abs + c_abs;
    att = att + (a1 - a2) * tot_abs;
    xt = xDs + ext * (xDs - xSs) / L;
    yt = yDs + ext * (yDs - ySs) / L;
    zt = zDs + ext * (zDs - zSs) / L;
    const R l_abs = chord_box(xSs, ySs, zSs, xt, yt, zt, ux1, uy1, uz1, ux2, uy2, uz2);
    // product in the reference's evaluation order (M_PI promotes from the 3rd factor on)
    W v = W(weight * phi / R(2.0f)) / W(M_PI);
    v = v * W(ratio);
    v = v * W(p_e);
    v = v * W(r_exp(-att));
    v = v * W(R(1.0f) - r_exp(-l_abs * tot_abs));
    v = v * W(pe_abs);
    v = v / W(tot_abs);
    return R(v);
}

template <class R>
static void voxel_point(const float* img, int vox, bool eps_counts, R& x, R& y, R& z) {
    const float e = eps_counts ? 0.001f : 0.0f;
    const int X = nfloor(img[0], e), Y = nfloor(img[1], e), Z = nfloor(img[2], e);
    const int iz = vox / (Y * X);
    const int rem = vox % (Y * X);
    const int iy = rem / X, ix = rem % X;
    R xi = (R((float)ix) - R((float)X) / R(2.0f) + R(0.5f)) * R(img[3]);
    R yi = (R((float)iy) - R((float)Y) / R(2.0f) + R(0.5f)) * R(img[4]);
    R zi = (R((float)iz) - R((float)Z) / R(2.0f) + R(0.5f)) * R(img[5]);
    xi = xi + R(img[8]);
    yi = yi + R(img[9]);
    zi = zi + R(img[10]);
    const R a = R(img[20]) * R(img[7]);
    const R ca = rot_cos(a), sa = rot_sin(a);
    x = xi * ca - yi * sa;
    y = xi * sa + yi * ca;
    z = zi;
}

// crystalScatterSysMatCuda over rows [det_begin, det_end) x all voxels (or the
// voxel sample o.vox, with pe holding those columns only).  bits == nullptr:
// the GeometryRelationShip criterion is evaluated per (pair, crystal) instead of
// read from the bitmap (the same FP32 test, so the same attenuator lists).
template <class R>
static void crystal_run(const float* det, const float* img, const float* phy, const float* pe, const uint32_t* bits,
                        float* dst, const ScOpts& o) {
    const int n = nfloor(det[0], 0.0001f);
    const int X = nfloor(img[0], 0.0f), Y = nfloor(img[1], 0.0f), Z = nfloor(img[2], 0.0f);
    const long V = (long)X * Y * Z;
    const int rows = o.det_end - o.det_begin;
    const int nth = o.threads > 0 ? o.threads : omp_get_max_threads();
    const bool use = nfloor(phy[0], 0.5f) == 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int rr = 0; rr < rows; ++rr) {
        const int r = o.det_begin + rr;
        Row<R> D = make_row<R>(det, img, phy, r);
        std::vector<std::vector<int>> lists(n);
        for (int k = 0; k < n; ++k) {
            if (k == r) continue;
            for (int a = 0; a < n; ++a) {
                if (a == r || a == k) continue;
                const bool on = bits ? test_bit(bits, crystal_pair_bit(k, r, a, n)) != 0
                                     : crystal_criterion(det, std::min(k, r), std::max(k, r), a);
                if (on) lists[k].push_back(a);
            }
        }
        const long nv = o.vox ? (long)o.n_vox : V;
        for (long j = 0; j < nv; ++j) {
            const long v = o.vox ? (long)o.vox[j] : j;
            R acc = R(0.0f);
            if (use && D.mu_t > R(0.01f)) {
                R xI, yI, zI;
                voxel_point<R>(img, (int)v, false, xI, yI, zI);
                for (int k = 0; k < n; ++k) {
                    if (k == r) continue;
                    const float* dk = det + 12 * k;
                    // mu_c / mu_pe of the detecting crystal (scatter.cu:1218-1219)
                    const R w = R(pe[(long)k * nv + j]) * D.mu_c / D.mu_pe;
                    acc += site_term<R>(D, det, img, xI, yI, zI, R(dk[1]), R(dk[2]) + R(img[11]), R(dk[3]), w,
                                        lists[k]);
                }
            }
            dst[(long)rr * nv + j] = (float)to_double(acc);
        }
    }
}

// collimatorScatterSysMatCuda (variant 0) / ScatterGen_Collimator (variant 1);
// voxel sample and bits == nullptr as crystal_run.
template <class R>
static void collimator_run(const float* col, const float* det, const float* img, const float* phy,
                           const uint32_t* bits, float* dst, const ScOpts& o) {
    using W = typename Wide<R>::type;
    const int n = nfloor(det[0], 0.0001f);
    const int nh = nfloor(col[10], 0.000001f);
    const bool eps_counts = (o.variant == 0);
    const float e = eps_counts ? 0.001f : 0.0f;
    const int X = nfloor(img[0], e), Y = nfloor(img[1], e), Z = nfloor(img[2], e);
    const long V = (long)X * Y * Z;
    const int rows = o.det_end - o.det_begin;
    const int nth = o.threads > 0 ? o.threads : omp_get_max_threads();
    const bool use = nfloor(phy[0], 0.5f) == 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int rr = 0; rr < rows; ++rr) {
        const int r = o.det_begin + rr;
        Row<R> D = make_row<R>(det, img, phy, r);
        std::vector<std::vector<int>> lists(nh);
        for (int h = 0; h < nh; ++h)
            for (int a = 0; a < n; ++a) {
                if (a == r) continue;
                const bool on = bits ? test_bit(bits, collimator_bit(h, r, a, n)) != 0
                                     : collimator_criterion(col, det, h, r, a);
                if (on) lists[h].push_back(a);
            }
        const long nv = o.vox ? (long)o.n_vox : V;
        for (long j = 0; j < nv; ++j) {
            const long v = o.vox ? (long)o.vox[j] : j;
            R acc = R(0.0f);
            if (use && D.mu_t > R(0.01f)) {
                R xI, yI, zI;
                voxel_point<R>(img, (int)v, eps_counts, xI, yI, zI);
                for (int h = 0; h < nh; ++h) {
                    const float* hp = col + 100 + 9 * h;
                    const R xS = R(hp[0]);
                    const R yS = (R(hp[1]) + R(hp[2])) / R(2.0f) + R(img[11]);
                    const R zS = R(hp[3]);
                    const R L = distance(xI, yI, zI, xS, yS, zS);
                    R solid;
                    if (o.variant == 0) {
                        const R vr = R(0.1f);
                        W num = W(vr * vr) * W(M_PI) + W((R(hp[2]) - R(hp[1])) * vr);
                        W den = W(R(4.0f)) * W(M_PI) * W(L) * W(L);
                        solid = R(num / den * W(r_fabs(yS - yI)) / W(L));
                    } else {
                        const R vr = R(0.2f);
                        solid = (vr * vr + (R(hp[2]) - R(hp[1])) * vr) / (R(4.0f) * L * L);
                    }
                    const R w = solid / R(col[15]) * R(col[17]);
                    acc += site_term<R>(D, det, img, xI, yI, zI, xS, yS, zS, w, lists[h]);
                }
            }
            dst[(long)rr * nv + j] = (float)to_double(acc);
        }
    }
}

// ---- collimator plate cells (a model of this build, no reference kernel) ---
// Plates, holes and cells from the collimator record; cells in (y, z, x) order.
// A cell of plate m is a site when its centre is outside every hole of m and,
// for keep >= 0, within keep mm (x-z) of one of m's hole walls.
struct OPlate {
    double x1, x2, y1, y2, z1, z2, mt, mpe, mc;
    int h0, h1;
};
struct OHole {
    double x, z, r, ya, yb, mt, mpe, mc;
};
struct OCell {
    double x, y, z, w;
};

static void read_collimator(const float* col, double fov2c, std::vector<OPlate>& P, std::vector<OHole>& H) {
    const int layers = (int)std::floor(col[0] + 0.000001f);
    int k = 0;
    for (int m = 0; m < layers; ++m) {
        const float* L = col + 10 * (m + 1);
        OPlate p;
        p.x1 = -0.5 * (double)L[1];
        p.x2 = 0.5 * (double)L[1];
        p.y1 = fov2c + (double)L[4] - 0.5 * (double)L[2];
        p.y2 = fov2c + (double)L[4] + 0.5 * (double)L[2];
        p.z1 = -0.5 * (double)L[3];
        p.z2 = 0.5 * (double)L[3];
        p.mt = L[5];
        p.mpe = L[6];
        p.mc = L[7];
        p.h0 = k;
        p.h1 = k + (int)std::floor(L[0]);
        for (; k < p.h1; ++k) {
            const float* q = col + 100 + 9 * k;
            OHole o;
            o.x = q[0];
            o.z = q[3];
            o.r = q[4];
            o.ya = std::fmin((double)q[1] + fov2c, (double)q[2] + fov2c);
            o.yb = std::fmax((double)q[1] + fov2c, (double)q[2] + fov2c);
            o.mt = q[5];
            o.mpe = q[6];
            o.mc = q[7];
            H.push_back(o);
        }
        P.push_back(p);
    }
}

static std::vector<OCell> make_cells(const std::vector<OPlate>& P, const std::vector<OHole>& H, double h,
                                     double keep) {
    std::vector<OCell> out;
    for (const OPlate& p : P) {
        if (!(p.mc > 0.0)) continue;
        const double sx = p.x2 - p.x1, sy = p.y2 - p.y1, sz = p.z2 - p.z1;
        const long nx = std::max(1L, (long)std::llround(sx / h)), ny = std::max(1L, (long)std::llround(sy / h));
        const long nz = std::max(1L, (long)std::llround(sz / h));
        const double cx = sx / (double)nx, cy = sy / (double)ny, cz = sz / (double)nz;
        const double w = p.mc * cx * cy * cz / (4.0 * M_PI);
        for (long j = 0; j < ny; ++j)
            for (long k = 0; k < nz; ++k)
                for (long i = 0; i < nx; ++i) {
                    const OCell c{p.x1 + (i + 0.5) * cx, p.y1 + (j + 0.5) * cy, p.z1 + (k + 0.5) * cz, w};
                    bool air = false, near = keep < 0.0;
                    for (int q = p.h0; q < p.h1; ++q) {
                        const double rho = std::sqrt((c.x - H[q].x) * (c.x - H[q].x) + (c.z - H[q].z) * (c.z - H[q].z));
                        if (rho < H[q].r && c.y >= H[q].ya && c.y <= H[q].yb) air = true;
                        if (rho - H[q].r <= keep) near = true;
                    }
                    if (!air && near) out.push_back(c);
                }
    }
    return out;
}

// mu-weighted path lengths (total, photoelectric, Compton) of the segment A -> B
// through the plates: every crossing of a plate face, hole wall or hole cap cuts
// the segment; each piece takes the material of its midpoint.
static void plate_path(const std::vector<OPlate>& P, const std::vector<OHole>& H, const double* a, const double* b,
                       double* out) {
    const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    std::vector<double> t{0.0, 1.0};
    auto plane = [&](int ax, double v) {
        if (d[ax] != 0.0) {
            const double u = (v - a[ax]) / d[ax];
            if (u > 0.0 && u < 1.0) t.push_back(u);
        }
    };
    for (const OPlate& p : P) {
        plane(0, p.x1); plane(0, p.x2); plane(1, p.y1); plane(1, p.y2); plane(2, p.z1); plane(2, p.z2);
    }
    for (const OHole& o : H) {
        plane(1, o.ya);
        plane(1, o.yb);
        const double qa = d[0] * d[0] + d[2] * d[2];
        if (qa > 0.0) {
            const double ex = a[0] - o.x, ez = a[2] - o.z;
            const double qb = 2.0 * (ex * d[0] + ez * d[2]), qc = ex * ex + ez * ez - o.r * o.r;
            const double disc = qb * qb - 4.0 * qa * qc;
            if (disc > 0.0) {
                for (double sgn : {-1.0, 1.0}) {
                    const double u = (-qb + sgn * std::sqrt(disc)) / (2.0 * qa);
                    if (u > 0.0 && u < 1.0) t.push_back(u);
                }
            }
        }
    }
    std::sort(t.begin(), t.end());
    out[0] = out[1] = out[2] = 0.0;
    for (size_t i = 0; i + 1 < t.size(); ++i) {
        const double dl = (t[i + 1] - t[i]) * len;
        if (!(dl > 0.0)) continue;
        const double u = 0.5 * (t[i] + t[i + 1]);
        const double x = a[0] + u * d[0], y = a[1] + u * d[1], z = a[2] + u * d[2];
        for (const OPlate& p : P) {
            if (x < p.x1 || x > p.x2 || y < p.y1 || y > p.y2 || z < p.z1 || z > p.z2) continue;
            double mt = p.mt, mpe = p.mpe, mc = p.mc;
            for (int q = p.h0; q < p.h1; ++q) {
                const OHole& o = H[q];
                if ((x - o.x) * (x - o.x) + (z - o.z) * (z - o.z) < o.r * o.r && y >= o.ya && y <= o.yb) {
                    mt = o.mt;
                    mpe = o.mpe;
                    mc = o.mc;
                    break;
                }
            }
            out[0] += mt * dl;
            out[1] += mpe * dl;
            out[2] += mc * dl;
        }
    }
}

// Cell-model collimator scatter of rows [det_begin, det_end): sites = cells,
// source weight mu_c vol / (4 pi d^2) exp(-plate path image -> cell), plate path
// cell -> detector centre added to the attenuation at E', crystals attenuating
// (cell, r) by the bounding-sphere criterion of scatter.cu:1960-2000 seen from
// the cell (FP64); the rest is site_term.
template <class R>
static void cells_run(const float* col, const float* det, const float* img, const float* phy, double h, double keep,
                      float* dst, const ScOpts& o) {
    const int n = nfloor(det[0], 0.0001f);
    const int X = nfloor(img[0], 0.001f), Y = nfloor(img[1], 0.001f), Z = nfloor(img[2], 0.001f);
    const long V = (long)X * Y * Z;
    const double fov2c = img[11];
    std::vector<OPlate> P;
    std::vector<OHole> H;
    read_collimator(col, fov2c, P, H);
    const std::vector<OCell> cells = make_cells(P, H, h, keep);
    const int nc = (int)cells.size();
    const int rows = o.det_end - o.det_begin;
    const int nth = o.threads > 0 ? o.threads : omp_get_max_threads();
    const bool use = nfloor(phy[0], 0.5f) == 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
    for (int rr = 0; rr < rows; ++rr) {
        const int r = o.det_begin + rr;
        Row<R> D = make_row<R>(det, img, phy, r);
        const float* dr = det + 12 * r;
        const double Q[3] = {(double)dr[1], (double)dr[2] + fov2c, (double)dr[3]};
        const double Rr = std::sqrt((double)dr[4] * dr[4] + (double)dr[6] * dr[6] + (double)dr[5] * dr[5]) / 2.0;
        std::vector<std::vector<int>> lists(nc);
        std::vector<double> out_pe(nc), out_c(nc);
        for (int c = 0; c < nc; ++c) {
            const double S[3] = {cells[c].x, cells[c].y, cells[c].z};
            double pp[3];
            plate_path(P, H, S, Q, pp);
            out_pe[c] = pp[1];
            out_c[c] = pp[2];
            const double L = std::sqrt((Q[0] - S[0]) * (Q[0] - S[0]) + (Q[1] - S[1]) * (Q[1] - S[1]) +
                                       (Q[2] - S[2]) * (Q[2] - S[2]));
            for (int a = 0; a < n; ++a) {
                if (a == r) continue;
                const float* da = det + 12 * a;
                const double K[3] = {(double)da[1], (double)da[2] + fov2c, (double)da[3]};
                const double lk = std::sqrt((K[0] - S[0]) * (K[0] - S[0]) + (K[1] - S[1]) * (K[1] - S[1]) +
                                            (K[2] - S[2]) * (K[2] - S[2]));
                const double Rk = std::sqrt((double)da[4] * da[4] + (double)da[5] * da[5] + (double)da[6] * da[6]) / 2.0;
                double e2 = 0.0;
                for (int q = 0; q < 3; ++q) {
                    const double e = (K[q] - S[q]) / lk - (Q[q] - S[q]) / L;
                    e2 += e * e;
                }
                if (std::sqrt(e2) <= Rk / lk + Rr / L) lists[c].push_back(a);
            }
        }
        for (long v = 0; v < V; ++v) {
            R acc = R(0.0f);
            if (use && D.mu_t > R(0.01f)) {
                double xI, yI, zI;
                voxel_point<double>(img, (int)v, true, xI, yI, zI);
                const double I[3] = {xI, yI, zI};
                for (int c = 0; c < nc; ++c) {
                    const double S[3] = {cells[c].x, cells[c].y, cells[c].z};
                    double pin[3];
                    plate_path(P, H, I, S, pin);
                    const double d2 = (S[0] - xI) * (S[0] - xI) + (S[1] - yI) * (S[1] - yI) + (S[2] - zI) * (S[2] - zI);
                    const double w = cells[c].w / d2 * std::exp(-pin[0]);
                    acc += site_term<R>(D, det, img, R(xI), R(yI), R(zI), R(S[0]), R(S[1]), R(S[2]), R(w), lists[c],
                                        R(out_pe[c]), R(out_c[c]));
                }
            }
            dst[(long)rr * V + v] = (float)to_double(acc);
        }
    }
}

}  // namespace oracle

using oracle::ScOpts;

extern "C" {

long long or_crystal_bitmap_words(int n) {
    long long pairs = (long long)n * (n + 1) / 2;
    return (pairs * n + 31) / 32;
}

long long or_collimator_bitmap_words(int n_holes, int n) { return ((long long)n * n * n_holes + 31) / 32; }

int or_bitmap_crystal(const float* det, uint32_t* words, long long n_words) {
    oracle::build_crystal_bitmap(det, words, n_words);
    return 0;
}

// The crystal-bitmap bits of listed pairs (i[p] < j[p]): out[p * n + k] = bit
// (pair(i, j), k) of the GeometryRelationShip_Crystal2Crystal layout.
int or_crystal_bits_for_pairs(const float* det, const int* pi, const int* pj, int n_pairs, unsigned char* out) {
    const int n = (int)det[0];
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < n_pairs; ++p)
        for (int k = 0; k < n; ++k)
            out[(long long)p * n + k] =
                (k != pi[p] && k != pj[p] && oracle::crystal_criterion(det, pi[p], pj[p], k)) ? 1 : 0;
    return 0;
}

int or_bitmap_collimator(const float* col, const float* det, uint32_t* words, long long n_words) {
    oracle::build_collimator_bitmap(col, det, words, n_words);
    return 0;
}

int or_scatter_crystal(const float* det, const float* img, const float* phy, const float* pe, const uint32_t* bits,
                       float* dst, const ScOpts* o, int f64) {
    if (f64)
        oracle::crystal_run<double>(det, img, phy, pe, bits, dst, *o);
    else
        oracle::crystal_run<float>(det, img, phy, pe, bits, dst, *o);
    return 0;
}

int or_scatter_collimator(const float* col, const float* det, const float* img, const float* phy,
                          const uint32_t* bits, float* dst, const ScOpts* o, int f64) {
    if (f64)
        oracle::collimator_run<double>(col, det, img, phy, bits, dst, *o);
    else
        oracle::collimator_run<float>(col, det, img, phy, bits, dst, *o);
    return 0;
}

long long or_cells(const float* col, float fov2c, double h, double keep, double* out, long long cap) {
    std::vector<oracle::OPlate> P;
    std::vector<oracle::OHole> H;
    oracle::read_collimator(col, fov2c, P, H);
    const auto cells = oracle::make_cells(P, H, h, keep);
    for (long long i = 0; i < (long long)cells.size() && i < cap; ++i) {
        out[4 * i + 0] = cells[i].x;
        out[4 * i + 1] = cells[i].y;
        out[4 * i + 2] = cells[i].z;
        out[4 * i + 3] = cells[i].w;
    }
    return (long long)cells.size();
}

int or_plate_path_f64(const float* col, float fov2c, const double* ab, double* out) {
    std::vector<oracle::OPlate> P;
    std::vector<oracle::OHole> H;
    oracle::read_collimator(col, fov2c, P, H);
    oracle::plate_path(P, H, ab, ab + 3, out);
    return 0;
}

int or_scatter_cells(const float* col, const float* det, const float* img, const float* phy, double h, double keep,
                     float* dst, const ScOpts* o, int f64) {
    if (f64)
        oracle::cells_run<double>(col, det, img, phy, h, keep, dst, *o);
    else
        oracle::cells_run<float>(col, det, img, phy, h, keep, dst, *o);
    return 0;
}

float or_kn_riemann_f32(float e0, float lo, float hi) { return oracle::kn_riemann<float>(e0, lo, hi, 0.01f); }
double or_kn_riemann_f64(double e0, double lo, double hi) {
    return oracle::kn_riemann<double>(e0, lo, hi, (double)0.01f);
}
float or_scatter_energy_f32(float theta, float e0) { return oracle::scatter_energy<float>(theta, e0); }
float or_window_probability_f32(float e, float res, float lo, float hi) {
    return oracle::window_probability<float>(e, res, lo, hi);
}
float or_cone_angle_f32(const float* a) {
    return oracle::cone_angle<float>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8]);
}

}  // extern "C"
