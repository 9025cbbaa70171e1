// Device-side ray / solid primitives for gfx950.
//
// Semantics follow the reference device functions (file:line below); the
// arithmetic is restructured for the GPU:
//   * everything that depends only on the sub-ray (length, reciprocal
//     directions, the parallel-axis pattern) is computed once per sub-ray and
//     reused for every box it is tested against;
//   * reciprocals / square roots use the hardware v_rcp_f32 / v_sqrt_f32
//     (1 ulp) instead of the IEEE division / square-root sequences;
//   * paths inside the detecting crystal are measured from the sub-point Q
//     (which lies inside both the crystal and its sub-cell), so they are
//     sub-millimetre numbers instead of differences of ~100 mm distances.
//
//   chord_world / chord_special  <- length_box_ray        PESysMatGen.cu:19-151
//   open_q                       <- length_box_ray_inside PESysMatGen.cu:153-307
//   absorb_q                     <- length_box_ray on the 1000 mm extension, PESysMatGen.cu:749-757
//   chord_hole64                 <- length_cylinder_ray   PESysMatGen.cu:309-373 (in FP64)
#pragma once
#include <hip/hip_runtime.h>

namespace sysmat {

constexpr float kEps = 0.001f;  // ray degeneracy / minimum chord (PESysMatGen.cu:22)

__device__ inline float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// FP64 1/sqrt(x), x > 0, and 1/x: the hardware approximations with two Newton steps
// (~1 ulp, against the ~22 and ~10 instructions of the correctly rounded sqrt and
// division).  rcp_nr(+-0) is +-inf, as the division gives.
__device__ inline double rsq_nr(double x) {
    double r = __builtin_amdgcn_rsq(x);
    r = r * fma(-0.5 * x * r, r, 1.5);
    r = r * fma(-0.5 * x * r, r, 1.5);
    return r;
}
__device__ inline double rcp_nr(double x) {
    const double r0 = __builtin_amdgcn_rcp(x);
    double r = fma(r0, fma(-x, r0, 1.0), r0);
    r = fma(r, fma(-x, r, 1.0), r);
    return x != 0.0 ? r : r0;
}

__device__ inline bool inside_closed(float v, float a, float b) { return v >= a && v <= b; }

// A segment P -> Q prepared for repeated slab tests (world frame).
struct WRay {
    float px, py, pz;
    float dx, dy, dz;   // Q - P
    float ix, iy, iz;   // t / d
    float t;            // |Q - P|
    int pat;            // bit 0/1/2: |d| < eps on x/y/z
};

__device__ inline WRay make_wray(float px, float py, float pz, float qx, float qy, float qz) {
    WRay r;
    r.px = px; r.py = py; r.pz = pz;
    r.dx = qx - px; r.dy = qy - py; r.dz = qz - pz;
    r.t = fsqrt(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
    r.pat = (fabsf(r.dx) < kEps ? 1 : 0) | (fabsf(r.dy) < kEps ? 2 : 0) | (fabsf(r.dz) < kEps ? 4 : 0);
    r.ix = r.t * frcp(r.dx);
    r.iy = r.t * frcp(r.dy);
    r.iz = r.t * frcp(r.dz);
    return r;
}

// Branches of length_box_ray for rays parallel to one, two or three axes.
// Coordinates are relative to the box frame the caller chose.
__device__ inline float chord_special(float px, float py, float pz, float qx, float qy, float qz, float ix,
                                      float iy, float iz, float t, int pat, float x1, float y1, float z1, float x2,
                                      float y2, float z2) {
    if (pat == 7) return 0.0f;
    if (pat == 3) {  // along z
        return (inside_closed(px, x1, x2) && inside_closed(py, y1, y2) &&
                ((pz <= z1 && qz >= z2) || (qz <= z1 && pz >= z2))) ? fabsf(z2 - z1) : 0.0f;
    }
    if (pat == 6) {  // along x
        return (inside_closed(pz, z1, z2) && inside_closed(py, y1, y2) &&
                ((px <= x1 && qx >= x2) || (qx <= x1 && px >= x2))) ? fabsf(x2 - x1) : 0.0f;
    }
    if (pat == 5) {  // along y
        return (inside_closed(px, x1, x2) && inside_closed(pz, z1, z2) &&
                ((py <= y1 && qy >= y2) || (qy <= y1 && py >= y2))) ? fabsf(y2 - y1) : 0.0f;
    }
    // one parallel axis: strictly inside that axis' range, slabs on the other two
    if ((pat == 1 && (px >= x2 || px <= x1)) || (pat == 2 && (py >= y2 || py <= y1)) ||
        (pat == 4 && (pz >= z2 || pz <= z1)))
        return 0.0f;
    float lo = -INFINITY, hi = INFINITY, a, b;
    if (pat != 1) { a = (x1 - px) * ix; b = (x2 - px) * ix; lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b)); }
    if (pat != 2) { a = (y1 - py) * iy; b = (y2 - py) * iy; lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b)); }
    if (pat != 4) { a = (z1 - pz) * iz; b = (z2 - pz) * iz; lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b)); }
    const float c = hi - lo;
    return (c >= kEps && lo > kEps && hi < t) ? c : 0.0f;
}

// length_box_ray for box [x1,x2] x [y1,y2] x [z1,z2] in the ray's frame: the
// chord counts only when the segment enters and leaves strictly between its
// end points (tmax - tmin >= eps, tmin > eps, tmax < t).
__device__ inline float chord_world(const WRay& r, float x1, float x2, float y1, float y2, float z1, float z2) {
    float a = (x1 - r.px) * r.ix, b = (x2 - r.px) * r.ix;
    float lo = fminf(a, b), hi = fmaxf(a, b);
    a = (y1 - r.py) * r.iy; b = (y2 - r.py) * r.iy;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    a = (z1 - r.pz) * r.iz; b = (z2 - r.pz) * r.iz;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    const float c = hi - lo;
    return (c >= kEps && lo > kEps && hi < r.t) ? c : 0.0f;
}

// chord_world with the slab parameters measured from the far end point Q
// instead of P (s = t - |Q - P|): the same chord and conditions (s_min > eps - t,
// s_max < 0), but the rounding of the parameters scales with the box's distance
// from Q -- a few mm for the crystals in front of the detecting one -- rather
// than with its distance from the voxel.
__device__ inline float chord_world_q(const WRay& r, float qx, float qy, float qz, float x1, float x2, float y1,
                                      float y2, float z1, float z2) {
    float a = (x1 - qx) * r.ix, b = (x2 - qx) * r.ix;
    float lo = fminf(a, b), hi = fmaxf(a, b);
    a = (y1 - qy) * r.iy; b = (y2 - qy) * r.iy;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    a = (z1 - qz) * r.iz; b = (z2 - qz) * r.iz;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    const float c = hi - lo;
    return (c >= kEps && lo > kEps - r.t && hi < 0.0f) ? c : 0.0f;
}

// Sub-ray in the detecting crystal's frame, parameterised from Q (inside the
// crystal): s = (param from P) - |Q - P|.
struct QRay {
    float px, py, pz;   // P in the crystal frame
    float qx, qy, qz;   // Q in the crystal frame
    float ix, iy, iz;   // |u| / u, u = Q - P
    float t;            // |u|
    int pat;            // |u| < eps per axis
    int pat_e;          // same test on the 1000 mm extension (PESysMatGen.cu:749-752)
    float t_ext;        // length of the extension beyond Q
};

__device__ inline QRay make_qray(float px, float py, float pz, float qx, float qy, float qz, float dist_world) {
    QRay r;
    r.px = px; r.py = py; r.pz = pz;
    r.qx = qx; r.qy = qy; r.qz = qz;
    const float ux = qx - px, uy = qy - py, uz = qz - pz;
    r.t = fsqrt(ux * ux + uy * uy + uz * uz);
    r.pat = (fabsf(ux) < kEps ? 1 : 0) | (fabsf(uy) < kEps ? 2 : 0) | (fabsf(uz) < kEps ? 4 : 0);
    r.ix = r.t * frcp(ux);
    r.iy = r.t * frcp(uy);
    r.iz = r.t * frcp(uz);
    // extended end point E = Q + 1000 (Q - P) / dist, as the reference builds it
    const float k = 1000.0f * frcp(dist_world);
    const float ex = (qx + k * ux) - px, ey = (qy + k * uy) - py, ez = (qz + k * uz) - pz;
    r.pat_e = (fabsf(ex) < kEps ? 1 : 0) | (fabsf(ey) < kEps ? 2 : 0) | (fabsf(ez) < kEps ? 4 : 0);
    r.t_ext = k * r.t;
    return r;
}

// make_qray for a crystal whose frame is the world translated (rotation 0): the
// direction, its length and reciprocals are those of the world ray w (the same
// difference vector, rounded once more or less), so only the extension is new.
__device__ inline QRay make_qray_axis(const WRay& w, float px, float py, float pz, float qx, float qy, float qz) {
    QRay r;
    r.px = px; r.py = py; r.pz = pz;
    r.qx = qx; r.qy = qy; r.qz = qz;
    r.t = w.t;
    r.pat = w.pat;
    r.ix = w.ix;
    r.iy = w.iy;
    r.iz = w.iz;
    const float k = 1000.0f * frcp(w.t);
    const float ex = (qx + k * w.dx) - px, ey = (qy + k * w.dy) - py, ez = (qz + k * w.dz) - pz;
    r.pat_e = (fabsf(ex) < kEps ? 1 : 0) | (fabsf(ey) < kEps ? 2 : 0) | (fabsf(ez) < kEps ? 4 : 0);
    r.t_ext = k * r.t;
    return r;
}

// Slab interval (relative to Q) of the box on one axis.
__device__ inline void qslab(float q, float inv, float b1, float b2, float& lo, float& hi) {
    const float a = (b1 - q) * inv, b = (b2 - q) * inv;
    lo = fminf(a, b);
    hi = fmaxf(a, b);
}

// length_box_ray_inside for a box that contains Q: path of P -> Q inside it.
__device__ inline float open_q(const QRay& r, float x1, float x2, float y1, float y2, float z1, float z2) {
    if (r.pat & 2) return 0.0f;  // |dy| < 0.001 -> 0 (PESysMatGen.cu:156)
    if ((r.pat & 5) == 5) {      // along y
        if (inside_closed(r.px, x1, x2) && inside_closed(r.pz, z1, z2)) {
            if (r.py < r.qy) return r.qy - y1;
            if (r.py > r.qy) return y2 - r.qy;
        }
        return 0.0f;
    }
    if (((r.pat & 1) && (r.px >= x2 || r.px <= x1)) || ((r.pat & 4) && (r.pz >= z2 || r.pz <= z1))) return 0.0f;
    float sl, sh, a, b;
    qslab(r.qy, r.iy, y1, y2, sl, sh);
    if (!(r.pat & 1)) { qslab(r.qx, r.ix, x1, x2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b); }
    if (!(r.pat & 4)) { qslab(r.qz, r.iz, z1, z2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b); }
    // closing rules of the reference with lo = t + sl, hi = t + sh
    if ((sh - sl) < kEps) return 0.0f;
    if (sl >= 0.0f) return 0.0f;
    if (r.t + sh <= kEps) return 0.0f;
    const bool lo_in = (r.t + sl) > kEps;
    if (sh >= 0.0f && lo_in) return -sl;
    if (!lo_in && sh <= 0.0f) return r.t + sh;
    if (!lo_in && sh >= 0.0f) return r.t;
    return sh - sl;
}

// length_box_ray of P -> E (E = Q extended 1000 mm) through a box that
// contains Q, measured from Q.
__device__ inline float absorb_q(const QRay& r, float x1, float x2, float y1, float y2, float z1, float z2) {
    const int pat = r.pat_e;
    if (pat == 7) return 0.0f;
    if (pat == 3 || pat == 5 || pat == 6) {
        const float f = r.t_ext * frcp(r.t);  // E = Q + f (Q - P)
        return chord_special(r.px, r.py, r.pz, r.qx + f * (r.qx - r.px), r.qy + f * (r.qy - r.py),
                             r.qz + f * (r.qz - r.pz), 0.0f, 0.0f, 0.0f, 0.0f, pat, x1, y1, z1, x2, y2, z2);
    }
    if ((pat == 1 && (r.px >= x2 || r.px <= x1)) || (pat == 2 && (r.py >= y2 || r.py <= y1)) ||
        (pat == 4 && (r.pz >= z2 || r.pz <= z1)))
        return 0.0f;
    float sl = -INFINITY, sh = INFINITY, a, b;
    if (pat != 1) { qslab(r.qx, r.ix, x1, x2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b); }
    if (pat != 2) { qslab(r.qy, r.iy, y1, y2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b); }
    if (pat != 4) { qslab(r.qz, r.iz, z1, z2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b); }
    const float c = sh - sl;
    return (c >= kEps && r.t + sl > kEps && sh < r.t_ext) ? c : 0.0f;
}

// Closing rules of length_box_ray_inside for a box containing Q, from its slab
// interval [sl, sh] relative to Q (as in open_q).
__device__ inline float open_rules(float sl, float sh, float t) {
    const bool lo_in = (t + sl) > kEps;
    float v = sh - sl;
    v = (!lo_in && sh >= 0.0f) ? t : v;
    v = (!lo_in && sh <= 0.0f) ? t + sh : v;
    v = (sh >= 0.0f && lo_in) ? -sl : v;
    return ((sh - sl) < kEps || sl >= 0.0f || t + sh <= kEps) ? 0.0f : v;
}

// open_q for the crystal box (a1) and the sub-cell box (a2) and absorb_q for the
// sub-cell box, sharing the slab intervals: the same operations as those three
// functions, without branches when no component of P -> Q or of the extension is
// degenerate; otherwise the functions themselves.
__device__ inline void self_paths(const QRay& r, float x1, float x2, float y1, float y2, float z1, float z2,
                                  float u1, float u2, float v1, float v2, float w1, float w2, float& a1, float& a2,
                                  float& l_abs) {
    if (r.pat | r.pat_e) {
        a1 = open_q(r, x1, x2, y1, y2, z1, z2);
        a2 = open_q(r, u1, u2, v1, v2, w1, w2);
        l_abs = absorb_q(r, u1, u2, v1, v2, w1, w2);
        return;
    }
    float sl, sh, a, b;
    qslab(r.qy, r.iy, y1, y2, sl, sh);
    qslab(r.qx, r.ix, x1, x2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b);
    qslab(r.qz, r.iz, z1, z2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b);
    a1 = open_rules(sl, sh, r.t);
    qslab(r.qy, r.iy, v1, v2, sl, sh);
    qslab(r.qx, r.ix, u1, u2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b);
    qslab(r.qz, r.iz, w1, w2, a, b); sl = fmaxf(sl, a); sh = fminf(sh, b);
    a2 = open_rules(sl, sh, r.t);
    const float c = sh - sl;
    l_abs = (c >= kEps && r.t + sl > kEps && sh < r.t_ext) ? c : 0.0f;
}

// self_paths for the usual geometry: the crystal box centred at the frame origin
// (half extents h) and the sub-cell box centred at Q (half extents s).  Per axis
// the crystal's slab interval is m -+ h |1/u| with m = -q / u, and the sub-cell's
// is -+ s |1/u|, so the intervals need no per-axis min / max and the sub-cell's
// is one min of three products.  The box faces differ from the tabulated ones by
// FP32 rounding only; degenerate components take self_paths on those faces.
__device__ inline void self_paths_centred(const QRay& r, float hx, float hy, float hz, float sx, float sy, float sz,
                                          float u1, float u2, float v1, float v2, float w1, float w2, float& a1,
                                          float& a2, float& l_abs) {
    if (r.pat | r.pat_e) {
        self_paths(r, -hx, hx, -hy, hy, -hz, hz, u1, u2, v1, v2, w1, w2, a1, a2, l_abs);
        return;
    }
    const float ax = fabsf(r.ix), ay = fabsf(r.iy), az = fabsf(r.iz);
    const float mx = -r.qx * r.ix, my = -r.qy * r.iy, mz = -r.qz * r.iz;
    const float sl = fmaxf(fmaxf(fmaf(-hx, ax, mx), fmaf(-hy, ay, my)), fmaf(-hz, az, mz));
    const float sh = fminf(fminf(fmaf(hx, ax, mx), fmaf(hy, ay, my)), fmaf(hz, az, mz));
    a1 = open_rules(sl, sh, r.t);
    const float w = fminf(fminf(sx * ax, sy * ay), sz * az);
    a2 = open_rules(-w, w, r.t);
    const float c = w + w;
    l_abs = (c >= kEps && r.t - w > kEps && w < r.t_ext) ? c : 0.0f;
}

// Per-sub-ray FP64 data of the collimator path.  The end points themselves are
// evaluated in FP64: rays nearly parallel to a hole wall amplify end-point
// rounding (~4e-6 mm in FP32 at 50 mm) by cot(angle) ~ 30-60x into the chord.
struct Ray64 {
    double px, py, pz, t;
    double kx, ky, kz;     // unit direction
    double sx, sz, iky;    // kx/ky, kz/ky, 1/ky
    double kk, ikk;        // kx^2 + kz^2 and its reciprocal
};

__device__ inline Ray64 make_ray64(double px, double py, double pz, double qx, double qy, double qz) {
    Ray64 r;
    r.px = px; r.py = py; r.pz = pz;
    double dx = qx - px, dy = qy - py, dz = qz - pz;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double it = rsq_nr(d2);  // 1 / |Q - P| (P != Q: the voxel lies before the collimator)
    r.t = d2 * it;
    r.kx = dx * it; r.ky = dy * it; r.kz = dz * it;
    r.iky = rcp_nr(r.ky);
    r.sx = r.kx * r.iky;
    r.sz = r.kz * r.iky;
    r.kk = r.kx * r.kx + r.kz * r.kz;
    r.ikk = rcp_nr(r.kk);
    return r;
}

// length_cylinder_ray in FP64: the wall-entry / wall-exit differences of rays
// grazing a hole are taken without FP32 cancellation.
__device__ inline double chord_hole64(const Ray64& r, double xc, double y1, double y2, double zc, double rad) {
    if (fabs(y1 - y2) < (double)0.001f) return 0.0;
    const double tsmall = (double)0.0001f;
    double d1 = y1 - r.py, d2 = y2 - r.py;
    double xl = r.px + r.sx * d1, xr = r.px + r.sx * d2;
    double zl = r.pz + r.sz * d1, zr = r.pz + r.sz * d2;
    double tlo = d1 * r.iky, thi = d2 * r.iky;
    double rr = rad * rad;
    bool in_l = (xl - xc) * (xl - xc) + (zl - zc) * (zl - zc) <= rr;
    bool in_r = (xr - xc) * (xr - xc) + (zr - zc) * (zr - zc) <= rr;
    if (in_l && in_r) {
        if (tlo <= tsmall || thi <= tsmall) return 0.0;
        if (tlo >= r.t || thi >= r.t) return 0.0;
        return fabs(thi - tlo);
    }
    double ox = r.px - xc, oz = r.pz - zc;
    double cr = r.kx * oz - r.kz * ox;
    double disc = r.kk * rr - cr * cr;
    if (disc <= (double)0.00001f) return 0.0;
    double b = -(r.kx * ox + r.kz * oz);
    double sq = disc * rsq_nr(disc);  // disc > 1e-5
    double t1 = (b - sq) * r.ikk, t2 = (b + sq) * r.ikk;
    if (!in_l && !in_r) {
        if ((t1 >= tlo) && (t2 <= thi)) {
            if (t2 <= tsmall || t1 <= tsmall) return 0.0;
            if (t2 >= r.t || t1 >= r.t) return 0.0;
            return 2.0 * sq * r.ikk;
        }
        return 0.0;
    }
    if (in_l) {
        if (t2 <= tsmall || tlo <= tsmall) return 0.0;
        if (t2 >= r.t || tlo >= r.t) return 0.0;
        return t2 - tlo;
    }
    if (thi <= tsmall || t1 <= tsmall) return 0.0;
    if (thi >= r.t || t1 >= r.t) return 0.0;
    return thi - t1;
}

}  // namespace sysmat

namespace sysmat {

// length_box_ray in FP64 with the reference's full branch structure
// (PESysMatGen.cu:19-151 / scatter.cu:31-323), for per-pair scatter chords.
__host__ __device__ inline double slab2_64(double ai, double bi, double da, double db, double a1, double a2, double b1,
                                           double b2) {
    const double eps = (double)0.001f;
    const double t = sqrt(da * da + db * db);
    const double ia = t / da, ib = t / db;
    double lo, hi, lo2, hi2;
    if (ia < 0) { lo = (a2 - ai) * ia; hi = (a1 - ai) * ia; } else { hi = (a2 - ai) * ia; lo = (a1 - ai) * ia; }
    if (ib < 0) { lo2 = (b2 - bi) * ib; hi2 = (b1 - bi) * ib; } else { hi2 = (b2 - bi) * ib; lo2 = (b1 - bi) * ib; }
    if (lo > hi2 || lo2 > hi) return 0.0;
    if (lo2 > lo) lo = lo2;
    if (hi2 < hi) hi = hi2;
    if ((hi - lo) < eps) return 0.0;
    if (lo >= t || hi >= t) return 0.0;
    if (lo <= eps || hi <= eps) return 0.0;
    return hi - lo;
}

__host__ __device__ inline double chord_box64(double xi, double yi, double zi, double xo, double yo, double zo,
                                              double x1, double y1, double z1, double x2, double y2, double z2) {
    const double eps = (double)0.001f;
    const double dx = xo - xi, dy = yo - yi, dz = zo - zi;
    const bool px = fabs(dx) < eps, py = fabs(dy) < eps, pz = fabs(dz) < eps;
    if (px && py && pz) return 0.0;
    if (px && py)
        return (xi >= x1 && xi <= x2 && yi >= y1 && yi <= y2 && ((zi <= z1 && zo >= z2) || (zo <= z1 && zi >= z2)))
                   ? fabs(z2 - z1) : 0.0;
    if (pz && py)
        return (zi >= z1 && zi <= z2 && yi >= y1 && yi <= y2 && ((xi <= x1 && xo >= x2) || (xo <= x1 && xi >= x2)))
                   ? fabs(x2 - x1) : 0.0;
    if (px && pz)
        return (xi >= x1 && xi <= x2 && zi >= z1 && zi <= z2 && ((yi <= y1 && yo >= y2) || (yo <= y1 && yi >= y2)))
                   ? fabs(y2 - y1) : 0.0;
    if (px) return (xi >= x2 || xi <= x1) ? 0.0 : slab2_64(yi, zi, dy, dz, y1, y2, z1, z2);
    if (py) return (yi >= y2 || yi <= y1) ? 0.0 : slab2_64(xi, zi, dx, dz, x1, x2, z1, z2);
    if (pz) return (zi >= z2 || zi <= z1) ? 0.0 : slab2_64(xi, yi, dx, dy, x1, x2, y1, y2);
    const double t = sqrt(dx * dx + dy * dy + dz * dz);
    const double ix = t / dx, iy = t / dy, iz = t / dz;
    double lo, hi, lo2, hi2;
    if (ix < 0) { lo = (x2 - xi) * ix; hi = (x1 - xi) * ix; } else { hi = (x2 - xi) * ix; lo = (x1 - xi) * ix; }
    if (iy < 0) { lo2 = (y2 - yi) * iy; hi2 = (y1 - yi) * iy; } else { hi2 = (y2 - yi) * iy; lo2 = (y1 - yi) * iy; }
    if (lo > hi2 || lo2 > hi) return 0.0;
    if (lo2 > lo) lo = lo2;
    if (hi2 < hi) hi = hi2;
    if (iz < 0) { lo2 = (z2 - zi) * iz; hi2 = (z1 - zi) * iz; } else { hi2 = (z2 - zi) * iz; lo2 = (z1 - zi) * iz; }
    if (lo > hi2 || lo2 > hi) return 0.0;
    if (lo2 > lo) lo = lo2;
    if (hi2 < hi) hi = hi2;
    if ((hi - lo) < eps) return 0.0;
    if (lo >= t || hi >= t) return 0.0;
    if (lo <= eps || hi <= eps) return 0.0;
    return hi - lo;
}

}  // namespace sysmat
