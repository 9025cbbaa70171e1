// Device-side ray / solid primitives for gfx950.
//
// Semantics follow the reference device functions (file:line below); the
// arithmetic is restructured for the GPU: everything that depends only on the
// sub-ray (length, reciprocal directions, p * inv) is computed once per
// sub-ray and reused for every box, so a slab test costs six FMAs plus
// min/max instead of a square root and three divisions.
//
//   chord_through  <- length_box_ray        PESysMatGen.cu:19-151
//   chord_open     <- length_box_ray_inside PESysMatGen.cu:153-307
//   chord_hole64   <- length_cylinder_ray   PESysMatGen.cu:309-373 (in FP64)
#pragma once
#include <hip/hip_runtime.h>

namespace sysmat {

constexpr float kEps = 0.001f;  // ray degeneracy / minimum chord (PESysMatGen.cu:22)

// A segment P -> Q prepared for repeated slab tests.
struct Ray {
    float px, py, pz;
    float dx, dy, dz;   // Q - P
    float t;            // |Q - P|
    float ix, iy, iz;   // t / d per axis (only meaningful on non-parallel axes)
    float qx, qy, qz;   // end point (for the along-axis special cases)
    int pat;            // bit 0/1/2: |d| < eps on x/y/z
};

__device__ inline Ray make_ray(float px, float py, float pz, float qx, float qy, float qz) {
    Ray r;
    r.px = px; r.py = py; r.pz = pz;
    r.qx = qx; r.qy = qy; r.qz = qz;
    r.dx = qx - px; r.dy = qy - py; r.dz = qz - pz;
    r.t = __fsqrt_rn(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
    r.pat = (fabsf(r.dx) < kEps ? 1 : 0) | (fabsf(r.dy) < kEps ? 2 : 0) | (fabsf(r.dz) < kEps ? 4 : 0);
    r.ix = __fdiv_rn(r.t, r.dx);
    r.iy = __fdiv_rn(r.t, r.dy);
    r.iz = __fdiv_rn(r.t, r.dz);
    return r;
}

__device__ inline bool inside_closed(float v, float a, float b) { return v >= a && v <= b; }

// Slab interval along one axis for box [b1, b2], starting point p.
__device__ inline void slab_iv(float p, float inv, float b1, float b2, float& lo, float& hi) {
    float a = (b1 - p) * inv, b = (b2 - p) * inv;
    lo = fminf(a, b);
    hi = fmaxf(a, b);
}

// Rare branches of length_box_ray: rays parallel to one, two or three axes.
__device__ __noinline__ float chord_through_special(const Ray& r, float x1, float y1, float z1, float x2, float y2,
                                                    float z2) {
    const int pat = r.pat;
    if (pat == 7) return 0.0f;
    if (pat == 3) {  // along z
        if (inside_closed(r.px, x1, x2) && inside_closed(r.py, y1, y2) &&
            ((r.pz <= z1 && r.qz >= z2) || (r.qz <= z1 && r.pz >= z2)))
            return fabsf(z2 - z1);
        return 0.0f;
    }
    if (pat == 6) {  // along x
        if (inside_closed(r.pz, z1, z2) && inside_closed(r.py, y1, y2) &&
            ((r.px <= x1 && r.qx >= x2) || (r.qx <= x1 && r.px >= x2)))
            return fabsf(x2 - x1);
        return 0.0f;
    }
    if (pat == 5) {  // along y
        if (inside_closed(r.px, x1, x2) && inside_closed(r.pz, z1, z2) &&
            ((r.py <= y1 && r.qy >= y2) || (r.qy <= y1 && r.py >= y2)))
            return fabsf(y2 - y1);
        return 0.0f;
    }
    // exactly one parallel axis: the start point must be strictly inside that
    // axis' range, the other two axes are slab-tested
    float lo = -INFINITY, hi = INFINITY, a, b;
    if (pat == 1) {
        if (r.px >= x2 || r.px <= x1) return 0.0f;
    } else {
        slab_iv(r.px, r.ix, x1, x2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    if (pat == 2) {
        if (r.py >= y2 || r.py <= y1) return 0.0f;
    } else {
        slab_iv(r.py, r.iy, y1, y2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    if (pat == 4) {
        if (r.pz >= z2 || r.pz <= z1) return 0.0f;
    } else {
        slab_iv(r.pz, r.iz, z1, z2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    return ((hi - lo) >= kEps && lo > kEps && hi < r.t) ? hi - lo : 0.0f;
}

// length_box_ray for box [x1..x2] x [y1..y2] x [z1..z2]: the chord counts only
// when the segment enters and leaves strictly between its end points.
__device__ inline float chord_through(const Ray& r, float x1, float y1, float z1, float x2, float y2, float z2) {
    if (r.pat != 0) return chord_through_special(r, x1, y1, z1, x2, y2, z2);
    float ax, bx, ay, by, az, bz;
    slab_iv(r.px, r.ix, x1, x2, ax, bx);
    slab_iv(r.py, r.iy, y1, y2, ay, by);
    slab_iv(r.pz, r.iz, z1, z2, az, bz);
    float lo = fmaxf(fmaxf(ax, ay), az);
    float hi = fminf(fminf(bx, by), bz);
    return ((hi - lo) >= kEps && lo > kEps && hi < r.t) ? hi - lo : 0.0f;
}

// Same test for a box given as centre +- half extent; the along-axis branches
// compare the start point in the box frame the way the reference does.
__device__ inline float chord_through_c(const Ray& r, float cx, float cy, float cz, float hx, float hy, float hz) {
    if (r.pat != 0) {
        Ray s = r;
        s.px = r.px - cx; s.py = r.py - cy; s.pz = r.pz - cz;
        s.qx = r.qx - cx; s.qy = r.qy - cy; s.qz = r.qz - cz;
        return chord_through_special(s, -hx, -hy, -hz, hx, hy, hz);
    }
    float ax, bx, ay, by, az, bz;
    slab_iv(r.px, r.ix, cx - hx, cx + hx, ax, bx);
    slab_iv(r.py, r.iy, cy - hy, cy + hy, ay, by);
    slab_iv(r.pz, r.iz, cz - hz, cz + hz, az, bz);
    float lo = fmaxf(fmaxf(ax, ay), az);
    float hi = fminf(fminf(bx, by), bz);
    return ((hi - lo) >= kEps && lo > kEps && hi < r.t) ? hi - lo : 0.0f;
}

// Closing rules of length_box_ray_inside (end point may be inside the box).
__device__ inline float finish_open(float lo, float hi, float t) {
    if ((hi - lo) < kEps) return 0.0f;
    if (lo >= t) return 0.0f;
    if (hi <= kEps) return 0.0f;
    if ((hi >= t) && (lo > kEps)) return t - lo;
    if ((lo <= kEps) && (hi <= t)) return hi;
    if ((lo <= kEps) && (hi >= t)) return t;
    return hi - lo;
}

// length_box_ray_inside: path of P -> Q inside the box when Q may lie inside.
__device__ inline float chord_open(const Ray& r, float x1, float y1, float z1, float x2, float y2, float z2) {
    if (r.pat & 2) return 0.0f;  // |dy| < 0.001 returns 0 first
    if (r.pat == 5) {            // along y
        if (inside_closed(r.px, x1, x2) && inside_closed(r.pz, z1, z2)) {
            if (r.py < r.qy) return r.qy - y1;
            if (r.py > r.qy) return y2 - r.qy;
        }
        return 0.0f;
    }
    float lo = -INFINITY, hi = INFINITY, a, b;
    if (r.pat & 1) {
        if (r.px >= x2 || r.px <= x1) return 0.0f;
    } else {
        slab_iv(r.px, r.ix, x1, x2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    slab_iv(r.py, r.iy, y1, y2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    if (r.pat & 4) {
        if (r.pz >= z2 || r.pz <= z1) return 0.0f;
    } else {
        slab_iv(r.pz, r.iz, z1, z2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    return finish_open(lo, hi, r.t);
}

// length_box_ray of the ray P -> Q extended 1000 mm past Q, through a box
// that contains Q (the sub-cell).  The reference evaluates it in distances
// from P (~1000 mm, FP32 ulp ~1e-4 mm against a ~0.2 mm chord); here the slab
// distances are taken from Q, which keeps them sub-millimetre and exact to
// ~1e-8 mm.  Branch rules are those of length_box_ray on the extended segment.
__device__ inline float chord_absorb(float px, float py, float pz, float qx, float qy, float qz, float dist,
                                     float x1, float y1, float z1, float x2, float y2, float z2) {
    const float ext = 1000.0f;
    float ex = qx + ext * (qx - px) / dist;
    float ey = qy + ext * (qy - py) / dist;
    float ez = qz + ext * (qz - pz) / dist;
    float dx = ex - px, dy = ey - py, dz = ez - pz;
    int pat = (fabsf(dx) < kEps ? 1 : 0) | (fabsf(dy) < kEps ? 2 : 0) | (fabsf(dz) < kEps ? 4 : 0);
    if (pat == 7) return 0.0f;
    if (pat == 3 || pat == 5 || pat == 6) {
        Ray s;
        s.px = px; s.py = py; s.pz = pz; s.qx = ex; s.qy = ey; s.qz = ez; s.pat = pat;
        return chord_through_special(s, x1, y1, z1, x2, y2, z2);
    }
    // direction of Q - P, normalised
    float ux = qx - px, uy = qy - py, uz = qz - pz;
    float tq = __fsqrt_rn(ux * ux + uy * uy + uz * uz);
    float lo = -INFINITY, hi = INFINITY, a, b;
    if (pat & 1) {
        if (px >= x2 || px <= x1) return 0.0f;
    } else {
        float inv = __fdiv_rn(tq, ux);
        slab_iv(qx, inv, x1, x2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    if (pat & 2) {
        if (py >= y2 || py <= y1) return 0.0f;
    } else {
        float inv = __fdiv_rn(tq, uy);
        slab_iv(qy, inv, y1, y2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    if (pat & 4) {
        if (pz >= z2 || pz <= z1) return 0.0f;
    } else {
        float inv = __fdiv_rn(tq, uz);
        slab_iv(qz, inv, z1, z2, a, b); lo = fmaxf(lo, a); hi = fminf(hi, b);
    }
    // entry strictly after P, exit strictly before the extended end
    float t_end = ext * tq / dist;
    return ((hi - lo) >= kEps && tq + lo > kEps && hi < t_end) ? hi - lo : 0.0f;
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
    r.t = sqrt(dx * dx + dy * dy + dz * dz);
    double it = 1.0 / r.t;
    r.kx = dx * it; r.ky = dy * it; r.kz = dz * it;
    r.iky = 1.0 / r.ky;
    r.sx = r.kx * r.iky;
    r.sz = r.kz * r.iky;
    r.kk = r.kx * r.kx + r.kz * r.kz;
    r.ikk = 1.0 / r.kk;
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
    double sq = sqrt(disc);
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
