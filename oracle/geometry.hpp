// Ray / solid primitives of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
//
// Each function restates one reference device function with the same branch
// structure, thresholds and operation order, so that the FP32 instantiation
// rounds like the reference's arithmetic:
//   chord_box        <- length_box_ray        PEGen_RayTracing_CircularHole/PESysMatGen.cu:19-151
//                                             (same semantics ScatterGen_RayTracing_CircularHole/scatter.cu:31-323)
//   chord_box_open   <- length_box_ray_inside PESysMatGen.cu:153-307 (scatter.cu:325-697)
//   chord_hole       <- length_cylinder_ray   PESysMatGen.cu:309-373
#pragma once
#include "numtypes.hpp"

namespace oracle {

template <class R> inline R rmax(R a, R b) { return (a >= b) ? a : b; }
template <class R> inline R rmin(R a, R b) { return (a <= b) ? a : b; }

// One slab of the slab method: parametric entry/exit along one axis.
template <class R>
inline void slab(R p, R inv, R b1, R b2, R& lo, R& hi) {
    if (inv < R(0.0f)) {
        lo = (b2 - p) * inv;
        hi = (b1 - p) * inv;
    } else {
        hi = (b2 - p) * inv;
        lo = (b1 - p) * inv;
    }
}

// Intersect [lo, hi] with [lo2, hi2]; false when they do not overlap.
template <class R>
inline bool clip(R& lo, R& hi, R lo2, R hi2) {
    if ((lo > hi2) || (lo2 > hi)) return false;
    if (lo2 > lo) lo = lo2;
    if (hi2 < hi) hi = hi2;
    return true;
}

// Chord accepted only when both crossings lie strictly inside the segment.
template <class R>
inline R finish_through(R lo, R hi, R t) {
    const R eps = R(0.001f);
    if ((hi - lo) < eps) return R(0.0f);
    if (lo >= t || hi >= t) return R(0.0f);
    if (lo <= eps || hi <= eps) return R(0.0f);
    return hi - lo;
}

// Chord of the part of the segment inside the box, the end point may be inside.
template <class R>
inline R finish_open(R lo, R hi, R t) {
    const R eps = R(0.001f);
    if ((hi - lo) < eps) return R(0.0f);
    if (lo >= t) return R(0.0f);
    if (hi <= eps) return R(0.0f);
    if ((hi >= t) && (lo > eps)) return t - lo;
    if ((lo <= eps) && (hi <= t)) return hi;
    if ((lo <= eps) && (hi >= t)) return t;
    return hi - lo;
}

// Two-axis slab for a ray parallel to the third axis (a, b are the live axes).
template <class R, bool Open>
inline R slab2(R ai, R bi, R da, R db, R a1, R a2, R b1, R b2) {
    R t = r_sqrt(da * da + db * db);
    R ia = t / da, ib = t / db;
    R lo, hi, lo2, hi2;
    slab(ai, ia, a1, a2, lo, hi);
    slab(bi, ib, b1, b2, lo2, hi2);
    if (!clip(lo, hi, lo2, hi2)) return R(0.0f);
    return Open ? finish_open(lo, hi, t) : finish_through(lo, hi, t);
}

template <class R, bool Open>
inline R slab3(R xi, R yi, R zi, R dx, R dy, R dz, R x1, R y1, R z1, R x2, R y2, R z2) {
    R t = r_sqrt(dx * dx + dy * dy + dz * dz);
    R ix = t / dx, iy = t / dy, iz = t / dz;
    R lo, hi, lo2, hi2;
    slab(xi, ix, x1, x2, lo, hi);
    slab(yi, iy, y1, y2, lo2, hi2);
    if (!clip(lo, hi, lo2, hi2)) return R(0.0f);
    slab(zi, iz, z1, z2, lo2, hi2);
    if (!clip(lo, hi, lo2, hi2)) return R(0.0f);
    return Open ? finish_open(lo, hi, t) : finish_through(lo, hi, t);
}

template <class R> inline bool within(R v, R a, R b) { return v >= a && v <= b; }

// length_box_ray: chord of segment (i -> o) through box [1, 2]; zero unless the
// segment enters and leaves the box strictly between its end points.
template <class R>
R chord_box(R xi, R yi, R zi, R xo, R yo, R zo, R x1, R y1, R z1, R x2, R y2, R z2) {
    const R eps = R(0.001f);
    R dx = xo - xi, dy = yo - yi, dz = zo - zi;
    bool px = r_fabs(dx) < eps, py = r_fabs(dy) < eps, pz = r_fabs(dz) < eps;
    if (px && py && pz) return R(0.0f);
    if (px && py) {  // along z
        if (within(xi, x1, x2) && within(yi, y1, y2) && ((zi <= z1 && zo >= z2) || (zo <= z1 && zi >= z2)))
            return r_fabs(z2 - z1);
        return R(0.0f);
    }
    if (pz && py) {  // along x
        if (within(zi, z1, z2) && within(yi, y1, y2) && ((xi <= x1 && xo >= x2) || (xo <= x1 && xi >= x2)))
            return r_fabs(x2 - x1);
        return R(0.0f);
    }
    if (px && pz) {  // along y
        if (within(xi, x1, x2) && within(zi, z1, z2) && ((yi <= y1 && yo >= y2) || (yo <= y1 && yi >= y2)))
            return r_fabs(y2 - y1);
        return R(0.0f);
    }
    if (px) {
        if (xi >= x2 || xi <= x1) return R(0.0f);
        return slab2<R, false>(yi, zi, dy, dz, y1, y2, z1, z2);
    }
    if (py) {
        if (yi >= y2 || yi <= y1) return R(0.0f);
        return slab2<R, false>(xi, zi, dx, dz, x1, x2, z1, z2);
    }
    if (pz) {
        if (zi >= z2 || zi <= z1) return R(0.0f);
        return slab2<R, false>(xi, yi, dx, dy, x1, x2, y1, y2);
    }
    return slab3<R, false>(xi, yi, zi, dx, dy, dz, x1, y1, z1, x2, y2, z2);
}

// length_box_ray_inside: chord of the segment inside the box when the end
// point o may lie inside it.  Zero for |dy| < 0.001 (first test of the
// reference), so only the y-live branches remain reachable.
template <class R>
R chord_box_open(R xi, R yi, R zi, R xo, R yo, R zo, R x1, R y1, R z1, R x2, R y2, R z2) {
    const R eps = R(0.001f);
    R dx = xo - xi, dy = yo - yi, dz = zo - zi;
    if (r_fabs(dy) < eps) return R(0.0f);
    bool px = r_fabs(dx) < eps, pz = r_fabs(dz) < eps;
    if (px && pz) {  // along y
        if (within(xi, x1, x2) && within(zi, z1, z2)) {
            if (yi < yo) return yo - y1;
            if (yi > yo) return y2 - yo;
        }
        return R(0.0f);
    }
    if (px) {
        if (xi >= x2 || xi <= x1) return R(0.0f);
        return slab2<R, true>(yi, zi, dy, dz, y1, y2, z1, z2);
    }
    if (pz) {
        if (zi >= z2 || zi <= z1) return R(0.0f);
        return slab2<R, true>(xi, yi, dx, dy, x1, x2, y1, y2);
    }
    return slab3<R, true>(xi, yi, zi, dx, dy, dz, x1, y1, z1, x2, y2, z2);
}

// length_cylinder_ray: chord of segment (i -> o) through the finite cylinder
// with axis along y at (xc, zc), faces y1..y2 and radius rad.
template <class R>
R chord_hole(R xi, R yi, R zi, R xo, R yo, R zo, R xc, R y1, R y2, R zc, R rad) {
    if (r_fabs(y1 - y2) < R(0.001f)) return R(0.0f);
    R dx = xo - xi, dy = yo - yi, dz = zo - zi;
    R t = r_sqrt(dx * dx + dy * dy + dz * dz);
    R kx = dx / t, ky = dy / t, kz = dz / t;
    R sx = kx / ky, sz = kz / ky;
    R d1 = y1 - yi, d2 = y2 - yi;
    R xl = xi + sx * d1, xr = xi + sx * d2;
    R zl = zi + sz * d1, zr = zi + sz * d2;
    R tlo = d1 / ky, thi = d2 / ky;
    R rr = rad * rad;
    bool in_l = (xl - xc) * (xl - xc) + (zl - zc) * (zl - zc) <= rr;
    bool in_r = (xr - xc) * (xr - xc) + (zr - zc) * (zr - zc) <= rr;
    const R tsmall = R(0.0001f);
    if (in_l && in_r) {
        if (tlo <= tsmall || thi <= tsmall) return R(0.0f);
        if (tlo >= t || thi >= t) return R(0.0f);
        return r_fabs(thi - tlo);
    }
    R ox = xi - xc, oz = zi - zc;
    R kk = kx * kx + kz * kz;
    R cr = kx * oz - kz * ox;
    R disc = kk * rad * rad - cr * cr;
    if (disc <= R(0.00001f)) return R(0.0f);
    R b = -(kx * ox + kz * oz);
    R sq = r_sqrt(disc);
    R t1 = (b - sq) / kk;
    R t2 = (b + sq) / kk;
    if (!in_l && !in_r) {
        if ((t1 >= tlo) && (t2 <= thi)) {
            if (t2 <= tsmall || t1 <= tsmall) return R(0.0f);
            if (t2 >= t || t1 >= t) return R(0.0f);
            return t2 - t1;
        }
        return R(0.0f);
    }
    if (in_l) {  // enters through the y1 face, leaves through the wall
        if (t2 <= tsmall || tlo <= tsmall) return R(0.0f);
        if (t2 >= t || tlo >= t) return R(0.0f);
        return t2 - tlo;
    }
    // enters through the wall, leaves through the y2 face
    if (thi <= tsmall || t1 <= tsmall) return R(0.0f);
    if (thi >= t || t1 >= t) return R(0.0f);
    return thi - t1;
}

}  // namespace oracle
