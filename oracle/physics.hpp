// Compton-scatter physics helpers of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
//
// Restated from ScatterGen_RayTracing_CircularHole/scatter.cu (same operation
// order; the FP32 instantiation rounds like the reference, including the spots
// where a double literal promotes an expression):
//   kn_shape          <- diffComptonSection        scatter.cu:961-976
//   kn_riemann        <- computeComptonIntegral    scatter.cu:979-993
//   cone_angle        <- calculateConeAngle        scatter.cu:996-1015
//   scatter_energy    <- calculateScatterEnergy    scatter.cu:1018-1024
//   window_probability<- calculategaussianIntegral scatter.cu:1026-1034
//   distance          <- calculateDist             scatter.cu:1036-1039
//   crystal_pair_bit / collimator_bit <- indexFrombitmap_* scatter.cu:1040-1074
#pragma once
#include <cstdint>

#include "geometry.hpp"

namespace oracle {

constexpr float kElectronRest = 511.0f;  // keV

// Klein-Nishina angular shape (unnormalised), alpha = E0/511.
template <class R>
inline R kn_shape(R theta, R e0) {
    R c = r_cos(theta);
    R alpha = e0 / R(511.0f);
    R f1 = alpha * (R(1.0f) - c);
    R f2 = R(1.0f) + c * c;
    return f2 / (R(1.0f) + f1) / (R(1.0f) + f1) * (R(1.0f) + f1 * f1 / f2 / (R(1.0f) + f1));
}

// Left Riemann sum over i = 0..floor((hi-lo)/step) inclusive of shape*sin*step.
template <class R>
inline R kn_riemann(R e0, R lo, R hi, R step) {
    R total = R(0.0f);
    int n = (int)to_double((hi - lo) / step);
    for (int i = 0; i <= n; ++i) {
        R theta = lo + R((float)i) * step;
        total += kn_shape(theta, e0) * r_sin(theta) * step;
    }
    return total;
}

template <class R>
inline R distance(R x1, R y1, R z1, R x2, R y2, R z2) {
    return r_sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
}

// Angle at A between the incoming direction S->A and the outgoing A->D.
template <class R>
inline R cone_angle(R xs, R ys, R zs, R xa, R ya, R za, R xd, R yd, R zd) {
    R ax = xa - xs, ay = ya - ys, az = za - zs;
    R bx = xd - xa, by = yd - ya, bz = zd - za;
    R dot = bx * ax + by * ay + bz * az;
    R ma = r_sqrt(ax * ax + ay * ay + az * az);
    R mb = r_sqrt(bx * bx + by * by + bz * bz);
    R c = dot / (ma * mb);
    if (c > R(1.0f)) c = R(1.0f);
    if (c < R(-1.0f)) c = R(-1.0f);
    return r_acos(c);
}

template <class R>
inline R scatter_energy(R theta, R e0) {
    R c = r_cos(theta);
    return e0 / (R(1.0f) + (e0 / R(511.0f)) * (R(1.0f) - c));
}

// Fraction of a Gaussian photopeak (FWHM = res * E) inside [lo, hi].
template <class R>
inline R window_probability(R e, R res, R lo, R hi) {
    using W = typename Wide<R>::type;
    R sigma = res / R(2.35482f) * e;
    R s2 = R(W(sigma) * W(std::sqrt(2.0)));
    R z1 = (lo - e) / s2;
    R z2 = (hi - e) / s2;
    return R(0.5f) * (r_erf(z2) - r_erf(z1));
}

// Bit layouts of the GeometryRelationShip files (LSB-first 32-bit words).
inline long long crystal_pair_bit(int i, int j, int k, int n) {
    if (i > j) {
        int t = i;
        i = j;
        j = t;
    }
    long long pair = (long long)i * n - ((long long)i * (i + 1)) / 2 + j;
    return pair * n + k;
}

inline long long collimator_bit(int h, int j, int k, int n) {
    return (long long)h * n * n + (long long)j * n + k;
}

inline int test_bit(const uint32_t* words, long long bit) { return (int)((words[bit >> 5] >> (bit & 31)) & 1u); }

}  // namespace oracle
