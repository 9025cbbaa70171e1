// Primary-Compton scatter kernels for gfx950.
//
//   bitmap_crystal_kernel     <- geometryRelationShip_Crystal2Crystal    scatter.cu:1440-1540
//   bitmap_collimator_kernel  <- geometryRelationShip_Collimator2Crystal scatter.cu:1939-2009
//   scatter_prep_kernel       per (detector r, site) constants; the reference
//                             recomputes them for every voxel (scatter.cu:1232-1417)
//   scatter_kernel            crystalScatterSysMatCuda / collimatorScatterSysMatCuda
//                             (scatter.cu:1077-1437, 1544-1936)
//
// The bitmap kernels reproduce the reference bits exactly: IEEE FP32 division
// and square root, no FMA contraction, the reference's operation order.  The
// scatter kernels evaluate everything that decides a branch (cone angle,
// scattered energy and the energy-window cut, the theta range and the
// Riemann step count of the KN weight, the "line through the detector" test)
// in FP64 and accumulate each element in FP64; the smooth factors (window
// probability, Range_phi, attenuation) are FP32.
#include <hip/hip_runtime.h>

#include "kn.hpp"
#include "scatter.hpp"
#include "trace.hpp"

namespace sysmat {

// ---------------------------------------------------------------------------
// GeometryRelationShip bitmaps
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)
__device__ inline float dist3(float x1, float y1, float z1, float x2, float y2, float z2) {
    const float a = x1 - x2, b = y1 - y2, c = z1 - z2;
    return __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(a, a), __fmul_rn(b, b)), __fmul_rn(c, c)));
}
__device__ inline float half_diag(const Crystal& K) {  // sqrt(w^2 + h^2 + t^2) / 2
    return __fdiv_rn(K.rbound, 2.0f);
}

// One thread = crystal pair (i < j) x 32 consecutive k.
__global__ void __launch_bounds__(256) bitmap_crystal_kernel(const Crystal* __restrict__ cry, int n,
                                                             unsigned* __restrict__ words) {
    const int i = blockIdx.y;
    const int nchunk = (n + 31) >> 5;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long njs = (long long)(n - 1 - i) * nchunk;
    if (t >= njs) return;
    const int j = i + 1 + (int)(t / nchunk);
    const int k0 = (int)(t % nchunk) * 32;
    const Crystal I = cry[i], J = cry[j];
    const float xi = I.cx, yi = I.y_raw, zi = I.cz, xj = J.cx, yj = J.y_raw, zj = J.cz;
    const float Ri = half_diag(I), Rj = half_diag(J);
    const float Lij = dist3(xi, yi, zi, xj, yj, zj);
    const float cij = __fdiv_rn(Ri, Lij), cji = __fdiv_rn(Rj, Lij);
    const float uijx = __fdiv_rn(xi - xj, Lij), uijy = __fdiv_rn(yi - yj, Lij), uijz = __fdiv_rn(zi - zj, Lij);
    const float ujix = __fdiv_rn(xj - xi, Lij), ujiy = __fdiv_rn(yj - yi, Lij), ujiz = __fdiv_rn(zj - zi, Lij);
    unsigned bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int k = k0 + b;
        if (k >= n || k == i || k == j) continue;
        const Crystal K = cry[k];
        const float xk = K.cx, yk = K.y_raw, zk = K.cz;
        const float Rk = half_diag(K);
        const float Lik = dist3(xk, yk, zk, xi, yi, zi);
        const float cki = __fdiv_rn(Rk, Lik);
        const float Ljk = dist3(xk, yk, zk, xj, yj, zj);
        const float ckj = __fdiv_rn(Rk, Ljk);
        const float ukix = __fdiv_rn(xk - xi, Lik), ukiy = __fdiv_rn(yk - yi, Lik), ukiz = __fdiv_rn(zk - zi, Lik);
        const float ukjx = __fdiv_rn(xk - xj, Ljk), ukjy = __fdiv_rn(yk - yj, Ljk), ukjz = __fdiv_rn(zk - zj, Ljk);
        bool cross = dist3(ukix, ukiy, ukiz, ujix, ujiy, ujiz) <= __fadd_rn(cki, cji);
        if (!cross) cross = dist3(ukjx, ukjy, ukjz, uijx, uijy, uijz) <= __fadd_rn(ckj, cij);
        if (cross) bits |= 1u << b;
    }
    if (!bits) return;
    const long long pair = (long long)i * n - ((long long)i * (i + 1)) / 2 + j;
    const long long bit0 = pair * n + k0;
    const long long w = bit0 >> 5;
    const int off = (int)(bit0 & 31);
    atomicOr(&words[w], bits << off);
    if (off) {
        const unsigned hi = bits >> (32 - off);
        if (hi) atomicOr(&words[w + 1], hi);
    }
}

// One thread = (hole h, detector j) x 32 consecutive k.
__global__ void __launch_bounds__(256) bitmap_collimator_kernel(const Hole* __restrict__ hol,
                                                                const Crystal* __restrict__ cry, int n,
                                                                unsigned* __restrict__ words) {
    const int h = blockIdx.y;
    const int nchunk = (n + 31) >> 5;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n * nchunk) return;
    const int j = (int)(t / nchunk);
    const int k0 = (int)(t % nchunk) * 32;
    const Hole H = hol[h];
    const float xh = H.x, yh = __fdiv_rn(__fadd_rn(H.y1_raw, H.y2_raw), 2.0f), zh = H.z;
    const Crystal J = cry[j];
    const float Rj = half_diag(J);
    const float Lij = dist3(xh, yh, zh, J.cx, J.y_raw, J.cz);
    const float cji = __fdiv_rn(Rj, Lij);
    const float ujix = __fdiv_rn(J.cx - xh, Lij), ujiy = __fdiv_rn(J.y_raw - yh, Lij), ujiz = __fdiv_rn(J.cz - zh, Lij);
    unsigned bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int k = k0 + b;
        if (k >= n) continue;
        const Crystal K = cry[k];
        const float Rk = half_diag(K);
        const float Lik = dist3(K.cx, K.y_raw, K.cz, xh, yh, zh);
        const float cki = __fdiv_rn(Rk, Lik);
        const float ukix = __fdiv_rn(K.cx - xh, Lik), ukiy = __fdiv_rn(K.y_raw - yh, Lik), ukiz = __fdiv_rn(K.cz - zh, Lik);
        if (dist3(ukix, ukiy, ukiz, ujix, ujiy, ujiz) <= __fadd_rn(cki, cji)) bits |= 1u << b;
    }
    if (!bits) return;
    const long long bit0 = (long long)h * n * n + (long long)j * n + k0;
    const long long w = bit0 >> 5;
    const int off = (int)(bit0 & 31);
    atomicOr(&words[w], bits << off);
    if (off) {
        const unsigned hi = bits >> (32 - off);
        if (hi) atomicOr(&words[w + 1], hi);
    }
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Per-pair constants
// ---------------------------------------------------------------------------
__device__ inline double ycen(const Crystal& K, double fov2c) { return (double)K.y_raw + fov2c; }

// Chord of S -> Q through crystal a, in a's frame (scatter.cu:1377-1397), FP64.
__device__ inline double att_chord(const Crystal& A, double fov2c, double sx, double sy, double sz, double qx,
                                   double qy, double qz) {
    const double ax = A.cx, ay = ycen(A, fov2c), az = A.cz;
    const double c = (A.rot == 0.0f) ? 1.0 : cos(-(double)A.rot), s = (A.rot == 0.0f) ? 0.0 : sin(-(double)A.rot);
    const double sxa = (sx - ax) * c - (sz - az) * s, sya = sy - ay, sza = (sx - ax) * s + (sz - az) * c;
    const double qxa = (qx - ax) * c - (qz - az) * s, qya = qy - ay, qza = (qx - ax) * s + (qz - az) * c;
    const double hx = 0.5 * (double)(2.0f * A.hx), hy = 0.5 * (double)(2.0f * A.hy), hz = 0.5 * (double)(2.0f * A.hz);
    return chord_box64(sxa, sya, sza, qxa, qya, qza, -hx, -hy, -hz, hx, hy, hz);
}

__device__ inline bool test_bit(const unsigned* __restrict__ w, long long bit) {
    return (w[bit >> 5] >> (bit & 31)) & 1u;
}

// One workgroup per detecting crystal r; lanes evaluate sites, survivors are
// compacted in index order into pairs[r * max_sites + slot].
__global__ void __launch_bounds__(256) scatter_prep_kernel(ScParams p, const ScRow* __restrict__ rows,
                                                           const Crystal* __restrict__ cry,
                                                           const Hole* __restrict__ hol,
                                                           const unsigned* __restrict__ bits,
                                                           ScPair* __restrict__ pairs, int* __restrict__ counts) {
    __shared__ int s_wcount[4];
    const int rr = blockIdx.x;
    const int r = p.det_begin + rr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const ScRow R = rows[r];
    const int n = p.n_det;
    const int n_sites = (p.mode == 0) ? n : p.max_sites;
    // image bounding sphere (rotated centre)
    const double icx0 = p.sx, icy0 = p.sy;
    const double icx = icx0 * p.rc - icy0 * p.rs, icy = icx0 * p.rs + icy0 * p.rc, icz = p.sz;
    const double rho = 0.5 * sqrt((double)p.X * p.X * p.wx * p.wx + (double)p.Y * p.Y * p.wy * p.wy +
                                  (double)p.Z * p.Z * p.wz * p.wz) + 1e-3;
    int count = 0;
    if (R.use != 0.0f) {
        for (int base = 0; base < n_sites; base += 256) {
            const int s = base + tid;
            bool alive = false;
            ScPair P{};
            if (s < n_sites && !(p.mode == 0 && s == r)) {
                double sx, sy, sz;
                if (p.mode == 0) {
                    const Crystal K = cry[s];
                    sx = K.cx; sy = ycen(K, p.fov2c); sz = K.cz;
                    alive = K.mu_t > 0.01f;  // a crystal with mu_t <= 0.01 has an all-zero PE row
                } else {
                    const Hole H = hol[s];
                    sx = H.x;
                    sy = ((double)H.y1_raw + (double)H.y2_raw) / 2.0 + p.fov2c;
                    sz = H.z;
                    alive = true;
                    P.hy1 = H.y1_raw;
                    P.hy2 = H.y2_raw;
                }
                const double bx = R.cx - sx, by = R.cy - sy, bz = R.cz - sz;
                const double L = sqrt(bx * bx + by * by + bz * bz);
                // pairs whose scatter angle is outside the window band for every voxel
                const double ax = sx - icx, ay = sy - icy, az = sz - icz;
                const double la = sqrt(ax * ax + ay * ay + az * az);
                if (alive && la > rho) {
                    const double c0 = (ax * bx + ay * by + az * bz) / (la * L);
                    const double th0 = acos(fmin(1.0, fmax(-1.0, c0)));
                    const double dth = asin(fmin(1.0, rho / la));
                    if (th0 - dth > R.th_hi + 1e-6 || th0 + dth < R.th_lo - 1e-6) alive = false;
                }
                if (alive) {
                    P.site = s;
                    P.sx = sx; P.sy = sy; P.sz = sz;
                    P.L = L;
                    P.bx = bx / L; P.by = by / L; P.bz = bz / L;
                    const double ssx = (sx - R.cx) * R.crn - (sz - R.cz) * R.srn, ssy = sy - R.cy;
                    const double ssz = (sx - R.cx) * R.srn + (sz - R.cz) * R.crn;
                    P.ssx = ssx; P.ssy = ssy; P.ssz = ssz;
                    for (int c = 0; c < 8; ++c) {
                        const double qx = (c & 1) ? R.hx : -R.hx, qy = (c & 2) ? R.hy : -R.hy, qz = (c & 4) ? R.hz : -R.hz;
                        const double ex = qx - ssx, ey = qy - ssy, ez = qz - ssz;
                        P.ib[c] = 1.0 / sqrt(ex * ex + ey * ey + ez * ez);
                    }
                    // absorption chord: S_self -> Q_self + 1000 (Q_self - S_self) / L, Q_self = 0
                    const double tx = -1000.0 * ssx / L, ty = -1000.0 * ssy / L, tz = -1000.0 * ssz / L;
                    P.l_abs = (float)chord_box64(ssx, ssy, ssz, tx, ty, tz, -R.hx, -R.hy, -R.hz, R.hx, R.hy, R.hz);
                    const double q = R.rdet / L;
                    P.c_fwd = (q >= 1.0) ? -2.0f : (float)(sqrt(1.0 - q * q) - 1e-6);
                    // attenuators listed by the bitmap (scatter.cu:1343-1401 / 1839-1897)
                    double ape = 0.0, ac = 0.0;
                    for (int a = 0; a < n; ++a) {
                        if (a == r) continue;
                        long long bit;
                        if (p.mode == 0) {
                            if (a == s) continue;
                            const int i0 = s < r ? s : r, j0 = s < r ? r : s;
                            bit = ((long long)i0 * n - ((long long)i0 * (i0 + 1)) / 2 + j0) * n + a;
                        } else {
                            bit = (long long)s * n * n + (long long)r * n + a;
                        }
                        if (!test_bit(bits, bit)) continue;
                        const Crystal A = cry[a];
                        const double c = att_chord(A, p.fov2c, sx, sy, sz, R.cx, R.cy, R.cz);
                        ape += c * (double)A.mu_pe;
                        ac += c * (double)A.mu_c;
                    }
                    P.A_pe = (float)ape;
                    P.A_c = (float)ac;
                }
            }
            const unsigned long long bal = __ballot(alive);
            if (lane == 0) s_wcount[wave] = __popcll(bal);
            __syncthreads();
            int off = count, all = 0;
            for (int w2 = 0; w2 < 4; ++w2) {
                const int c = s_wcount[w2];
                if (w2 < wave) off += c;
                all += c;
            }
            if (alive) pairs[(long long)rr * p.max_sites + off + __popcll(bal & below)] = P;
            count += all;
            __syncthreads();
        }
    }
    if (tid == 0) counts[rr] = count;
}

// ---------------------------------------------------------------------------
// Per-voxel accumulation
// ---------------------------------------------------------------------------
__device__ inline float window_probability(double e, double res, double lo, double hi) {
    // 0.5 (erf(z2) - erf(z1)), z = (bound - E') / (sqrt(2) sigma); tails via erfc
    const double s2 = res / (double)2.35482f * e * 1.4142135623730951;
    const float z1 = (float)((lo - e) / s2), z2 = (float)((hi - e) / s2);
    if (z1 >= 0.0f) return 0.5f * (erfcf(z1) - erfcf(z2));
    if (z2 <= 0.0f) return 0.5f * (erfcf(-z2) - erfcf(-z1));
    return 0.5f * (erff(z2) - erff(z1));
}

__global__ void __launch_bounds__(256) scatter_kernel(ScParams p, const ScRow* __restrict__ rows,
                                                      const ScPair* __restrict__ pairs,
                                                      const int* __restrict__ counts, const float* __restrict__ pe,
                                                      float* __restrict__ dst) {
    // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so give
    // each XCD a contiguous range of (voxel tile, detector group) work where
    // consecutive groups share the voxel tile (its PE columns stay in that XCD's L2).
    const long long nb = p.n_blocks;
    const long long padded = (nb + 7) / 8 * 8;
    const long long b = blockIdx.x;
    const long long idx = (b % 8) * (padded / 8) + b / 8;
    if (idx >= nb) return;
    const int tile = (int)(idx / p.n_rgroups);
    const int rg = (int)(idx % p.n_rgroups);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = rg * 4 + wave;
    if (rr >= p.det_count) return;
    const int r = p.det_begin + rr;
    const long long v = (long long)tile * 64 + lane;
    const bool active = v < p.V;
    const ScRow R = rows[r];
    double acc = 0.0;
    if (R.use != 0.0f && active) {
        // voxel point (scatter.cu:1153-1166), FP64 from the FP32 parameters
        const int X = p.X, Y = p.Y;
        const int iz = (int)(v / ((long long)X * Y));
        const int rem = (int)(v % ((long long)X * Y));
        const int iy = rem / X, ix = rem % X;
        const double x0 = ((double)ix - (double)X / 2.0 + 0.5) * p.wx + p.sx;
        const double y0 = ((double)iy - (double)Y / 2.0 + 0.5) * p.wy + p.sy;
        const double zI = ((double)iz - (double)p.Z / 2.0 + 0.5) * p.wz + p.sz;
        const double xI = x0 * p.rc - y0 * p.rs, yI = x0 * p.rs + y0 * p.rc;
        const double xIs = (xI - R.cx) * R.crn - (zI - R.cz) * R.srn, yIs = yI - R.cy;
        const double zIs = (xI - R.cx) * R.srn + (zI - R.cz) * R.crn;
        const KnCoef kc = kn_coef(p.e0);
        const double e0 = p.e0, ea = p.e0 / 511.0;
        const double two_pi = 2.0 * M_PI;
        const float mu_ratio = (float)(R.mu_c / R.mu_pe);
        const int n_sites = counts[rr];
        const ScPair* P = pairs + (long long)rr * p.max_sites;
        for (int j = 0; j < n_sites; ++j) {
            const ScPair S = P[j];
            // source weight
            double w;
            if (p.mode == 0) {
                w = (double)(pe[(long long)S.site * p.V + v] * mu_ratio);
            } else {
                const double lx = xI - S.sx, ly = yI - S.sy, lz = zI - S.sz;
                const double Li = sqrt(lx * lx + ly * ly + lz * lz);
                const double dyh = (double)S.hy2 - (double)S.hy1;
                double solid;
                if (p.mode == 1) {
                    const double vr = (double)0.1f;
                    solid = (vr * vr * M_PI + dyh * vr) / (4.0 * M_PI * Li * Li) * fabs(S.sy - yI) / Li;
                } else {
                    const double vr = (double)0.2f;
                    solid = (vr * vr + dyh * vr) / (4.0 * Li * Li);
                }
                w = solid * p.w_col;
            }
            // cone angle at the site (scatter.cu:996-1015, 1246-1248)
            const double ax = S.sx - xI, ay = S.sy - yI, az = S.sz - zI;
            const double ma = sqrt(ax * ax + ay * ay + az * az);
            const double ima = 1.0 / ma;
            double c = (ax * S.bx + ay * S.by + az * S.bz) * ima;
            c = fmin(1.0, fmax(-1.0, c));
            const double es = e0 / (1.0 + ea * (1.0 - c));
            const double res_s = R.res * sqrt(es / e0);
            const double k2 = 2.0 * res_s / (double)2.35482f;
            if ((es * (1.0 + k2)) <= R.lo || (es * (1.0 - k2)) >= R.hi) continue;
            const float p_e = window_probability(es, res_s, R.lo, R.hi);
            const double sn = sqrt(fmax(0.0, 1.0 - c * c));
            float phi;
            if (S.L * sn * 2.0 <= R.rdet)
                phi = (float)two_pi;
            else
                phi = 4.0f * asinf(fminf((float)(R.rdet / 2.0 / S.L / sn), 1.0f));
            // theta range over the corners of r (scatter.cu:1303-1336)
            const double asx = S.ssx - xIs, asy = S.ssy - yIs, asz = S.ssz - zIs;
            const double d0 = asx * S.ssx + asy * S.ssy + asz * S.ssz;
            const double qx = R.hx * asx, qy = R.hy * asy, qz = R.hz * asz;
            double cmin = 2.0, cmax = -2.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double dq = ((q & 1) ? qx : -qx) + ((q & 2) ? qy : -qy) + ((q & 4) ? qz : -qz) - d0;
                const double cq = fmin(1.0, fmax(-1.0, dq * S.ib[q] * ima));
                cmin = fmin(cmin, cq);
                cmax = fmax(cmax, cq);
            }
            double th_min = acos(cmax);
            const double th_max = acos(cmin);
            if (c >= (double)S.c_fwd || c <= -(double)S.c_fwd) {
                // the extended image -> site line through crystal r (scatter.cu:1298-1301, 1332-1335)
                const double f = 1000.0 * ima;
                const double tx = S.ssx + f * asx, ty = S.ssy + f * asy, tz = S.ssz + f * asz;
                const double through = chord_box64(xIs, yIs, zIs, tx, ty, tz, -R.hx, -R.hy, -R.hz, R.hx, R.hy, R.hz);
                if (through > (double)0.001f) th_min = 0.0;
            }
            const int nstep = (int)((th_max - th_min) / p.h);
            const double ratio = kn_sum(kc, th_min, nstep, p.h) / p.kn_total;
            // attenuation with mu scaled to E' (pe ~ E^-3, Compton ~ E^-1)
            const double x = e0 / es, x3 = x * x * x;
            const float att = (float)(x3 * (double)S.A_pe + x * (double)S.A_c);
            const double pe_abs = R.mu_pe * x3, tot_abs = pe_abs + R.mu_c * x;
            const float absorb = -expm1f(-(float)((double)S.l_abs * tot_abs));
            const double term = w * ((double)phi / two_pi) * ratio * (double)p_e * (double)expf(-att) *
                                (double)absorb * (pe_abs / tot_abs);
            acc += term;
        }
    }
    if (active) dst[(long long)rr * p.row_stride + v] = (float)acc;
}

hipError_t launch_bitmap_crystal(const Crystal* cry, int n, unsigned* words, hipStream_t stream) {
    if (n < 2) return hipSuccess;
    const int nchunk = (n + 31) / 32;
    const long long maxt = (long long)(n - 1) * nchunk;
    dim3 grid((unsigned)((maxt + 255) / 256), (unsigned)n);
    hipLaunchKernelGGL(bitmap_crystal_kernel, grid, dim3(256), 0, stream, cry, n, words);
    return hipGetLastError();
}

hipError_t launch_bitmap_collimator(const Hole* hol, int n_holes, const Crystal* cry, int n, unsigned* words,
                                    hipStream_t stream) {
    if (n_holes <= 0 || n <= 0) return hipSuccess;
    const int nchunk = (n + 31) / 32;
    const long long maxt = (long long)n * nchunk;
    dim3 grid((unsigned)((maxt + 255) / 256), (unsigned)n_holes);
    hipLaunchKernelGGL(bitmap_collimator_kernel, grid, dim3(256), 0, stream, hol, cry, n, words);
    return hipGetLastError();
}

hipError_t launch_scatter_prep(const ScParams& p, const ScRow* rows, const Crystal* cry, const Hole* hol,
                               const unsigned* bits, ScPair* pairs, int* counts, hipStream_t stream) {
    if (p.det_count <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_prep_kernel, dim3((unsigned)p.det_count), dim3(256), 0, stream, p, rows, cry, hol, bits,
                       pairs, counts);
    return hipGetLastError();
}

hipError_t launch_scatter(const ScParams& p, const ScRow* rows, const ScPair* pairs, const int* counts,
                          const float* pe, float* dst, hipStream_t stream) {
    if (p.n_blocks <= 0) return hipSuccess;
    const long long padded = (p.n_blocks + 7) / 8 * 8;
    hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)padded), dim3(256), 0, stream, p, rows, pairs, counts, pe, dst);
    return hipGetLastError();
}

}  // namespace sysmat
