// Primary-Compton scatter kernels for gfx950.
//
//   bitmap_crystal_kernel     <- geometryRelationShip_Crystal2Crystal    scatter.cu:1440-1540
//   bitmap_collimator_kernel  <- geometryRelationShip_Collimator2Crystal scatter.cu:1939-2009
//   scatter_prep_kernel       per (detector r, site) constants; the reference
//                             recomputes them for every voxel (scatter.cu:1232-1417)
//   scatter_kernel            crystalScatterSysMatCuda / collimatorScatterSysMatCuda
//                             (scatter.cu:1077-1437, 1544-1936), and the plate-cell
//                             collimator model (sites = cells, weights from
//                             cell_weight_kernel)
//
// The bitmap kernels evaluate the reference's criterion in the reference's
// operation order with IEEE FP32 division and square root and no FMA
// contraction, and match the FP32 restatement of it bit for bit.  nvcc contracts
// the reference's sums of products into FMAs by default, so a criterion value
// that ties its bound within an ulp may round the other way there: bit parity
// with files the reference's own build writes is unpinned.  The
// scatter kernels evaluate everything that decides a branch (cone angle,
// scattered energy and the energy-window cut, the theta range and the
// Riemann step count of the KN weight, the "line through the detector" test)
// in FP64 and accumulate each element in FP64; the smooth factors (window
// probability, Range_phi, attenuation) are FP32.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "approx.hpp"
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

// Attenuator criteria of the GeometryRelationShip bitmaps (angular bounding
// spheres; coordinates without FOV2C).  Shared by the bitmap builders and the
// bitmap-free pair preparation, so both decide identically.
struct CrossIJ {
    float xi, yi, zi, xj, yj, zj, cij, cji, uijx, uijy, uijz, ujix, ujiy, ujiz;
};
__device__ inline CrossIJ cross_ij(const Crystal& I, const Crystal& J) {
    CrossIJ c;
    c.xi = I.cx; c.yi = I.y_raw; c.zi = I.cz;
    c.xj = J.cx; c.yj = J.y_raw; c.zj = J.cz;
    const float Ri = half_diag(I), Rj = half_diag(J);
    const float Lij = dist3(c.xi, c.yi, c.zi, c.xj, c.yj, c.zj);
    c.cij = __fdiv_rn(Ri, Lij);
    c.cji = __fdiv_rn(Rj, Lij);
    c.uijx = __fdiv_rn(c.xi - c.xj, Lij); c.uijy = __fdiv_rn(c.yi - c.yj, Lij); c.uijz = __fdiv_rn(c.zi - c.zj, Lij);
    c.ujix = __fdiv_rn(c.xj - c.xi, Lij); c.ujiy = __fdiv_rn(c.yj - c.yi, Lij); c.ujiz = __fdiv_rn(c.zj - c.zi, Lij);
    return c;
}
// crystal K between crystals i < j (scatter.cu:1460-1530)
__device__ inline bool cross_k(const CrossIJ& c, const Crystal& K) {
    const float xk = K.cx, yk = K.y_raw, zk = K.cz;
    const float Rk = half_diag(K);
    const float Lik = dist3(xk, yk, zk, c.xi, c.yi, c.zi);
    const float cki = __fdiv_rn(Rk, Lik);
    const float Ljk = dist3(xk, yk, zk, c.xj, c.yj, c.zj);
    const float ckj = __fdiv_rn(Rk, Ljk);
    const float ukix = __fdiv_rn(xk - c.xi, Lik), ukiy = __fdiv_rn(yk - c.yi, Lik), ukiz = __fdiv_rn(zk - c.zi, Lik);
    const float ukjx = __fdiv_rn(xk - c.xj, Ljk), ukjy = __fdiv_rn(yk - c.yj, Ljk), ukjz = __fdiv_rn(zk - c.zj, Ljk);
    if (dist3(ukix, ukiy, ukiz, c.ujix, c.ujiy, c.ujiz) <= __fadd_rn(cki, c.cji)) return true;
    return dist3(ukjx, ukjy, ukjz, c.uijx, c.uijy, c.uijz) <= __fadd_rn(ckj, c.cij);
}
struct CrossHJ {
    float xh, yh, zh, cji, ujix, ujiy, ujiz;
};
__device__ inline CrossHJ cross_hj(const Hole& H, const Crystal& J) {
    CrossHJ c;
    c.xh = H.x; c.yh = __fdiv_rn(__fadd_rn(H.y1_raw, H.y2_raw), 2.0f); c.zh = H.z;
    const float Rj = half_diag(J);
    const float Lij = dist3(c.xh, c.yh, c.zh, J.cx, J.y_raw, J.cz);
    c.cji = __fdiv_rn(Rj, Lij);
    c.ujix = __fdiv_rn(J.cx - c.xh, Lij); c.ujiy = __fdiv_rn(J.y_raw - c.yh, Lij); c.ujiz = __fdiv_rn(J.cz - c.zh, Lij);
    return c;
}
// crystal K between layer-0 hole h and crystal j (scatter.cu:1960-2000)
__device__ inline bool cross_hk(const CrossHJ& c, const Crystal& K) {
    const float Rk = half_diag(K);
    const float Lik = dist3(K.cx, K.y_raw, K.cz, c.xh, c.yh, c.zh);
    const float cki = __fdiv_rn(Rk, Lik);
    const float ukix = __fdiv_rn(K.cx - c.xh, Lik), ukiy = __fdiv_rn(K.y_raw - c.yh, Lik), ukiz = __fdiv_rn(K.cz - c.zh, Lik);
    return dist3(ukix, ukiy, ukiz, c.ujix, c.ujiy, c.ujiz) <= __fadd_rn(cki, c.cji);
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
    const CrossIJ c = cross_ij(cry[i], cry[j]);
    unsigned bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int k = k0 + b;
        if (k >= n || k == i || k == j) continue;
        if (cross_k(c, cry[k])) bits |= 1u << b;
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
    const CrossHJ c = cross_hj(hol[h], cry[j]);
    unsigned bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int k = k0 + b;
        if (k >= n) continue;
        if (cross_hk(c, cry[k])) bits |= 1u << b;
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

// Image bounding sphere (rotated centre, radius + 1e-3) used to prune pairs.
struct ImageSphere {
    double cx, cy, cz, rho;
};
__device__ inline ImageSphere image_sphere(const ScParams& p) {
    ImageSphere b;
    b.cx = p.sx * p.rc - p.sy * p.rs;
    b.cy = p.sx * p.rs + p.sy * p.rc;
    b.cz = p.sz;
    b.rho = 0.5 * sqrt((double)p.X * p.X * p.wx * p.wx + (double)p.Y * p.Y * p.wy * p.wy +
                       (double)p.Z * p.Z * p.wz * p.wz) + 1e-3;
    return b;
}

// Every crystal whose box the segment A -> B can cross (a superset; others have a
// zero chord): per slab grid, the segment clipped to the slab's y-range, then the
// grid rows its z-extent covers and, per row, the columns its x-extent covers,
// all widened by the grid's margin.  Visits slabs, rows, columns and cell
// members in ascending order, each crystal once.
template <class F>
__device__ inline void segment_crystals(const ScParams& p, double ax, double ay, double az, double bx, double by,
                                        double bz, F&& visit) {
    const double dx = bx - ax, dy = by - ay, dz = bz - az;
    for (int g = 0; g < p.n_grids; ++g) {
        const SlabGrid G = p.grids[g];
        double t0 = 0.0, t1 = 1.0;
        const double ylo = G.y1 - 1e-3, yhi = G.y2 + 1e-3;
        if (dy != 0.0) {
            double ta = (ylo - ay) / dy, tb = (yhi - ay) / dy;
            if (ta > tb) { const double q = ta; ta = tb; tb = q; }
            t0 = fmax(t0, ta);
            t1 = fmin(t1, tb);
        } else if (ay < ylo || ay > yhi) {
            continue;
        }
        if (!(t1 >= t0)) continue;
        const double xa = ax + t0 * dx, za = az + t0 * dz, xb = ax + t1 * dx, zb = az + t1 * dz;
        const double m = G.m, icx = 1.0 / G.cx, icz = 1.0 / G.cz;
        const double xmin = fmin(xa, xb) - m, xmax = fmax(xa, xb) + m;
        const double zmin = fmin(za, zb) - m, zmax = fmax(za, zb) + m;
        const int ix0 = max(0, (int)floor((xmin - G.x0) * icx)), ix1 = min(G.nx - 1, (int)floor((xmax - G.x0) * icx));
        const int iz0 = max(0, (int)floor((zmin - G.z0) * icz)), iz1 = min(G.nz - 1, (int)floor((zmax - G.z0) * icz));
        if (ix0 > ix1 || iz0 > iz1) continue;
        const double ddz = zb - za, ddx = xb - xa;
        for (int iz = iz0; iz <= iz1; ++iz) {
            double u0 = xmin, u1 = xmax;
            if (ddz != 0.0) {
                // part of the clipped segment whose z lies within the row's widened window
                const double w0 = G.z0 + (double)iz * G.cz - m, w1 = G.z0 + (double)(iz + 1) * G.cz + m;
                double s0 = (w0 - za) / ddz, s1 = (w1 - za) / ddz;
                if (s0 > s1) { const double q = s0; s0 = s1; s1 = q; }
                s0 = fmax(s0, 0.0);
                s1 = fmin(s1, 1.0);
                if (s0 > s1) continue;
                const double x0 = xa + s0 * ddx, x1 = xa + s1 * ddx;
                u0 = fmin(x0, x1) - m;
                u1 = fmax(x0, x1) + m;
            }
            const int jx0 = max(ix0, (int)floor((u0 - G.x0) * icx)), jx1 = min(ix1, (int)floor((u1 - G.x0) * icx));
            const int* cellp = p.grid_cell + G.cell_base + (long long)iz * G.nx;
            for (int ix = jx0; ix <= jx1; ++ix) {
                const int e1 = cellp[ix + 1];
                for (int e = cellp[ix]; e < e1; ++e) visit(p.grid_item[e]);
            }
        }
    }
}

// Per-(detector r, site s) constants.  Returns false when the pair cannot
// contribute for any voxel (dead site, or scatter angle outside the window band
// over the whole image).  Attenuators come from the bitmap when `bits` is given,
// else from the same criterion evaluated inline.
__device__ inline bool make_pair(const ScParams& p, const ScRow& R, int r, int s, const Crystal* __restrict__ cry,
                                 const Hole* __restrict__ hol, const unsigned* __restrict__ bits,
                                 const ImageSphere& B, ScPair& P) {
    const int n = p.n_det;
    P = ScPair{};
    double sx, sy, sz;
    bool alive;
    if (p.mode == 0) {
        const Crystal K = cry[s];
        sx = K.cx; sy = ycen(K, p.fov2c); sz = K.cz;
        alive = K.mu_t > 0.01f;  // a crystal with mu_t <= 0.01 has an all-zero PE row
    } else if (p.mode == 3) {
        const Cell C = p.cells[s];
        sx = C.x; sy = C.y; sz = C.z;
        alive = true;
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
    const double ax = sx - B.cx, ay = sy - B.cy, az = sz - B.cz;
    const double la = sqrt(ax * ax + ay * ay + az * az);
    if (alive && la > B.rho) {
        const double c0 = (ax * bx + ay * by + az * bz) / (la * L);
        const double th0 = acos(fmin(1.0, fmax(-1.0, c0)));
        const double dth = asin(fmin(1.0, B.rho / la));
        if (th0 - dth > R.th_hi + 1e-6 || th0 + dth < R.th_lo - 1e-6) alive = false;
    }
    if (!alive) return false;
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
    // attenuators (scatter.cu:1343-1401 / 1839-1897): crystals that pass the
    // GeometryRelationShip criterion and have a non-zero chord on S -> Q.  With
    // slab grids only crystals the segment can cross are examined (every other
    // crystal has a zero chord), else all of them.
    double ape = 0.0, ac = 0.0;
    if (p.mode == 3) {  // plate path of a cell site to the detector centre
        const PathAtt o = plate_path(p.cplates, p.n_cplates, p.choles, sx, sy, sz, R.cx, R.cy, R.cz);
        ape = o.pe;
        ac = o.c;
    }
    const int i0 = s < r ? s : r, j0 = s < r ? r : s;
    const long long pair_bit = ((long long)i0 * n - ((long long)i0 * (i0 + 1)) / 2 + j0) * n;
    CrossIJ cij{};
    CrossHJ chj{};
    if (!bits && p.mode == 0) cij = cross_ij(cry[i0], cry[j0]);
    if (!bits && (p.mode == 1 || p.mode == 2)) chj = cross_hj(hol[s], cry[r]);
    const double ujx = bx / L, ujy = by / L, ujz = bz / L, cj = R.rdet / L;  // cell criterion (FP64)
    auto criterion = [&](int a, const Crystal& A) -> bool {
        if (p.mode == 0) return bits ? test_bit(bits, pair_bit + a) : cross_k(cij, A);
        if (p.mode == 3) {
            const double kx = (double)A.cx - sx, ky = ycen(A, p.fov2c) - sy, kz = (double)A.cz - sz;
            const double lk = sqrt(kx * kx + ky * ky + kz * kz);
            const double hx = A.hx, hy = A.hy, hz = A.hz;
            const double ck = sqrt(hx * hx + hy * hy + hz * hz) / lk;
            const double ex = kx / lk - ujx, ey = ky / lk - ujy, ez = kz / lk - ujz;
            return sqrt(ex * ex + ey * ey + ez * ez) <= ck + cj;
        }
        return bits ? test_bit(bits, (long long)s * n * n + (long long)r * n + a) : cross_hk(chj, A);
    };
    auto visit = [&](int a) {
        if (a == r || (p.mode == 0 && a == s)) return;
        const Crystal A = cry[a];
        double c;
        if (bits) {  // cheap test first
            if (!criterion(a, A)) return;
            c = att_chord(A, p.fov2c, sx, sy, sz, R.cx, R.cy, R.cz);
        } else {
            c = att_chord(A, p.fov2c, sx, sy, sz, R.cx, R.cy, R.cz);
            if (c == 0.0 || !criterion(a, A)) return;
        }
        ape += c * (double)A.mu_pe;
        ac += c * (double)A.mu_c;
    };
    if (p.n_grids > 0) {
        segment_crystals(p, sx, sy, sz, R.cx, R.cy, R.cz, visit);
    } else {
        for (int a = 0; a < n; ++a) visit(a);
    }
    P.A_pe = (float)ape;
    P.A_c = (float)ac;
    return true;
}

__device__ inline int prep_row(const ScParams& p, int rr) { return p.row_list ? p.row_list[rr] : p.det_begin + rr; }

// One workgroup per (detector row, chunk of kPrepChunk sites); lanes evaluate
// sites, survivors are compacted in index order into
// pairs[rr * max_sites + chunk * kPrepChunk + slot], their number into
// counts[rr * chunks + chunk].
__global__ void __launch_bounds__(256) scatter_prep_kernel(ScParams p, const ScRow* __restrict__ rows,
                                                           const Crystal* __restrict__ cry,
                                                           const Hole* __restrict__ hol,
                                                           const unsigned* __restrict__ bits,
                                                           ScPair* __restrict__ pairs, int* __restrict__ counts) {
    __shared__ int s_wcount[4];
    const int nch = site_chunks(p);
    const int rr = (int)(blockIdx.x / nch), ch = (int)(blockIdx.x % nch);
    const int r = prep_row(p, rr);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const ScRow R = rows[r];
    const int n_sites = (p.mode == 0) ? p.n_det : p.max_sites;
    const ImageSphere B = image_sphere(p);
    int count = 0;
    if (R.use != 0.0f) {
        const int s = ch * kPrepChunk + tid;
        ScPair P;
        const bool alive = s < n_sites && !(p.mode == 0 && s == r) && make_pair(p, R, r, s, cry, hol, bits, B, P);
        const unsigned long long bal = __ballot(alive);
        if (lane == 0) s_wcount[wave] = __popcll(bal);
        __syncthreads();
        int off = 0;
        for (int w2 = 0; w2 < 4; ++w2) {
            const int c = s_wcount[w2];
            if (w2 < wave) off += c;
            count += c;
        }
        if (alive) pairs[(long long)rr * p.max_sites + (long long)ch * kPrepChunk + off + __popcll(bal & below)] = P;
    }
    if (tid == 0) counts[(long long)rr * nch + ch] = count;
}

// ---------------------------------------------------------------------------
// Per-voxel accumulation
// ---------------------------------------------------------------------------
__device__ inline float window_probability(double e, double res, double lo, double hi) {
    // 0.5 (erf(z2) - erf(z1)), z = (bound - E') / (sqrt(2) sigma); tails via erfc
    constexpr double kFwhm = (double)2.35482f;  // FWHM / sigma (reference constant)
    const double s2 = res * (1.0 / kFwhm) * e * 1.4142135623730951;
    const double is2 = rcp_nr(s2);  // one reciprocal for both bounds
    const float z1 = (float)((lo - e) * is2), z2 = (float)((hi - e) * is2);
    // both tails through erfc of |z| (branch-free; erf(z2) - erf(z1) = 2 - erfc(-z1) - erfc(z2)
    // when z1 < 0 < z2)
    const float a = erfc_nr(fabsf(z1)), b = erfc_nr(fabsf(z2));
    if (z1 >= 0.0f) return 0.5f * (a - b);
    if (z2 <= 0.0f) return 0.5f * (b - a);
    return 0.5f * ((2.0f - a) - b);
}

// Voxel point of element v, world and crystal-r frames (scatter.cu:1153-1166),
// FP64 from the FP32 parameters.
struct VoxelPt {
    double xI, yI, zI, xIs, yIs, zIs;
};
__device__ inline VoxelPt voxel_point(const ScParams& p, const ScRow& R, double x0, double y0, double zI) {
    VoxelPt q;
    q.xI = x0 * p.rc - y0 * p.rs;
    q.yI = x0 * p.rs + y0 * p.rc;
    q.zI = zI;
    q.xIs = (q.xI - R.cx) * R.crn - (q.zI - R.cz) * R.srn;
    q.yIs = q.yI - R.cy;
    q.zIs = (q.xI - R.cx) * R.srn + (q.zI - R.cz) * R.crn;
    return q;
}
__device__ inline VoxelPt voxel_of(const ScParams& p, const ScRow& R, long long v) {
    const int X = p.X, Y = p.Y;
    const int iz = (int)(v / ((long long)X * Y));
    const int rem = (int)(v % ((long long)X * Y));
    const int iy = rem / X, ix = rem % X;
    const double x0 = ((double)ix - (double)X / 2.0 + 0.5) * p.wx + p.sx;
    const double y0 = ((double)iy - (double)Y / 2.0 + 0.5) * p.wy + p.sy;
    const double zI = ((double)iz - (double)p.Z / 2.0 + 0.5) * p.wz + p.sz;
    return voxel_point(p, R, x0, y0, zI);
}

// Source weight of a site for the voxel: PE[site, v] mu_c/mu_pe (crystal mode,
// pe_v = PE[site, v]) or the virtual-disc solid angle of the hole.
__device__ inline double site_weight(const ScParams& p, const ScRow& R, const ScPair& S, const VoxelPt& q,
                                     float pe_v) {
    if (p.mode == 0) return (double)(pe_v * (float)(R.mu_c / R.mu_pe));
    if (p.mode == 3) return (double)pe_v;
    const double lx = q.xI - S.sx, ly = q.yI - S.sy, lz = q.zI - S.sz;
    const double iL = rsq_nr(lx * lx + ly * ly + lz * lz);  // 1 / |I - S|
    const double iL2 = iL * iL;
    const double dyh = (double)S.hy2 - (double)S.hy1;
    double solid;
    if (p.mode == 1) {
        const double vr = (double)0.1f;
        solid = (vr * vr * M_PI + dyh * vr) * (0.25 / M_PI) * iL2 * fabs(S.sy - q.yI) * iL;
    } else {
        const double vr = (double)0.2f;
        solid = (vr * vr + dyh * vr) * 0.25 * iL2;
    }
    return solid * p.w_col;
}

// sqrt(x), x >= 0, as x / sqrt(x) through rsq_nr (~1 ulp).
__device__ inline double sqrt_nr(double x) { return x > 0.0 ? x * rsq_nr(x) : 0.0; }

// Contribution of one site to element (r, voxel) with source weight w
// (scatter.cu:1240-1427 / 1736-1923).
// kn: the (cos n h, sin n h) table (a workgroup's LDS copy in the scatter kernel).
template <bool kSeries>
__device__ inline double site_term(const ScParams& p, const ScRow& R, const ScPair& S, const VoxelPt& q,
                                   const KnCoef& kc, double w, const double2* __restrict__ kn) {
    const double e0 = p.e0, ea = p.e0 / 511.0;
    const double two_pi = 2.0 * M_PI;
    // cone angle at the site (scatter.cu:996-1015, 1246-1248)
    const double ax = S.sx - q.xI, ay = S.sy - q.yI, az = S.sz - q.zI;
    // 1 / |S - I| (relative error ~1e-16, as sqrt followed by a division)
    const double ima = rsq_nr(ax * ax + ay * ay + az * az);
    double c = (ax * S.bx + ay * S.by + az * S.bz) * ima;
    c = fmin(1.0, fmax(-1.0, c));
    const double wc = 1.0 + ea * (1.0 - c);  // E0 / E'
    // 1 / sqrt(wc) gives both E' = E0 / wc and the resolution at E', res sqrt(E' / E0)
    const double rwc = rsq_nr(wc);
    const double es = e0 * (rwc * rwc);
    const double res_s = R.res * rwc;
    const double k2 = res_s * (2.0 / (double)2.35482f);
    if ((es * (1.0 + k2)) <= R.lo || (es * (1.0 - k2)) >= R.hi) return 0.0;
    const float p_e = window_probability(es, res_s, R.lo, R.hi);
    const double sn = sqrt_nr(fmax(0.0, 1.0 - c * c));
    float phi;
    if (S.L * sn * 2.0 <= R.rdet)
        phi = (float)two_pi;
    else
        phi = 4.0f * asinf(fminf((float)(R.rdet * 0.5 * rcp_nr(S.L * sn)), 1.0f));
    // theta range over the corners of r (scatter.cu:1303-1336)
    const double asx = S.ssx - q.xIs, asy = S.ssy - q.yIs, asz = S.ssz - q.zIs;
    const double d0 = asx * S.ssx + asy * S.ssy + asz * S.ssz;
    const double qx = R.hx * asx, qy = R.hy * asy, qz = R.hz * asz;
    double cmin = INFINITY, cmax = -INFINITY;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const double dq = ((k & 1) ? qx : -qx) + ((k & 2) ? qy : -qy) + ((k & 4) ? qz : -qz) - d0;
        const double cq = dq * S.ib[k];
        cmin = fmin(cmin, cq);
        cmax = fmax(cmax, cq);
    }
    // scaled by 1 / |S - I| > 0 and clamped once: min / max commute with both monotone
    // maps, and with the same rounding as scaling each corner
    cmin = fmin(1.0, fmax(-1.0, cmin * ima));
    cmax = fmin(1.0, fmax(-1.0, cmax * ima));
    // theta_min = acos(cmax), theta_max = acos(cmin); theta_min = 0 when the extended
    // image -> site line passes through crystal r (scatter.cu:1298-1301, 1332-1335)
    bool through = false;
    if (c >= (double)S.c_fwd || c <= -(double)S.c_fwd) {
        const double f = 1000.0 * ima;
        const double tx = S.ssx + f * asx, ty = S.ssy + f * asy, tz = S.ssz + f * asz;
        through = chord_box64(q.xIs, q.yIs, q.zIs, tx, ty, tz, -R.hx, -R.hy, -R.hz, R.hx, R.hy, R.hz) >
                  (double)0.001f;
    }
    // cos / sin of theta_min directly, theta_max - theta_min from one atan2
    const double smax = sqrt_nr(fmax(0.0, 1.0 - cmin * cmin));  // sin theta_max
    double ua, sa, dy, dx;  // theta_max - theta_min = atan2(dy, dx)
    if (through) {
        ua = 1.0;
        sa = 0.0;
        dy = smax;
        dx = cmin;
    } else {
        ua = cmax;
        sa = sqrt_nr(fmax(0.0, 1.0 - cmax * cmax));
        dy = smax * cmax - cmin * sa;
        dx = cmin * cmax + smax * sa;
    }
    // step count floor((theta_max - theta_min) / h): an FP32 atan2 (error below 3e-7 rad,
    // 3e-5 of a step) decides it unless the quotient lies within 5e-4 of a step boundary,
    // where the FP64 atan2 does -- the same count as FP64 throughout (dy >= 0: the angle
    // difference lies in [0, pi])
    const float nf = atan2_upper((float)dy, (float)dx) * p.ihf;
    const float fr = nf - floorf(nf);
    int nraw;
    if (fr > 5e-4f && fr < 1.0f - 5e-4f)
        nraw = (int)nf;
    else
        nraw = (int)(atan2(dy, dx) / p.h);
    const int nstep = min(kKnSteps, max(0, nraw));
    const double2 cs = kn[nstep];
    const double ratio = kn_sum_cs<kSeries>(kc, ua, sa, cs.x, cs.y, p.h) * p.ikn_total;
    // attenuation with mu scaled to E' (pe ~ E^-3, Compton ~ E^-1)
    const double x = wc, x3 = x * x * x;
    const float att = (float)(x3 * (double)S.A_pe + x * (double)S.A_c);
    const double pe_abs = R.mu_pe * x3, tot_abs = pe_abs + R.mu_c * x;
    const float absorb = -expm1f(-(float)((double)S.l_abs * tot_abs));
    return w * ((double)phi * (1.0 / two_pi)) * ratio * (double)p_e * (double)expf(-att) * (double)absorb *
           (pe_abs * rcp_nr(tot_abs));
}

// Four waves per SIMD (128 VGPRs, ~160 B/lane spilled): the per-site PE gather and
// the FP64 chains of the site term need the extra waves to hide behind; measured
// against 2 (no or few spills): -29 % (and again -38 % in a later build) on the C3
// crystal scatter, -37 % on rows of the 16384-crystal detector; 3 waves were in
// between, 5-8 waves spill more and lose again.
// kSeries: the KN logarithm by its series (photon energies up to 236 keV; launch_scatter
// picks the instance), which keeps the library log1p and its registers out of the kernel.
template <bool kSeries>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) scatter_kernel(ScParams p, const ScRow* __restrict__ rows,
                                                      const ScPair* __restrict__ pairs,
                                                      const int* __restrict__ counts, const float* __restrict__ pe,
                                                      float* __restrict__ dst) {
    // the KN end-angle table in LDS: its per-site lookup depends on the voxel's theta
    // range, so from global memory it is a dependent gather in every site iteration
    // (staged before any wave can leave, so every wave reaches the barrier)
    __shared__ double2 s_kn[kKnSteps + 1];
    for (int i = threadIdx.x; i <= kKnSteps; i += blockDim.x) s_kn[i] = p.kn_tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int rr, tile;
    if (p.items) {
        const long long it = (long long)blockIdx.x * 4 + wave;
        if (it >= p.n_items) return;
        const int2 item = p.items[it];
        rr = item.x;
        tile = item.y;
    } else {
        // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so give
        // each XCD a contiguous range of (voxel tile, detector group) work where
        // consecutive groups share the voxel tile (its PE columns stay in that XCD's L2).
        const long long nb = p.n_blocks;
        const long long padded = (nb + 7) / 8 * 8;
        const long long b = blockIdx.x;
        const long long idx = (b % 8) * (padded / 8) + b / 8;
        if (idx >= nb) return;
        tile = (int)(idx / p.n_rgroups);
        rr = (int)(idx % p.n_rgroups) * 4 + wave;
        if (rr >= p.det_count) return;
    }
    // one detector row and one voxel tile per wave: tell the compiler, so the pair records
    // are read through the scalar cache and the per-site skip below is a scalar branch
    rr = __builtin_amdgcn_readfirstlane(rr);
    tile = __builtin_amdgcn_readfirstlane(tile);
    const int r = __builtin_amdgcn_readfirstlane(prep_row(p, rr));
    const long long drow = (p.row_list && p.dst_by_row) ? (long long)(r - p.det_begin) : (long long)rr;
    long long v;
    bool active;
    if (p.brick) {  // 4 x 4 x 4 voxel brick, x fastest
        const int nbx = (p.X + 3) >> 2, nby = (p.Y + 3) >> 2;
        const int bx = tile % nbx, t2 = tile / nbx, by = t2 % nby, bz = t2 / nby;
        const int ix = 4 * bx + (lane & 3), iy = 4 * by + ((lane >> 2) & 3), iz = 4 * bz + (lane >> 4);
        active = ix < p.X && iy < p.Y && iz < p.Z;
        v = ((long long)iz * p.Y + iy) * p.X + ix;
    } else {        // 64 consecutive voxels
        v = (long long)tile * 64 + lane;
        active = v < p.V;
    }
    const ScRow R = rows[r];
    double acc = 0.0;
    if (R.use != 0.0f) {  // uniform: one detector row per wave
        const VoxelPt q = voxel_of(p, R, active ? v : 0);
        // bounding sphere of the wave's voxel centres (FP32, widened by 1e-3 mm)
        float mnx = active ? (float)q.xI : INFINITY, mxx = active ? (float)q.xI : -INFINITY;
        float mny = active ? (float)q.yI : INFINITY, mxy = active ? (float)q.yI : -INFINITY;
        float mnz = active ? (float)q.zI : INFINITY, mxz = active ? (float)q.zI : -INFINITY;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            mnx = fminf(mnx, __shfl_xor(mnx, o, 64)); mxx = fmaxf(mxx, __shfl_xor(mxx, o, 64));
            mny = fminf(mny, __shfl_xor(mny, o, 64)); mxy = fmaxf(mxy, __shfl_xor(mxy, o, 64));
            mnz = fminf(mnz, __shfl_xor(mnz, o, 64)); mxz = fmaxf(mxz, __shfl_xor(mxz, o, 64));
        }
        const float tcx = 0.5f * (mnx + mxx), tcy = 0.5f * (mny + mxy), tcz = 0.5f * (mnz + mxz);
        const float trad = 0.5f * sqrtf((mxx - mnx) * (mxx - mnx) + (mxy - mny) * (mxy - mny) +
                                        (mxz - mnz) * (mxz - mnz)) + 1e-3f;
        // scatter angles outside [th_lo, th_hi] fail the energy-window test of the site term;
        // 2 mrad covers the FP32 rounding of the bound below (at most ~5e-4 rad near cos = 1)
        const float thl = (float)R.th_lo - 2e-3f, thh = (float)R.th_hi + 2e-3f;
        // the same test in cosines (no acos / asin per site) where the band edges lie in
        // (0, pi): cos is strictly decreasing there
        const bool cos_test = thl < 3.1f && thh > 0.05f;
        const bool hi_open = thh >= 3.14159265f, lo_open = thl <= 0.0f;
        const float chh = cosf(fminf(thh, 3.14159265f)), cll = cosf(fmaxf(thl, 0.0f));
        const KnCoef kc = kn_coef(p.e0);
        const int nch = site_chunks(p), csz = site_chunk_size(p);
        int cap = (p.site_cap > 0) ? p.site_cap : 0x7fffffff;
        const bool table = p.mode == 0 || p.mode == 3;
        for (int c = 0; c < nch && cap > 0; ++c) {
            int n_c = counts[(long long)rr * nch + c];
            if (n_c > cap) n_c = cap;
            cap -= n_c;
            const ScPair* P = pairs + (long long)rr * p.max_sites + (long long)c * csz;
            for (int j0 = 0; j0 < n_c; j0 += 64) {
                // the window-band test of 64 sites at once, one per lane: a site is dropped
                // for the whole wave when every voxel's scatter angle lies outside the band
                // (for I within trad of the centre T the angle between S - I and S - T is at
                // most asin(trad / |S - T|)); the kept sites then run in index order
                bool keep = false;
                if (j0 + lane < n_c) {
                    const ScPair& S = P[j0 + lane];
                    keep = true;
                    const float ax = (float)S.sx - tcx, ay = (float)S.sy - tcy, az = (float)S.sz - tcz;
                    const float la2 = ax * ax + ay * ay + az * az;
                    if (la2 > trad * trad) {
                        const float ila = __builtin_amdgcn_rsqf(la2);
                        const float c0 = fminf(
                            1.0f, fmaxf(-1.0f, (ax * (float)S.bx + ay * (float)S.by + az * (float)S.bz) * ila));
                        const float s1 = fminf(1.0f, trad * ila * 1.0001f);  // sin dth
                        if (cos_test) {
                            // th0 - dth > thh  <=>  th0 > dth and cos(th0 - dth) < cos thh;
                            // th0 + dth < thl  <=>  th0 + dth < pi and cos(th0 + dth) > cos thl
                            const float c1 = fsqrt(fmaxf(0.0f, 1.0f - s1 * s1));  // cos dth
                            const float s0 = fsqrt(fmaxf(0.0f, 1.0f - c0 * c0));  // sin th0
                            const bool above = !hi_open && c0 < c1 && c0 * c1 + s0 * s1 < chh;
                            const bool below = !lo_open && s0 * c1 + c0 * s1 > 0.0f && c0 * c1 - s0 * s1 > cll;
                            keep = !(above || below);
                        } else {
                            const float th0 = acosf(c0);
                            const float dth = asinf(s1);
                            keep = !(th0 - dth > thh || th0 + dth < thl);
                        }
                    }
                }
                unsigned long long m = __ballot(keep);
                if (!active) continue;
                // the PE gather of the next kept site is issued before the current site term
                float pv = 0.0f;
                if (table && m) pv = pe[(long long)P[j0 + __builtin_ctzll(m)].site * p.V + v];
                while (m) {
                    const int j = j0 + __builtin_ctzll(m);
                    m &= m - 1;
                    const float pv_j = pv;
                    if (table && m) pv = pe[(long long)P[j0 + __builtin_ctzll(m)].site * p.V + v];
                    const ScPair& S = P[j];
                    const double w = site_weight(p, R, S, q, pv_j);
                    acc += site_term<kSeries>(p, R, S, q, kc, w, s_kn);
                }
            }
        }
    }
    if (active) {
        float* o = dst + drow * p.row_stride + v;
        if (p.accumulate) *o += (float)acc;
        else *o = (float)acc;
    }
}

// Source weights of the collimator cells: W[c][v] = w_c / |c - I|^2 exp(-A_t(I -> c)),
// I the voxel centre (world).  Block b covers cell b / tiles, voxels 256 (b % tiles) + lane.
__global__ void __launch_bounds__(256) cell_weight_kernel(ScParams p, long long tiles, float* __restrict__ W) {
    const long long b = blockIdx.x;
    const int c = (int)(b / tiles);
    const long long v = (b % tiles) * 256 + threadIdx.x;
    if (v >= p.V) return;
    const int X = p.X, Y = p.Y;
    const int iz = (int)(v / ((long long)X * Y));
    const int rem = (int)(v % ((long long)X * Y));
    const int iy = rem / X, ix = rem % X;
    const double x0 = ((double)ix - (double)X / 2.0 + 0.5) * p.wx + p.sx;
    const double y0 = ((double)iy - (double)Y / 2.0 + 0.5) * p.wy + p.sy;
    const double zI = ((double)iz - (double)p.Z / 2.0 + 0.5) * p.wz + p.sz;
    const double xI = x0 * p.rc - y0 * p.rs, yI = x0 * p.rs + y0 * p.rc;
    const Cell C = p.cells[c];
    const double dx = C.x - xI, dy = C.y - yI, dz = C.z - zI;
    const PathAtt a = plate_path(p.cplates, p.n_cplates, p.choles, xI, yI, zI, C.x, C.y, C.z);
    W[(long long)c * p.V + v] = (float)(C.w / (dx * dx + dy * dy + dz * dz) * exp(-a.t));
}

// Scatter-surrogate ranking: one workgroup per detector row; every candidate
// site's pair constants and its single-site weight at the image centre (unit
// PE); a bitonic sort in LDS keeps the best n_keep, best first.
__global__ void __launch_bounds__(256) surrogate_rank_kernel(ScParams p, SurrogateRank k,
                                                             const ScRow* __restrict__ rows,
                                                             const Crystal* __restrict__ cry,
                                                             const unsigned* __restrict__ bits,
                                                             ScPair* __restrict__ keep, int* __restrict__ counts) {
    __shared__ float s_key[256];
    __shared__ int s_idx[256];
    __shared__ int s_rank[256];
    const int rr = blockIdx.x;
    const int r = p.det_begin + rr;
    const int t = threadIdx.x;
    const ScRow R = rows[r];
    const ImageSphere B = image_sphere(p);
    ScPair P;
    float key = -1.0f;
    if (R.use != 0.0f && t < k.n_cand) {
        const int s = k.cand[(long long)r * k.n_cand + t];
        if (s >= 0 && s != r && make_pair(p, R, r, s, cry, nullptr, bits, B, P)) {
            const VoxelPt q = voxel_point(p, R, p.sx, p.sy, p.sz);
            const double wpe = (double)(float)(R.mu_c / R.mu_pe);
            const double g = (p.e0 / 511.0 <= kKnSeriesMaxAlpha)
                                 ? site_term<true>(p, R, P, q, kn_coef(p.e0), wpe, p.kn_tab)
                                 : site_term<false>(p, R, P, q, kn_coef(p.e0), wpe, p.kn_tab);
            key = g > 0.0 ? (float)g : 0.0f;
            P.pad = key;
            // weight without its angular factor at the image centre (voxel-dependent ring prior)
            const double ax = P.sx - q.xI, ay = P.sy - q.yI, az = P.sz - q.zI;
            const float c0 = (float)((ax * P.bx + ay * P.by + az * P.bz) / sqrt(ax * ax + ay * ay + az * az));
            const float f0 = k.ftab ? sg_angular(k.ftab + (long long)r * (kSgTab + 1), c0) : 0.0f;
            P.hy1 = f0 > 1e-30f ? key / f0 : 0.0f;
        }
    }
    s_key[t] = key;
    s_idx[t] = t;
    __syncthreads();
    // bitonic sort, descending by key (ties: lower candidate index first)
    for (int size = 2; size <= 256; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            const int o = t ^ stride;
            if (o > t) {
                const bool desc = (t & size) == 0;
                const float ka = s_key[t], kb = s_key[o];
                const int ia = s_idx[t], ib = s_idx[o];
                const bool a_first = ka > kb || (ka == kb && ia < ib);
                if (desc != a_first) {
                    s_key[t] = kb; s_key[o] = ka;
                    s_idx[t] = ib; s_idx[o] = ia;
                }
            }
            __syncthreads();
        }
    }
    s_rank[s_idx[t]] = t;
    __syncthreads();
    const int rank = s_rank[t];
    if (key >= 0.0f && rank < k.n_keep) keep[(long long)rr * k.n_keep + rank] = P;
    if (t == 0) {
        int c = 0;
        while (c < k.n_keep && c < 256 && s_key[c] >= 0.0f) ++c;
        counts[rr] = c;
    }
}

hipError_t launch_cell_weights(const ScParams& p, int n_cells, float* weights, hipStream_t stream) {
    if (n_cells <= 0 || p.V <= 0) return hipSuccess;
    const long long tiles = (p.V + 255) / 256;
    const long long max_blocks = (1LL << 32) / 256 - 8;
    const int per = (int)std::max(1LL, std::min<long long>(n_cells, max_blocks / tiles));
    for (int c0 = 0; c0 < n_cells; c0 += per) {
        ScParams q = p;
        q.cells = p.cells + c0;
        const int cnt = std::min(per, n_cells - c0);
        hipLaunchKernelGGL(cell_weight_kernel, dim3((unsigned)(tiles * cnt)), dim3(256), 0, stream, q, tiles,
                           weights + (long long)c0 * p.V);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
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
    if (p.site_chunk != kPrepChunk) return hipErrorInvalidValue;  // pair tables come in kPrepChunk chunks
    const long long blocks = (long long)p.det_count * site_chunks(p);
    hipLaunchKernelGGL(scatter_prep_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, rows, cry, hol, bits,
                       pairs, counts);
    return hipGetLastError();
}

hipError_t launch_scatter(const ScParams& p, const ScRow* rows, const ScPair* pairs, const int* counts,
                          const float* pe, float* dst, hipStream_t stream) {
    const long long max_blocks = (1LL << 32) / 256 - 8;  // gridDim.x * blockDim.x < 2^32
    const auto scatter_kernel_e = (p.e0 / 511.0 <= kKnSeriesMaxAlpha) ? scatter_kernel<true> : scatter_kernel<false>;
    if (p.items) {
        for (long long i0 = 0; i0 < p.n_items; i0 += max_blocks * 4) {
            ScParams q = p;
            q.items = p.items + i0;
            q.n_items = std::min(p.n_items - i0, max_blocks * 4);
            hipLaunchKernelGGL(scatter_kernel_e, dim3((unsigned)((q.n_items + 3) / 4)), dim3(256), 0, stream, q, rows,
                               pairs, counts, pe, dst);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    if (p.n_blocks <= 0) return hipSuccess;
    // detector chunks (multiples of 4 rows) small enough for one grid
    const long long rg_max = std::max(1LL, max_blocks / std::max(1, p.n_tiles));
    const int chunk = (int)std::min<long long>(p.det_count, rg_max * 4);
    for (int d0 = 0; d0 < p.det_count; d0 += chunk) {
        ScParams q = p;
        q.det_count = std::min(chunk, p.det_count - d0);
        if (p.row_list) {
            q.row_list = p.row_list + d0;
        } else {
            q.det_begin = p.det_begin + d0;
        }
        q.n_rgroups = (q.det_count + 3) / 4;
        q.n_blocks = (long long)q.n_tiles * q.n_rgroups;
        const bool shift_dst = !(p.row_list && p.dst_by_row);
        const long long grid = (q.n_blocks + 7) / 8 * 8;
        hipLaunchKernelGGL(scatter_kernel_e, dim3((unsigned)grid), dim3(256), 0, stream, q, rows,
                           pairs + (long long)d0 * p.max_sites, counts + (long long)d0 * site_chunks(p), pe,
                           shift_dst ? dst + (long long)d0 * p.row_stride : dst);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_surrogate_rank(const ScParams& p, const SurrogateRank& k, const ScRow* rows, const Crystal* cry,
                                 const unsigned* bits, ScPair* keep, int* counts, hipStream_t stream) {
    if (p.det_count <= 0) return hipSuccess;
    hipLaunchKernelGGL(surrogate_rank_kernel, dim3((unsigned)p.det_count), dim3(256), 0, stream, p, k, rows, cry, bits,
                       keep, counts);
    return hipGetLastError();
}

}  // namespace sysmat
