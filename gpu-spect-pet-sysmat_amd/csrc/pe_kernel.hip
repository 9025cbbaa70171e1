// Photoelectric (PE) system-matrix tracer for gfx950.
//
// Semantics: photodetectorCudaMe (PEGen_RayTracing_CircularHole/PESysMatGen.cu:375-764).
// One matrix element (detector d, voxel v) = sum over sub-rays from the voxel
// (centre, or n_v^3 averaged sub-points) to the crystal sub-cell centres of
// solid angle x exp(-sum mu L) x absorption, where the attenuators are the
// plates, the holes and the other crystals that pass the reference's
// per-element angular-sphere cull (PESysMatGen.cu:504-590).
//
// Schedule (MI355X-first, not the reference's one-thread-per-element grid):
//   * once per plan, each detector's candidate crystals / holes for the whole
//     image slab (pe_detector_cull_kernel);
//   * one workgroup = 4 wave64s = one detector x a group of 2 x 2 x 2 voxel
//     tiles (one tile when a launch has few tiles); the workgroup culls the
//     detector's list for the group into LDS, then per tile that list;
//   * one wave = one element at a time, its 64 lanes = 64 sub-rays.  Per
//     element the wave filters the tile candidates lane-parallel and compacts
//     the survivors' boxes into LDS, so the attenuator loop is uniform across
//     the wave;
//   * every level (image slab, tile group, tile, element) applies the
//     angular-sphere bound (a proven superset of the reference's per-element
//     cull) and the "bundle" bound (bundle_boxes: per y-slab lateral limits of
//     every sub-ray segment; outside them a crystal or hole has a zero chord
//     for every sub-ray); the element level then applies the reference cull
//     predicate itself.  Candidate lists stay in index order, so each element
//     sums its attenuators in the same order at every tile shape;
//   * per-workgroup LDS table of the sub-ray index decomposition and crystal
//     sub-cell geometry (S <= 256);
//   * FP32 tracing; FP64 for the element accumulator and the collimator path
//     (hole chords and plate length from FP64 end points).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pe_kernel.hpp"
#include "trace.hpp"

namespace sysmat {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kListCap = 64;    // per-wave compacted attenuator boxes (beyond: bitmask path)
constexpr int kHoleCap = 48;    // per-wave compacted holes (beyond: bitmask path)
constexpr int kSubTab = 256;    // sub-rays tabulated per workgroup with their geometry (S <= kSubTab)
constexpr int kIdxTab = 2048;   // sub-ray indices only, in the same LDS (kSubTab < S <= kIdxTab)
constexpr int kElTab = 64;      // tile elements whose voxel-centre constants are tabulated per tile

// Per-sub-ray constants of a workgroup (one detecting crystal): the index
// decomposition and the crystal sub-cell geometry (PESysMatGen.cu:606-616, 727-759).
struct alignas(16) SubCell {
    unsigned idx0, idx1;                 // ny | nx << 16, nz | va << 8 | vb << 16 | vc << 24
    float xDs, yDs, zDs;                 // sub-cell centre, crystal frame
    float xQ, yQ, zQ;                    // the same in the world
    float ux1, ux2, uy1, uy2, uz1, uz2;  // sub-cell box, crystal frame
    float pad0, pad1;
};

__device__ inline void sub_cell_geometry(const Crystal& D, int nx, int ny, int nz, float wD, float tD, float hD,
                                         float rcx, float rcy, float rcz, float cD, float sD, SubCell& c) {
    c.xDs = -wD / 2.0f + ((float)nx + 0.5f) * rcx * wD;
    c.zDs = -hD / 2.0f + ((float)nz + 0.5f) * rcz * hD;
    c.yDs = -tD / 2.0f + ((float)ny + 0.5f) * rcy * tD;
    c.xQ = D.cx + (c.xDs * cD - c.zDs * sD);
    c.zQ = D.cz + (c.xDs * sD + c.zDs * cD);
    c.yQ = D.cy + c.yDs;
    c.ux1 = ((float)nx * rcx - 0.5f) * wD;
    c.ux2 = (((float)nx + 1.0f) * rcx - 0.5f) * wD;
    c.uy1 = ((float)ny * rcy - 0.5f) * tD;
    c.uy2 = (((float)ny + 1.0f) * rcy - 0.5f) * tD;
    c.uz1 = ((float)nz * rcz - 0.5f) * hD;
    c.uz2 = (((float)nz + 1.0f) * rcz - 0.5f) * hD;
}

__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// a / b from rb = 1 / b (correctly rounded): q = a rb, then one residual
// correction with FMA (Markstein's final step), for |a| and b well inside the
// normal range.
__device__ inline double div_rcp(double a, double b, double rb) {
    const double q = a * rb;
    return fma(fma(-q, b, a), rb, q);
}

// Wave-uniform copy of an integer (bit-preserving for unsigned).
__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline unsigned uniform(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }

// q = s / d, r = s % d for 0 <= s < 2^24 via a float reciprocal and one correction.
__device__ inline int divmod(int s, int d, float rd, int& r) {
    int q = (int)((float)s * rd);
    int m = s - q * d;
    if (m < 0) { q -= 1; m += d; }
    else if (m >= d) { q += 1; m -= d; }
    r = m;
    return q;
}

// Voxel centre (before image rotation) of voxel (ix, iy, iz) -- PESysMatGen.cu:438-444.
__device__ inline void voxel_centre(const PeParams& p, int ix, int iy, int iz, float& x, float& y, float& z) {
    x = ((float)ix - (float)p.X / 2.0f + 0.5f) * p.wx + p.sx;
    y = ((float)iy - (float)p.Y / 2.0f + 0.5f) * p.wy + p.sy;
    z = ((float)iz - (float)p.Zfull / 2.0f + 0.5f) * p.wz + p.sz;
}

__device__ inline void rotate_image(const PeParams& p, float x, float y, float& xr, float& yr) {
    xr = x * p.rc - y * p.rs;
    yr = x * p.rs + y * p.rc;
}

// Voxel-centre quantities of element e of a tile (nvx x nvy x .. voxels from
// (vx0, vy0, vz0)) seen from detecting crystal D: local voxel index, centre
// before and after the image rotation, and the reference's cull terms -- the
// unit vector towards D and r_D / (2 |D - I|) (PESysMatGen.cu:438-444, 506-589).
struct ElementConsts {
    int lx, ly, lz;
    float xc0, yc0, zc0, xI, yI;
    float crit_self, ux, uy, uz;
};
__device__ inline void element_consts(const PeParams& p, const Crystal& D, int e, int nvx, int nvy, int vx0, int vy0,
                                      int vz0, ElementConsts& c) {
    c.lx = e % nvx;
    c.ly = (e / nvx) % nvy;
    c.lz = e / (nvx * nvy);
    voxel_centre(p, vx0 + c.lx, vy0 + c.ly, p.z0 + vz0 + c.lz, c.xc0, c.yc0, c.zc0);
    rotate_image(p, c.xc0, c.yc0, c.xI, c.yI);
    const float zI = c.zc0;
    const float dist0 =
        sqrtf((D.cy - c.yI) * (D.cy - c.yI) + (D.cx - c.xI) * (D.cx - c.xI) + (D.cz - zI) * (D.cz - zI));
    c.crit_self = D.rbound / dist0 / 2.0f;
    c.ux = (D.cx - c.xI) / dist0;
    c.uy = (D.cy - c.yI) / dist0;
    c.uz = (D.cz - zI) / dist0;
}

// Compacted attenuator (32 B in LDS): world box and mu for crystals with no
// rotation; rotated crystals carry their index and are traced in their frame.
struct Att {
    float x1, x2, y1, y2, z1, z2, mu;
    int k;  // crystal index; negative (-(k+1)) when the crystal is rotated
};

// Fast form of the reference cull predicate |e/|e| - u| <= r/|e| + crit_self
// (PESysMatGen.cu:506-589) with hardware rsqrt: +1 accept, -1 reject, 0 when
// the two sides are closer than the error bound of this form (2e-6 absolute +
// 1e-5 relative, against < 1.5e-6 from the approximations: unit-vector
// components of magnitude <= 1 with a few ulp each), in which case the caller
// evaluates the reference's operations.
__device__ inline int cull_fast(float ex, float ey, float ez, float r, float ux, float uy, float uz, float crit_self) {
    const float d2 = ex * ex + ey * ey + ez * ez;
    const float id = __builtin_amdgcn_rsqf(d2);
    const float vx = ex * id - ux, vy = ey * id - uy, vz = ez * id - uz;
    const float l2 = vx * vx + vy * vy + vz * vz;
    const float rhs = r * id + crit_self;
    const float m = 1e-5f * rhs + 2e-6f;
    const float lo = rhs - m, hi = rhs + m;
    if (lo > 0.0f && l2 < lo * lo) return 1;
    if (l2 > hi * hi) return -1;
    return 0;
}

// Chord of the sub-ray P -> Q through crystal k (reference order: PESysMatGen.cu:687-725).
__device__ inline float crystal_chord(const WRay& R, float qx, float qy, float qz, const Crystal* __restrict__ cry,
                                      int k, float x1, float x2, float y1, float y2, float z1, float z2,
                                      bool rotated) {
    if (!rotated && R.pat == 0) return chord_world_q(R, qx, qy, qz, x1, x2, y1, y2, z1, z2);
    const Crystal K = cry[k];
    if (!rotated) {  // along-axis branches, compared in the crystal frame
        const float px = R.px - K.cx, py = R.py - K.cy, pz = R.pz - K.cz;
        return chord_special(px, py, pz, px + R.dx, py + R.dy, pz + R.dz, R.ix, R.iy, R.iz, R.t, R.pat, -K.hx, -K.hy,
                             -K.hz, K.hx, K.hy, K.hz);
    }
    const float xPk = (R.px - K.cx) * K.cr - (R.pz - K.cz) * K.sr, yPk = R.py - K.cy;
    const float zPk = (R.px - K.cx) * K.sr + (R.pz - K.cz) * K.cr;
    const float xQk = (qx - K.cx) * K.cr - (qz - K.cz) * K.sr, yQk = qy - K.cy;
    const float zQk = (qx - K.cx) * K.sr + (qz - K.cz) * K.cr;
    const WRay Rk = make_wray(xPk, yPk, zPk, xQk, yQk, zQk);
    if (Rk.pat != 0)
        return chord_special(xPk, yPk, zPk, xQk, yQk, zQk, Rk.ix, Rk.iy, Rk.iz, Rk.t, Rk.pat, -K.hx, -K.hy, -K.hz,
                             K.hx, K.hy, K.hz);
    return chord_world_q(Rk, xQk, yQk, zQk, -K.hx, K.hx, -K.hy, K.hy, -K.hz, K.hz);
}

// Superset cull of a sphere of voxel centres (centre t, radius rho) against
// detector D.  For any voxel centre I within rho of t, the reference predicate
// |u_k - u_d| <= r_k/(2 D_k) + r_d/(2 D_d) implies
//   |u_k(t) - u_d(t)| <= r_k/(2 (D_k0 - rho)) + r_d/(2 (D_d0 - rho))
//                        + 2 rho/(2 D_k0 - rho) + 2 rho/(2 D_d0 - rho)
// by the Dunkl-Williams inequality |a/|a| - b/|b|| <= 2 |a - b| / (|a| + |b|)
// with |a - b| <= rho, |a| >= D0 - rho, |b| = D0.
// y-range: sub-points of crystal d lie at y <= cy + (t/2)(1 - 1/n_y); voxel
// sub-points at y >= (voxel centre y) - pey.  When every sub-ray runs toward +y
// by at least 0.01 mm (so no sub-ray takes the y-parallel branches), a crystal
// whose y-range lies beyond the sub-points, or before the voxel sub-points, by
// more than the 1e-3 mm margin (>> FP32 rounding of the slab parameters at these
// distances) has a zero chord for every sub-ray: its y-slab interval starts
// after |PQ| or ends before 0 (length_box_ray keeps only chords strictly inside
// the segment, PESysMatGen.cu:19-151).
struct SphereCull {
    float tcx, tcy, tcz, rho;
    float udx, udy, udz, crit_d;
    float q_ymax, pymin;
    bool yclip;
};

__device__ inline SphereCull sphere_cull(const PeParams& p, const Crystal& D, float tcx, float tcy, float tcz,
                                         float rho) {
    SphereCull c;
    c.tcx = tcx; c.tcy = tcy; c.tcz = tcz; c.rho = rho;
    const float hvx0 = 0.5f * p.wx * (1.0f - 1.0f / (float)p.nv);
    const float hvy0 = 0.5f * p.wy * (1.0f - 1.0f / (float)p.nv);
    const float pey0 = fabsf(p.rs) * hvx0 + fabsf(p.rc) * hvy0;
    c.q_ymax = D.cy + D.hy * (1.0f - 1.0f / (float)p.cy) + 1e-3f;
    c.pymin = tcy - rho - pey0 - 1e-3f;
    c.yclip = (D.cy - D.hy) - (tcy + rho + pey0) > 0.01f;
    const float ddx = D.cx - tcx, ddy = D.cy - tcy, ddz = D.cz - tcz;
    const float Dd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
    c.udx = ddx / Dd; c.udy = ddy / Dd; c.udz = ddz / Dd;
    c.crit_d = (Dd > rho) ? 0.5f * D.rbound / (Dd - rho) + 2.0f * rho / (2.0f * Dd - rho) : 4.0f;
    return c;
}

__device__ inline bool sphere_take_crystal(const SphereCull& c, const Crystal* __restrict__ cry, int k) {
    const float4 c0 = reinterpret_cast<const float4*>(cry + k)[0];
    const float rb = cry[k].rbound;
    const float khy = cry[k].hy;
    const float kx = c0.x - c.tcx, ky = c0.y - c.tcy, kz = c0.z - c.tcz;
    const float Dk = sqrtf(kx * kx + ky * ky + kz * kz);
    if (c.yclip && (c0.y - khy >= c.q_ymax || c0.y + khy <= c.pymin)) return false;
    if (Dk <= c.rho) return true;
    const float iD = 1.0f / Dk;
    const float bound = 0.5f * rb / (Dk - c.rho) + 2.0f * c.rho / (2.0f * Dk - c.rho) + c.crit_d;
    const float vx = kx * iD - c.udx, vy = ky * iD - c.udy, vz = kz * iD - c.udz;
    return sqrtf(vx * vx + vy * vy + vz * vz) <= bound * 1.0001f + 1e-5f;
}

__device__ inline bool sphere_take_hole(const SphereCull& c, const Hole& H) {
    const float hx = H.x - c.tcx, hy = 0.5f * (H.y1 + H.y2) - c.tcy, hz = H.z - c.tcz;
    const float Dh = sqrtf(hx * hx + hy * hy + hz * hz);
    if (Dh <= c.rho) return true;
    const float iD = 1.0f / Dh;
    const float bound = H.rbound / (Dh - c.rho) + 2.0f * c.rho / (2.0f * Dh - c.rho) + c.crit_d;
    const float vx = hx * iD - c.udx, vy = hy * iD - c.udy, vz = hz * iD - c.udz;
    return sqrtf(vx * vx + vy * vy + vz * vz) <= bound * 1.0001f + 1e-5f;
}

// Bounding sphere of the voxel centres of the box [v0, v1) (slab-local z),
// in the rotated image frame.
__device__ inline void box_sphere(const PeParams& p, int vx0, int vy0, int vz0, int vx1, int vy1, int vz1, float& tcx,
                                  float& tcy, float& tcz, float& rho) {
    float ax, ay, az, bx, by, bz;
    voxel_centre(p, vx0, vy0, p.z0 + vz0, ax, ay, az);
    voxel_centre(p, vx1 - 1, vy1 - 1, p.z0 + vz1 - 1, bx, by, bz);
    rotate_image(p, 0.5f * (ax + bx), 0.5f * (ay + by), tcx, tcy);
    tcz = 0.5f * (az + bz);
    rho = 0.5f * sqrtf((bx - ax) * (bx - ax) + (by - ay) * (by - ay) + (bz - az) * (bz - az)) * 1.0001f + 1e-4f;
}

// Rotated-frame centre and half extents of the voxel centres of the box [v0, v1)
// (slab-local z).
__device__ inline void box_extent(const PeParams& p, int vx0, int vy0, int vz0, int vx1, int vy1, int vz1, float& cx,
                                  float& cy, float& cz, float& hx, float& hy, float& hz) {
    float ax, ay, az, bx, by, bz;
    voxel_centre(p, vx0, vy0, p.z0 + vz0, ax, ay, az);
    voxel_centre(p, vx1 - 1, vy1 - 1, p.z0 + vz1 - 1, bx, by, bz);
    rotate_image(p, 0.5f * (ax + bx), 0.5f * (ay + by), cx, cy);
    cz = 0.5f * (az + bz);
    const float ux = 0.5f * (bx - ax), uy = 0.5f * (by - ay);
    hx = fabsf(p.rc) * ux + fabsf(p.rs) * uy;
    hy = fabsf(p.rs) * ux + fabsf(p.rc) * uy;
    hz = 0.5f * (bz - az);
}

// Lateral (x, z) bounds, per y-slab, of every segment from a point of box P
// (centre pc, half extents ph) to a point of box Q (centre qc, half extents qh),
// for P entirely before Q in y.  A segment's lateral position at height y is
// a + (b - a) t, t = clamp((y - Py) / (Qy - Py), 0, 1), with a, b its lateral end
// points: non-decreasing in a and b, and t non-increasing in Py and Qy and
// non-decreasing in y (a segment covers no y outside [Py, Qy]).  So the minimum
// over the boxes and the slab takes the low lateral ends and, when they rise
// toward Q, the smallest t (high Py, Qy, lower face) -- else the largest; the
// maximum likewise with the high ends.  One lane per (slab, axis): lanes 0..15
// do x, lanes 32..47 z.  The 2e-3 mm + 1e-5 relative margin covers the rounding
// of these evaluations and of the points themselves.  The whole wave calls this;
// box[s] = (x lo, x hi, z lo, z hi).  A crystal or hole of slab s outside box[s]
// has a zero chord on every segment.
__device__ inline void bundle_boxes(const PeParams& p, const Slab* __restrict__ slb, int lane, float pcx, float pcy,
                                    float pcz, float phx, float phy, float phz, float qcx, float qcy, float qcz,
                                    float qhx, float qhy, float qhz, float4* box) {
    const bool zaxis = lane >= 32;
    const int s = lane & 31;
    if (s >= p.n_slabs) return;
    const Slab SL = slb[s];
    const float pc = zaxis ? pcz : pcx, ph = zaxis ? phz : phx;
    const float qc = zaxis ? qcz : qcx, qh = zaxis ? qhz : qhx;
    const float Pylo = pcy - phy, Pyhi = pcy + phy, Qylo = qcy - qhy, Qyhi = qcy + qhy;
    const float Pl = pc - ph, Ql = qc - qh, Ph = pc + ph, Qh = qc + qh;
    const bool up_l = Ql >= Pl, up_h = Qh >= Ph;
    const float ay = up_l ? Pyhi : Pylo, aq = up_l ? Qyhi : Qylo, af = up_l ? SL.y1 : SL.y2;
    const float by = up_h ? Pylo : Pyhi, bq = up_h ? Qylo : Qyhi, bf = up_h ? SL.y2 : SL.y1;
    // slopes through the hardware reciprocal (<= 2.4e-7 relative; the bound moves by at
    // most |Q - P| x 2.4e-7, far inside the margin)
    const float mn = Pl + ((Ql - Pl) * frcp(aq - ay)) * (fminf(fmaxf(af, ay), aq) - ay);
    const float mx = Ph + ((Qh - Ph) * frcp(bq - by)) * (fminf(fmaxf(bf, by), bq) - by);
    const float mg = 2e-3f + 1e-5f * fmaxf(fabsf(mn), fabsf(mx));
    float* b = reinterpret_cast<float*>(box + s) + (zaxis ? 2 : 0);
    b[0] = mn - mg;
    b[1] = mx + mg;
}

// Crystal k / hole H against the bundle boxes (true: may have a non-zero chord).
__device__ inline bool bundle_take_crystal(const Crystal* __restrict__ cry, int k, const float4* box) {
    const float4 c0 = reinterpret_cast<const float4*>(cry + k)[0];  // cx, cy, cz, mu_t
    const float4 c1 = reinterpret_cast<const float4*>(cry + k)[1];  // hx, hy, hz, rot
    const float4 c3 = reinterpret_cast<const float4*>(cry + k)[3];  // mu_c, res, slab, axis
    if (c3.z < 0.0f) return true;
    const float4 bb = box[(int)c3.z];
    const float ex = (c3.w != 0.0f) ? c1.x : cry[k].hxz;
    const float ez = (c3.w != 0.0f) ? c1.z : ex;
    return !(c0.x + ex < bb.x || c0.x - ex > bb.y || c0.z + ez < bb.z || c0.z - ez > bb.w);
}
__device__ inline bool bundle_take_hole(const Hole& H, const float4* box) {
    if (H.slab < 0.0f) return true;
    const float4 bb = box[(int)H.slab];
    return !(H.x + H.r < bb.x || H.x - H.r > bb.y || H.z + H.r < bb.z || H.z - H.r > bb.w);
}

// Tile group grp: tiles [tx0, tx1) x [ty0, ty1) x [tz0, tz1), voxels [vx0, vx1) x ...
struct GroupBox {
    int tx0, ty0, tz0, tx1, ty1, tz1;
    int vx0, vy0, vz0, vx1, vy1, vz1;
};

__device__ inline GroupBox group_box(const PeParams& p, long long grp) {
    GroupBox G;
    G.tx0 = (int)(grp % p.ngx) * p.gx;
    G.ty0 = (int)((grp / p.ngx) % p.ngy) * p.gy;
    G.tz0 = (int)(grp / ((long long)p.ngx * p.ngy)) * p.gz;
    G.tx1 = min(G.tx0 + p.gx, p.ntx);
    G.ty1 = min(G.ty0 + p.gy, p.nty);
    G.tz1 = min(G.tz0 + p.gz, p.ntz);
    G.vx0 = G.tx0 * p.tx;
    G.vy0 = G.ty0 * p.ty;
    G.vz0 = G.tz0 * p.tz;
    G.vx1 = min(G.tx1 * p.tx, p.X);
    G.vy1 = min(G.ty1 * p.ty, p.Y);
    G.vz1 = min(G.tz1 * p.tz, p.Zslab);
    return G;
}

// Half extents of the sub-point clouds: voxel sub-points around a voxel centre
// (rotated frame) and the sub-cell centres of crystal D around its centre (world
// AABB of the rotated cloud).
struct SubBoxes {
    float pex, pey, pez;  // voxel sub-points
    float qex, qhy, qez;  // sub-cell centres of D
};

__device__ inline SubBoxes sub_boxes(const PeParams& p, const Crystal& D, float cD, float sD) {
    SubBoxes b;
    const float hvx = 0.5f * p.wx * (1.0f - 1.0f / (float)p.nv);
    const float hvy = 0.5f * p.wy * (1.0f - 1.0f / (float)p.nv);
    b.pez = 0.5f * p.wz * (1.0f - 1.0f / (float)p.nv);
    b.pex = fabsf(p.rc) * hvx + fabsf(p.rs) * hvy;
    b.pey = fabsf(p.rs) * hvx + fabsf(p.rc) * hvy;
    const float qhx = D.hx * (1.0f - 1.0f / (float)p.cx), qhz = D.hz * (1.0f - 1.0f / (float)p.cz);
    b.qhy = D.hy * (1.0f - 1.0f / (float)p.cy);
    b.qex = fabsf(cD) * qhx + fabsf(sD) * qhz;
    b.qez = fabsf(sD) * qhx + fabsf(cD) * qhz;
    return b;
}

// Bundle boxes of every sub-ray from the voxel sub-points of the box [v0, v1) to
// the sub-cell centres of crystal D (half extents pex/pey/pez, qex/qhy/qez) into
// box, computed by wave 0; returns whether they apply (P entirely more than 1 mm
// before Q in y).  Every thread of the block calls it (one barrier).
__device__ inline bool group_bundle(const PeParams& p, const Slab* __restrict__ slb, const Crystal& D, int vx0,
                                    int vy0, int vz0, int vx1, int vy1, int vz1, float pex, float pey, float pez,
                                    float qex, float qhy, float qez, int lane, int wave, float4* box) {
    float cx, cy, cz, hx, hy, hz;
    box_extent(p, vx0, vy0, vz0, vx1, vy1, vz1, cx, cy, cz, hx, hy, hz);
    const bool ok = p.use_bundle && ((D.cy - qhy) - (cy + hy + pey) > 1.0f);
    if (ok && wave == 0)
        bundle_boxes(p, slb, lane, cx, cy, cz, hx + pex, hy + pey, hz + pez, D.cx, D.cy, D.cz, qex, qhy, qez, box);
    __syncthreads();
    return ok;
}

// Ordered block-wide compaction step: every thread offers `take`; returns the
// slot of a taken item (from `base_count`) and adds the block's count to it.
// Needs s_wcount[kWaves] and all threads of the block.
__device__ inline int block_compact(bool take, int& base_count, int* s_wcount, int lane, int wave,
                                    unsigned long long below) {
    const unsigned long long bal = __ballot(take);
    if (lane == 0) s_wcount[wave] = __popcll(bal);
    __syncthreads();
    int off = base_count, all = 0;
    for (int w2 = 0; w2 < kWaves; ++w2) {
        const int c = s_wcount[w2];
        if (w2 < wave) off += c;
        all += c;
    }
    base_count += all;
    __syncthreads();
    return off + __popcll(bal & below);
}

// Per-detector candidate lists for the whole image slab (one workgroup per
// detector): the crystals and holes that pass the sphere cull of the slab's
// voxel-centre sphere, in index order.  Every tile sphere lies inside the slab
// sphere, and both are supersets of every element's exact cull list, so the
// tile cull may scan this list instead of the whole table and still keep every
// element's list (and its index order) unchanged.  fill = 0 counts into
// counts[2 d], counts[2 d + 1]; fill = 1 writes the lists at the offsets.
__global__ void __launch_bounds__(kThreads) pe_detector_cull_kernel(PeParams p, const Crystal* __restrict__ cry,
                                                                    const Hole* __restrict__ hol,
                                                                    const Slab* __restrict__ slb, int fill,
                                                                    int* __restrict__ counts, PeLists L) {
    __shared__ int s_wcount[kWaves];
    __shared__ float4 s_box[kMaxSlabs];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int d = blockIdx.x;
    const Crystal D = cry[d];
    int nc = 0, nh = 0;
    if (D.mu_t > 0.01f) {
        float tcx, tcy, tcz, rho;
        box_sphere(p, 0, 0, 0, p.X, p.Y, p.Zslab, tcx, tcy, tcz, rho);
        const SphereCull C = sphere_cull(p, D, tcx, tcy, tcz, rho);
        const SubBoxes B = sub_boxes(p, D, cosf(D.rot), sinf(D.rot));
        const bool b_ok = group_bundle(p, slb, D, 0, 0, 0, p.X, p.Y, p.Zslab, B.pex, B.pey, B.pez, B.qex, B.qhy,
                                       B.qez, lane, wave, s_box);
        unsigned short* cl = fill ? L.cry_list + L.cry_ptr[d] : nullptr;
        unsigned short* hl = fill ? L.hol_list + L.hol_ptr[d] : nullptr;
        for (int base = 0; base < p.n_det; base += kThreads) {
            const int k = base + tid;
            const bool take = k < p.n_det && k != d && sphere_take_crystal(C, cry, k) &&
                              (!b_ok || bundle_take_crystal(cry, k, s_box));
            const int slot = block_compact(take, nc, s_wcount, lane, wave, below);
            if (take && fill) cl[slot] = (unsigned short)k;
        }
        for (int base = 0; base < p.n_holes; base += kThreads) {
            const int h = base + tid;
            bool take = false;
            if (h < p.n_holes) {
                const Hole H = hol[h];
                take = sphere_take_hole(C, H) && (!b_ok || bundle_take_hole(H, s_box));
            }
            const int slot = block_compact(take, nh, s_wcount, lane, wave, below);
            if (take && fill) hl[slot] = (unsigned short)h;
        }
    }
    if (!fill && tid == 0) {
        counts[2 * d] = nc;
        counts[2 * d + 1] = nh;
    }
}

__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3)))
pe_tile_kernel(PeParams p, const Crystal* __restrict__ cry, const Hole* __restrict__ hol, const Plate* __restrict__ pla,
               const Slab* __restrict__ slb, PeLists L, float* __restrict__ dst, int* __restrict__ status) {
    __shared__ unsigned short s_grp[kCapGroup];
    __shared__ unsigned short s_hgrp[kCapHoles];
    __shared__ unsigned short s_cand[kCapCrystals];
    __shared__ unsigned short s_hcand[kCapHoles];
    __shared__ unsigned s_mask[kWaves][kCapCrystals / 32];
    __shared__ unsigned s_hmask[kWaves][kCapHoles / 32];
    __shared__ Att s_att[kWaves][kListCap];
    __shared__ int s_hol[kWaves][kHoleCap];
    __shared__ float4 s_box[kWaves][kMaxSlabs];
    __shared__ float4 s_tbox[kMaxSlabs];  // bundle boxes of the tile group, then of each tile
    __shared__ int s_wcount[kWaves];
    __shared__ int s_meta[8];  // tile-group state, re-read per tile (keeps it out of scalar registers)
    __shared__ SubCell s_sub[kSubTab];
    __shared__ float4 s_elA[kElTab];  // per-element constants of the current tile (element_consts)
    __shared__ float4 s_elB[kElTab];
    __shared__ float2 s_elC[kElTab];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    // workgroup = one detector row x tile groups [grp_begin, grp_end): the detector's
    // set-up below (sub-cell table, bundle end boxes, plate sums) and the register
    // spills of that set-up are paid once for all of them
    const long long wg = blockIdx.x;
    const int drow = (int)(wg % p.det_count);
    const long long grp_begin = (wg / p.det_count) * p.grp_per_wg;
    const long long grp_end = min(grp_begin + p.grp_per_wg, (long long)p.ngx * p.ngy * p.ngz);
    const int d = p.det_begin + drow;

    const Crystal D = cry[d];
    float* row = dst + (long long)drow * p.row_stride;

    if (!(D.mu_t > 0.01f)) {  // no culling, no sub-rays: the element is 0 (PESysMatGen.cu:504, 598)
        for (long long grp = grp_begin; grp < grp_end; ++grp) {
            const GroupBox G = group_box(p, grp);
            const int nvx = G.vx1 - G.vx0, nvy = G.vy1 - G.vy0, nvox = nvx * nvy * (G.vz1 - G.vz0);
            for (int e = tid; e < nvox; e += kThreads) {
                const int lx = e % nvx, ly = (e / nvx) % nvy, lz = e / (nvx * nvy);
                row[(long long)(G.vz0 + lz) * p.X * p.Y + (G.vy0 + ly) * p.X + (G.vx0 + lx)] = 0.0f;
            }
        }
        return;
    }

    const int S = p.S;
    const float inv_nv3 = 1.0f / (float)(p.nv * p.nv * p.nv);
    // crystal-d sub-point geometry
    const float wD = 2.0f * D.hx, tD = 2.0f * D.hy, hD = 2.0f * D.hz;
    const float cD = cosf(D.rot), sD = sinf(D.rot);      // sub-points into the world
    const float cDn = cosf(-D.rot), sDn = sinf(-D.rot);  // world into the crystal frame
    const double cD64 = (D.rot == 0.0f) ? 1.0 : cos((double)D.rot);
    const double sD64 = (D.rot == 0.0f) ? 0.0 : sin((double)D.rot);
    const float fcx = (float)p.cx, fcy = (float)p.cy, fcz = (float)p.cz;
    const float rcx = 1.0f / fcx, rcy = 1.0f / fcy, rcz = 1.0f / fcz, rnv = 1.0f / (float)p.nv;
    const float A_x = tD * hD / fcz / fcy, A_z = wD * tD / fcx / fcy, A_y = wD * hD / fcx / fcz;
    const float absorb_scale = D.mu_pe / D.mu_t;
    // half extents of a crystal sub-cell (its centre is the sub-point Q)
    const float sub_hx = 0.5f * rcx * wD, sub_hy = 0.5f * rcy * tD, sub_hz = 0.5f * rcz * hD;
    const bool axis_d = D.rot == 0.0f;  // crystal frame = translated world frame
    // sum of mu * thickness over the plates (FP64 and FP32) for the plate path of
    // elements without candidate holes; plates under 1e-5 mm are handled per sub-ray
    double plate_mt = 0.0;
    float plate_mtf = 0.0f;
    bool thin_plates = false;
    for (int m = 0; m < p.n_layers; ++m) {
        const Plate PL = pla[m];
        if (PL.thick >= 0.00001f) {
            plate_mt += (double)PL.mu_t * (double)PL.thick;
            plate_mtf += PL.mu_t * PL.thick;
        } else {
            thin_plates = true;
        }
    }
    // bundle filter end boxes: voxel sub-points, sub-cell centres of crystal d
    const SubBoxes B = sub_boxes(p, D, cD, sD);
    const float pex = B.pex, pey = B.pey, hvz = B.pez, qex = B.qex, qhy = B.qhy, qez = B.qez;

    // sub-ray s -> (crystal sub-cell, voxel sub-point) indices and sub-cell
    // geometry; the group cull's barriers order these writes before the passes
    const bool sub_tab = S <= kSubTab;
    // indices only: nx, ny < 2^16 and nz, va, vb, vc < 2^8 packed in 8 bytes
    const bool idx_tab = !sub_tab && S <= kIdxTab && p.cx < 65536 && p.cy < 65536 && p.cz < 256 && p.nv < 256;
    uint2* s_idx = reinterpret_cast<uint2*>(s_sub);
    static_assert(sizeof(SubCell) * kSubTab >= sizeof(uint2) * kIdxTab, "index table shares the sub-cell table");
    if (sub_tab || idx_tab) {
        for (int s = tid; s < S; s += kThreads) {
            int ny, nx, nz, va = 0, vb = 0;
            int q = divmod(s, p.cy, rcy, ny);
            q = divmod(q, p.cx, rcx, nx);
            const int iv = divmod(q, p.cz, rcz, nz);
            const int t1 = divmod(iv, p.nv, rnv, va);
            const int vc = divmod(t1, p.nv, rnv, vb);
            SubCell c;
            c.idx0 = (unsigned)ny | ((unsigned)nx << 16);
            c.idx1 = (unsigned)nz | ((unsigned)va << 8) | ((unsigned)vb << 16) | ((unsigned)vc << 24);
            if (idx_tab) {
                s_idx[s] = make_uint2(c.idx0, c.idx1);
                continue;
            }
            sub_cell_geometry(D, nx, ny, nz, wD, tD, hD, rcx, rcy, rcz, cD, sD, c);
            c.pad0 = c.pad1 = 0.0f;
            s_sub[s] = c;
        }
    }

    for (long long grp = grp_begin; grp < grp_end; ++grp) {
        // ---- group cull: the group's voxel-centre sphere over the detector's
        //      slab-level candidates (pe_detector_cull_kernel) into LDS; each tile
        //      then culls these shorter lists (a group of one tile writes its tile
        //      lists directly).  Ordered compaction keeps every list sorted by index.
        //      The previous group's last barrier orders these LDS writes after its
        //      reads, and group_bundle's barrier orders the s_meta write below after
        //      every wave's last read of the previous group's s_meta.
        const GroupBox G = group_box(p, grp);
        const int tx0 = G.tx0, ty0 = G.ty0, tz0 = G.tz0, tx1 = G.tx1, ty1 = G.ty1, tz1 = G.tz1;
        const int gvx0 = G.vx0, gvy0 = G.vy0, gvz0 = G.vz0, gvx1 = G.vx1, gvy1 = G.vy1, gvz1 = G.vz1;
        const bool grouped = (tx1 - tx0) * (ty1 - ty0) * (tz1 - tz0) > 1;
        {
            int ng = 0, nhg = 0;
            float tcx, tcy, tcz, rho;
            box_sphere(p, gvx0, gvy0, gvz0, gvx1, gvy1, gvz1, tcx, tcy, tcz, rho);
            const SphereCull C = sphere_cull(p, D, tcx, tcy, tcz, rho);
            const bool gb_ok = group_bundle(p, slb, D, gvx0, gvy0, gvz0, gvx1, gvy1, gvz1, pex, pey, hvz, qex, qhy, qez,
                                            lane, wave, s_tbox);
            const long long c0 = L.cry_ptr[d], c1 = L.cry_ptr[d + 1];
            for (long long base = c0; base < c1; base += kThreads) {
                const long long i = base + tid;
                int k = 0;
                bool take = false;
                if (i < c1) {
                    k = (int)L.cry_list[i];
                    take = sphere_take_crystal(C, cry, k) && (!gb_ok || bundle_take_crystal(cry, k, s_tbox));
                }
                const int slot = block_compact(take, ng, s_wcount, lane, wave, below);
                if (take && slot < kCapGroup) {
                    if (grouped) s_grp[slot] = (unsigned short)k;
                    else s_cand[slot] = (unsigned short)k;
                }
            }
            const long long h0 = L.hol_ptr[d], h1 = L.hol_ptr[d + 1];
            for (long long base = h0; base < h1; base += kThreads) {
                const long long i = base + tid;
                int h = 0;
                bool take = false;
                if (i < h1) {
                    h = (int)L.hol_list[i];
                    const Hole H = hol[h];
                    take = sphere_take_hole(C, H) && (!gb_ok || bundle_take_hole(H, s_tbox));
                }
                const int slot = block_compact(take, nhg, s_wcount, lane, wave, below);
                if (take && slot < kCapHoles) {
                    if (grouped) s_hgrp[slot] = (unsigned short)h;
                    else s_hcand[slot] = (unsigned short)h;
                }
            }
            if (ng > kCapGroup || nhg > kCapHoles) {  // host retries with smaller tiles
                if (tid == 0) atomicOr(status, 1);
                return;
            }
            if (tid == 0) {
                s_meta[0] = tx0; s_meta[1] = ty0; s_meta[2] = tz0;
                s_meta[3] = tx1 - tx0; s_meta[4] = ty1 - ty0; s_meta[5] = tz1 - tz0;
                s_meta[6] = ng; s_meta[7] = nhg;
            }
            __syncthreads();
        }

        for (int t = 0; t < uniform(s_meta[3] * s_meta[4] * s_meta[5]); ++t) {
            const int gnx = uniform(s_meta[3]), gny = uniform(s_meta[4]);
            const int tix = uniform(s_meta[0]) + t % gnx, tiy = uniform(s_meta[1]) + (t / gnx) % gny;
            const int tiz = uniform(s_meta[2]) + t / (gnx * gny);
            const int vx0 = tix * p.tx, vy0 = tiy * p.ty, vz0 = tiz * p.tz;
            const int vx1 = min(vx0 + p.tx, p.X), vy1 = min(vy0 + p.ty, p.Y), vz1 = min(vz0 + p.tz, p.Zslab);
            const int nvx = vx1 - vx0, nvy = vy1 - vy0;
            const int nvox = nvx * nvy * (vz1 - vz0);

            // ---- tile cull: superset of every element's list ---------------------
            int nc = uniform(s_meta[6]), nh = uniform(s_meta[7]);
            if (grouped) {
                nc = 0;
                nh = 0;
                float tcx, tcy, tcz, rho;
                box_sphere(p, vx0, vy0, vz0, vx1, vy1, vz1, tcx, tcy, tcz, rho);
                const SphereCull C = sphere_cull(p, D, tcx, tcy, tcz, rho);
                const bool tb_ok = group_bundle(p, slb, D, vx0, vy0, vz0, vx1, vy1, vz1, pex, pey, hvz, qex, qhy, qez,
                                                lane, wave, s_tbox);
                const int ng = uniform(s_meta[6]), nhg = uniform(s_meta[7]);
                for (int base = 0; base < ng; base += kThreads) {
                    const int i = base + tid;
                    int k = 0;
                    bool take = false;
                    if (i < ng) {
                        k = (int)s_grp[i];
                        take = sphere_take_crystal(C, cry, k) && (!tb_ok || bundle_take_crystal(cry, k, s_tbox));
                    }
                    const int slot = block_compact(take, nc, s_wcount, lane, wave, below);
                    if (take && slot < kCapCrystals) s_cand[slot] = (unsigned short)k;
                }
                for (int base = 0; base < nhg; base += kThreads) {
                    const int i = base + tid;
                    int h = 0;
                    bool take = false;
                    if (i < nhg) {
                        h = (int)s_hgrp[i];
                        const Hole H = hol[h];
                        take = sphere_take_hole(C, H) && (!tb_ok || bundle_take_hole(H, s_tbox));
                    }
                    const int slot = block_compact(take, nh, s_wcount, lane, wave, below);
                    if (take && slot < kCapHoles) s_hcand[slot] = (unsigned short)h;
                }
            }
            if (nc > kCapCrystals || nh > kCapHoles) {  // host retries with smaller tiles
                if (tid == 0) atomicOr(status, 1);
                return;
            }
            // per-element constants of the tile, one thread per element (tiles of up to kElTab
            // voxels): the voxel-centre terms of element_consts once per element instead of on
            // every lane of its wave; larger tiles evaluate them in the element loop
            const bool el_tab = nvox <= kElTab;
            if (el_tab && tid < nvox) {
                ElementConsts c;
                element_consts(p, D, tid, nvx, nvy, vx0, vy0, vz0, c);
                s_elA[tid] = make_float4(c.xc0, c.yc0, c.zc0, c.xI);
                s_elB[tid] = make_float4(c.yI, c.crit_self, c.ux, c.uy);
                s_elC[tid] = make_float2(c.uz, __int_as_float(c.lx | (c.ly << 10) | (c.lz << 20)));
            }
            __syncthreads();  // the candidate lists and element constants are complete before any wave reads them
            const int ncw = (nc + 31) >> 5, nhw = (nh + 31) >> 5;

            for (int e = wave; e < nvox; e += kWaves) {
                ElementConsts ec;
                if (el_tab) {
                    const float4 a = s_elA[e], b = s_elB[e];
                    const float2 c = s_elC[e];
                    const int pk = __float_as_int(c.y);
                    ec.lx = pk & 0x3FF;
                    ec.ly = (pk >> 10) & 0x3FF;
                    ec.lz = pk >> 20;
                    ec.xc0 = a.x; ec.yc0 = a.y; ec.zc0 = a.z; ec.xI = a.w;
                    ec.yI = b.x; ec.crit_self = b.y; ec.ux = b.z; ec.uy = b.w; ec.uz = c.x;
                } else {
                    element_consts(p, D, e, nvx, nvy, vx0, vy0, vz0, ec);
                }
                const int ix = vx0 + ec.lx, iy = vy0 + ec.ly, izl = vz0 + ec.lz;
                const float xc0 = ec.xc0, yc0 = ec.yc0, zc0 = ec.zc0;
                const float xI = ec.xI, yI = ec.yI, zI = zc0;
                const float crit_self = ec.crit_self, ux = ec.ux, uy = ec.uy, uz = ec.uz;

                // -- bundle bounds per y-slab: P box (voxel sub-points), Q box (sub-cell centres of d)
                const float P_yhi = yI + pey, Q_ylo = D.cy - qhy;
                const bool bundle_ok = p.use_bundle && (Q_ylo - P_yhi > 1.0f);
                if (bundle_ok)
                    bundle_boxes(p, slb, lane, xI, yI, zI, pex, pey, hvz, D.cx, D.cy, D.cz, qex, qhy, qez, s_box[wave]);
                __builtin_amdgcn_wave_barrier();

                // -- crystals: bundle overlap, then the exact predicate; survivors are
                //    kept as a bitmask and, when they fit, as compacted boxes
                int n_att = 0;
                bool any_rotated = false;  // some listed attenuator needs its own frame
                for (int w = 0; w < ncw; w += 2) {
                    const int j = w * 32 + lane;
                    bool take = false;
                    Crystal K;
                    int k = 0;
                    if (j < nc) {
                        k = (int)s_cand[j];
                        K = cry[k];
                        take = true;
                        if (bundle_ok && K.slab >= 0.0f) {
                            const float4 bb = s_box[wave][(int)K.slab];
                            const float ex = (K.axis != 0.0f) ? K.hx : K.hxz;
                            const float ez = (K.axis != 0.0f) ? K.hz : ex;
                            take = !(K.cx + ex < bb.x || K.cx - ex > bb.y || K.cz + ez < bb.z || K.cz - ez > bb.w);
                        }
                        if (take) {
                            const int fast =
                                cull_fast(K.cx - xI, K.cy - yI, K.cz - zI, 0.5f * K.rbound, ux, uy, uz, crit_self);
                            if (fast != 0) {
                                take = fast > 0;
                            } else {  // near the boundary: the reference's operations
                                float crit = K.rbound;
                                const float dk =
                                    sqrtf((xI - K.cx) * (xI - K.cx) + (yI - K.cy) * (yI - K.cy) +
                                          (zI - K.cz) * (zI - K.cz));
                                crit = crit / dk / 2.0f;
                                const float vx = (K.cx - xI) / dk, vy = (K.cy - yI) / dk, vz = (K.cz - zI) / dk;
                                take = sqrtf((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <=
                                       crit + crit_self;
                            }
                        }
                    }
                    const unsigned long long bal = __ballot(take);
                    if (__ballot(take && K.axis == 0.0f)) any_rotated = true;
                    if (take) {
                        const int slot = n_att + __popcll(bal & below);
                        if (slot < kListCap) {
                            Att A;
                            A.x1 = K.cx - K.hx; A.x2 = K.cx + K.hx;
                            A.y1 = K.cy - K.hy; A.y2 = K.cy + K.hy;
                            A.z1 = K.cz - K.hz; A.z2 = K.cz + K.hz;
                            A.mu = K.mu_t;
                            A.k = (K.axis != 0.0f) ? k : -(k + 1);
                            s_att[wave][slot] = A;
                        }
                    }
                    n_att += __popcll(bal);
                    if (lane == 0) {
                        s_mask[wave][w] = (unsigned)bal;
                        if (w + 1 < (kCapCrystals / 32)) s_mask[wave][w + 1] = (unsigned)(bal >> 32);
                    }
                }
                // -- holes: bundle overlap, exact predicate; layer_holes bit m set when
                //    some hole of layer m survives
                unsigned layer_holes = 0;
                int n_hol = 0;
                for (int w = 0; w < nhw; w += 2) {
                    const int j = w * 32 + lane;
                    bool take = false;
                    int h = 0;
                    if (j < nh) {
                        h = (int)s_hcand[j];
                        const Hole H = hol[h];
                        take = true;
                        if (bundle_ok && H.slab >= 0.0f) {
                            const float4 bb = s_box[wave][(int)H.slab];
                            take = !(H.x + H.r < bb.x || H.x - H.r > bb.y || H.z + H.r < bb.z || H.z - H.r > bb.w);
                        }
                        if (take) {
                            const float yh = (H.y1 + H.y2) / 2.0f;
                            const int fast = cull_fast(H.x - xI, yh - yI, H.z - zI, H.rbound, ux, uy, uz, crit_self);
                            if (fast != 0) {
                                take = fast > 0;
                            } else {  // near the boundary: the reference's operations
                                float crit = H.rbound;
                                const float dh =
                                    sqrtf((xI - H.x) * (xI - H.x) + (yI - yh) * (yI - yh) + (zI - H.z) * (zI - H.z));
                                crit = crit / dh;
                                const float vx = (H.x - xI) / dh, vy = (yh - yI) / dh, vz = (H.z - zI) / dh;
                                take = sqrtf((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <=
                                       crit + crit_self;
                            }
                        }
                        if (take) layer_holes |= 1u << (int)H.layer;
                    }
                    const unsigned long long bal = __ballot(take);
                    if (take) {
                        const int slot = n_hol + __popcll(bal & below);
                        if (slot < kHoleCap) s_hol[wave][slot] = h;
                    }
                    n_hol += __popcll(bal);
                    if (lane == 0) {
                        s_hmask[wave][w] = (unsigned)bal;
                        if (w + 1 < (kCapHoles / 32)) s_hmask[wave][w + 1] = (unsigned)(bal >> 32);
                    }
                }
                {  // OR over the wave: one ballot per layer
                    unsigned any = 0;
                    for (int m = 0; m < p.n_layers; ++m)
                        if (__ballot((layer_holes >> m) & 1u) != 0ull) any |= 1u << m;
                    layer_holes = uniform(any);
                }
                n_att = uniform(n_att);
                n_hol = uniform(n_hol);
                // -- plate classes: test per sub-ray, every sub-ray passes, every sub-ray misses
                unsigned plate_test = 0;  // bit m: per-sub-ray test needed
                bool element_blocked = false;
                for (int m = 0; m < p.n_layers; ++m) {
                    const Plate PL = pla[m];
                    int cls = 0;
                    if (bundle_ok && PL.slab >= 0.0f) {
                        const float4 bb = s_box[wave][(int)PL.slab];
                        const bool crosses = (P_yhi < PL.y1 - 0.01f) && (Q_ylo > PL.y2 + 0.01f) && PL.thick >= 0.2f;
                        if (crosses && bb.x >= PL.x1 + 0.01f && bb.y <= PL.x2 - 0.01f && bb.z >= PL.z1 + 0.01f &&
                            bb.w <= PL.z2 - 0.01f)
                            cls = 1;
                        else if (bb.y < PL.x1 || bb.x > PL.x2 || bb.w < PL.z1 || bb.z > PL.z2)
                            cls = 2;
                    }
                    if (cls == 0) plate_test |= 1u << m;
                    if (cls == 2) element_blocked = true;
                }
                __builtin_amdgcn_wave_barrier();

                double acc = 0.0;
                if (!element_blocked) {
                    // FP64 voxel centre (before rotation) for the collimator path
                    const double xc64 = ((double)ix - p.hX64 + 0.5) * p.wx64 + p.sx64;
                    const double yc64 = ((double)iy - p.hY64 + 0.5) * p.wy64 + p.sy64;
                    const double zc64 = ((double)(p.z0 + izl) - p.hZ64 + 0.5) * p.wz64 + p.sz64;
                    const bool att_fast = n_att <= kListCap;
                    const bool hol_fast = n_hol <= kHoleCap;
                    for (int base = 0; base < S; base += 64) {
                        const int s = base + lane;
                        if (s >= S) break;
                        // sub-ray index: voxel sub-point outer, then NumZ, NumX, NumY (reference loop order)
                        int ny, nx, nz, va = 0, vb = 0, vc = 0;
                        SubCell SC;
                        if (sub_tab) {
                            SC = s_sub[s];
                            ny = (int)(SC.idx0 & 0xFFFFu);
                            nx = (int)(SC.idx0 >> 16);
                            nz = (int)(SC.idx1 & 0xFFu);
                            va = (int)((SC.idx1 >> 8) & 0xFFu);
                            vb = (int)((SC.idx1 >> 16) & 0xFFu);
                            vc = (int)(SC.idx1 >> 24);
                        } else if (idx_tab) {
                            const uint2 e = s_idx[s];
                            ny = (int)(e.x & 0xFFFFu);
                            nx = (int)(e.x >> 16);
                            nz = (int)(e.y & 0xFFu);
                            va = (int)((e.y >> 8) & 0xFFu);
                            vb = (int)((e.y >> 16) & 0xFFu);
                            vc = (int)(e.y >> 24);
                            sub_cell_geometry(D, nx, ny, nz, wD, tD, hD, rcx, rcy, rcz, cD, sD, SC);
                        } else {
                            int q = divmod(s, p.cy, rcy, ny);
                            q = divmod(q, p.cx, rcx, nx);
                            const int iv = divmod(q, p.cz, rcz, nz);
                            if (p.nv > 1) {
                                const int t1 = divmod(iv, p.nv, rnv, va);
                                vc = divmod(t1, p.nv, rnv, vb);
                            }
                            sub_cell_geometry(D, nx, ny, nz, wD, tD, hD, rcx, rcy, rcz, cD, sD, SC);
                        }
                        float att = 0.0f;
                        double att64 = 0.0;
                        // collimator layers, plate length minus hole chords (PESysMatGen.cu:661-685),
                        // when some hole is a candidate: first, because this FP64 path needs only the
                        // sub-ray's indices, so its registers are free again before the FP32 ray is built
                        if (layer_holes != 0) {
                            // end points in FP64 with the same formulas (quotients by the
                            // sampling counts through their reciprocals, correctly rounded)
                            double px = xc64, py = yc64, pz = zc64;
                            if (p.nv > 1) {
                                const double fnv = (double)p.nv;
                                px += (div_rcp((double)va + 0.5, fnv, p.rnv64) - 0.5) * p.wx64;
                                py += (div_rcp((double)vb + 0.5, fnv, p.rnv64) - 0.5) * p.wy64;
                                pz += (div_rcp((double)vc + 0.5, fnv, p.rnv64) - 0.5) * p.wz64;
                            }
                            const double rpx = px * p.rc64 - py * p.rs64, rpy = px * p.rs64 + py * p.rc64;
                            const double w64 = 2.0 * (double)D.hx, h64 = 2.0 * (double)D.hz, t64 = 2.0 * (double)D.hy;
                            const double xs = -w64 / 2.0 + div_rcp((double)nx + 0.5, (double)p.cx, p.rcx64) * w64;
                            const double zs = -h64 / 2.0 + div_rcp((double)nz + 0.5, (double)p.cz, p.rcz64) * h64;
                            const double ys = -t64 / 2.0 + div_rcp((double)ny + 0.5, (double)p.cy, p.rcy64) * t64;
                            const double qx = (double)D.cx + (xs * cD64 - zs * sD64);
                            const double qz = (double)D.cz + (xs * sD64 + zs * cD64);
                            const double qy = ((double)D.y_raw + (double)p.fov2c) + ys;
                            const Ray64 R64 = make_ray64(rpx, rpy, pz, qx, qy, qz);
                            for (int m = 0; m < p.n_layers; ++m) {
                                const Plate PL = pla[m];
                                double holes_len = 0.0;
                                if ((layer_holes >> m) & 1u) {
                                    auto one_hole = [&](int h) {
                                        const Hole H = hol[h];
                                        if ((int)H.layer != m) return;
                                        const double c = chord_hole64(R64, H.x, (double)H.y1_raw + (double)p.fov2c,
                                                                      (double)H.y2_raw + (double)p.fov2c, H.z, H.r);
                                        holes_len += c;
                                        att64 += (double)H.mu_t * c;
                                    };
                                    if (hol_fast) {
                                        for (int j = 0; j < n_hol; ++j) one_hole(uniform(s_hol[wave][j]));
                                    } else {
                                        for (int w = 0; w < nhw; ++w) {
                                            unsigned msk = uniform(s_hmask[wave][w]);
                                            while (msk) {
                                                const int b = __builtin_ctz(msk);
                                                msk &= msk - 1;
                                                one_hole(uniform((int)s_hcand[w * 32 + b]));
                                            }
                                        }
                                    }
                                }
                                const double in_plate = div_rcp((double)PL.thick, R64.ky, R64.iky);
                                if ((in_plate - holes_len) >= (double)0.00001f)
                                    att64 += (double)PL.mu_t * (in_plate - holes_len);
                            }
                        }
                        float xP = xI, yP = yI, zP = zI;
                        if (p.nv > 1) {
                            const float ox = (((float)va + 0.5f) * rnv - 0.5f) * p.wx;
                            const float oy = (((float)vb + 0.5f) * rnv - 0.5f) * p.wy;
                            const float oz = (((float)vc + 0.5f) * rnv - 0.5f) * p.wz;
                            rotate_image(p, xc0 + ox, yc0 + oy, xP, yP);
                            zP = zc0 + oz;
                        }
                        const float xPs = (xP - D.cx) * cDn - (zP - D.cz) * sDn;
                        const float yPs = yP - D.cy;
                        const float zPs = (xP - D.cx) * sDn + (zP - D.cz) * cDn;
                        // crystal sub-point, crystal frame then world (PESysMatGen.cu:606-616)
                        const float xDs = SC.xDs, yDs = SC.yDs, zDs = SC.zDs;
                        const float xQ = SC.xQ, yQ = SC.yQ, zQ = SC.zQ;

                        const WRay R = make_wray(xP, yP, zP, xQ, yQ, zQ);
                        const float dist = R.t;
                        // solid angle: largest projected sub-face (PESysMatGen.cu:618-639)
                        const float fx = A_x * fabsf(R.dx * cD - R.dz * sD);
                        const float fz = A_z * fabsf(R.dx * sD + R.dz * cD);
                        const float fy = A_y * fabsf(R.dy);
                        const float solid = fmaxf(fmaxf(fx, fy), fz) * p.inv4pi * frcp(dist * dist * dist);

                        bool blocked = false;
                        // plate miss => blocked (PESysMatGen.cu:645-659)
                        for (unsigned mm = plate_test; mm; mm &= mm - 1) {
                            const int m = __builtin_ctz(mm);
                            const Plate PL = pla[m];
                            const float c =
                                (R.pat == 0) ? chord_world(R, PL.x1, PL.x2, PL.y1, PL.y2, PL.z1, PL.z2)
                                             : chord_special(R.px, R.py, R.pz, xQ, yQ, zQ, R.ix, R.iy, R.iz, R.t, R.pat,
                                                             PL.x1, PL.y1, PL.z1, PL.x2, PL.y2, PL.z2);
                            if (c < 0.1f) blocked = true;
                        }
                        // collimator layers without candidate holes (PESysMatGen.cu:661-685)
                        if (layer_holes == 0) {
                            // plate path thick * |PQ| / dy = thick (1 + q), q = r^2 / (1 + sqrt(1 + r^2)),
                            // r = lateral / dy: mu * thick summed in FP64 and only the small q term
                            // rounded in FP32, so the exponent of strongly attenuated elements (tens
                            // of mean free paths) keeps ~1e-7 absolute accuracy.  dy <= 0 gives a
                            // negative length, which the reference does not add.
                            if (R.dy > 0.0f) {
                                const float idy = frcp(R.dy);
                                const float r2 = (R.dx * R.dx + R.dz * R.dz) * (idy * idy);
                                const float qv = r2 * frcp(1.0f + fsqrt(1.0f + r2));
                                att64 += plate_mt + (double)(plate_mtf * qv);
                                if (thin_plates) {  // plates under 1e-5 mm count only along very oblique paths
                                    for (int m = 0; m < p.n_layers; ++m) {
                                        const Plate PL = pla[m];
                                        if (PL.thick < 0.00001f && PL.thick * (1.0f + qv) >= 0.00001f)
                                            att64 += (double)PL.mu_t * (double)PL.thick +
                                                     (double)((PL.mu_t * PL.thick) * qv);
                                    }
                                }
                            }
                        }
                        // other crystals on the way (PESysMatGen.cu:687-725)
                        if (att_fast && !any_rotated && __ballot(R.pat != 0) == 0ull) {
                            // common case, branch-free: axis-aligned boxes, no axis-parallel sub-ray in the
                            // wave; parameters measured from Q (the attenuators lie near the detecting
                            // crystal); unrolled so the LDS reads of several boxes are in flight together
    #pragma unroll 4
                            for (int j = 0; j < n_att; ++j) {
                                const Att A = s_att[wave][j];
                                att += chord_world_q(R, xQ, yQ, zQ, A.x1, A.x2, A.y1, A.y2, A.z1, A.z2) * A.mu;
                            }
                        } else if (att_fast) {
                            for (int j = 0; j < n_att; ++j) {
                                const Att A = s_att[wave][j];
                                const int k = uniform(A.k);
                                const bool rotated = k < 0;
                                att += crystal_chord(R, xQ, yQ, zQ, cry, rotated ? -k - 1 : k, A.x1, A.x2, A.y1, A.y2,
                                                     A.z1, A.z2, rotated) * A.mu;
                            }
                        } else {
                            for (int w = 0; w < ncw; ++w) {
                                unsigned msk = uniform(s_mask[wave][w]);
                                while (msk) {
                                    const int b = __builtin_ctz(msk);
                                    msk &= msk - 1;
                                    const int k = uniform((int)s_cand[w * 32 + b]);
                                    const Crystal K = cry[k];
                                    att += crystal_chord(R, xQ, yQ, zQ, cry, k, K.cx - K.hx, K.cx + K.hx, K.cy - K.hy,
                                                         K.cy + K.hy, K.cz - K.hz, K.cz + K.hz, K.axis == 0.0f) *
                                           K.mu_t;
                                }
                            }
                        }
                        // path inside crystal d before the sub-cell, and absorption in the
                        // sub-cell along the 1000 mm extension (PESysMatGen.cu:727-759)
                        const float ux1 = SC.ux1, ux2 = SC.ux2, uy1 = SC.uy1, uy2 = SC.uy2, uz1 = SC.uz1, uz2 = SC.uz2;
                        const QRay Rq = axis_d ? make_qray_axis(R, xPs, yPs, zPs, xDs, yDs, zDs)
                                               : make_qray(xPs, yPs, zPs, xDs, yDs, zDs, dist);
                        float a1, a2, l_abs;
                        self_paths_centred(Rq, D.hx, D.hy, D.hz, sub_hx, sub_hy, sub_hz, ux1, ux2, uy1, uy2, uz1, uz2,
                                           a1, a2, l_abs);
                        att += (a1 - a2) * D.mu_t;
                        // exp(-a) = 2^-(a log2 e): the integer part of the exponent applied exactly
                        // (ldexp), the fraction in FP32, so the relative error stays ~2 ulp for any
                        // attenuation (rounding a itself to FP32 would cost |a| * 6e-8)
                        const double y2 = (att64 + (double)att) * 1.4426950408889634;
                        const double yi = floor(y2);
                        const float eff =
                            (y2 < 160.0) ? ldexpf(__builtin_amdgcn_exp2f(-(float)(y2 - yi)), -(int)yi) : 0.0f;
                        const float absorb = -expm1f(-l_abs * D.mu_t) * absorb_scale;
                        const float contrib = eff * solid * absorb;
                        if (!blocked) acc += (double)contrib;
                    }
                }
                acc = wave_sum64(acc);
                if (lane == 0) {
                    row[(long long)izl * p.X * p.Y + iy * p.X + ix] = (float)(acc * (double)inv_nv3);
                }
                __builtin_amdgcn_wave_barrier();
            }
            __syncthreads();  // every wave is done with this tile's lists before the next tile rewrites them
        }
    }
}

// Host launcher: grid = detectors x runs of up to kGroupsPerWorkgroup tile
// groups, one workgroup per run.  Tile groups share one cull over the detector
// list; a launch that would have too few workgroups runs shorter runs, then one
// tile per workgroup.  Either way every element's result is the same.
hipError_t launch_pe(const PeParams& p0, const Crystal* cry, const Hole* hol, const Plate* pla, const Slab* slb,
                     const PeLists& L, float* dst, int* status, hipStream_t stream) {
    PeParams p = p0;
    if ((long long)p.ngx * p.ngy * p.ngz * p.det_count < kGroupMinWorkgroups) {
        p.gx = p.gy = p.gz = 1;
        p.ngx = p.ntx;
        p.ngy = p.nty;
        p.ngz = p.ntz;
    }
    const long long groups = (long long)p.ngx * p.ngy * p.ngz;
    if (groups <= 0 || p.det_count <= 0) return hipSuccess;
    p.grp_per_wg = kGroupsPerWorkgroup;
    while (p.grp_per_wg > 1 && (groups + p.grp_per_wg - 1) / p.grp_per_wg * p.det_count < kLoopMinWorkgroups)
        p.grp_per_wg /= 2;
    const long long runs = (groups + p.grp_per_wg - 1) / p.grp_per_wg;  // workgroups per detector row
    // gridDim.x * blockDim.x must stay below 2^32: launch detector chunks
    const long long max_blocks = (1LL << 32) / kThreads - 1;
    const int chunk = (int)std::max(1LL, std::min<long long>(p.det_count, max_blocks / runs));
    for (int d0 = 0; d0 < p.det_count; d0 += chunk) {
        PeParams q = p;
        q.det_begin = p.det_begin + d0;
        q.det_count = std::min(chunk, p.det_count - d0);
        hipLaunchKernelGGL(pe_tile_kernel, dim3((unsigned)(runs * q.det_count)), dim3(kThreads), 0, stream, q, cry,
                           hol, pla, slb, L, dst + (long long)d0 * p.row_stride, status);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace sysmat

namespace sysmat {

hipError_t launch_detector_cull(const PeParams& p, const Crystal* cry, const Hole* hol, const Slab* slb, int fill,
                                int* counts, const PeLists& L, hipStream_t stream) {
    if (p.n_det <= 0) return hipSuccess;
    hipLaunchKernelGGL(pe_detector_cull_kernel, dim3((unsigned)p.n_det), dim3(kThreads), 0, stream, p, cry, hol, slb,
                       fill, counts, L);
    return hipGetLastError();
}

}  // namespace sysmat
