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
//   * one workgroup = 4 wave64s = one detector x one voxel tile;
//   * the workgroup culls all crystals / holes once against the tile's
//     bounding sphere (a proven superset of every element's cull list) into LDS;
//   * one wave = one element at a time, its 64 lanes = 64 sub-rays; the exact
//     per-element cull predicate is evaluated lane-parallel over the tile
//     candidates and kept as a wave-uniform bitmask in LDS, so the attenuator
//     loop is uniform (no divergence) and reads crystal records with scalar loads;
//   * a per-element "bundle" filter bounds every sub-ray of the element
//     laterally in each y-slab (exact by per-variable monotonicity, evaluated
//     at the 32 vertex rays), dropping attenuators no sub-ray can touch --
//     they would contribute a zero chord, so results are unchanged;
//   * FP32 tracing; FP64 for the element accumulator and the collimator path
//     (hole chords and plate length), per the precision plan of the design.
#include <hip/hip_runtime.h>

#include "pe_kernel.hpp"
#include "trace.hpp"

namespace sysmat {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;

__device__ inline float wave_min32(float v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wave_max32(float v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <class T>
__device__ inline T uniform(T v) {
    return __builtin_amdgcn_readfirstlane(v);
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

__global__ void __launch_bounds__(kThreads) pe_tile_kernel(PeParams p, const Crystal* __restrict__ cry,
                                                           const Hole* __restrict__ hol,
                                                           const Plate* __restrict__ pla,
                                                           const Slab* __restrict__ slb, float* __restrict__ dst,
                                                           int* __restrict__ status) {
    __shared__ unsigned s_cand[kCapCrystals];
    __shared__ unsigned s_hcand[kCapHoles];
    __shared__ unsigned s_mask[kWaves][kCapCrystals / 32];
    __shared__ unsigned s_hmask[kWaves][kCapHoles / 32];
    __shared__ float4 s_box[kWaves][kMaxSlabs];
    __shared__ int s_wcount[kWaves];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    const long long wg = blockIdx.x;
    const int drow = (int)(wg % p.det_count);
    const long long tile = wg / p.det_count;
    const int d = p.det_begin + drow;
    const int tix = (int)(tile % p.ntx);
    const int tiy = (int)((tile / p.ntx) % p.nty);
    const int tiz = (int)(tile / ((long long)p.ntx * p.nty));
    const int vx0 = tix * p.tx, vy0 = tiy * p.ty, vz0 = tiz * p.tz;
    const int vx1 = min(vx0 + p.tx, p.X), vy1 = min(vy0 + p.ty, p.Y), vz1 = min(vz0 + p.tz, p.Zslab);
    const int nvox = (vx1 - vx0) * (vy1 - vy0) * (vz1 - vz0);

    const Crystal D = cry[d];
    float* row = dst + (long long)drow * p.row_stride;

    if (!(D.mu_t > 0.01f)) {  // no culling, no sub-rays: the element is 0
        for (int e = tid; e < nvox; e += kThreads) {
            int lx = e % (vx1 - vx0), ly = (e / (vx1 - vx0)) % (vy1 - vy0), lz = e / ((vx1 - vx0) * (vy1 - vy0));
            row[(long long)(vz0 + lz) * p.X * p.Y + (vy0 + ly) * p.X + (vx0 + lx)] = 0.0f;
        }
        return;
    }

    // ---- tile bounding sphere of the voxel centres -----------------------
    float ax, ay, az, bx, by, bz;
    voxel_centre(p, vx0, vy0, p.z0 + vz0, ax, ay, az);
    voxel_centre(p, vx1 - 1, vy1 - 1, p.z0 + vz1 - 1, bx, by, bz);
    float tcx0 = 0.5f * (ax + bx), tcy0 = 0.5f * (ay + by), tcz = 0.5f * (az + bz);
    float tcx, tcy;
    rotate_image(p, tcx0, tcy0, tcx, tcy);
    const float rho = 0.5f * sqrtf((bx - ax) * (bx - ax) + (by - ay) * (by - ay) + (bz - az) * (bz - az)) * 1.0001f + 1e-4f;

    // ---- workgroup cull: superset of every element's list ----------------
    // Ordered compaction (ballot + wave prefix), so candidate lists are sorted
    // by index and the kernel is deterministic run to run.
    int nc = 0, nh = 0;
    {
        const float ddx = D.cx - tcx, ddy = D.cy - tcy, ddz = D.cz - tcz;
        const float Dd = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
        const float udx = ddx / Dd, udy = ddy / Dd, udz = ddz / Dd;
        const float crit_d = (Dd > rho) ? 0.5f * D.rbound / (Dd - rho) + 2.0f * rho / Dd : 4.0f;
        const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        for (int base = 0; base < p.n_det; base += kThreads) {
            const int k = base + tid;
            bool take = false;
            if (k < p.n_det && k != d) {
                const Crystal K = cry[k];
                float kx = K.cx - tcx, ky = K.cy - tcy, kz = K.cz - tcz;
                float Dk = sqrtf(kx * kx + ky * ky + kz * kz);
                if (Dk <= rho) {
                    take = true;
                } else {
                    float bound = 0.5f * K.rbound / (Dk - rho) + 2.0f * rho / Dk + crit_d;
                    float vx = kx / Dk - udx, vy = ky / Dk - udy, vz = kz / Dk - udz;
                    take = sqrtf(vx * vx + vy * vy + vz * vz) <= bound * 1.0001f + 1e-5f;
                }
            }
            const unsigned long long bal = __ballot(take);
            if (lane == 0) s_wcount[wave] = __popcll(bal);
            __syncthreads();
            int off = nc, all = 0;
            for (int w2 = 0; w2 < kWaves; ++w2) {
                const int c = s_wcount[w2];
                if (w2 < wave) off += c;
                all += c;
            }
            const int slot = off + __popcll(bal & below);
            if (take && slot < kCapCrystals) s_cand[slot] = (unsigned)k;
            nc += all;
            __syncthreads();
        }
        for (int base = 0; base < p.n_holes; base += kThreads) {
            const int h = base + tid;
            bool take = false;
            if (h < p.n_holes) {
                const Hole H = hol[h];
                float hx = H.x - tcx, hy = 0.5f * (H.y1 + H.y2) - tcy, hz = H.z - tcz;
                float Dh = sqrtf(hx * hx + hy * hy + hz * hz);
                if (Dh <= rho) {
                    take = true;
                } else {
                    float bound = H.rbound / (Dh - rho) + 2.0f * rho / Dh + crit_d;
                    float vx = hx / Dh - udx, vy = hy / Dh - udy, vz = hz / Dh - udz;
                    take = sqrtf(vx * vx + vy * vy + vz * vz) <= bound * 1.0001f + 1e-5f;
                }
            }
            const unsigned long long bal = __ballot(take);
            if (lane == 0) s_wcount[wave] = __popcll(bal);
            __syncthreads();
            int off = nh, all = 0;
            for (int w2 = 0; w2 < kWaves; ++w2) {
                const int c = s_wcount[w2];
                if (w2 < wave) off += c;
                all += c;
            }
            const int slot = off + __popcll(bal & below);
            if (take && slot < kCapHoles) s_hcand[slot] = (unsigned)h;
            nh += all;
            __syncthreads();
        }
    }
    if (nc > kCapCrystals || nh > kCapHoles) {  // host retries with smaller tiles
        if (tid == 0) atomicOr(status, 1);
        return;
    }

    const int ncw = (nc + 31) >> 5, nhw = (nh + 31) >> 5;
    const int nsub_v = p.nv * p.nv * p.nv;
    const int S = p.S;
    const float inv_nv3 = 1.0f / (float)nsub_v;
    // crystal-d sub-point geometry
    const float wD = 2.0f * D.hx, tD = 2.0f * D.hy, hD = 2.0f * D.hz;
    const float cD = cosf(D.rot), sD = sinf(D.rot);  // rotation of the sub-points into the world
    const double cD64 = (D.rot == 0.0f) ? 1.0 : cos((double)D.rot);
    const double sD64 = (D.rot == 0.0f) ? 0.0 : sin((double)D.rot);
    const float fcx = (float)p.cx, fcy = (float)p.cy, fcz = (float)p.cz;
    const float A_x = tD * hD / fcz / fcy, A_z = wD * tD / fcx / fcy, A_y = wD * hD / fcx / fcz;
    // world AABB of crystal d (bundle filter end box)
    const float qex = fabsf(cD) * D.hx + fabsf(sD) * D.hz, qez = fabsf(sD) * D.hx + fabsf(cD) * D.hz;

    for (int e = wave; e < nvox; e += kWaves) {
        const int lx = e % (vx1 - vx0), ly = (e / (vx1 - vx0)) % (vy1 - vy0), lz = e / ((vx1 - vx0) * (vy1 - vy0));
        const int ix = vx0 + lx, iy = vy0 + ly, izl = vz0 + lz;
        float xc0, yc0, zc0;
        voxel_centre(p, ix, iy, p.z0 + izl, xc0, yc0, zc0);
        float xI, yI;
        rotate_image(p, xc0, yc0, xI, yI);
        const float zI = zc0;

        // -- exact cull predicate from the voxel centre (PESysMatGen.cu:506-589)
        const float dist0 = sqrtf((D.cy - yI) * (D.cy - yI) + (D.cx - xI) * (D.cx - xI) + (D.cz - zI) * (D.cz - zI));
        const float crit_self = D.rbound / dist0 / 2.0f;
        const float ux = (D.cx - xI) / dist0, uy = (D.cy - yI) / dist0, uz = (D.cz - zI) / dist0;

        // -- bundle bounds per y-slab: P box (voxel sub-points), Q box (crystal d)
        const float hvx = 0.5f * p.wx * (1.0f - 1.0f / (float)p.nv);
        const float hvy = 0.5f * p.wy * (1.0f - 1.0f / (float)p.nv);
        const float hvz = 0.5f * p.wz * (1.0f - 1.0f / (float)p.nv);
        const float pex = fabsf(p.rc) * hvx + fabsf(p.rs) * hvy, pey = fabsf(p.rs) * hvx + fabsf(p.rc) * hvy;
        const float P_ylo = yI - pey, P_yhi = yI + pey, Q_ylo = D.cy - D.hy, Q_yhi = D.cy + D.hy;
        const bool bundle_ok = p.use_bundle && (Q_ylo - P_yhi > 1.0f);
        if (bundle_ok) {
            // lane bits: 0 Pa, 1 Py, 2 Qa, 3 Qy, 4 slab face; lanes >= 32 do z
            const bool zaxis = lane >= 32;
            const int v = lane & 31;
            const float Pa = zaxis ? ((v & 1) ? zI + hvz : zI - hvz) : ((v & 1) ? xI + pex : xI - pex);
            const float Py = (v & 2) ? P_yhi : P_ylo;
            const float Qa = zaxis ? ((v & 4) ? D.cz + qez : D.cz - qez) : ((v & 4) ? D.cx + qex : D.cx - qex);
            const float Qy = (v & 8) ? Q_yhi : Q_ylo;
            for (int s = 0; s < p.n_slabs; ++s) {
                const Slab SL = slb[s];
                const float y = (v & 16) ? SL.y2 : SL.y1;
                const float val = Pa + (Qa - Pa) * ((y - Py) / (Qy - Py));
                const float mn = wave_min32(val), mx = wave_max32(val);
                const float m = 2e-3f + 1e-5f * fmaxf(fabsf(mn), fabsf(mx));
                if (lane == 0) { s_box[wave][s].x = mn - m; s_box[wave][s].y = mx + m; }
                if (lane == 32) { s_box[wave][s].z = mn - m; s_box[wave][s].w = mx + m; }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");

        // -- crystal mask: exact predicate && bundle overlap
        for (int w = 0; w < ncw; w += 2) {
            const int j = w * 32 + lane;
            bool take = false;
            if (j < nc) {
                const int k = (int)s_cand[j];
                const Crystal K = cry[k];
                float crit = K.rbound;
                float dk = sqrtf((xI - K.cx) * (xI - K.cx) + (yI - K.cy) * (yI - K.cy) + (zI - K.cz) * (zI - K.cz));
                crit = crit / dk / 2.0f;
                float vx = (K.cx - xI) / dk, vy = (K.cy - yI) / dk, vz = (K.cz - zI) / dk;
                take = sqrtf((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <= crit + crit_self;
                if (take && bundle_ok && K.slab >= 0.0f) {
                    const float4 bb = s_box[wave][(int)K.slab];
                    const float ex = (K.axis != 0.0f) ? K.hx : sqrtf(K.hx * K.hx + K.hz * K.hz);
                    const float ez = (K.axis != 0.0f) ? K.hz : ex;
                    take = !(K.cx + ex < bb.x || K.cx - ex > bb.y || K.cz + ez < bb.z || K.cz - ez > bb.w);
                }
            }
            const unsigned long long bal = __ballot(take);
            if (lane == 0) {
                s_mask[wave][w] = (unsigned)bal;
                if (w + 1 < (kCapCrystals / 32)) s_mask[wave][w + 1] = (unsigned)(bal >> 32);
            }
        }
        // -- hole mask: exact predicate && bundle overlap; layer_holes bit m set
        //    when some hole of layer m survives
        unsigned layer_holes = 0;
        for (int w = 0; w < nhw; w += 2) {
            const int j = w * 32 + lane;
            bool take = false;
            if (j < nh) {
                const int h = (int)s_hcand[j];
                const Hole H = hol[h];
                const float yh = (H.y1 + H.y2) / 2.0f;
                float crit = H.rbound;
                float dh = sqrtf((xI - H.x) * (xI - H.x) + (yI - yh) * (yI - yh) + (zI - H.z) * (zI - H.z));
                crit = crit / dh;
                float vx = (H.x - xI) / dh, vy = (yh - yI) / dh, vz = (H.z - zI) / dh;
                take = sqrtf((vz - uz) * (vz - uz) + (vy - uy) * (vy - uy) + (vx - ux) * (vx - ux)) <= crit + crit_self;
                if (take && bundle_ok && H.slab >= 0.0f) {
                    const float4 bb = s_box[wave][(int)H.slab];
                    take = !(H.x + H.r < bb.x || H.x - H.r > bb.y || H.z + H.r < bb.z || H.z - H.r > bb.w);
                }
                if (take) layer_holes |= 1u << (int)H.layer;
            }
            const unsigned long long bal = __ballot(take);
            if (lane == 0) {
                s_hmask[wave][w] = (unsigned)bal;
                if (w + 1 < (kCapHoles / 32)) s_hmask[wave][w + 1] = (unsigned)(bal >> 32);
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) layer_holes |= (unsigned)__shfl_xor((int)layer_holes, o, 64);
        layer_holes = uniform(layer_holes);
        // -- plate classes: 0 test per sub-ray, 1 every sub-ray passes, 2 every sub-ray misses
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
            // image point in the crystal-d frame (PESysMatGen.cu:460-464), voxel centre
            const float cDn = cosf(-D.rot), sDn = sinf(-D.rot);
            // FP64 voxel centre (before rotation) for the collimator path
            const double xc64 = ((double)ix - (double)p.X / 2.0 + 0.5) * (double)p.wx + (double)p.sx;
            const double yc64 = ((double)iy - (double)p.Y / 2.0 + 0.5) * (double)p.wy + (double)p.sy;
            const double zc64 = ((double)(p.z0 + izl) - (double)p.Zfull / 2.0 + 0.5) * (double)p.wz + (double)p.sz;
            for (int base = 0; base < S; base += 64) {
                const int s = base + lane;
                if (s >= S) break;
                // sub-ray index: voxel sub-point outer, then NumZ, NumX, NumY (reference loop order)
                int rr = s;
                const int ny = rr % p.cy; rr /= p.cy;
                const int nx = rr % p.cx; rr /= p.cx;
                const int nz = rr % p.cz; rr /= p.cz;
                const int iv = rr;
                float xP = xI, yP = yI, zP = zI;
                if (p.nv > 1) {
                    const int a = iv % p.nv, b = (iv / p.nv) % p.nv, c = iv / (p.nv * p.nv);
                    const float fnv = (float)p.nv;
                    const float ox = (((float)a + 0.5f) / fnv - 0.5f) * p.wx;
                    const float oy = (((float)b + 0.5f) / fnv - 0.5f) * p.wy;
                    const float oz = (((float)c + 0.5f) / fnv - 0.5f) * p.wz;
                    rotate_image(p, xc0 + ox, yc0 + oy, xP, yP);
                    zP = zc0 + oz;
                }
                const float xPs = (xP - D.cx) * cDn - (zP - D.cz) * sDn;
                const float yPs = yP - D.cy;
                const float zPs = (xP - D.cx) * sDn + (zP - D.cz) * cDn;
                // crystal sub-point, crystal frame then world (PESysMatGen.cu:606-616)
                const float xDs = -wD / 2.0f + ((float)nx + 0.5f) / fcx * wD;
                const float zDs = -hD / 2.0f + ((float)nz + 0.5f) / fcz * hD;
                const float yDs = -tD / 2.0f + ((float)ny + 0.5f) / fcy * tD;
                const float xQ = D.cx + (xDs * cD - zDs * sD);
                const float zQ = D.cz + (xDs * sD + zDs * cD);
                const float yQ = D.cy + yDs;

                const Ray R = make_ray(xP, yP, zP, xQ, yQ, zQ);
                const float dx = R.dx, dy = R.dy, dz = R.dz;
                const float dsq = dy * dy + dx * dx + dz * dz;
                const float dist = R.t;
                // solid angle: largest projected sub-face (PESysMatGen.cu:618-639)
                const float fx = A_x * fabsf(dx * cD - dz * sD);
                const float fz = A_z * fabsf(dx * sD + dz * cD);
                const float fy = A_y * fabsf(dy);
                const float solid = fmaxf(fmaxf(fx, fy), fz) * (p.inv4pi / (dsq * dist));

                bool blocked = false;
                for (unsigned mm = plate_test; mm; mm &= mm - 1) {
                    const int m = __builtin_ctz(mm);
                    const Plate PL = pla[m];
                    if (chord_through(R, PL.x1, PL.y1, PL.z1, PL.x2, PL.y2, PL.z2) < 0.1f) blocked = true;
                }
                float att = 0.0f;
                double att64 = 0.0;
                // collimator layers: plate length minus hole chords
                {
                    bool r64_ready = false;
                    Ray64 R64;
                    for (int m = 0; m < p.n_layers; ++m) {
                        const Plate PL = pla[m];
                        if (!((layer_holes >> m) & 1u)) {  // no hole of this layer can be hit
                            const float in_plate = PL.thick / (dy / dist);
                            if (in_plate >= 0.00001f) att += PL.mu_t * in_plate;
                            continue;
                        }
                        if (!r64_ready) {
                            // end points in FP64 with the same formulas
                            double px = xc64, py = yc64, pz = zc64;
                            if (p.nv > 1) {
                                const int a = iv % p.nv, b = (iv / p.nv) % p.nv, c = iv / (p.nv * p.nv);
                                const double fnv = (double)p.nv;
                                px += (((double)a + 0.5) / fnv - 0.5) * (double)p.wx;
                                py += (((double)b + 0.5) / fnv - 0.5) * (double)p.wy;
                                pz += (((double)c + 0.5) / fnv - 0.5) * (double)p.wz;
                            }
                            const double rpx = px * p.rc64 - py * p.rs64, rpy = px * p.rs64 + py * p.rc64;
                            const double w64 = 2.0 * (double)D.hx, h64 = 2.0 * (double)D.hz, t64 = 2.0 * (double)D.hy;
                            const double xs = -w64 / 2.0 + ((double)nx + 0.5) / (double)p.cx * w64;
                            const double zs = -h64 / 2.0 + ((double)nz + 0.5) / (double)p.cz * h64;
                            const double ys = -t64 / 2.0 + ((double)ny + 0.5) / (double)p.cy * t64;
                            const double qx = (double)D.cx + (xs * cD64 - zs * sD64);
                            const double qz = (double)D.cz + (xs * sD64 + zs * cD64);
                            const double qy = ((double)D.y_raw + (double)p.fov2c) + ys;
                            R64 = make_ray64(rpx, rpy, pz, qx, qy, qz);
                            r64_ready = true;
                        }
                        double holes_len = 0.0;
                        for (int w = 0; w < nhw; ++w) {
                            unsigned msk = uniform(s_hmask[wave][w]);
                            while (msk) {
                                const int b = __builtin_ctz(msk);
                                msk &= msk - 1;
                                const int h = uniform((int)s_hcand[w * 32 + b]);
                                if (h < PL.hole_begin || h >= PL.hole_end) continue;
                                const Hole H = hol[h];
                                const double c = chord_hole64(R64, H.x, (double)H.y1_raw + (double)p.fov2c,
                                                              (double)H.y2_raw + (double)p.fov2c, H.z, H.r);
                                holes_len += c;
                                att64 += (double)H.mu_t * c;
                            }
                        }
                        const double in_plate = (double)PL.thick / R64.ky;
                        if ((in_plate - holes_len) >= (double)0.00001f) att64 += (double)PL.mu_t * (in_plate - holes_len);
                    }
                }
                // other crystals on the way (PESysMatGen.cu:687-725)
                for (int w = 0; w < ncw; ++w) {
                    unsigned msk = uniform(s_mask[wave][w]);
                    while (msk) {
                        const int b = __builtin_ctz(msk);
                        msk &= msk - 1;
                        const int k = uniform((int)s_cand[w * 32 + b]);
                        const Crystal K = cry[k];
                        float c;
                        if (K.axis != 0.0f) {
                            c = chord_through_c(R, K.cx, K.cy, K.cz, K.hx, K.hy, K.hz);
                        } else {
                            const float xPk = (xP - K.cx) * K.cr - (zP - K.cz) * K.sr, yPk = yP - K.cy;
                            const float zPk = (xP - K.cx) * K.sr + (zP - K.cz) * K.cr;
                            const float xQk = (xQ - K.cx) * K.cr - (zQ - K.cz) * K.sr, yQk = yQ - K.cy;
                            const float zQk = (xQ - K.cx) * K.sr + (zQ - K.cz) * K.cr;
                            const Ray Rk = make_ray(xPk, yPk, zPk, xQk, yQk, zQk);
                            c = chord_through(Rk, -K.hx, -K.hy, -K.hz, K.hx, K.hy, K.hz);
                        }
                        att += c * K.mu_t;
                    }
                }
                // path inside crystal d before the sub-cell (PESysMatGen.cu:727-746)
                const float ux1 = ((float)nx / fcx - 0.5f) * wD, ux2 = (((float)nx + 1.0f) / fcx - 0.5f) * wD;
                const float uy1 = ((float)ny / fcy - 0.5f) * tD, uy2 = (((float)ny + 1.0f) / fcy - 0.5f) * tD;
                const float uz1 = ((float)nz / fcz - 0.5f) * hD, uz2 = (((float)nz + 1.0f) / fcz - 0.5f) * hD;
                const Ray Rs = make_ray(xPs, yPs, zPs, xDs, yDs, zDs);
                const float a1 = chord_open(Rs, -D.hx, -D.hy, -D.hz, D.hx, D.hy, D.hz);
                const float a2 = chord_open(Rs, ux1, uy1, uz1, ux2, uy2, uz2);
                att += (a1 - a2) * D.mu_t;
                const float eff = expf(-(float)(att64 + (double)att));
                // absorption in the sub-cell along the extended ray (PESysMatGen.cu:749-759)
                const float l_abs = chord_absorb(xPs, yPs, zPs, xDs, yDs, zDs, dist, ux1, uy1, uz1, ux2, uy2, uz2);
                const float absorb = -expm1f(-l_abs * D.mu_t) * D.mu_pe / D.mu_t;
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
}

// Host launcher: grid = detectors x tiles, one workgroup each.
hipError_t launch_pe(const PeParams& p, const Crystal* cry, const Hole* hol, const Plate* pla, const Slab* slb,
                     float* dst, int* status, hipStream_t stream) {
    const long long tiles = (long long)p.ntx * p.nty * p.ntz;
    const long long blocks = tiles * p.det_count;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(pe_tile_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, stream, p, cry, hol, pla, slb, dst,
                       status);
    return hipGetLastError();
}

}  // namespace sysmat
