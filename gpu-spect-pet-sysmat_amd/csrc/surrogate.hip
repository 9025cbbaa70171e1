// Scatter-surrogate kernels: features and the fused features -> MLP ensemble
// -> far-field prediction, written over the exact near field in place.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlp_device.hpp"
#include "surrogate.hpp"

namespace sysmat {

// Image-frame position of voxel v (rotated like the scatter kernels).
__device__ inline void sg_position(const SgParams& q, long long v, float& xI, float& yI, float& zI) {
    const long long XY = (long long)q.X * q.Y;
    const int iz = (int)(v / XY);
    const int rem = (int)(v % XY);
    const int iy = rem / q.X, ix = rem % q.X;
    const float x0 = ((float)ix - (float)q.X * 0.5f + 0.5f) * q.wx + q.sx;
    const float y0 = ((float)iy - (float)q.Y * 0.5f + 0.5f) * q.wy + q.sy;
    zI = ((float)iz - (float)q.Z * 0.5f + 0.5f) * q.wz + q.sz;
    xI = x0 * q.rc - y0 * q.rs;
    yI = x0 * q.rs + y0 * q.rc;
}

// Ring prior over sites j = j0, j0 + 2, ... < cnt of the row's ranked list (rt:
// the row's two float4 per site): each site's weight re-evaluated at this
// voxel's scatter angle.  Unrolled so several PE gathers are in flight (no
// branch: every factor is finite and >= 0, so a zero-weight site adds +0, as if
// skipped).  The prior is the
// even-site sum plus the odd-site sum, so the fused kernel can give the two
// halves to the two half-waves.
constexpr int kRingUnroll = 4;
__device__ inline float sg_ring(const SgParams& q, const float4* __restrict__ rt, int j0, int cnt,
                                const float* __restrict__ tab, const float* __restrict__ pe, long long v, float xI,
                                float yI, float zI) {
    float ring = 0.0f;
#pragma unroll kRingUnroll
    for (int j = j0; j < cnt; j += 2) {
        const float4 a = rt[2 * j], b = rt[2 * j + 1];
        const float wg = a.w;
        const float ax = a.x - xI, ay = a.y - yI, az = a.z - zI;
        const float c = (ax * b.x + ay * b.y + az * b.z) * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az);
        const float pv = pe[(long long)__float_as_int(b.w) * q.V + v];
        // the select keeps the product rounded before the sum (no FMA contraction), as in
        // sg_ring_half
        const float t = pv * wg * sg_angular(tab, c);
        ring += wg != 0.0f ? t : 0.0f;
    }
    return ring;
}

// sg_ring for the fused kernel: the half-wave `half` takes the sites of its parity,
// in a loop over site pairs whose trip count is wave-uniform (so it unrolls and the
// gathers of several sites are in flight); the lane past the list's end adds +0.
// Same sums, in the same order, as sg_ring(q, rt, n_near + half, ...).
__device__ inline float sg_ring_half(const SgParams& q, const float4* __restrict__ rt, int half, int cnt,
                                     const float* tab, const float* __restrict__ pe, long long v, float xI, float yI,
                                     float zI) {
    float ring = 0.0f;
#pragma unroll kRingUnroll
    for (int jp = q.n_near; jp < cnt; jp += 2) {
        const bool ok = jp + half < cnt;
        const int j = ok ? jp + half : jp;
        const float4 a = rt[2 * j], b = rt[2 * j + 1];
        const float ax = a.x - xI, ay = a.y - yI, az = a.z - zI;
        const float c = (ax * b.x + ay * b.y + az * b.z) * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az);
        const float pv = pe[(long long)__float_as_int(b.w) * q.V + v];
        const float t = pv * a.w * sg_angular(tab, c);
        ring += ok ? t : 0.0f;
    }
    return ring;
}

__global__ void surrogate_ring_table_kernel(const ScPair* __restrict__ keep, const int* __restrict__ keep_cnt,
                                            int n_det, int n_keep, float4* __restrict__ ring) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n_det * n_keep) return;
    const int r = (int)(i / n_keep), j = (int)(i % n_keep);
    float4 a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
    if (j < keep_cnt[r]) {
        const ScPair& S = keep[i];
        a = float4{(float)S.sx, (float)S.sy, (float)S.sz, S.hy1};
        b = float4{(float)S.bx, (float)S.by, (float)S.bz, __int_as_float(S.site)};
    }
    ring[2 * i] = a;
    ring[2 * i + 1] = b;
}

// Per row: 1 + rank of the last site (rank >= n_near) with a non-zero ring weight.
// Sites past it add +0 to the prior, so stopping there changes no sum.
__global__ void surrogate_ring_count_kernel(const ScPair* __restrict__ keep, const int* __restrict__ keep_cnt,
                                            int n_det, int n_keep, int n_near, int* __restrict__ ring_cnt) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_det) return;
    int c = n_near;
    for (int j = keep_cnt[r] - 1; j >= n_near; --j) {
        if (keep[(long long)r * n_keep + j].hy1 != 0.0f) {
            c = j + 1;
            break;
        }
    }
    ring_cnt[r] = c;
}

hipError_t launch_surrogate_ring_table(const ScPair* keep, const int* keep_cnt, int n_det, int n_keep, int n_near,
                                      float4* ring, int* ring_cnt, hipStream_t s) {
    const long long n = (long long)n_det * n_keep;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(surrogate_ring_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, keep, keep_cnt,
                       n_det, n_keep, ring);
    hipLaunchKernelGGL(surrogate_ring_count_kernel, dim3((unsigned)((n_det + 255) / 256)), dim3(256), 0, s, keep,
                       keep_cnt, n_det, n_keep, n_near, ring_cnt);
    return hipGetLastError();
}

// Features of element (row r, voxel v) from its position and ring prior.
__device__ inline void sg_finish(const SgParams& q, const SgRow& R, int r, long long v, const float* __restrict__ pe,
                                 float s_near, float ring, float xI, float yI, float zI, float* f) {
    const float inv_rho = 1.0f / q.irho;
    f[0] = (xI - q.icx) * inv_rho;
    f[1] = (yI - q.icy) * inv_rho;
    f[2] = (zI - q.icz) * inv_rho;
    f[3] = R.f3;
    f[4] = R.f4;
    f[5] = R.f5;
    const float inv_s = 1.0f / q.s_scale;
    f[6] = 0.1f * __logf(s_near * inv_s + 1e-6f);
    f[7] = 0.1f * __logf(ring * inv_s + 1e-6f);
    f[8] = 0.05f * __logf(pe[(long long)r * q.V + v] / q.pe_scale + 1e-9f);
    const float dx = R.cx - xI, dy = R.cy - yI, dz = R.cz - zI;
    // explicit FMAs: the same rounding in every kernel this is inlined into (left to
    // contraction, the features and the fused apply kernels rounded it differently)
    const float d = sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
    f[9] = dx / d;
    f[10] = dz / d;
    f[11] = __logf(d / q.d_ref);
}

__global__ void __launch_bounds__(256) surrogate_features_kernel(SgParams q, const SgRow* __restrict__ rows,
                                                                 const ScPair* __restrict__ keep,
                                                                 const int* __restrict__ keep_cnt,
                                                                 const float* __restrict__ pe,
                                                                 const float* __restrict__ near, long long near_stride,
                                                                 const int* __restrict__ vox, int n_vox,
                                                                 float* __restrict__ x) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)q.det_count * n_vox) return;
    const int rr = (int)(i / n_vox), j = (int)(i % n_vox);
    const int r = q.det_begin + rr;
    const long long v = vox ? (long long)vox[j] : (long long)j;
    const float4* kr = q.ring + 2LL * r * q.n_keep;
    const int cnt = q.ring_cnt[r];
    const float* tab = q.ftab + (long long)r * (kSgTab + 1);
    float xI, yI, zI;
    sg_position(q, v, xI, yI, zI);
    const float ring = sg_ring(q, kr, q.n_near, cnt, tab, pe, v, xI, yI, zI) +
                       sg_ring(q, kr, q.n_near + 1, cnt, tab, pe, v, xI, yI, zI);
    float f[kSgFeatures];
    sg_finish(q, rows[r], r, v, pe, near[rr * near_stride + v], ring, xI, yI, zI, f);
#pragma unroll
    for (int k = 0; k < kSgFeatures; ++k) x[i * kSgFeatures + k] = f[k];
}

// Voxels per wave of the apply kernel: three 16-element MFMA tiles.
constexpr int kSgRun = 48;

// Voxel i (0 .. kSgRun-1) of wave chunk `ch` of a row: the run of kSgRun consecutive
// voxels (row order, x fastest) from kSgRun ch, so each ring-prior site's PE gather
// of a wave reads whole cache lines.
__device__ inline long long sg_voxel(const SgParams& q, long long ch, int i, bool& ok) {
    const long long v = (long long)kSgRun * ch + i;
    ok = i < kSgRun && v < q.V;
    return v;
}

// 4 x 4 x 4 brick (exact-fallback tile) of voxel v.
__device__ inline long long sg_brick(const SgParams& q, long long v) {
    const long long XY = (long long)q.X * q.Y;
    const int iz = (int)(v / XY);
    const int rem = (int)(v % XY);
    const int iy = rem / q.X, ix = rem % q.X;
    const int nbx = (q.X + 3) >> 2, nby = (q.Y + 3) >> 2;
    return ((long long)(iz >> 2) * nby + (iy >> 2)) * nbx + (ix >> 2);
}

// One wave = a run of 48 consecutive voxels of one detector row.
// Features go through LDS into the MFMA B-operand layout; every member runs
// with register-resident activations and weight tiles streamed from L2 (each
// wave on its own: sharing weight slabs through LDS cost two workgroup barriers
// per slab, more than the L2 reads it saved; a double-buffered form with one
// barrier per slab needed more than 256 VGPRs for its staging and was 6 % slower).
// Disagreeing members flag the brick (tile) for the exact kernel, which prunes
// scatter sites per brick.  Two waves per SIMD, no spill: three element tiles per
// wave with each tile pair's weight tiles requested two K blocks ahead of their
// MFMAs move 1.5x fewer weight bytes per element than two tiles with all 16
// requested up front (MLP alone: 59.3 against 70.4 ms on 6.7e7 elements,
// identical output; the member statistics live in LDS, which keeps the kernel at
// 256 VGPRs without spills: 71.7 -> 61.8 ms on 256 rows of a 64^3 image); the
// features take two rounds of 32 lanes for 48 voxels.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) surrogate_apply_kernel(SgParams q, const SgRow* __restrict__ rows,
                                                              const ScPair* __restrict__ keep,
                                                              const int* __restrict__ keep_cnt,
                                                              const uint8_t* __restrict__ packed,
                                                              const float* __restrict__ pe, float* __restrict__ dst,
                                                              uint8_t* __restrict__ flags) {
    __shared__ float s_f[4][kSgRun][kSgFeatures + 1];
    __shared__ float s_tab[4][kSgTab + 1];  // the wave's row angular table
    // the wave index through readfirstlane: what derives from it is known to be wave-uniform
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int g = lane >> 4, j16 = lane & 15;
    constexpr int kT = kSgRun / 16;  // element tiles per wave
    const long long chunks = (q.V + kSgRun - 1) / kSgRun;
    const long long gw = (long long)blockIdx.x * 4 + w;
    const int rr = (int)(gw / chunks);
    const long long ch = gw % chunks;
    const bool valid_wave = rr < q.det_count;
    const int r = q.det_begin + (valid_wave ? rr : 0);
    const SgRow R = rows[r];
    const bool use = valid_wave && R.use != 0.0f;
    {
        const float* tab = q.ftab + (long long)r * (kSgTab + 1);
        for (int i = lane; i <= kSgTab; i += 64) s_tab[w][i] = tab[i];
        __syncthreads();  // the rows' angular tables are in LDS
        // lanes e and e + 32 share element e: even and odd ring sites; two rounds
        // (elements 0..31, then 32..47)
        for (int rnd = 0; rnd < 2; ++rnd) {
            const int e = 32 * rnd + (lane & 31), half = lane >> 5;
            bool ok;
            const long long v = sg_voxel(q, ch, e, ok);
            const bool live = use && ok;
            float xI = 0.0f, yI = 0.0f, zI = 0.0f, ring = 0.0f;
            if (live) {
                sg_position(q, v, xI, yI, zI);
                ring = sg_ring_half(q, q.ring + 2LL * r * q.n_keep, half, q.ring_cnt[r], s_tab[w], pe, v, xI, yI, zI);
            }
            const float other = __shfl_xor(ring, 32, 64);
            if (half == 0 && e < kSgRun) {
                float f[kSgFeatures];
                if (live) {
                    sg_finish(q, R, r, v, pe, dst[rr * q.row_stride + v], ring + other, xI, yI, zI, f);
                } else {
#pragma unroll
                    for (int k = 0; k < kSgFeatures; ++k) f[k] = 0.0f;
                }
#pragma unroll
                for (int k = 0; k < kSgFeatures; ++k) s_f[w][e][k] = f[k];
            }
        }
    }
    __syncthreads();
    bf16x8 xf[kT];
#pragma unroll
    for (int t = 0; t < kT; ++t)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 8 * g + s;
            xf[t][s] = (bf16)(k < kSgFeatures ? s_f[w][16 * t + j16][k] : 0.0f);
        }
    // per-element member statistics (sum, min, max) in the wave's LDS feature rows, whose
    // features are in xf now
    float* st = &s_f[w][0][0];
    if (g == 0) {
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            st[(16 * t + j16) * (kSgFeatures + 1) + 0] = 0.0f;
            st[(16 * t + j16) * (kSgFeatures + 1) + 1] = 3.0e38f;
            st[(16 * t + j16) * (kSgFeatures + 1) + 2] = 0.0f;
        }
    }
    for (int m = 0; m < q.n_members; ++m) {
        float y[kT];
        mlp_wave_forward<kT, 2>(packed + (long long)m * kMlpPacked, xf, y);
        if (g == 0) {
#pragma unroll
            for (int t = 0; t < kT; ++t) {
                const float fr = fmaxf(__expf(y[t] * q.sig_y + q.mu_y) - q.phi, 0.0f);
                float* o = st + (16 * t + j16) * (kSgFeatures + 1);
                o[0] += fr;
                o[1] = fminf(o[1], fr);
                o[2] = fmaxf(o[2], fr);
            }
        }
    }
    float far_sum[kT], far_min[kT], far_max[kT];
#pragma unroll
    for (int t = 0; t < kT; ++t) {
        const float* o = st + (16 * t + j16) * (kSgFeatures + 1);
        far_sum[t] = o[0];
        far_min[t] = o[1];
        far_max[t] = o[2];
    }
    if (g == 0 && use) {
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            bool ok;
            const long long v = sg_voxel(q, ch, 16 * t + j16, ok);
            if (!ok) continue;
            float* o = dst + rr * q.row_stride + v;
            const float s_hat = *o + far_sum[t] / (float)q.n_members;
            // members disagree: the voxel's brick goes to the exact kernel (lanes of one
            // brick store the same byte)
            if (q.n_members > 1 && s_hat > q.floor_abs && (far_max[t] - far_min[t]) > q.tol * s_hat && flags)
                flags[(long long)rr * q.n_tiles + sg_brick(q, v)] = 1;
            *o = s_hat;
        }
    } else if (g == 0 && valid_wave) {
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            bool ok;
            const long long v = sg_voxel(q, ch, 16 * t + j16, ok);
            if (ok) dst[rr * q.row_stride + v] = 0.0f;
        }
    }
}

hipError_t launch_surrogate_features(const SgParams& q, const SgRow* rows, const ScPair* keep, const int* keep_cnt,
                                     const float* pe, const float* near, long long near_stride, const int* vox,
                                     int n_vox, float* x, hipStream_t s) {
    const long long n = (long long)q.det_count * n_vox;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(surrogate_features_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, q, rows, keep,
                       keep_cnt, pe, near, near_stride, vox, n_vox, x);
    return hipGetLastError();
}

hipError_t launch_surrogate_apply(const SgParams& q, const SgRow* rows, const ScPair* keep, const int* keep_cnt,
                                  const uint8_t* packed, const float* pe, float* dst, uint8_t* flags, hipStream_t s) {
    const long long per_row = (q.V + kSgRun - 1) / kSgRun;  // waves per detector row: runs of 48 voxels
    if (per_row <= 0 || q.det_count <= 0) return hipSuccess;
    // gridDim.x * blockDim.x must stay below 2^32: launch detector chunks
    const long long max_waves = ((1LL << 32) / 256 - 1) * 4;
    const int chunk = (int)std::max(1LL, std::min<long long>(q.det_count, max_waves / per_row));
    for (int d0 = 0; d0 < q.det_count; d0 += chunk) {
        SgParams c = q;
        c.det_begin = q.det_begin + d0;
        c.det_count = std::min(chunk, q.det_count - d0);
        const long long waves = (long long)c.det_count * per_row;
        hipLaunchKernelGGL(surrogate_apply_kernel, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, c, rows, keep,
                           keep_cnt, packed, pe, dst + (long long)d0 * q.row_stride,
                           flags ? flags + (long long)d0 * q.n_tiles : nullptr);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace sysmat
