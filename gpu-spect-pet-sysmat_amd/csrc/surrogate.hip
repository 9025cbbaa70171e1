// Scatter-surrogate kernels: features and the fused features -> MLP ensemble
// -> far-field prediction, written over the exact near field in place.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlp_device.hpp"
#include "surrogate.hpp"

namespace sysmat {

// Features of element (row r, voxel v); keep/cnt: the row's ranked sites.
__device__ inline void sg_features(const SgParams& q, const SgRow& R, const ScPair* __restrict__ keep, int cnt,
                                   int r, long long v, const float* __restrict__ pe, float s_near, float* f) {
    const long long XY = (long long)q.X * q.Y;
    const int iz = (int)(v / XY);
    const int rem = (int)(v % XY);
    const int iy = rem / q.X, ix = rem % q.X;
    const float x0 = ((float)ix - (float)q.X * 0.5f + 0.5f) * q.wx + q.sx;
    const float y0 = ((float)iy - (float)q.Y * 0.5f + 0.5f) * q.wy + q.sy;
    const float zI = ((float)iz - (float)q.Z * 0.5f + 0.5f) * q.wz + q.sz;
    const float xI = x0 * q.rc - y0 * q.rs, yI = x0 * q.rs + y0 * q.rc;
    const float inv_rho = 1.0f / q.irho;
    f[0] = (xI - q.icx) * inv_rho;
    f[1] = (yI - q.icy) * inv_rho;
    f[2] = (zI - q.icz) * inv_rho;
    f[3] = R.f3;
    f[4] = R.f4;
    f[5] = R.f5;
    // ring prior: sites n_near.. with their weight re-evaluated at this voxel's scatter angle
    float ring = 0.0f;
    const float* tab = q.ftab + (long long)r * (kSgTab + 1);
    for (int j = q.n_near; j < cnt; ++j) {
        const ScPair& S = keep[j];
        const float wg = S.hy1;
        if (wg == 0.0f) continue;
        const float ax = (float)S.sx - xI, ay = (float)S.sy - yI, az = (float)S.sz - zI;
        const float c = (ax * (float)S.bx + ay * (float)S.by + az * (float)S.bz) * __builtin_amdgcn_rsqf(ax * ax + ay * ay + az * az);
        ring += pe[(long long)S.site * q.V + v] * wg * sg_angular(tab, c);
    }
    const float inv_s = 1.0f / q.s_scale;
    f[6] = 0.1f * __logf(s_near * inv_s + 1e-6f);
    f[7] = 0.1f * __logf(ring * inv_s + 1e-6f);
    f[8] = 0.05f * __logf(pe[(long long)r * q.V + v] / q.pe_scale + 1e-9f);
    const float dx = R.cx - xI, dy = R.cy - yI, dz = R.cz - zI;
    const float d = sqrtf(dx * dx + dy * dy + dz * dz);
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
    float f[kSgFeatures];
    sg_features(q, rows[r], keep + (long long)r * q.n_keep, keep_cnt[r], r, v, pe, near[rr * near_stride + v], f);
#pragma unroll
    for (int k = 0; k < kSgFeatures; ++k) x[i * kSgFeatures + k] = f[k];
}

// Voxel i (0..31) of wave chunk `ch` of a row: half (ch & 1) of the 4 x 4 x 4
// brick ch / 2 (x fastest, the half holds z-layers 2 (ch & 1) and 2 (ch & 1) + 1).
__device__ inline long long sg_voxel(const SgParams& q, long long ch, int i, bool& ok) {
    const int nbx = (q.X + 3) >> 2, nby = (q.Y + 3) >> 2;
    const long long b = ch >> 1;
    const int bx = (int)(b % nbx), b2 = (int)(b / nbx), by = b2 % nby, bz = b2 / nby;
    const int ix = 4 * bx + (i & 3), iy = 4 * by + ((i >> 2) & 3), iz = 4 * bz + 2 * (int)(ch & 1) + (i >> 4);
    ok = ix < q.X && iy < q.Y && iz < q.Z;
    return ((long long)iz * q.Y + iy) * q.X + ix;
}

// One wave = half a 4 x 4 x 4 voxel brick (32 voxels) of one detector row.
// Features go through LDS into the MFMA B-operand layout; every member runs
// with register-resident activations and weight slabs shared by the
// workgroup's four waves through LDS (every wave runs every member).
// Disagreeing members flag the brick (tile) for the exact kernel, which prunes
// scatter sites per brick.
__global__ void __launch_bounds__(256) surrogate_apply_kernel(SgParams q, const SgRow* __restrict__ rows,
                                                              const ScPair* __restrict__ keep,
                                                              const int* __restrict__ keep_cnt,
                                                              const uint8_t* __restrict__ packed,
                                                              const float* __restrict__ pe, float* __restrict__ dst,
                                                              uint8_t* __restrict__ flags) {
    __shared__ float s_f[4][32][kSgFeatures + 1];
    __shared__ uint4 s_w[16 * 64];  // one weight slab, shared by the four waves
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, j16 = lane & 15;
    const long long chunks = 2 * q.n_tiles;
    const long long gw = (long long)blockIdx.x * 4 + w;
    const int rr = (int)(gw / chunks);
    const long long ch = gw % chunks;
    const bool valid_wave = rr < q.det_count;
    const int r = q.det_begin + (valid_wave ? rr : 0);
    const SgRow R = rows[r];
    const bool use = valid_wave && R.use != 0.0f;
    if (lane < 32) {
        bool ok;
        const long long v = sg_voxel(q, ch, lane, ok);
        float f[kSgFeatures];
        if (use && ok) {
            sg_features(q, R, keep + (long long)r * q.n_keep, keep_cnt[r], r, v, pe, dst[rr * q.row_stride + v], f);
        } else {
#pragma unroll
            for (int k = 0; k < kSgFeatures; ++k) f[k] = 0.0f;
        }
#pragma unroll
        for (int k = 0; k < kSgFeatures; ++k) s_f[w][lane][k] = f[k];
    }
    __syncthreads();
    bf16x8 xf[2];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 8 * g + s;
            xf[t][s] = (bf16)(k < kSgFeatures ? s_f[w][16 * t + j16][k] : 0.0f);
        }
    float far_sum[2] = {0.f, 0.f}, far_min[2] = {3.0e38f, 3.0e38f}, far_max[2] = {0.f, 0.f};
    for (int m = 0; m < q.n_members; ++m) {
        float y[2];
        mlp_block_forward(packed + (long long)m * kMlpPacked, xf, y, s_w);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const float fr = fmaxf(__expf(y[t] * q.sig_y + q.mu_y) - q.phi, 0.0f);
            far_sum[t] += fr;
            far_min[t] = fminf(far_min[t], fr);
            far_max[t] = fmaxf(far_max[t], fr);
        }
    }
    bool flag = false;
    if (g == 0 && use) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            bool ok;
            const long long v = sg_voxel(q, ch, 16 * t + j16, ok);
            if (!ok) continue;
            float* o = dst + rr * q.row_stride + v;
            const float s_hat = *o + far_sum[t] / (float)q.n_members;
            if (q.n_members > 1 && s_hat > q.floor_abs && (far_max[t] - far_min[t]) > q.tol * s_hat) flag = true;
            *o = s_hat;
        }
    } else if (g == 0 && valid_wave) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            bool ok;
            const long long v = sg_voxel(q, ch, 16 * t + j16, ok);
            if (ok) dst[rr * q.row_stride + v] = 0.0f;
        }
    }
    const unsigned long long any = __ballot(flag);
    if (lane == 0 && any && flags) flags[(long long)rr * q.n_tiles + (ch >> 1)] = 1;
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
    const long long per_row = 2 * q.n_tiles;  // waves per detector row: two per brick
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
