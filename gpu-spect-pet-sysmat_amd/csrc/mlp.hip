// Scatter-surrogate MLP kernels for gfx950: training (BF16 MFMA GEMMs with
// fused SiLU epilogues, head, reductions, Adam) and register-resident
// inference.  The host-side step sequencing and the C ABI are at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/sysmat.h"
#include "common.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"

namespace sysmat {

// ---------------------------------------------------------------------------
// GEMM: C(m,n) = sum_k A(m,k) B(k,n), bf16 operands staged through LDS in
// k-stages of up to 128 (k-contiguous rows, padded by 8 elements), 64x64 block
// tile, 4 waves of 32x32 (2x2 MFMA 16x16x32), FP32 accumulation, split-K over
// blockIdx.z.  The training GEMMs are small (K = 32 or 256, or 128 per split),
// so a wide stage means one or two global round trips per tile instead of
// eight.  K ranges are multiples of 32.
// ---------------------------------------------------------------------------
constexpr int kGemmKS = 128;
constexpr int kGemmLd = kGemmKS + 8;

template <bool AK, bool BK, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) bf16 As[64][kGemmLd];
    __shared__ __attribute__((aligned(16))) bf16 Bs[64][kGemmLd];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int n0 = blockIdx.x * 64, m0 = blockIdx.y * 64;
    const int kbeg = blockIdx.z * g.ksplit, kend = min(g.K, kbeg + g.ksplit);
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int l16 = lane & 15, g4 = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += kGemmKS) {
        const int nk = min(kGemmKS, kend - k0);
        // 64 x nk of each operand: (64 * 128 / 8) / 256 = 4 uint4 per thread
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int idx = tid + 256 * it;
            if (AK) {
                const int row = idx >> 4, k8 = (idx & 15) * 8, m = m0 + row;
                if (k8 < nk) {
                    uint4 v = {0u, 0u, 0u, 0u};
                    if (m < g.M) v = *reinterpret_cast<const uint4*>(g.A + (long long)m * g.lda + k0 + k8);
                    *reinterpret_cast<uint4*>(&As[row][k8]) = v;
                }
            } else {
                const int k = idx >> 3, m8 = (idx & 7) * 8, m = m0 + m8;
                if (k < nk) {
                    uint4 v = {0u, 0u, 0u, 0u};
                    if (m < g.M) v = *reinterpret_cast<const uint4*>(g.A + (long long)(k0 + k) * g.lda + m);
                    const bf16x8 e = __builtin_bit_cast(bf16x8, v);
#pragma unroll
                    for (int i = 0; i < 8; ++i) As[m8 + i][k] = e[i];
                }
            }
            if (BK) {
                const int col = idx >> 4, k8 = (idx & 15) * 8, n = n0 + col;
                if (k8 < nk) {
                    uint4 v = {0u, 0u, 0u, 0u};
                    if (n < g.N) v = *reinterpret_cast<const uint4*>(g.B + (long long)n * g.ldb + k0 + k8);
                    *reinterpret_cast<uint4*>(&Bs[col][k8]) = v;
                }
            } else {
                const int k = idx >> 3, n8 = (idx & 7) * 8, n = n0 + n8;
                if (k < nk) {
                    uint4 v = {0u, 0u, 0u, 0u};
                    if (n < g.N) v = *reinterpret_cast<const uint4*>(g.B + (long long)(k0 + k) * g.ldb + n);
                    const bf16x8 e = __builtin_bit_cast(bf16x8, v);
#pragma unroll
                    for (int i = 0; i < 8; ++i) Bs[n8 + i][k] = e[i];
                }
            }
        }
        __syncthreads();
        for (int kb = 0; kb < nk; kb += 32) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * i + l16][kb + 8 * g4]);
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * j + l16][kb + 8 * g4]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    // acc[i][j][r]: row m0 + wm + 16i + 4*g4 + r, column n0 + wn + 16j + l16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + wn + 16 * j + l16;
            if (n >= g.N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + 16 * i + 4 * g4 + r;
                if (m >= g.M) continue;
                const long long o = (long long)m * g.ldc + n;
                const float v = acc[i][j][r];
                if (EPI == EPI_FWD) {
                    const float z = v + g.bias[n];
                    g.Z[o] = z;
                    g.H[o] = (bf16)silu_f(z);
                } else if (EPI == EPI_BWD) {
                    g.D[o] = (bf16)(v * dsilu_f(g.Zin[o]));
                } else {
                    g.C[(long long)blockIdx.z * g.M * g.ldc + o] = v;
                }
            }
        }
}

hipError_t launch_gemm(int ak, int bk, int epi, const GemmArgs& g, int splits, hipStream_t s) {
    dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64), (unsigned)splits);
#define SM_GEMM(A_, B_, E_)                                                                  \
    if (ak == A_ && bk == B_ && epi == E_) {                                                \
        hipLaunchKernelGGL((gemm_kernel<A_ != 0, B_ != 0, E_>), grid, dim3(256), 0, s, g); \
        return hipGetLastError();                                                           \
    }
    SM_GEMM(1, 1, EPI_FWD)
    SM_GEMM(1, 0, EPI_BWD)
    SM_GEMM(0, 0, EPI_PART)
#undef SM_GEMM
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// element-wise and reduction kernels
// ---------------------------------------------------------------------------
__global__ void to_bf16_kernel(const float* __restrict__ p, bf16* __restrict__ q, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = (bf16)p[i];
}

// x [B][12] FP32 -> X [B][32] bf16 (zero padded); rows >= n are zero
__global__ void pack_input_kernel(const float* __restrict__ x, long long n, long long B, bf16* __restrict__ X) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B * kMlpK1) return;
    const long long b = i / kMlpK1;
    const int k = (int)(i % kMlpK1);
    X[i] = (bf16)((b < n && k < kMlpIn) ? x[b * kMlpIn + k] : 0.0f);
}

// y[b] = w5 . silu(Z4[b]) + b5; with targets: dy = 2 w (y - t), loss partial
// (one wave per row; rows >= n give y = 0, dy = 0).
// wpart (nullable): n_wpart partial sums of the raw weights wt, added in order;
// the weights are then wt / total (normalised batch weights).
__global__ void __launch_bounds__(256) head_fwd_kernel(const float* __restrict__ P, const float* __restrict__ Z4,
                                                       long long n, long long B, const float* __restrict__ t,
                                                       const float* __restrict__ wt, float* __restrict__ y,
                                                       float* __restrict__ dy, float* __restrict__ loss_part,
                                                       const float* __restrict__ wpart, int n_wpart) {
    __shared__ float s_loss[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long b = (long long)blockIdx.x * 4 + w;
    float l = 0.0f;
    if (b < B) {
        float s = 0.0f;
        if (b < n) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = lane + 64 * q;
                s += P[kOffW5 + k] * silu_f(Z4[b * kMlpH + k]);
            }
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) {
            const float yy = b < n ? s + P[kOffB5] : 0.0f;
            if (y && b < n) y[b] = yy;
            if (t) {
                float d = 0.0f;
                if (b < n) {
                    const float r = yy - t[b];
                    float ww = wt ? wt[b] : 1.0f / (float)n;
                    if (wpart) {
                        float tot = 0.0f;
                        for (int q = 0; q < n_wpart; ++q) tot += wpart[q];
                        ww *= 1.0f / tot;
                    }
                    d = 2.0f * ww * r;
                    l = ww * r * r;
                }
                dy[b] = d;
            }
        }
    }
    if (lane == 0) s_loss[w] = l;
    __syncthreads();
    if (threadIdx.x == 0 && loss_part) loss_part[blockIdx.x] = s_loss[0] + s_loss[1] + s_loss[2] + s_loss[3];
}

// dZ4 = bf16(dy w5 silu'(Z4)); partials of dw5 = sum dy h4, db5 = sum dy
// (block of 256 threads = columns, rows [blockIdx.x*R, +R)).
__global__ void __launch_bounds__(256) head_bwd_kernel(const float* __restrict__ P, const float* __restrict__ Z4,
                                                       const float* __restrict__ dy, long long B, int R,
                                                       bf16* __restrict__ dZ4, float* __restrict__ part) {
    const int k = threadIdx.x;
    const float wk = P[kOffW5 + k];
    float gw = 0.0f, gb = 0.0f;
    const long long b0 = (long long)blockIdx.x * R;
    const int nr = (int)std::min<long long>(R, B - b0);
    // unrolled so the row loads are in flight together (the sums keep their order)
#pragma unroll 8
    for (int i = 0; i < nr; ++i) {
        const long long b = b0 + i;
        const float z = Z4[b * kMlpH + k], d = dy[b];
        const float sg = 1.0f / (1.0f + __expf(-z));
        gw += d * (z * sg);
        gb += d;
        dZ4[b * kMlpH + k] = (bf16)(d * wk * (sg * (1.0f + z * (1.0f - sg))));
    }
    part[(long long)blockIdx.x * 257 + k] = gw;
    if (k == 0) part[(long long)blockIdx.x * 257 + 256] = gb;
}

// column sums of bf16 D [B][256] over row blocks -> part [blocks][256]
__global__ void __launch_bounds__(256) colsum_kernel(const bf16* __restrict__ D, long long B, int R,
                                                     float* __restrict__ part) {
    const int k = threadIdx.x;
    float s = 0.0f;
    const long long b0 = (long long)blockIdx.x * R;
    const int nr = (int)std::min<long long>(R, B - b0);
#pragma unroll 16
    for (int i = 0; i < nr; ++i) s += (float)D[(b0 + i) * kMlpH + k];
    part[(long long)blockIdx.x * kMlpH + k] = s;
}

// out[dst_off + i*dst_stride... ] : out[i] = sum_s part[s*stride + i], i < n
__global__ void reduce_parts_kernel(const float* __restrict__ part, int S, long long stride, long long n,
                                    float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
#pragma unroll 16
    for (int q = 0; q < S; ++q) s += part[(long long)q * stride + i];
    out[i] = s;
}

// dW partials [S][M][ldc] -> grad rows: out[m*ld_out + n] for n < ncols
__global__ void reduce_w_kernel(const float* __restrict__ part, int S, int M, long long ldc, int ncols,
                                float* __restrict__ out, long long ld_out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)M * ncols) return;
    const long long m = i / ncols, c = i % ncols;
    float s = 0.0f;
#pragma unroll 16
    for (int q = 0; q < S; ++q) s += part[(long long)q * M * ldc + m * ldc + c];
    out[m * ld_out + c] = s;
}

// pb (nullable): the bf16 copy of the updated parameters for the next step's GEMMs
__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr_t, float b1, float b2, float eps,
                            bf16* __restrict__ pb) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = i <= kOffB5 ? gr[i] : 0.0f;  // the padding after b5 has no gradient
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pn = p[i] - lr_t * mi / (sqrtf(vi) + eps);
    p[i] = pn;
    if (pb) pb[i] = (bf16)pn;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Batch of training step `step` drawn on the device: row b takes pair
// mix64(mix64(seed ^ K step) + b) mod n (with replacement).  One thread per row
// gathers the features into the bf16 GEMM layout (12 of 32 columns, rest zero),
// the target and the raw weight (w NULL: 1); each 256-row block leaves the sum of
// its weights (wave xor-butterflies, then the four waves in order) in wpart,
// which the head kernel adds in block order to normalise the weights.
constexpr int kSampleRows = 256;
__global__ void __launch_bounds__(kSampleRows) sample_batch_kernel(const float* __restrict__ x,
                                                                   const float* __restrict__ t,
                                                                   const float* __restrict__ w, long long n, int B,
                                                                   unsigned long long seed, long long step,
                                                                   bf16* __restrict__ X, float* __restrict__ tb,
                                                                   float* __restrict__ wb, float* __restrict__ wpart) {
    __shared__ float s_w[kSampleRows / 64];
    const int b = blockIdx.x * kSampleRows + threadIdx.x;
    float wi = 0.0f;
    if (b < B) {
        const unsigned long long h = mix64(mix64(seed ^ (0x632BE59BD9B4E019ull * (unsigned long long)step)) +
                                           (unsigned long long)b);
        const long long i = (long long)(h % (unsigned long long)n);
        const float* xi = x + i * kMlpIn;
        bf16x8 e[4];
#pragma unroll
        for (int k = 0; k < kMlpK1; ++k) e[k >> 3][k & 7] = (bf16)(k < kMlpIn ? xi[k] : 0.0f);
        uint4* dst = reinterpret_cast<uint4*>(X + (long long)b * kMlpK1);
#pragma unroll
        for (int q = 0; q < 4; ++q) dst[q] = __builtin_bit_cast(uint4, e[q]);
        tb[b] = t[i];
        wi = w ? w[i] : 1.0f;
        wb[b] = wi;
    }
    float ws = wi;
    for (int o = 32; o > 0; o >>= 1) ws += __shfl_xor(ws, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        float tot = 0.0f;
#pragma unroll
        for (int q = 0; q < kSampleRows / 64; ++q) tot += s_w[q];
        wpart[blockIdx.x] = tot;
    }
}

// uniform +-sqrt(6/fan_in) weights, zero biases, w5 uniform +-1/16, zero W1 padding
__global__ void init_kernel(float* __restrict__ p, unsigned long long seed) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kMlpParams) return;
    const float u = (float)((mix64(seed * 0x100000001B3ull + (unsigned long long)i) >> 40) * (1.0 / 16777216.0)) * 2.0f -
                    1.0f;
    float v = 0.0f;
    if (i < kOffB1) {
        v = (i % kMlpK1) < kMlpIn ? u * sqrtf(6.0f / kMlpIn) : 0.0f;
    } else if ((i >= kOffW2 && i < kOffB2) || (i >= kOffW3 && i < kOffB3) || (i >= kOffW4 && i < kOffB4)) {
        v = u * sqrtf(6.0f / kMlpH);
    } else if (i >= kOffW5 && i < kOffB5) {
        v = u / 16.0f;
    }
    p[i] = v;
}

// flat FP32 parameters -> packed inference layout (lane-major MFMA A tiles)
__global__ void pack_kernel(const float* __restrict__ P, uint8_t* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // one bf16 slot of a tile, or a vector entry
    constexpr long long n_tile_slots = (16 + 3 * 128) * 512;                // 512 bf16 per tile
    if (i < n_tile_slots) {
        const long long tile = i / 512;
        const int within = (int)(i % 512), lane = within / 8, s = within % 8, g = lane >> 4, r = lane & 15;
        bf16* dst = reinterpret_cast<bf16*>(out);
        float v;
        if (tile < 16) {  // layer 1: m = tile, direct feature labels
            const int row = 16 * (int)tile + r, k = 8 * g + s;
            v = P[kOffW1 + row * kMlpK1 + k];
            dst[kPkL1 / 2 + tile * 512 + within] = (bf16)v;
        } else {
            const long long t2 = tile - 16;
            const int layer = (int)(t2 / 128), mt = (int)((t2 % 128) / 8), kb = (int)(t2 % 8);
            const long long off = layer == 0 ? kOffW2 : (layer == 1 ? kOffW3 : kOffW4);
            const int row = 16 * mt + r, k = 32 * kb + k_perm(g, s);
            v = P[off + (long long)row * kMlpH + k];
            const long long base = (layer == 0 ? kPkL2 : (layer == 1 ? kPkL3 : kPkL4)) / 2;
            dst[base + (long long)(mt * 8 + kb) * 512 + within] = (bf16)v;
        }
        return;
    }
    const long long j = i - n_tile_slots;
    float* f = reinterpret_cast<float*>(out);
    if (j < 1024) {
        const long long boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
        f[kPkBias / 4 + j] = P[boff[j / 256] + (j % 256)];
    } else if (j < 1280) {
        f[kPkW5 / 4 + (j - 1024)] = P[kOffW5 + (j - 1024)];
    } else if (j == 1280) {
        f[kPkB5 / 4] = P[kOffB5];
    }
}

// generic inference on feature rows: y[m*n + e] for members m; a wave takes 48
// rows (three MFMA element tiles, the fused apply kernel's form; outputs are the
// same for any tiling)
constexpr int kInferTiles = 3;
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
mlp_infer_kernel(const uint8_t* __restrict__ packed, int n_members, const float* __restrict__ x, long long n,
                 float* __restrict__ y) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4;
    const long long e0 = ((long long)blockIdx.x * 4 + w) * (16 * kInferTiles);
    if (e0 >= n) return;
    bf16x8 xf[kInferTiles];
#pragma unroll
    for (int t = 0; t < kInferTiles; ++t) {
        const long long e = e0 + 16 * t + (lane & 15);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int k = 8 * g + s;
            xf[t][s] = (bf16)((e < n && k < kMlpIn) ? x[e * kMlpIn + k] : 0.0f);
        }
    }
    for (int m = 0; m < n_members; ++m) {
        float yy[kInferTiles];
        mlp_wave_forward<kInferTiles, 2>(packed + (long long)m * kMlpPacked, xf, yy);
        if (g == 0) {
#pragma unroll
            for (int t = 0; t < kInferTiles; ++t) {
                const long long e = e0 + 16 * t + lane;
                if (e < n) y[(long long)m * n + e] = yy[t];
            }
        }
    }
}

hipError_t launch_mlp_infer(const uint8_t* packed, int n_members, const float* x, long long n, float* y,
                            hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long per_block = 4 * 16 * kInferTiles;
    const long long blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)blocks), dim3(256), 0, s, packed, n_members, x, n, y);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host sequencing
// ---------------------------------------------------------------------------
namespace {

constexpr int kHeadRows = 64;  // rows per block of the head / column-sum reductions
constexpr float kAdamB1 = 0.9f, kAdamB2 = 0.999f, kAdamEps = 1e-8f;

// Adam step size with the bias corrections of step `step` (counting from 1)
float adam_lr(float lr, int step) {
    return lr * std::sqrt(1.0f - std::pow(kAdamB2, (float)step)) / (1.0f - std::pow(kAdamB1, (float)step));
}

struct Ws {
    bf16* Pb;
    bf16* X;
    float* Z;   // [4][B][256]
    bf16* A;    // [4][B][256]  (a1..a4)
    float* y;
    float* dy;
    bf16* dZ;   // [2][B][256]
    float* part;
    float* loss_part;
    float* grad;
    float* tb;  // sampled batch targets / raw weights / per-block weight sums (multi-step training)
    float* wb;
    float* wpart;
    long long B;
    int splits;
};

long long align256(long long x) { return (x + 255) / 256 * 256; }

// K splits of the weight-gradient GEMMs (K = batch): 128 rows per split, so a
// 256 x 256 gradient at batch 4096 runs as 16 tiles x 32 splits = 512 workgroups
// (two per CU) instead of 128, and each walks 4 instead of 16 k-steps.
int wgrad_splits(long long B) {
    const long long s = std::max(1LL, std::min(64LL, B / 128));
    return (int)s;
}

long long ws_bytes(long long B, Ws* w, uint8_t* base) {
    const int S = wgrad_splits(B);
    const long long hb = (B + kHeadRows - 1) / kHeadRows;
    long long off = 0;
    auto take = [&](long long bytes) {
        const long long o = off;
        off = align256(off + bytes);
        return base ? base + o : nullptr;
    };
    uint8_t* pb = take(kMlpParams * 2);
    uint8_t* x = take(B * kMlpK1 * 2);
    uint8_t* z = take(4 * B * kMlpH * 4);
    uint8_t* a = take(4 * B * kMlpH * 2);
    uint8_t* y = take(B * 4);
    uint8_t* dy = take(B * 4);
    uint8_t* dz = take(2 * B * kMlpH * 2);
    const long long part_floats = std::max<long long>((long long)S * kMlpH * kMlpH, hb * 257);
    uint8_t* part = take(part_floats * 4);
    uint8_t* lp = take(((B + 3) / 4) * 4);
    uint8_t* gr = take(kMlpParams * 4);
    uint8_t* tb = take(B * 4);
    uint8_t* wb = take(B * 4);
    uint8_t* wp = take(((B + kSampleRows - 1) / kSampleRows) * 4);
    if (w) {
        w->Pb = (bf16*)pb;
        w->X = (bf16*)x;
        w->Z = (float*)z;
        w->A = (bf16*)a;
        w->y = (float*)y;
        w->dy = (float*)dy;
        w->dZ = (bf16*)dz;
        w->part = (float*)part;
        w->loss_part = (float*)lp;
        w->grad = (float*)gr;
        w->tb = (float*)tb;
        w->wb = (float*)wb;
        w->wpart = (float*)wp;
        w->B = B;
        w->splits = S;
    }
    return off;
}

unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

int check_batch(int n, long long B) {
    if (n <= 0 || B <= 0 || B % 64 != 0 || n > B) return set_error(SM_E_ARG, "mlp: batch must be a positive multiple of 64 >= n");
    return SM_OK;
}

// forward with activations kept for backward; x: [n][12] FP32 (rows n..B-1 zero);
// x NULL: w.X and w.Pb already hold the packed batch and the bf16 parameters
int forward_train(const float* P, const float* x, int n, Ws& w, hipStream_t s) {
    const long long B = w.B;
    if (x) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(kMlpParams)), dim3(256), 0, s, P, w.Pb, kMlpParams);
        hipLaunchKernelGGL(pack_input_kernel, dim3(nblk(B * kMlpK1)), dim3(256), 0, s, x, (long long)n, B, w.X);
        HIP_TRY(hipGetLastError());
    }
    const long long woff[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
    const long long boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    for (int l = 0; l < 4; ++l) {
        GemmArgs g{};
        g.A = l == 0 ? w.X : w.A + (long long)(l - 1) * B * kMlpH;
        g.lda = l == 0 ? kMlpK1 : kMlpH;
        g.B = w.Pb + woff[l];
        g.ldb = l == 0 ? kMlpK1 : kMlpH;
        g.M = (int)B;
        g.N = kMlpH;
        g.K = l == 0 ? kMlpK1 : kMlpH;
        g.ksplit = g.K;
        g.ldc = kMlpH;
        g.bias = P + boff[l];
        g.Z = w.Z + (long long)l * B * kMlpH;
        g.H = w.A + (long long)l * B * kMlpH;
        HIP_TRY(launch_gemm(1, 1, EPI_FWD, g, 1, s));
    }
    return SM_OK;
}

// backward from dy (in w.dy) into w.grad
int backward(const float* P, int n, Ws& w, hipStream_t s) {
    const long long B = w.B;
    const int hb = (int)((B + kHeadRows - 1) / kHeadRows);
    const float* Z4 = w.Z + 3 * B * kMlpH;
    bf16* dZ = w.dZ;
    hipLaunchKernelGGL(head_bwd_kernel, dim3(hb), dim3(256), 0, s, P, Z4, w.dy, B, kHeadRows, dZ, w.part);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(nblk(257)), dim3(256), 0, s, w.part, hb, 257LL, 257LL,
                       w.grad + kOffW5);
    HIP_TRY(hipGetLastError());
    const long long woff[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
    const long long boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    int cur = 0;
    for (int l = 3; l >= 0; --l) {
        bf16* D = dZ + (long long)cur * B * kMlpH;
        // bias gradient
        hipLaunchKernelGGL(colsum_kernel, dim3(hb), dim3(256), 0, s, D, B, kHeadRows, w.part);
        hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(256), 0, s, w.part, hb, (long long)kMlpH,
                           (long long)kMlpH, w.grad + boff[l]);
        HIP_TRY(hipGetLastError());
        // weight gradient: dW[m][k] = sum_b D[b][m] a[b][k]
        const int ncols = l == 0 ? kMlpK1 : kMlpH;
        GemmArgs g{};
        g.A = D;
        g.lda = kMlpH;
        g.B = l == 0 ? w.X : w.A + (long long)(l - 1) * B * kMlpH;
        g.ldb = ncols;
        g.M = kMlpH;
        g.N = ncols;
        g.K = (int)B;
        const int S = w.splits;
        g.ksplit = (int)(((B + S - 1) / S + 31) / 32 * 32);
        const int S_eff = (int)((B + g.ksplit - 1) / g.ksplit);
        g.ldc = ncols;
        g.C = w.part;
        HIP_TRY(launch_gemm(0, 0, EPI_PART, g, S_eff, s));
        hipLaunchKernelGGL(reduce_w_kernel, dim3(nblk((long long)kMlpH * ncols)), dim3(256), 0, s, w.part, S_eff,
                           kMlpH, (long long)ncols, ncols, w.grad + woff[l], (long long)ncols);
        HIP_TRY(hipGetLastError());
        if (l > 0) {
            // dZ_{l-1} = bf16((D W_l) * silu'(Z_{l-1}))
            GemmArgs h{};
            h.A = D;
            h.lda = kMlpH;
            h.B = w.Pb + woff[l];
            h.ldb = kMlpH;
            h.M = (int)B;
            h.N = kMlpH;
            h.K = kMlpH;
            h.ksplit = kMlpH;
            h.ldc = kMlpH;
            h.Zin = w.Z + (long long)(l - 1) * B * kMlpH;
            h.D = dZ + (long long)(1 - cur) * B * kMlpH;
            HIP_TRY(launch_gemm(1, 0, EPI_BWD, h, 1, s));
            cur = 1 - cur;
        }
    }
    return SM_OK;
}

}  // namespace
}  // namespace sysmat

using namespace sysmat;

extern "C" {

long long sm_mlp_param_count(void) { return kMlpParams; }
long long sm_mlp_packed_bytes(void) { return kMlpPacked; }

long long sm_mlp_workspace_bytes(int batch) {
    if (batch <= 0 || batch % 64 != 0) return set_error(SM_E_ARG, "mlp: batch must be a positive multiple of 64");
    return ws_bytes(batch, nullptr, nullptr);
}

int sm_mlp_init(float* params_dev, unsigned long long seed, void* stream) {
    clear_error();
    if (!params_dev) return set_error(SM_E_ARG, "null params");
    hipLaunchKernelGGL(init_kernel, dim3(nblk(kMlpParams)), dim3(256), 0, (hipStream_t)stream, params_dev, seed);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

int sm_mlp_forward(const float* params_dev, const float* x_dev, int n, int batch, void* ws, long long ws_size,
                   float* y_dev, void* stream) {
    clear_error();
    int rc = check_batch(n, batch);
    if (rc < 0) return rc;
    if (!params_dev || !x_dev || !ws || !y_dev) return set_error(SM_E_ARG, "null argument");
    Ws w{};
    if (ws_bytes(batch, &w, (uint8_t*)ws) > ws_size) return set_error(SM_E_ARG, "mlp: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    if ((rc = forward_train(params_dev, x_dev, n, w, s)) < 0) return rc;
    hipLaunchKernelGGL(head_fwd_kernel, dim3(nblk(batch, 4)), dim3(256), 0, s, params_dev,
                       w.Z + 3LL * batch * kMlpH, (long long)n, (long long)batch, nullptr, nullptr, y_dev, nullptr,
                       nullptr, nullptr, 0);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

int sm_mlp_backward(const float* params_dev, const float* dy_dev, int n, int batch, void* ws, long long ws_size,
                    float* grad_dev, void* stream) {
    clear_error();
    int rc = check_batch(n, batch);
    if (rc < 0) return rc;
    if (!params_dev || !dy_dev || !ws || !grad_dev) return set_error(SM_E_ARG, "null argument");
    Ws w{};
    if (ws_bytes(batch, &w, (uint8_t*)ws) > ws_size) return set_error(SM_E_ARG, "mlp: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(w.dy, dy_dev, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    if (batch > n) HIP_TRY(hipMemsetAsync(w.dy + n, 0, sizeof(float) * (batch - n), s));
    HIP_TRY(hipMemsetAsync(w.grad, 0, sizeof(float) * kMlpParams, s));
    if ((rc = backward(params_dev, n, w, s)) < 0) return rc;
    HIP_TRY(hipMemcpyAsync(grad_dev, w.grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice, s));
    return SM_OK;
}

int sm_mlp_train_step(float* params_dev, float* adam_m, float* adam_v, const float* x_dev, const float* t_dev,
                      const float* w_dev, int n, int batch, float lr, int step, void* ws, long long ws_size,
                      float* loss_dev, void* stream) {
    clear_error();
    int rc = check_batch(n, batch);
    if (rc < 0) return rc;
    if (!params_dev || !adam_m || !adam_v || !x_dev || !t_dev || !ws) return set_error(SM_E_ARG, "null argument");
    if (step < 1) return set_error(SM_E_ARG, "mlp: Adam step counts from 1");
    Ws w{};
    if (ws_bytes(batch, &w, (uint8_t*)ws) > ws_size) return set_error(SM_E_ARG, "mlp: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    if ((rc = forward_train(params_dev, x_dev, n, w, s)) < 0) return rc;
    const unsigned hb = nblk(batch, 4);
    hipLaunchKernelGGL(head_fwd_kernel, dim3(hb), dim3(256), 0, s, params_dev, w.Z + 3LL * batch * kMlpH,
                       (long long)n, (long long)batch, t_dev, w_dev, w.y, w.dy, w.loss_part, nullptr, 0);
    HIP_TRY(hipGetLastError());
    if (loss_dev) {
        hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(1), 0, s, w.loss_part, (int)hb, 1LL, 1LL, loss_dev);
        HIP_TRY(hipGetLastError());
    }
    // every gradient entry up to b5 is written (not accumulated) by the reductions of
    // backward(); Adam reads the padding after b5 as zero
    if ((rc = backward(params_dev, n, w, s)) < 0) return rc;
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(kMlpParams)), dim3(256), 0, s, params_dev, w.grad, adam_m, adam_v,
                       kMlpParams, adam_lr(lr, step), kAdamB1, kAdamB2, kAdamEps, (bf16*)nullptr);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

int sm_mlp_train_run(int n_members, float* const* params_dev, float* const* adam_m, float* const* adam_v,
                     const float* x_dev, const float* t_dev, const float* w_dev, long long n, int batch, int steps,
                     float lr, float lr_floor, const unsigned long long* seeds, int step0, void* const* ws,
                     long long ws_size, float* const* loss_dev, void* stream) {
    clear_error();
    if (n_members < 1 || n_members > 8) return set_error(SM_E_ARG, "mlp: 1 to 8 members per run");
    if (n <= 0 || batch <= 0 || batch % 64 != 0) return set_error(SM_E_ARG, "mlp: need n > 0 and a batch that is a positive multiple of 64");
    if (steps < 0 || step0 < 0) return set_error(SM_E_ARG, "mlp: negative step count");
    if (!params_dev || !adam_m || !adam_v || !x_dev || !t_dev || !seeds || !ws) return set_error(SM_E_ARG, "null argument");
    Ws w[8];
    for (int m = 0; m < n_members; ++m) {
        if (!params_dev[m] || !adam_m[m] || !adam_v[m] || !ws[m]) return set_error(SM_E_ARG, "null argument");
        if (ws_bytes(batch, &w[m], (uint8_t*)ws[m]) > ws_size) return set_error(SM_E_ARG, "mlp: workspace too small");
    }
    // member m runs on its own stream (member 0 on the caller's), forked from and joined
    // back to the caller's stream: the members' short, latency-bound kernels overlap
    hipStream_t s0 = (hipStream_t)stream;
    hipStream_t st[8] = {s0};
    hipEvent_t fork = nullptr, done[8] = {};
    // (destroying a stream or event with work pending is deferred by the runtime until
    // that work completes, so cleanup never waits)
    auto cleanup = [&]() {
        for (int m = 1; m < n_members; ++m) {
            if (done[m]) (void)hipEventDestroy(done[m]);
            if (st[m]) (void)hipStreamDestroy(st[m]);
        }
        if (fork) (void)hipEventDestroy(fork);
    };
    auto fail = [&](int rc) {
        cleanup();
        return rc;
    };
    if (n_members > 1) {
        if (hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: event"));
        if (hipEventRecord(fork, s0) != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: event record"));
        for (int m = 1; m < n_members; ++m) {
            if (hipStreamCreateWithFlags(&st[m], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&done[m], hipEventDisableTiming) != hipSuccess ||
                hipStreamWaitEvent(st[m], fork, 0) != hipSuccess)
                return fail(set_error(SM_E_DEVICE, "mlp: member stream"));
        }
    }
    const unsigned hb = nblk(batch, 4);
    const int nwp = (batch + kSampleRows - 1) / kSampleRows;
    int rc;
    for (int m = 0; m < n_members; ++m) {
        hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(kMlpParams)), dim3(256), 0, st[m], params_dev[m], w[m].Pb,
                           kMlpParams);
        if (hipGetLastError() != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: launch"));
    }
    for (int it = 0; it < steps; ++it) {
        const int step = step0 + it + 1;
        // cosine schedule from lr over the run, held at lr_floor from below
        const float lr_it = std::max(lr * 0.5f * (1.0f + std::cos(3.14159265358979f * (float)it / (float)steps)),
                                     lr_floor);
        for (int m = 0; m < n_members; ++m) {
            Ws& W = w[m];
            hipStream_t s = st[m];
            hipLaunchKernelGGL(sample_batch_kernel, dim3(nwp), dim3(kSampleRows), 0, s, x_dev, t_dev, w_dev, n,
                               batch, seeds[m], (long long)step, W.X, W.tb, W.wb, W.wpart);
            if (hipGetLastError() != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: launch"));
            if ((rc = forward_train(params_dev[m], nullptr, batch, W, s)) < 0) return fail(rc);
            hipLaunchKernelGGL(head_fwd_kernel, dim3(hb), dim3(256), 0, s, params_dev[m],
                               W.Z + 3LL * batch * kMlpH, (long long)batch, (long long)batch, W.tb, W.wb, W.y, W.dy,
                               W.loss_part, W.wpart, nwp);
            if (loss_dev && loss_dev[m] && it == steps - 1)
                hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(1), 0, s, W.loss_part, (int)hb, 1LL, 1LL,
                                   loss_dev[m]);
            if (hipGetLastError() != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: launch"));
            if ((rc = backward(params_dev[m], batch, W, s)) < 0) return fail(rc);
            hipLaunchKernelGGL(adam_kernel, dim3(nblk(kMlpParams)), dim3(256), 0, s, params_dev[m], W.grad, adam_m[m],
                               adam_v[m], kMlpParams, adam_lr(lr_it, step), kAdamB1, kAdamB2, kAdamEps, W.Pb);
            if (hipGetLastError() != hipSuccess) return fail(set_error(SM_E_DEVICE, "mlp: launch"));
        }
    }
    for (int m = 1; m < n_members; ++m) {
        if (hipEventRecord(done[m], st[m]) != hipSuccess || hipStreamWaitEvent(s0, done[m], 0) != hipSuccess)
            return fail(set_error(SM_E_DEVICE, "mlp: member join"));
    }
    cleanup();
    return SM_OK;
}

int sm_mlp_pack(const float* params_dev, void* packed_dev, void* stream) {
    clear_error();
    if (!params_dev || !packed_dev) return set_error(SM_E_ARG, "null argument");
    const long long total = (16 + 3 * 128) * 512 + 1281;
    hipLaunchKernelGGL(pack_kernel, dim3(nblk(total)), dim3(256), 0, (hipStream_t)stream, params_dev,
                       (uint8_t*)packed_dev);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

int sm_mlp_infer(const void* packed_dev, int n_members, const float* x_dev, long long n, float* y_dev, void* stream) {
    clear_error();
    if (!packed_dev || !x_dev || !y_dev || n_members < 1 || n < 0) return set_error(SM_E_ARG, "bad argument");
    HIP_TRY(launch_mlp_infer((const uint8_t*)packed_dev, n_members, x_dev, n, y_dev, (hipStream_t)stream));
    return SM_OK;
}

}  // extern "C"
