// Register-resident MLP forward for one wavefront (device code, header-only so
// the generic inference kernel and the fused scatter-surrogate kernel share it).
//
// A wave evaluates 32 elements as two 16-column tiles t.  Activations never
// leave the VGPRs: the accumulators of output tiles (2kb, 2kb+1) of a layer,
// rounded to bf16, are exactly the B operand of K block kb of the next layer
// (k_perm), so each layer is 8 x 8 x 4 MFMAs with weights streamed as 1 KiB
// lane-major tiles (L2/MALL resident, 406 KiB per member).
#pragma once
#include "mlp.hpp"

namespace sysmat {

__device__ __forceinline__ bf16x8 ld_tile(const uint8_t* __restrict__ base, long long tile, int lane) {
    const uint4 u = reinterpret_cast<const uint4*>(base + tile * 1024)[lane];
    return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bias + SiLU on output tile pair kb and repack as bf16 B operand slots.
__device__ __forceinline__ void act_pack(const f32x4 (&acc)[2][2], const float* __restrict__ bias, int kb, int g,
                                         bf16x8 (&out)[2]) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float z = acc[mm][t][i] + bias[32 * kb + 16 * mm + 4 * g + i];
                out[t][mm * 4 + i] = (bf16)silu_f(z);
            }
        }
    }
}

// One hidden 256 -> 256 layer: in[kb][t] -> out[kb][t].
__device__ __forceinline__ void hidden_layer(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                             const bf16x8 (&in)[8][2], bf16x8 (&out)[8][2], int lane) {
    const int g = lane >> 4;
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][2];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[mm][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kbi = 0; kbi < 8; ++kbi) {
            const bf16x8 a0 = ld_tile(W, (2 * kbo) * 8 + kbi, lane);
            const bf16x8 a1 = ld_tile(W, (2 * kbo + 1) * 8 + kbi, lane);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                acc[0][t] = mfma16(a0, in[kbi][t], acc[0][t]);
                acc[1][t] = mfma16(a1, in[kbi][t], acc[1][t]);
            }
        }
        act_pack(acc, bias, kbo, g, out[kbo]);
    }
}

__device__ __forceinline__ void mlp_wave_forward(const uint8_t* __restrict__ P, const bf16x8 (&x)[2], float (&y)[2]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const float* bias = reinterpret_cast<const float*>(P + kPkBias);
    const float* w5 = reinterpret_cast<const float*>(P + kPkW5);
    bf16x8 A[8][2], B[8][2];
    // layer 1: one K block (features), output tile pairs kb = 0..7
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        const bf16x8 a0 = ld_tile(P + kPkL1, 2 * kb, lane), a1 = ld_tile(P + kPkL1, 2 * kb + 1, lane);
        f32x4 acc[2][2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[0][t] = mfma16(a0, x[t], f32x4{0.f, 0.f, 0.f, 0.f});
            acc[1][t] = mfma16(a1, x[t], f32x4{0.f, 0.f, 0.f, 0.f});
        }
        act_pack(acc, bias, kb, g, A[kb]);
    }
    hidden_layer(P + kPkL2, bias + 256, A, B, lane);
    hidden_layer(P + kPkL3, bias + 512, B, A, lane);
    // layer 4 + head: y = w5 . silu(z4) + b5, reduced over the four lane groups
    float yp[2] = {0.f, 0.f};
    const float* b4 = bias + 768;
    const uint8_t* W4 = P + kPkL4;
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][2];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[mm][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kbi = 0; kbi < 8; ++kbi) {
            const bf16x8 a0 = ld_tile(W4, (2 * kbo) * 8 + kbi, lane);
            const bf16x8 a1 = ld_tile(W4, (2 * kbo + 1) * 8 + kbi, lane);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                acc[0][t] = mfma16(a0, A[kbi][t], acc[0][t]);
                acc[1][t] = mfma16(a1, A[kbi][t], acc[1][t]);
            }
        }
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int n = 32 * kbo + 16 * mm + 4 * g + i;
                const float wn = w5[n], bn = b4[n];
#pragma unroll
                for (int t = 0; t < 2; ++t) yp[t] += wn * silu_f(acc[mm][t][i] + bn);
            }
    }
    const float b5 = *reinterpret_cast<const float*>(P + kPkB5);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        float v = yp[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[t] = v + b5;
    }
}

// Block-cooperative forward for a 256-thread workgroup (4 waves x 32 elements):
// each 16-tile weight slab (output tile pair kbo of a layer: tiles 16 kbo ..
// 16 kbo + 15, contiguous in the packed layout) is staged once into LDS and
// read by all four waves, so the L2 reads per element drop fourfold.  The MFMAs
// and operands are those of mlp_wave_forward, so the outputs are identical.
// Every thread of the workgroup must call it (barriers); s_w holds 16 KiB.
__device__ __forceinline__ void stage_slab(const uint8_t* __restrict__ base, long long tile0, uint4* s_w) {
    __syncthreads();  // the previous slab has been read by every wave
    const uint4* src = reinterpret_cast<const uint4*>(base + tile0 * 1024);
#pragma unroll
    for (int i = 0; i < 4; ++i) s_w[threadIdx.x + 256 * i] = src[threadIdx.x + 256 * i];
    __syncthreads();
}

__device__ __forceinline__ bf16x8 lds_tile(const uint4* s_w, int t, int lane) {
    return __builtin_bit_cast(bf16x8, s_w[t * 64 + lane]);
}

__device__ __forceinline__ void hidden_layer_block(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                                   const bf16x8 (&in)[8][2], bf16x8 (&out)[8][2], int lane,
                                                   uint4* s_w) {
    const int g = lane >> 4;
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        stage_slab(W, 16 * kbo, s_w);
        f32x4 acc[2][2];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[mm][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kbi = 0; kbi < 8; ++kbi) {
            const bf16x8 a0 = lds_tile(s_w, kbi, lane);
            const bf16x8 a1 = lds_tile(s_w, 8 + kbi, lane);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                acc[0][t] = mfma16(a0, in[kbi][t], acc[0][t]);
                acc[1][t] = mfma16(a1, in[kbi][t], acc[1][t]);
            }
        }
        act_pack(acc, bias, kbo, g, out[kbo]);
    }
}

__device__ __forceinline__ void mlp_block_forward(const uint8_t* __restrict__ P, const bf16x8 (&x)[2],
                                                  float (&y)[2], uint4* s_w) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const float* bias = reinterpret_cast<const float*>(P + kPkBias);
    const float* w5 = reinterpret_cast<const float*>(P + kPkW5);
    bf16x8 A[8][2], B[8][2];
    stage_slab(P + kPkL1, 0, s_w);
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        const bf16x8 a0 = lds_tile(s_w, 2 * kb, lane), a1 = lds_tile(s_w, 2 * kb + 1, lane);
        f32x4 acc[2][2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[0][t] = mfma16(a0, x[t], f32x4{0.f, 0.f, 0.f, 0.f});
            acc[1][t] = mfma16(a1, x[t], f32x4{0.f, 0.f, 0.f, 0.f});
        }
        act_pack(acc, bias, kb, g, A[kb]);
    }
    hidden_layer_block(P + kPkL2, bias + 256, A, B, lane, s_w);
    hidden_layer_block(P + kPkL3, bias + 512, B, A, lane, s_w);
    float yp[2] = {0.f, 0.f};
    const float* b4 = bias + 768;
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        stage_slab(P + kPkL4, 16 * kbo, s_w);
        f32x4 acc[2][2];
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int t = 0; t < 2; ++t) acc[mm][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kbi = 0; kbi < 8; ++kbi) {
            const bf16x8 a0 = lds_tile(s_w, kbi, lane);
            const bf16x8 a1 = lds_tile(s_w, 8 + kbi, lane);
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                acc[0][t] = mfma16(a0, A[kbi][t], acc[0][t]);
                acc[1][t] = mfma16(a1, A[kbi][t], acc[1][t]);
            }
        }
#pragma unroll
        for (int mm = 0; mm < 2; ++mm)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int n = 32 * kbo + 16 * mm + 4 * g + i;
                const float wn = w5[n], bn = b4[n];
#pragma unroll
                for (int t = 0; t < 2; ++t) yp[t] += wn * silu_f(acc[mm][t][i] + bn);
            }
    }
    const float b5 = *reinterpret_cast<const float*>(P + kPkB5);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        float v = yp[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[t] = v + b5;
    }
}

}  // namespace sysmat
