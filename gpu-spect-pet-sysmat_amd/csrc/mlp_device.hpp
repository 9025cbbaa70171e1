// Register-resident MLP forward for one wavefront (device code, header-only so
// the generic inference kernel and the fused scatter-surrogate kernel share it).
//
// A wave evaluates 32 elements as two 16-column tiles t.  Activations never
// leave the VGPRs: the accumulators of output tiles (2kb, 2kb+1) of a layer,
// rounded to bf16, are exactly the B operand of K block kb of the next layer
// (k_perm), so each layer is 8 x 8 x 4 MFMAs with weights streamed as 1 KiB
// lane-major tiles (L2 resident, 406 KiB per member).
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

// Accumulators of output tile pair kb start at the bias: lane (g, j) of tile mm
// holds rows 32 kb + 16 mm + 4 g + i (i = 0..3) of column j, the same four
// rows for both element tiles, so one 16-byte load per mm serves both.
__device__ __forceinline__ void bias_acc(const float* __restrict__ bias, int kb, int g, f32x4 (&acc)[2][2]) {
#pragma unroll
    for (int mm = 0; mm < 2; ++mm) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 32 * kb + 16 * mm + 4 * g);
        acc[mm][0] = b;
        acc[mm][1] = b;
    }
}

// SiLU on output tile pair kb (bias already in the accumulators) and repack as
// bf16 B operand slots.
__device__ __forceinline__ void act_pack(const f32x4 (&acc)[2][2], bf16x8 (&out)[2]) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
#pragma unroll
            for (int i = 0; i < 4; ++i) out[t][mm * 4 + i] = (bf16)silu_f(acc[mm][t][i]);
        }
    }
}

// z = W in + bias for output tile pair kbo of a 256 -> 256 layer; its 16 weight tiles
// are requested together before the first MFMA, so their L2 latencies overlap.
__device__ __forceinline__ void tile_pair(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                          const bf16x8 (&in)[8][2], int kbo, int lane, f32x4 (&acc)[2][2]) {
    bf16x8 a0[8], a1[8];
#pragma unroll
    for (int kbi = 0; kbi < 8; ++kbi) {
        a0[kbi] = ld_tile(W, (2 * kbo) * 8 + kbi, lane);
        a1[kbi] = ld_tile(W, (2 * kbo + 1) * 8 + kbi, lane);
    }
    bias_acc(bias, kbo, lane >> 4, acc);
#pragma unroll
    for (int kbi = 0; kbi < 8; ++kbi) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[0][t] = mfma16(a0[kbi], in[kbi][t], acc[0][t]);
            acc[1][t] = mfma16(a1[kbi], in[kbi][t], acc[1][t]);
        }
    }
}

// One hidden 256 -> 256 layer: in[kb][t] -> out[kb][t].
__device__ __forceinline__ void hidden_layer(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                             const bf16x8 (&in)[8][2], bf16x8 (&out)[8][2], int lane) {
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][2];
        tile_pair(W, bias, in, kbo, lane, acc);
        act_pack(acc, out[kbo]);
    }
}

// Weight tiles stream from L2 straight into the MFMA A operands (each feeds the
// wave's two element tiles); there is no LDS staging and no barrier, so the
// waves of a workgroup run independently.
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
        bias_acc(bias, kb, g, acc);
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            acc[0][t] = mfma16(a0, x[t], acc[0][t]);
            acc[1][t] = mfma16(a1, x[t], acc[1][t]);
        }
        act_pack(acc, A[kb]);
    }
    hidden_layer(P + kPkL2, bias + 256, A, B, lane);
    hidden_layer(P + kPkL3, bias + 512, B, A, lane);
    // layer 4 + head: y = w5 . silu(z4) + b5, reduced over the four lane groups
    float yp[2] = {0.f, 0.f};
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][2];
        tile_pair(P + kPkL4, bias + 768, A, kbo, lane, acc);
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const f32x4 wn = *reinterpret_cast<const f32x4*>(w5 + 32 * kbo + 16 * mm + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < 2; ++t) yp[t] += wn[i] * silu_f(acc[mm][t][i]);
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
