// Register-resident MLP forward for one wavefront (device code, header-only so
// the generic inference kernel and the fused scatter-surrogate kernel share it).
//
// A wave evaluates 16 kT elements as kT 16-column tiles t.  Activations never
// leave the VGPRs: the accumulators of output tiles (2kb, 2kb+1) of a layer,
// rounded to bf16, are exactly the B operand of K block kb of the next layer
// (k_perm), so each layer is 8 x 8 x 2kT MFMAs with weights streamed as 1 KiB
// lane-major tiles (L2 resident, 406 KiB per member).  Every weight tile a wave
// loads feeds kT MFMAs, so the L2 traffic per element falls as 1 / kT; the
// registers of the activations grow as kT (kT = 3 is the most that keeps two
// waves per SIMD, at 256 VGPRs in the fused scatter-surrogate kernel).  Every
// element's arithmetic is the same for any kT and prefetch distance, so the
// outputs are bit-identical.
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
// rows for every element tile, so one 16-byte load per mm serves all of them.
template <int kT>
__device__ __forceinline__ void bias_acc(const float* __restrict__ bias, int kb, int g, f32x4 (&acc)[2][kT]) {
#pragma unroll
    for (int mm = 0; mm < 2; ++mm) {
        const f32x4 b = *reinterpret_cast<const f32x4*>(bias + 32 * kb + 16 * mm + 4 * g);
#pragma unroll
        for (int t = 0; t < kT; ++t) acc[mm][t] = b;
    }
}

// SiLU on output tile pair kb (bias already in the accumulators) and repack as
// bf16 B operand slots.
template <int kT>
__device__ __forceinline__ void act_pack(const f32x4 (&acc)[2][kT], bf16x8 (&out)[kT]) {
#pragma unroll
    for (int t = 0; t < kT; ++t) {
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
#pragma unroll
            for (int i = 0; i < 4; ++i) out[t][mm * 4 + i] = (bf16)silu_f(acc[mm][t][i]);
        }
    }
}

// z = W in + bias for output tile pair kbo of a 256 -> 256 layer.  Its weight
// tiles for K blocks 0 .. kD-1 are requested before the first MFMA and block
// kbi + kD while block kbi is multiplied, so kD blocks' L2 latencies overlap
// (kD = 8: all 16 tiles up front).
template <int kT, int kD>
__device__ __forceinline__ void tile_pair(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                          const bf16x8 (&in)[8][kT], int kbo, int lane, f32x4 (&acc)[2][kT]) {
    bf16x8 a0[8], a1[8];
#pragma unroll
    for (int kbi = 0; kbi < kD; ++kbi) {
        a0[kbi] = ld_tile(W, (2 * kbo) * 8 + kbi, lane);
        a1[kbi] = ld_tile(W, (2 * kbo + 1) * 8 + kbi, lane);
    }
    bias_acc<kT>(bias, kbo, lane >> 4, acc);
#pragma unroll
    for (int kbi = 0; kbi < 8; ++kbi) {
        if (kbi + kD < 8) {
            a0[kbi + kD] = ld_tile(W, (2 * kbo) * 8 + kbi + kD, lane);
            a1[kbi + kD] = ld_tile(W, (2 * kbo + 1) * 8 + kbi + kD, lane);
        }
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            acc[0][t] = mfma16(a0[kbi], in[kbi][t], acc[0][t]);
            acc[1][t] = mfma16(a1[kbi], in[kbi][t], acc[1][t]);
        }
    }
}

// One hidden 256 -> 256 layer: in[kb][t] -> out[kb][t].
template <int kT, int kD>
__device__ __forceinline__ void hidden_layer(const uint8_t* __restrict__ W, const float* __restrict__ bias,
                                             const bf16x8 (&in)[8][kT], bf16x8 (&out)[8][kT], int lane) {
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][kT];
        tile_pair<kT, kD>(W, bias, in, kbo, lane, acc);
        act_pack<kT>(acc, out[kbo]);
    }
}

// Weight tiles stream from L2 straight into the MFMA A operands (each feeds the
// wave's kT element tiles); there is no LDS staging and no barrier, so the
// waves of a workgroup run independently.  kT = 2 with kD = 8 (32 elements, all
// tiles of a pair up front) is the generic inference kernel's form;
// kT = 3 with kD = 2 (48 elements) moves 1.5x fewer weight bytes per element.
template <int kT, int kD = 8>
__device__ __forceinline__ void mlp_wave_forward(const uint8_t* __restrict__ P, const bf16x8 (&x)[kT], float (&y)[kT]) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const float* bias = reinterpret_cast<const float*>(P + kPkBias);
    const float* w5 = reinterpret_cast<const float*>(P + kPkW5);
    bf16x8 A[8][kT], B[8][kT];
    // layer 1: one K block (features), output tile pairs kb = 0..7
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        const bf16x8 a0 = ld_tile(P + kPkL1, 2 * kb, lane), a1 = ld_tile(P + kPkL1, 2 * kb + 1, lane);
        f32x4 acc[2][kT];
        bias_acc<kT>(bias, kb, g, acc);
#pragma unroll
        for (int t = 0; t < kT; ++t) {
            acc[0][t] = mfma16(a0, x[t], acc[0][t]);
            acc[1][t] = mfma16(a1, x[t], acc[1][t]);
        }
        act_pack<kT>(acc, A[kb]);
    }
    hidden_layer<kT, kD>(P + kPkL2, bias + 256, A, B, lane);
    hidden_layer<kT, kD>(P + kPkL3, bias + 512, B, A, lane);
    // layer 4 + head: y = w5 . silu(z4) + b5, reduced over the four lane groups
    float yp[kT];
#pragma unroll
    for (int t = 0; t < kT; ++t) yp[t] = 0.f;
#pragma unroll
    for (int kbo = 0; kbo < 8; ++kbo) {
        f32x4 acc[2][kT];
        tile_pair<kT, kD>(P + kPkL4, bias + 768, A, kbo, lane, acc);
#pragma unroll
        for (int mm = 0; mm < 2; ++mm) {
            const f32x4 wn = *reinterpret_cast<const f32x4*>(w5 + 32 * kbo + 16 * mm + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int t = 0; t < kT; ++t) yp[t] += wn[i] * silu_f(acc[mm][t][i]);
        }
    }
    const float b5 = *reinterpret_cast<const float*>(P + kPkB5);
#pragma unroll
    for (int t = 0; t < kT; ++t) {
        float v = yp[t];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[t] = v + b5;
    }
}

}  // namespace sysmat
