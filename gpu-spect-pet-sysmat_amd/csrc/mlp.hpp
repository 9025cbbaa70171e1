// Scatter-surrogate MLP: 12 -> 256 -> 256 -> 256 -> 256 -> 1, SiLU, bfloat16
// operands on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32 accumulation.
//
// Network (training and inference kernels compute exactly this):
//   a0 = bf16(x padded to 32),  z_l = bf16(W_l) a_{l-1} + b_l,  h_l = silu(z_l),
//   a_l = bf16(h_l) (l = 1..4),  y = w5 . h4 + b5   (h4 in FP32).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sysmat {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 12;
constexpr int kMlpK1 = 32;
constexpr int kMlpH = 256;

// flat FP32 parameter layout (floats)
constexpr long long kOffW1 = 0, kOffB1 = 8192, kOffW2 = 8448, kOffB2 = 73984, kOffW3 = 74240, kOffB3 = 139776,
                    kOffW4 = 140032, kOffB4 = 205568, kOffW5 = 205824, kOffB5 = 206080;
constexpr long long kMlpParams = 206144;

// packed inference layout (bytes): MFMA A-operand tiles, lane-major, 1 KiB per
// (16 output rows x 32 k) tile, then FP32 biases and the head.
constexpr long long kPkL1 = 0;                // 16 tiles
constexpr long long kPkL2 = 16 * 1024;        // 16 x 8 tiles
constexpr long long kPkL3 = kPkL2 + 128 * 1024;
constexpr long long kPkL4 = kPkL3 + 128 * 1024;
constexpr long long kPkBias = kPkL4 + 128 * 1024;  // 4 x 256 floats
constexpr long long kPkW5 = kPkBias + 4 * 1024;    // 256 floats
constexpr long long kPkB5 = kPkW5 + 1024;          // 1 float (+pad)
constexpr long long kMlpPacked = kPkB5 + 256;

// SiLU and its derivative with the logistic through v_rcp_f32 (1 ulp) instead of an
// IEEE division (~10 instructions): the activations are rounded to bf16 next, so the
// ulp does not show, and the epilogues, not the MFMAs, set the inference kernel's
// VALU time.
__device__ __forceinline__ float sigmoid_f(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_f(float z) {
    const float s = sigmoid_f(z);
    return s * (1.0f + z * (1.0f - s));
}

// k label (lane group g, slot s) of a 32-wide K block -> neuron offset within
// the block, chosen so that two consecutive 16-row accumulator tiles of one
// layer are, lane for lane, the B operand of the next layer.
__host__ __device__ __forceinline__ int k_perm(int g, int s) { return s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4); }

struct GemmArgs {
    const bf16* A;
    long long lda;  // A(m,k): AK ? A[m*lda + k] : A[k*lda + m]
    const bf16* B;
    long long ldb;  // B(k,n): BK ? B[n*ldb + k] : B[k*ldb + n]
    int M, N, K, ksplit;
    long long ldc;        // output (m,n) at m*ldc + n
    float* C;             // EPI_PART: C + split*M*ldc
    const float* bias;    // EPI_FWD: bias[n]
    float* Z;             // EPI_FWD: z out
    bf16* H;              // EPI_FWD: bf16(silu(z)) out
    const float* Zin;     // EPI_BWD: z of the layer below
    bf16* D;              // EPI_BWD: bf16(acc * silu'(z)) out
};

enum { EPI_FWD = 0, EPI_BWD = 1, EPI_PART = 2 };

hipError_t launch_gemm(int ak, int bk, int epi, const GemmArgs& g, int splits, hipStream_t s);
hipError_t launch_mlp_infer(const uint8_t* packed, int n_members, const float* x, long long n, float* y,
                            hipStream_t s);

}  // namespace sysmat
