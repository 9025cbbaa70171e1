// Beam masks and multiplexing index (MPXI) of a sparse system matrix.
//
// Consumers of the CSR output: the reference computes the same quantities from
// dense .sysmat files on the host (PEGen_RayTracing_CircularHole/
// extract_beam_prop_masks.py: a detector's beam = voxels above a fraction of its
// PPDF maximum; compute_mpxi_3d.py: per voxel, the number of detector beams that
// contain it).  Here they run on the device-resident CSR rows the PE tracer
// produces, one workgroup per row:
//
//   row max   = max of the row's stored values (the CSR keeps every element >=
//               rel * row max, so the maximum itself is always stored)
//   in beam   = (double)v > frac * (double)(row max)
//   mpxi[col] += 1 for every element in a beam   (global atomics, uint32)
//   beam_size[row] = number of elements in the row's beam
#include <hip/hip_runtime.h>

#include "../../include/sysmat.h"
#include "common.hpp"

namespace sysmat {

constexpr int kBeamThreads = 256;

__global__ void __launch_bounds__(kBeamThreads) csr_beams_kernel(const long long* __restrict__ row_ptr,
                                                                 const int* __restrict__ col,
                                                                 const float* __restrict__ val, int rows,
                                                                 double frac, unsigned* __restrict__ mpxi,
                                                                 int* __restrict__ beam_size) {
    __shared__ float s_max[kBeamThreads / 64];
    __shared__ int s_cnt[kBeamThreads / 64];
    const int r = blockIdx.x;
    if (r >= rows) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long b = row_ptr[r], e = row_ptr[r + 1];
    float m = 0.0f;
    for (long long j = b + tid; j < e; j += kBeamThreads) m = fmaxf(m, val[j]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (lane == 0) s_max[wave] = m;
    __syncthreads();
    for (int w = 0; w < kBeamThreads / 64; ++w) m = fmaxf(m, s_max[w]);
    const double thr = frac * (double)m;
    int cnt = 0;
    if (m > 0.0f) {
        for (long long j = b + tid; j < e; j += kBeamThreads) {
            if ((double)val[j] > thr) {
                atomicAdd(&mpxi[col[j]], 1u);
                ++cnt;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) s_cnt[wave] = cnt;
    __syncthreads();
    if (tid == 0 && beam_size) {
        int t = 0;
        for (int w = 0; w < kBeamThreads / 64; ++w) t += s_cnt[w];
        beam_size[r] = t;
    }
}

}  // namespace sysmat

using namespace sysmat;

extern "C" {

int sm_csr_beams_device(const long long* row_ptr_dev, const int* col_dev, const float* val_dev, int rows,
                        double frac, unsigned* mpxi_dev, int* beam_size_dev, void* stream) {
    clear_error();
    if (!row_ptr_dev || !col_dev || !val_dev || !mpxi_dev) return set_error(SM_E_ARG, "null argument");
    if (rows < 0) return set_error(SM_E_ARG, "negative row count");
    if (!(frac >= 0.0)) return set_error(SM_E_ARG, "beam fraction must be >= 0");
    if (rows == 0) return SM_OK;
    hipLaunchKernelGGL(csr_beams_kernel, dim3((unsigned)rows), dim3(kBeamThreads), 0, (hipStream_t)stream,
                       row_ptr_dev, col_dev, val_dev, rows, frac, mpxi_dev, beam_size_dev);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

}  // extern "C"
