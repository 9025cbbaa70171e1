// Sparse output stage: dense det-major rows -> CSR, keeping the elements that
// reach a fraction `rel` of their row's maximum.
//
//   keep(v) = v > 0  and  (double)v >= rel * (double)rowmax,   rowmax = max(0, max_j v_j)
//
// Three HBM-bound kernels, all one workgroup (4 waves) per row except the scan:
//   csr_count_kernel   row max, then the kept count (second read hits the MALL)
//   csr_scan_kernel    in-place inclusive scan of the counts -> row_ptr (int64)
//   csr_fill_kernel    ordered compaction: per 64-lane ballot + mbcnt prefix,
//                      wave totals through LDS, column indices ascending in a row
// Rows come from the PE / scatter kernels in device memory; a whole CSR result
// can be kept host-side in an SmCsr handle and written to a .csr file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sysmat.h"
#include "common.hpp"

namespace sysmat {

constexpr int kCsrThreads = 256;
constexpr int kCsrWaves = kCsrThreads / 64;

__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ long long wave_sum(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ bool keep(float v, double thr) { return v > 0.0f && (double)v >= thr; }

// VEC: the row base is 16-byte aligned and cols % 4 == 0, so float4 loads stay
// inside the row.
template <bool VEC>
__global__ void __launch_bounds__(kCsrThreads) csr_count_kernel(const float* __restrict__ dense, long long cols,
                                                                 long long row_stride, double rel,
                                                                 float* __restrict__ row_max,
                                                                 long long* __restrict__ row_ptr) {
    __shared__ float s_max[kCsrWaves];
    __shared__ long long s_cnt[kCsrWaves];
    const int r = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const float* row = dense + (long long)r * row_stride;
    float m = 0.0f;
    if constexpr (VEC) {
        const float4* r4 = reinterpret_cast<const float4*>(row);
        for (long long i = t; i < cols / 4; i += kCsrThreads) {
            float4 v = r4[i];
            m = fmaxf(m, fmaxf(fmaxf(v.x, v.y), fmaxf(v.z, v.w)));
        }
    } else {
        for (long long i = t; i < cols; i += kCsrThreads) m = fmaxf(m, row[i]);
    }
    m = wave_max(m);
    if (lane == 0) s_max[w] = m;
    __syncthreads();
    m = s_max[0];
    for (int k = 1; k < kCsrWaves; ++k) m = fmaxf(m, s_max[k]);
    const double thr = rel * (double)m;
    long long c = 0;
    if constexpr (VEC) {
        const float4* r4 = reinterpret_cast<const float4*>(row);
        for (long long i = t; i < cols / 4; i += kCsrThreads) {
            float4 v = r4[i];
            c += (int)keep(v.x, thr) + (int)keep(v.y, thr) + (int)keep(v.z, thr) + (int)keep(v.w, thr);
        }
    } else {
        for (long long i = t; i < cols; i += kCsrThreads) c += (int)keep(row[i], thr);
    }
    c = wave_sum(c);
    if (lane == 0) s_cnt[w] = c;
    __syncthreads();
    if (t == 0) {
        long long s = 0;
        for (int k = 0; k < kCsrWaves; ++k) s += s_cnt[k];
        row_ptr[r + 1] = s;
        row_max[r] = m;
    }
}

// One workgroup of 1024 threads: row_ptr[1..rows] holds counts on entry and
// their inclusive prefix sums on exit; row_ptr[0] = 0.
__global__ void __launch_bounds__(1024) csr_scan_kernel(long long* __restrict__ row_ptr, int rows) {
    __shared__ long long s_tot[16];
    __shared__ long long s_carry;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t == 0) {
        s_carry = 0;
        row_ptr[0] = 0;
    }
    __syncthreads();
    for (int base = 0; base < rows; base += 1024) {
        const int i = base + t;
        long long v = i < rows ? row_ptr[i + 1] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            long long u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane == 63) s_tot[w] = v;
        __syncthreads();
        long long off = s_carry;
        for (int k = 0; k < w; ++k) off += s_tot[k];
        if (i < rows) row_ptr[i + 1] = v + off;
        __syncthreads();
        if (t == 1023) s_carry = v + off;
        __syncthreads();
    }
}

template <bool VEC>
__global__ void __launch_bounds__(kCsrThreads) csr_fill_kernel(const float* __restrict__ dense, long long cols,
                                                                long long row_stride, double rel,
                                                                const float* __restrict__ row_max,
                                                                const long long* __restrict__ row_ptr,
                                                                int* __restrict__ col_idx, float* __restrict__ val) {
    constexpr int PER = VEC ? 4 : 1;
    constexpr int CHUNK = kCsrThreads * PER;
    __shared__ unsigned s_tot[2][kCsrWaves];
    const int r = blockIdx.x;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const float* row = dense + (long long)r * row_stride;
    const double thr = rel * (double)row_max[r];
    long long out = row_ptr[r];
    int buf = 0;
    for (long long c0 = 0; c0 < cols; c0 += CHUNK, buf ^= 1) {
        const long long j0 = c0 + (long long)t * PER;
        float v[PER];
        if constexpr (VEC) {
            const float4 q =
                j0 < cols ? *reinterpret_cast<const float4*>(row + j0) : make_float4(0.f, 0.f, 0.f, 0.f);
            v[0] = q.x;
            v[1] = q.y;
            v[2] = q.z;
            v[3] = q.w;
        } else {
            v[0] = j0 < cols ? row[j0] : 0.0f;
        }
        bool k[PER];
        unsigned below = 0, wave_tot = 0;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            k[e] = keep(v[e], thr);
            const unsigned long long b = __ballot(k[e]);
            below += lanes_below(b);
            wave_tot += (unsigned)__popcll(b);
        }
        if (lane == 0) s_tot[buf][w] = wave_tot;
        __syncthreads();
        unsigned woff = 0, ctot = 0;
#pragma unroll
        for (int q = 0; q < kCsrWaves; ++q) {
            const unsigned x = s_tot[buf][q];
            woff += q < w ? x : 0u;
            ctot += x;
        }
        long long pos = out + woff + below;
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            if (k[e]) {
                col_idx[pos] = (int)(j0 + e);
                val[pos] = v[e];
                ++pos;
            }
        }
        out += ctot;
    }
}

static bool vec_ok(const float* dense, long long cols, long long row_stride) {
    return ((uintptr_t)dense % 16 == 0) && (cols % 4 == 0) && (row_stride % 4 == 0);
}

static int check_dims(const float* dense, int rows, long long cols, long long& row_stride, double rel) {
    if (!dense || rows <= 0 || cols <= 0) return set_error(SM_E_ARG, "csr: empty or null dense input");
    if (cols > 0x7fffffffLL) return set_error(SM_E_ARG, "csr: more than 2^31-1 columns");
    if (row_stride == 0) row_stride = cols;
    if (row_stride < cols) return set_error(SM_E_ARG, "csr: row_stride < cols");
    if (!(rel >= 0.0) || rel > 1.0) return set_error(SM_E_ARG, "csr: threshold must be in [0, 1]");
    return SM_OK;
}

static int csr_count(const float* dense, int rows, long long cols, long long row_stride, double rel, float* row_max,
                     long long* row_ptr, long long* nnz, hipStream_t s) {
    if (vec_ok(dense, cols, row_stride))
        hipLaunchKernelGGL(csr_count_kernel<true>, dim3(rows), dim3(kCsrThreads), 0, s, dense, cols, row_stride, rel,
                           row_max, row_ptr);
    else
        hipLaunchKernelGGL(csr_count_kernel<false>, dim3(rows), dim3(kCsrThreads), 0, s, dense, cols, row_stride,
                           rel, row_max, row_ptr);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(csr_scan_kernel, dim3(1), dim3(1024), 0, s, row_ptr, rows);
    HIP_TRY(hipGetLastError());
    long long n = 0;
    HIP_TRY(hipMemcpyAsync(&n, row_ptr + rows, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *nnz = n;
    return SM_OK;
}

static int csr_fill(const float* dense, int rows, long long cols, long long row_stride, double rel,
                    const float* row_max, const long long* row_ptr, int* col_idx, float* val, hipStream_t s) {
    if (vec_ok(dense, cols, row_stride))
        hipLaunchKernelGGL(csr_fill_kernel<true>, dim3(rows), dim3(kCsrThreads), 0, s, dense, cols, row_stride, rel,
                           row_max, row_ptr, col_idx, val);
    else
        hipLaunchKernelGGL(csr_fill_kernel<false>, dim3(rows), dim3(kCsrThreads), 0, s, dense, cols, row_stride, rel,
                           row_max, row_ptr, col_idx, val);
    HIP_TRY(hipGetLastError());
    return SM_OK;
}

}  // namespace sysmat

// Host-side CSR result: rows x cols, int64 row pointers, int32 columns, float32 values.
struct SmCsr {
    long long rows = 0, cols = 0;
    double rel = 0.0;
    std::vector<long long> row_ptr{0};
    std::vector<int> col;
    std::vector<float> val;
};

using namespace sysmat;

namespace {
// .csr file: 64-byte header, then row_ptr[rows+1] int64, col[nnz] int32, val[nnz] float32
// (little endian).  Header: "SMCSR\0\0\1" magic, int64 rows, int64 cols, int64 nnz,
// float64 rel threshold, 24 reserved bytes (zero).
const char kMagic[8] = {'S', 'M', 'C', 'S', 'R', 0, 0, 1};
}  // namespace

extern "C" {

int sm_csr_count_device(const float* dense_dev, int rows, long long cols, long long row_stride, double rel,
                        float* row_max_dev, long long* row_ptr_dev, long long* nnz, void* stream) {
    clear_error();
    int rc = check_dims(dense_dev, rows, cols, row_stride, rel);
    if (rc < 0) return rc;
    if (!row_max_dev || !row_ptr_dev || !nnz) return set_error(SM_E_ARG, "csr: null output");
    return csr_count(dense_dev, rows, cols, row_stride, rel, row_max_dev, row_ptr_dev, nnz, (hipStream_t)stream);
}

int sm_csr_fill_device(const float* dense_dev, int rows, long long cols, long long row_stride, double rel,
                       const float* row_max_dev, const long long* row_ptr_dev, int* col_dev, float* val_dev,
                       void* stream) {
    clear_error();
    int rc = check_dims(dense_dev, rows, cols, row_stride, rel);
    if (rc < 0) return rc;
    if (!row_max_dev || !row_ptr_dev || !col_dev || !val_dev) return set_error(SM_E_ARG, "csr: null output");
    return csr_fill(dense_dev, rows, cols, row_stride, rel, row_max_dev, row_ptr_dev, col_dev, val_dev,
                    (hipStream_t)stream);
}

int sm_pe_compute_csr(const float* col, const float* det, const float* img, const SmPeOptions* o, double rel,
                      SmCsr** out) {
    clear_error();
    if (!col || !det || !img || !o || !out) return set_error(SM_E_ARG, "null argument");
    *out = nullptr;
    if (!(rel >= 0.0) || rel > 1.0) return set_error(SM_E_ARG, "csr: threshold must be in [0, 1]");
    SmPePlan* plan = nullptr;
    int rc = sm_pe_plan_create(col, det, img, o, &plan);
    if (rc < 0) return rc;
    struct PlanGuard {
        SmPePlan* p;
        ~PlanGuard() { sm_pe_plan_destroy(p); }
    } guard{plan};
    const int X = (int)std::floor(img[0] + 0.001f), Y = (int)std::floor(img[1] + 0.001f);
    const int Z = (int)std::floor(img[2] + 0.001f);
    const int N = (int)std::floor(det[0] + 0.001f);
    const long long zc = o->z_count > 0 ? o->z_count : Z - o->z_begin;
    const long long V = (long long)X * Y * zc;
    const int d0 = o->det_begin, dn = o->det_count > 0 ? o->det_count : N - o->det_begin;
    if (d0 < 0 || dn <= 0 || d0 + dn > N) return set_error(SM_E_ARG, "detector range outside the detector");
    // dense block of at most 1 GiB (at least one row)
    const int block = (int)std::max(1LL, std::min<long long>(dn, (1LL << 28) / std::max(1LL, V)));
    DeviceBuf<float> dense, rmax, val;
    DeviceBuf<long long> rptr;
    DeviceBuf<int> cidx;
    if ((rc = dense.alloc((size_t)block * V)) < 0) return rc;
    if ((rc = rmax.alloc(block)) < 0) return rc;
    if ((rc = rptr.alloc(block + 1)) < 0) return rc;
    SmCsr* C = new SmCsr();
    C->rows = dn;
    C->cols = V;
    C->rel = rel;
    C->row_ptr.assign(1, 0);
    C->row_ptr.reserve(dn + 1);
    std::vector<long long> rp(block + 1);
    auto fail = [&](int code) {
        delete C;
        return code;
    };
    for (int b = 0; b < dn; b += block) {
        const int cnt = std::min(block, dn - b);
        if ((rc = sm_pe_plan_launch(plan, d0 + b, cnt, dense.ptr, V, nullptr)) < 0) return fail(rc);
        if ((rc = sm_pe_plan_sync(plan)) < 0) return fail(rc);
        long long nnz = 0;
        if ((rc = csr_count(dense.ptr, cnt, V, V, rel, rmax.ptr, rptr.ptr, &nnz, nullptr)) < 0) return fail(rc);
        if ((size_t)nnz > cidx.n) {
            if ((rc = cidx.alloc((size_t)nnz)) < 0) return fail(rc);
            if ((rc = val.alloc((size_t)nnz)) < 0) return fail(rc);
        }
        if (nnz > 0 && (rc = csr_fill(dense.ptr, cnt, V, V, rel, rmax.ptr, rptr.ptr, cidx.ptr, val.ptr, nullptr)) < 0)
            return fail(rc);
        const long long base = C->row_ptr.back();
        if (hipMemcpy(rp.data(), rptr.ptr, sizeof(long long) * (cnt + 1), hipMemcpyDeviceToHost) != hipSuccess)
            return fail(set_error(SM_E_DEVICE, "csr: row_ptr copy failed"));
        for (int i = 1; i <= cnt; ++i) C->row_ptr.push_back(base + rp[i]);
        C->col.resize(base + nnz);
        C->val.resize(base + nnz);
        if (nnz > 0) {
            if (hipMemcpy(C->col.data() + base, cidx.ptr, sizeof(int) * nnz, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(C->val.data() + base, val.ptr, sizeof(float) * nnz, hipMemcpyDeviceToHost) != hipSuccess)
                return fail(set_error(SM_E_DEVICE, "csr: value copy failed"));
        }
    }
    *out = C;
    return SM_OK;
}

int sm_csr_info(const SmCsr* c, long long* rows, long long* cols, long long* nnz) {
    clear_error();
    if (!c) return set_error(SM_E_ARG, "null csr");
    if (rows) *rows = c->rows;
    if (cols) *cols = c->cols;
    if (nnz) *nnz = (long long)c->col.size();
    return SM_OK;
}

int sm_csr_copy(const SmCsr* c, long long* row_ptr, int* col_idx, float* val) {
    clear_error();
    if (!c) return set_error(SM_E_ARG, "null csr");
    if (row_ptr) std::memcpy(row_ptr, c->row_ptr.data(), sizeof(long long) * c->row_ptr.size());
    if (col_idx && !c->col.empty()) std::memcpy(col_idx, c->col.data(), sizeof(int) * c->col.size());
    if (val && !c->val.empty()) std::memcpy(val, c->val.data(), sizeof(float) * c->val.size());
    return SM_OK;
}

int sm_csr_write(const char* path, const SmCsr* const* parts, int n_parts) {
    clear_error();
    if (!path || !parts || n_parts <= 0) return set_error(SM_E_ARG, "csr write: no parts");
    long long rows = 0, nnz = 0;
    const long long cols = parts[0] ? parts[0]->cols : 0;
    const double rel = parts[0] ? parts[0]->rel : 0.0;
    for (int i = 0; i < n_parts; ++i) {
        if (!parts[i]) return set_error(SM_E_ARG, "csr write: null part");
        if (parts[i]->cols != cols) return set_error(SM_E_ARG, "csr write: parts differ in column count");
        rows += parts[i]->rows;
        nnz += (long long)parts[i]->col.size();
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(SM_E_IO, std::string("cannot create ") + path);
    unsigned char head[64] = {0};
    std::memcpy(head, kMagic, 8);
    std::memcpy(head + 8, &rows, 8);
    std::memcpy(head + 16, &cols, 8);
    std::memcpy(head + 24, &nnz, 8);
    std::memcpy(head + 32, &rel, 8);
    bool ok = std::fwrite(head, 1, 64, f) == 64;
    long long base = 0;
    const long long zero = 0;
    ok = ok && std::fwrite(&zero, 8, 1, f) == 1;
    for (int i = 0; i < n_parts && ok; ++i) {
        const auto& rp = parts[i]->row_ptr;
        std::vector<long long> shifted(rp.size() - 1);
        for (size_t j = 1; j < rp.size(); ++j) shifted[j - 1] = rp[j] + base;
        ok = std::fwrite(shifted.data(), 8, shifted.size(), f) == shifted.size();
        base += (long long)parts[i]->col.size();
    }
    for (int i = 0; i < n_parts && ok; ++i)
        ok = std::fwrite(parts[i]->col.data(), 4, parts[i]->col.size(), f) == parts[i]->col.size();
    for (int i = 0; i < n_parts && ok; ++i)
        ok = std::fwrite(parts[i]->val.data(), 4, parts[i]->val.size(), f) == parts[i]->val.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(SM_E_IO, std::string("short write to ") + path);
    return SM_OK;
}

int sm_csr_destroy(SmCsr* c) {
    delete c;
    return SM_OK;
}

}  // extern "C"
