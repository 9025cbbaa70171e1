// C ABI and host orchestration of the PE tracer.
//
// Replaces PESysMatGen (PEGen_RayTracing_CircularHole/PESysMatGen.cu:766-833):
// the reference re-allocates and re-uploads fixed 80000-float parameter blocks
// for every Z-slab; here the geometry is repacked once into HBM-resident
// tables (tables.hpp) and slabs / detector blocks are launched against them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sysmat.h"
#include "common.hpp"
#include "pe_kernel.hpp"

namespace sysmat {

struct PePlanImpl {
    int device = 0;
    Tables T;
    DeviceBuf<Crystal> cry;
    DeviceBuf<Hole> hol;
    DeviceBuf<Plate> pla;
    DeviceBuf<Slab> slb;
    DeviceBuf<int> status;
    DeviceBuf<long long> cry_ptr, hol_ptr;      // per-detector candidate lists (launch_detector_cull)
    DeviceBuf<unsigned short> cry_list, hol_list;
    PeParams base{};
    hipStream_t last_stream = nullptr;
    int tile_fixed = 0;
    struct Launch {
        int det_begin, det_count;
        float* dst;
        long long row_stride;
        hipStream_t stream;
    };
    std::vector<Launch> pending;  // launches since the last sync (re-run with smaller tiles on overflow)
    DeviceArena* arena = nullptr; // set: every device buffer is a piece of this region (no allocation)
    PeLists lists() const { return PeLists{cry_ptr.ptr, cry_list.ptr, hol_ptr.ptr, hol_list.ptr}; }
};

// Device buffer of the plan: hipMalloc'd, or carved from the plan's arena.
template <class T>
static int plan_alloc(PePlanImpl& P, DeviceBuf<T>& b, size_t n) {
    if (!P.arena) return b.alloc(n);
    T* p = P.arena->take<T>(n);
    if (!p)
        return set_error(SM_E_ARG, "geometry tables and candidate lists exceed the " +
                                       std::to_string(P.arena->cap >> 20) + " MiB allocation-free scratch");
    b.attach(p, n);
    return SM_OK;
}

template <class T>
static int plan_upload(PePlanImpl& P, DeviceBuf<T>& b, const std::vector<T>& v) {
    int rc = plan_alloc(P, b, v.size());
    if (rc < 0) return rc;
    if (!v.empty()) HIP_TRY(hipMemcpy(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return SM_OK;
}

__global__ void zero_ints_kernel(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Clears the status word with a kernel (no memset node if the caller captures the stream).
static hipError_t clear_status(int* status, hipStream_t stream) {
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, stream, status, 1);
    return hipGetLastError();
}

static int fill_params(const float* img, const SmPeOptions& o, const Tables& T, PeParams& p) {
    const int X = (int)std::floor(img[0] + 0.001f), Y = (int)std::floor(img[1] + 0.001f);
    const int Z = (int)std::floor(img[2] + 0.001f);
    if (X <= 0 || Y <= 0 || Z <= 0) return set_error(SM_E_ARG, "image voxel counts must be positive");
    if (o.voxel_sub < 1 || o.sub_x < 1 || o.sub_y < 1 || o.sub_z < 1)
        return set_error(SM_E_ARG, "sampling counts must be >= 1");
    const int zb = o.z_begin, zc = o.z_count > 0 ? o.z_count : Z - o.z_begin;
    if (zb < 0 || zc <= 0 || zb + zc > Z) return set_error(SM_E_ARG, "z range outside the image");
    p = PeParams{};
    p.X = X;
    p.Y = Y;
    p.Zslab = zc;
    p.Zfull = Z;
    p.z0 = zb;
    p.wx = img[3];
    p.wy = img[4];
    p.wz = img[5];
    p.sx = img[8];
    p.sy = img[9];
    p.sz = img[10];
    const float ang = (float)o.rotation * img[7];
    p.rc = std::cos(ang);
    p.rs = std::sin(ang);
    const double ang64 = (double)o.rotation * (double)img[7];
    p.rc64 = std::cos(ang64);
    p.rs64 = std::sin(ang64);
    p.fov2c = img[11];
    p.nv = o.voxel_sub;
    p.cx = o.sub_x;
    p.cy = o.sub_y;
    p.cz = o.sub_z;
    long long S = (long long)p.nv * p.nv * p.nv * p.cx * p.cy * p.cz;
    if (S > (1 << 24)) return set_error(SM_E_ARG, "too many sub-rays per element");
    p.S = (int)S;
    p.rnv64 = 1.0 / (double)p.nv;
    p.hX64 = (double)p.X / 2.0;
    p.hY64 = (double)p.Y / 2.0;
    p.hZ64 = (double)p.Zfull / 2.0;
    p.wx64 = (double)p.wx;
    p.wy64 = (double)p.wy;
    p.wz64 = (double)p.wz;
    p.sx64 = (double)p.sx;
    p.sy64 = (double)p.sy;
    p.sz64 = (double)p.sz;
    p.rcx64 = 1.0 / (double)p.cx;
    p.rcy64 = 1.0 / (double)p.cy;
    p.rcz64 = 1.0 / (double)p.cz;
    p.n_det = T.n_det;
    p.n_layers = T.n_layers;
    p.n_holes = T.n_holes;
    p.n_slabs = (int)T.slabs.size();
    p.use_bundle = (o.bundle_filter && p.n_slabs > 0) ? 1 : 0;
    p.inv4pi = (float)(1.0 / (4.0 * M_PI));
    return SM_OK;
}

// Voxel tiles of tx x ty x tz, grouped g x g x g per workgroup when the whole
// plan (every detector of the table x every tile) is large enough.
static void set_tiles(PeParams& p, int tx, int ty, int tz, int g = kGroupTiles) {
    // the tracer packs a voxel's offset in its tile into 10 bits per axis
    p.tx = std::max(1, std::min({tx, p.X, 1024}));
    p.ty = std::max(1, std::min({ty, p.Y, 1024}));
    p.tz = std::max(1, std::min({tz, p.Zslab, 1024}));
    p.ntx = (p.X + p.tx - 1) / p.tx;
    p.nty = (p.Y + p.ty - 1) / p.ty;
    p.ntz = (p.Zslab + p.tz - 1) / p.tz;
    if ((long long)p.n_det * p.ntx * p.nty * p.ntz < kGroupMinTiles) g = 1;
    p.gx = std::max(1, std::min(g, p.ntx));
    p.gy = std::max(1, std::min(g, p.nty));
    p.gz = std::max(1, std::min(g, p.ntz));
    p.ngx = (p.ntx + p.gx - 1) / p.gx;
    p.ngy = (p.nty + p.gy - 1) / p.gy;
    p.ngz = (p.ntz + p.gz - 1) / p.gz;
}

static int upload(PePlanImpl& P, const float* col, const float* det, const float* img) {
    std::string err = build_tables(col, det, img[11], P.T);
    if (!err.empty()) return set_error(SM_E_ARG, err);
    int rc;
    if ((rc = plan_upload(P, P.cry, P.T.crystals)) < 0) return rc;
    if ((rc = plan_upload(P, P.hol, P.T.holes)) < 0) return rc;
    if ((rc = plan_upload(P, P.pla, P.T.plates)) < 0) return rc;
    if ((rc = plan_upload(P, P.slb, P.T.slabs)) < 0) return rc;
    if ((rc = plan_alloc(P, P.status, 1)) < 0) return rc;
    return SM_OK;
}

// Per-detector candidate lists for the plan's image slab and rotation: count
// pass, host prefix sums, fill pass (synchronous, once per plan).
static int build_lists(PePlanImpl& P) {
    const PeParams& p = P.base;
    const int n = p.n_det;
    DeviceBuf<int> counts;
    int rc;
    if ((rc = plan_alloc(P, counts, 2 * (size_t)n)) < 0) return rc;
    if ((rc = plan_alloc(P, P.cry_ptr, (size_t)n + 1)) < 0) return rc;
    if ((rc = plan_alloc(P, P.hol_ptr, (size_t)n + 1)) < 0) return rc;
    HIP_TRY(launch_detector_cull(p, P.cry.ptr, P.hol.ptr, P.slb.ptr, 0, counts.ptr, P.lists(), nullptr));
    std::vector<int> c(2 * (size_t)n);
    HIP_TRY(hipMemcpy(c.data(), counts.ptr, c.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<long long> cp((size_t)n + 1, 0), hp((size_t)n + 1, 0);
    for (int d = 0; d < n; ++d) {
        cp[d + 1] = cp[d] + c[2 * (size_t)d];
        hp[d + 1] = hp[d] + c[2 * (size_t)d + 1];
    }
    HIP_TRY(hipMemcpy(P.cry_ptr.ptr, cp.data(), cp.size() * sizeof(long long), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(P.hol_ptr.ptr, hp.data(), hp.size() * sizeof(long long), hipMemcpyHostToDevice));
    if ((rc = plan_alloc(P, P.cry_list, (size_t)cp[n])) < 0) return rc;
    if ((rc = plan_alloc(P, P.hol_list, (size_t)hp[n])) < 0) return rc;
    HIP_TRY(launch_detector_cull(p, P.cry.ptr, P.hol.ptr, P.slb.ptr, 1, counts.ptr, P.lists(), nullptr));
    HIP_TRY(hipDeviceSynchronize());
    return SM_OK;
}

// Launch with the plan's tile; on cull-list overflow halve the tile (at a
// single voxel: drop the tile grouping) and relaunch (synchronous).  Returns
// SM_OK or an error.
static int run_sync(PePlanImpl& P, PeParams p, float* dst, hipStream_t stream) {
    int t[3] = {p.tx, p.ty, p.tz};
    int g = kGroupTiles;
    for (;;) {
        set_tiles(p, t[0], t[1], t[2], g);
        HIP_TRY(clear_status(P.status.ptr, stream));
        HIP_TRY(launch_pe(p, P.cry.ptr, P.hol.ptr, P.pla.ptr, P.slb.ptr, P.lists(), dst, P.status.ptr, stream));
        int st = 0;
        HIP_TRY(hipMemcpyAsync(&st, P.status.ptr, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (!(st & 1)) return SM_OK;
        if (t[0] == 1 && t[1] == 1 && t[2] == 1) {
            if (g == 1)
                return set_error(SM_E_OVERFLOW, "cull list exceeds 4096 crystals / 1024 holes for a single voxel");
            g = 1;
            continue;
        }
        for (int& v : t) v = std::max(1, v / 2);
    }
}

static void auto_tiles(const SmPeOptions& o, int& tx, int& ty, int& tz) {
    tx = o.tile_x > 0 ? o.tile_x : 4;
    ty = o.tile_y > 0 ? o.tile_y : 4;
    tz = o.tile_z > 0 ? o.tile_z : 4;
}

// Plan on the current device: tables, candidate lists and the status word in
// hipMalloc'd buffers, or (arena != nullptr) carved from that region.
static int plan_create(const float* col, const float* det, const float* img, const SmPeOptions& o,
                       DeviceArena* arena, PePlanImpl** out) {
    auto* P = new PePlanImpl();
    P->device = o.device;
    P->arena = arena;
    int rc = upload(*P, col, det, img);
    if (rc >= 0) rc = fill_params(img, o, P->T, P->base);
    if (rc >= 0) rc = build_lists(*P);
    if (rc >= 0 && (clear_status(P->status.ptr, nullptr) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess))
        rc = set_error(SM_E_DEVICE, "status reset failed");
    if (rc < 0) {
        delete P;
        return rc;
    }
    int tx, ty, tz;
    auto_tiles(o, tx, ty, tz);
    set_tiles(P->base, tx, ty, tz);
    P->base.det_begin = o.det_begin;
    P->base.det_count = o.det_count > 0 ? o.det_count : P->T.n_det - o.det_begin;
    *out = P;
    return SM_OK;
}

// ---- allocation-free variant of sm_pe_compute_device -----------------------
// Every device buffer of its plan (geometry tables, per-detector candidate
// lists, status word) is a piece of this file-scope array; each device holds
// its own copy, whose address is looked up once per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kScratchDevices];

// Device address of `scratch` on `device` (which must be current).
static int scratch_region(int device, unsigned char** out) {
    if (device < 0 || device >= kScratchDevices) return set_error(SM_E_DEVICE, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SM_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return SM_OK;
}

}  // namespace sysmat

using namespace sysmat;

extern "C" {

int sm_pe_default_options(SmPeOptions* o) {
    if (!o) return set_error(SM_E_ARG, "null options");
    std::memset(o, 0, sizeof(*o));
    o->voxel_sub = 1;
    o->sub_x = 8;
    o->sub_y = 16;
    o->sub_z = 8;
    o->det_begin = 0;
    o->det_count = 0;
    o->z_begin = 0;
    o->z_count = 0;
    o->rotation = 0;
    o->bundle_filter = 1;
    o->device = 0;
    return SM_OK;
}

int sm_pe_plan_create(const float* col, const float* det, const float* img, const SmPeOptions* o, SmPePlan** plan) {
    clear_error();
    if (!col || !det || !img || !o || !plan) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(o->device);
    if (rc < 0) return rc;
    PePlanImpl* P = nullptr;
    if ((rc = plan_create(col, det, img, *o, nullptr, &P)) < 0) return rc;
    *plan = reinterpret_cast<SmPePlan*>(P);
    return SM_OK;
}

int sm_pe_plan_launch(SmPePlan* plan, int det_begin, int det_count, float* dst_dev, long long row_stride,
                      void* stream) {
    clear_error();
    auto* P = reinterpret_cast<PePlanImpl*>(plan);
    if (!P || !dst_dev) return set_error(SM_E_ARG, "null argument");
    if (det_begin < 0 || det_count <= 0 || det_begin + det_count > P->T.n_det)
        return set_error(SM_E_ARG, "detector range outside the detector table");
    int rc = use_device(P->device);
    if (rc < 0) return rc;
    PeParams p = P->base;
    p.det_begin = det_begin;
    p.det_count = det_count;
    p.row_stride = row_stride > 0 ? row_stride : (long long)p.X * p.Y * p.Zslab;
    hipStream_t s = (hipStream_t)stream;
    P->last_stream = s;
    HIP_TRY(launch_pe(p, P->cry.ptr, P->hol.ptr, P->pla.ptr, P->slb.ptr, P->lists(), dst_dev, P->status.ptr, s));
    P->pending.push_back({det_begin, det_count, dst_dev, p.row_stride, s});
    return SM_OK;
}

int sm_pe_plan_sync(SmPePlan* plan) {
    clear_error();
    auto* P = reinterpret_cast<PePlanImpl*>(plan);
    if (!P) return set_error(SM_E_ARG, "null plan");
    int rc = use_device(P->device);
    if (rc < 0) return rc;
    int st = 0;
    for (const auto& l : P->pending) HIP_TRY(hipStreamSynchronize(l.stream));
    HIP_TRY(hipStreamSynchronize(P->last_stream));
    HIP_TRY(hipMemcpy(&st, P->status.ptr, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<PePlanImpl::Launch> todo;
    todo.swap(P->pending);
    if (st & 1) {
        // some tile's cull lists overflowed: recompute those launches with smaller tiles
        int tx = std::max(1, P->base.tx / 2), ty = std::max(1, P->base.ty / 2), tz = std::max(1, P->base.tz / 2);
        for (const auto& l : todo) {
            PeParams p = P->base;
            p.det_begin = l.det_begin;
            p.det_count = l.det_count;
            p.row_stride = l.row_stride;
            set_tiles(p, tx, ty, tz);
            if ((rc = run_sync(*P, p, l.dst, l.stream)) < 0) return rc;
        }
    }
    HIP_TRY(hipMemset(P->status.ptr, 0, sizeof(int)));
    return SM_OK;
}

int sm_pe_plan_destroy(SmPePlan* plan) {
    auto* P = reinterpret_cast<PePlanImpl*>(plan);
    if (P) {
        use_device(P->device);
        delete P;
    }
    return SM_OK;
}

long long sm_pe_plan_subrays(const SmPePlan* plan) {
    auto* P = reinterpret_cast<const PePlanImpl*>(plan);
    return P ? (long long)P->base.S : 0;
}

int sm_pe_compute_device(const float* col, const float* det, const float* img, const SmPeOptions* o,
                         float* dst_dev, long long row_stride, void* stream) {
    SmPePlan* plan = nullptr;
    int rc = sm_pe_plan_create(col, det, img, o, &plan);
    if (rc < 0) return rc;
    auto* P = reinterpret_cast<PePlanImpl*>(plan);
    PeParams p = P->base;
    p.row_stride = row_stride > 0 ? row_stride : (long long)p.X * p.Y * p.Zslab;
    if (p.det_begin < 0 || p.det_count <= 0 || p.det_begin + p.det_count > P->T.n_det) {
        sm_pe_plan_destroy(plan);
        return set_error(SM_E_ARG, "detector range outside the detector table");
    }
    rc = run_sync(*P, p, dst_dev, (hipStream_t)stream);
    sm_pe_plan_destroy(plan);
    return rc;
}

long long sm_pe_noalloc_scratch_bytes(void) { return (long long)kNoAllocScratchBytes; }

int sm_pe_compute_device_noalloc(const float* col, const float* det, const float* img, const SmPeOptions* o,
                                 float* dst_dev, long long row_stride, void* stream) {
    clear_error();
    if (!col || !det || !img || !o || !dst_dev) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(o->device);
    if (rc < 0) return rc;
    unsigned char* base = nullptr;
    if ((rc = scratch_region(o->device, &base)) < 0) return rc;
    DeviceArena arena;
    arena.base = base;
    arena.cap = kNoAllocScratchBytes;
    PePlanImpl* P = nullptr;
    if ((rc = plan_create(col, det, img, *o, &arena, &P)) < 0) return rc;
    PeParams p = P->base;
    p.row_stride = row_stride > 0 ? row_stride : (long long)p.X * p.Y * p.Zslab;
    if (p.det_begin < 0 || p.det_count <= 0 || p.det_begin + p.det_count > P->T.n_det) {
        delete P;
        return set_error(SM_E_ARG, "detector range outside the detector table");
    }
    rc = run_sync(*P, p, dst_dev, (hipStream_t)stream);
    delete P;  // views of the scratch array only: nothing is freed
    return rc;
}

int sm_pe_compute(const float* col, const float* det, const float* img, const SmPeOptions* o, float* dst) {
    clear_error();
    if (!col || !det || !img || !o || !dst) return set_error(SM_E_ARG, "null argument");
    int rc = use_device(o->device);
    if (rc < 0) return rc;
    const int X = (int)std::floor(img[0] + 0.001f), Y = (int)std::floor(img[1] + 0.001f);
    const int Z = (int)std::floor(img[2] + 0.001f);
    const int N = (int)std::floor(det[0] + 0.000001f);
    const int zc = o->z_count > 0 ? o->z_count : Z - o->z_begin;
    const int dc = o->det_count > 0 ? o->det_count : N - o->det_begin;
    if (X <= 0 || Y <= 0 || zc <= 0 || dc <= 0) return set_error(SM_E_ARG, "empty output range");
    const size_t n = (size_t)dc * X * Y * zc;
    DeviceBuf<float> out;
    if ((rc = out.alloc(n)) < 0) return rc;
    rc = sm_pe_compute_device(col, det, img, o, out.ptr, 0, nullptr);
    if (rc < 0) return rc;
    HIP_TRY(hipMemcpy(dst, out.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
    return SM_OK;
}

int sm_PESysMatGen(const float* col, const float* det, const float* img, float* dst, int cuda_id) {
    clear_error();
    if (!col || !det || !img || !dst) return set_error(SM_E_ARG, "null argument");
    // in-band slab description (PEGen_CircularHole.cpp:165,177,189-190).  The
    // reference sizes the caller's buffer with floor(v + 0.001) (driver) and
    // floor(v) (PESysMatGen.cu:396-405); the tables here use floor(v + 0.001).
    // A count where the two differ (e.g. 63.9995) would make rows one voxel
    // longer than the caller's buffer, so it is rejected.
    for (int k : {0, 1, 2}) {
        if (std::floor(img[k]) != std::floor(img[k] + 0.001f))
            return set_error(SM_E_ARG, "image voxel count Img[" + std::to_string(k) + "] = " +
                                           std::to_string(img[k]) + " is not integral");
    }
    const int X = (int)std::floor(img[0]), Y = (int)std::floor(img[1]);
    const int slabZ = (int)std::floor(img[2]);
    const int fullZ = (int)std::floor(img[30]);
    const int z0 = (int)std::floor(img[31] + 0.001f);
    if (X <= 0 || Y <= 0 || slabZ <= 0) return set_error(SM_E_ARG, "image voxel counts must be positive");
    SmPeOptions o;
    sm_pe_default_options(&o);
    o.device = cuda_id;
    o.rotation = (int)std::floor(img[20] + 0.001f);
    // The kernel positions voxels with the full Z (Img[30]) and computes the
    // slab's planes z0 .. z0+slabZ-1; present that as a full image of fullZ.
    float img2[32];
    std::memcpy(img2, img, sizeof(img2));
    img2[2] = (float)std::max(fullZ, z0 + slabZ);
    o.z_begin = z0;
    o.z_count = slabZ;
    // The reference uses Img[30] for voxel positions; keep it when it differs
    // from the slab range (fullZ == 0 means "unset" in the reference too).
    SmPePlan* plan = nullptr;
    int rc = sm_pe_plan_create(col, det, img2, &o, &plan);
    if (rc < 0) return rc;
    auto* P = reinterpret_cast<PePlanImpl*>(plan);
    if (P->base.Zfull != fullZ) {
        // Img[30] unset or short of the slab: the candidate lists follow the
        // voxel positions, so cull them again for the reference's positions
        P->base.Zfull = fullZ;
        P->base.hZ64 = (double)fullZ / 2.0;  // the FP64 centre of the collimator path follows it
        if ((rc = build_lists(*P)) < 0) {
            sm_pe_plan_destroy(plan);
            return rc;
        }
    }
    if (P->base.X != X || P->base.Y != Y || P->base.Zslab != slabZ) {
        sm_pe_plan_destroy(plan);
        return set_error(SM_E_ARG, "slab shape of the tables differs from the in-band slab fields");
    }
    const size_t n = (size_t)P->base.det_count * X * Y * slabZ;
    DeviceBuf<float> out;
    if ((rc = out.alloc(n)) < 0) {
        sm_pe_plan_destroy(plan);
        return rc;
    }
    PeParams p = P->base;
    p.row_stride = (long long)X * Y * slabZ;
    rc = run_sync(*P, p, out.ptr, nullptr);
    sm_pe_plan_destroy(plan);
    if (rc < 0) return rc;
    HIP_TRY(hipMemcpy(dst, out.ptr, n * sizeof(float), hipMemcpyDeviceToHost));
    return X * Y * slabZ;
}

}  // extern "C"

// C++-linkage drop-in with the reference's exact signature (PESysMatGen.h:4-9).
int PESysMatGen(float* parameter_Collimator, float* parameter_Detector, float* parameter_Image, float* dst,
                int cuda_id) {
    int rc = sm_PESysMatGen(parameter_Collimator, parameter_Detector, parameter_Image, dst, cuda_id);
    if (rc < 0) std::fprintf(stderr, "PESysMatGen: %s\n", sm_last_error());
    return rc;
}
