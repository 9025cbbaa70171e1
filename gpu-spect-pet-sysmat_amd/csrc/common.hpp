// Shared host helpers: per-thread error text, device selection, RAII device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/sysmat.h"

namespace sysmat {

int set_error(int code, const std::string& msg);
void clear_error();
int use_device(int device);

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess)                                                                   \
            return ::sysmat::set_error(_e == hipErrorOutOfMemory ? SM_E_NOMEM : SM_E_DEVICE,    \
                                       std::string(#expr) + ": " + hipGetErrorString(_e));      \
    } while (0)

template <class T>
struct DeviceBuf {
    T* ptr = nullptr;
    size_t n = 0;
    bool owned = true;  // false: a view of memory someone else owns (never freed here)
    DeviceBuf() = default;
    DeviceBuf(const DeviceBuf&) = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    ~DeviceBuf() { release(); }
    void release() {
        if (ptr && owned) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
        owned = true;
    }
    void attach(T* p, size_t count) {
        release();
        ptr = p;
        n = count;
        owned = false;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&ptr, count * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            return set_error(SM_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        n = count;
        owned = true;
        return 0;
    }
    int upload(const std::vector<T>& v) {
        int rc = alloc(v.size());
        if (rc < 0) return rc;
        if (!v.empty()) {
            hipError_t e = hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
            if (e != hipSuccess) return set_error(SM_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        }
        return 0;
    }
};

// Bump allocator over a device region that someone else owns (no allocation):
// consecutive 256-byte-aligned pieces, nullptr once the region is full.
struct DeviceArena {
    unsigned char* base = nullptr;
    size_t cap = 0, used = 0;
    template <class T>
    T* take(size_t count) {
        if (count == 0) count = 1;
        const size_t off = (used + 255) & ~size_t(255);
        if (off > cap || count > (cap - off) / sizeof(T)) return nullptr;
        used = off + count * sizeof(T);
        return reinterpret_cast<T*>(base + off);
    }
};

}  // namespace sysmat
