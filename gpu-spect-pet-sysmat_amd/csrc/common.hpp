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
    DeviceBuf() = default;
    DeviceBuf(const DeviceBuf&) = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    ~DeviceBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
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

}  // namespace sysmat
