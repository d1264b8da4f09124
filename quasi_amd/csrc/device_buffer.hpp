// device_buffer.hpp -- a typed device allocation that frees itself, for the
// host code of the GPU scene build (bvh_gpu.hip, bvh_layout.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace quasi_hip {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    bool alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)) != hipSuccess) { p = nullptr; return false; }
        n = count;
        return true;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Hands the allocation to the caller, who frees it with hipFree.
    T* detach() {
        T* q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    ~DevBuf() { release(); }
};

}  // namespace quasi_hip
