// hip_util.hpp -- HIP error checks and an owning device buffer, shared by
// the host runtime (runtime.hip) and the device-side matrix builder
// (device_build.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace xp {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPX(expr) ::xp::hip_check((expr), #expr)

// The library's caching device allocator (device_pool.cpp): blocks of at
// least 256 KB are kept after release and reused (a released block is
// reused only after a device synchronisation), smaller ones go straight to
// hipMalloc / hipFree.  pool_trim returns every cached block to the runtime.
void* pool_alloc(size_t bytes);
void pool_free(void* p, size_t bytes);
void pool_trim();
size_t pool_cached_bytes();

// Owning device allocation of `n` elements of T (through the caching
// allocator).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = static_cast<T*>(pool_alloc(count * sizeof(T)));
        n = count;
    }
    void ensure(size_t count) { if (count > n) alloc(std::max(count, n + n / 2)); }
    void release() {
        if (p != nullptr) pool_free(p, n * sizeof(T));
        p = nullptr;
        n = 0;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        if (count) HIPX(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T* dst, size_t count, hipStream_t s) const {
        if (count) HIPX(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

template <typename T>
DevBuf<T> to_device(const std::vector<T>& v, hipStream_t s) {
    DevBuf<T> d(v.size());
    d.upload(v.data(), v.size(), s);
    return d;
}

}  // namespace xp
