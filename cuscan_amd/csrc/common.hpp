// Shared host-side helpers: status/error plumbing and HIP checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/umaphip.h"

namespace umaphip {

void set_error(const char *fmt, ...);

struct Fail {
    umaphip_status status;
};

#define UMAPHIP_REQUIRE(cond, code, ...)          \
    do {                                          \
        if (!(cond)) {                            \
            ::umaphip::set_error(__VA_ARGS__);    \
            throw ::umaphip::Fail{code};          \
        }                                         \
    } while (0)

#define HIP_CHECK(expr)                                                              \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) {                                                      \
            ::umaphip::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                                 __FILE__, __LINE__);                                \
            throw ::umaphip::Fail{_e == hipErrorOutOfMemory ? UMAPHIP_MEMORY          \
                                                            : UMAPHIP_BACKEND};      \
        }                                                                            \
    } while (0)

// Runs `body` and converts failures into a status (the C ABI never throws).
template <class F>
umaphip_status guarded(F &&body) {
    try {
        body();
        return UMAPHIP_OK;
    } catch (const Fail &f) {
        return f.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return UMAPHIP_MEMORY;
    } catch (...) {
        set_error("unexpected C++ exception");
        return UMAPHIP_BACKEND;
    }
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Device buffer owned by host code (plans and drivers only).
template <class T>
struct DeviceArray {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { alloc(n); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    ~DeviceArray() {
        if (ptr) (void)hipFree(ptr);
    }
    void alloc(size_t n) {
        if (ptr) {
            (void)hipFree(ptr);
            ptr = nullptr;
        }
        count = n;
        if (n) HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T)));
    }
    size_t bytes() const { return count * sizeof(T); }
};

}  // namespace umaphip
