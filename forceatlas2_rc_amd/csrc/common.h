// Shared host-side helpers of libfa2hip.so: error reporting, HIP checks,
// per-device streams.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "message.h"

namespace fa2 {

void set_hip_error(const char *what, hipError_t e);

// Stream of a device, created on first use and kept until
// fa2_gpu_shutdown() (the reference caches its OpenCL context and queue the
// same way, src/gpu.c:236-321). Sets the device current. nullptr on error.
hipStream_t device_stream(int device);

// Number of devices (0 when the runtime is unusable), cached.
int device_count();

// Sticky-error helper: the first failing HIP call is kept.
struct HipStatus {
    hipError_t err = hipSuccess;
    const char *what = nullptr;
    bool ok() const { return err == hipSuccess; }
    void check(hipError_t e, const char *w)
    {
        if (err == hipSuccess && e != hipSuccess) {
            err = e;
            what = w;
        }
    }
};

// Device buffer that frees itself.
// With a stream, the memory is allocated and freed in that stream's order
// (hipMallocAsync / hipFreeAsync): freeing it then waits for no other work
// on the device, where hipFree would wait for every stream.
template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t count = 0;
    hipStream_t st = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t n, hipStream_t stream = nullptr)
    {
        release();
        count = n;
        st = stream;
        const size_t bytes = (n > 0 ? n : 1) * sizeof(T);
        return st ? hipMallocAsync((void **) &p, bytes, st) : hipMalloc((void **) &p, bytes);
    }
    void release()
    {
        if (p) (void) (st ? hipFreeAsync(p, st) : hipFree(p));
        p = nullptr;
        count = 0;
        st = nullptr;
    }
    size_t bytes() const { return count * sizeof(T); }
};

// Pinned host buffer that frees itself.
template <typename T> struct PinBuf {
    T *p = nullptr;
    size_t count = 0;
    PinBuf() = default;
    PinBuf(const PinBuf &) = delete;
    PinBuf &operator=(const PinBuf &) = delete;
    ~PinBuf() { release(); }
    hipError_t alloc(size_t n)
    {
        release();
        count = n;
        return hipHostMalloc((void **) &p, (n > 0 ? n : 1) * sizeof(T), hipHostMallocDefault);
    }
    void release()
    {
        if (p) (void) hipHostFree(p);
        p = nullptr;
        count = 0;
    }
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned) ((a + b - 1) / b); }

} // namespace fa2
