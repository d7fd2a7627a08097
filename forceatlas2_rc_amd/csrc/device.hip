// Devices, errors, clock and the per-device stream cache of libfa2hip.so.
// Replaces the OpenCL binding of the reference (src/ocl_loader.c,
// src/gpu_internal.h, device part of src/gpu.c:129-351): HIP is linked
// directly, there is nothing to load at run time and no program to build.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <mutex>

#include "common.h"
#include "../../include/fa2.h"

namespace fa2 {

static thread_local char g_err[2048] = "";

void set_message(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void set_hip_error(const char *what, hipError_t e)
{
    set_message("%s (HIP error %d: %s)", what, (int) e, hipGetErrorString(e));
}

static std::mutex g_mu;
static int g_ndev = -1;
static constexpr int kMaxDevices = 64;
static hipStream_t g_stream[kMaxDevices] = {};

int device_count()
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ndev >= 0) return g_ndev;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        set_hip_error("no usable HIP runtime or device was found", e);
        return 0; // not cached: a device may appear later (src/gpu.c:161-165)
    }
    g_ndev = n < kMaxDevices ? n : kMaxDevices;
    return g_ndev;
}

hipStream_t device_stream(int device)
{
    if (device < 0 || device >= device_count()) {
        set_message("there is no GPU device %d", device + 1);
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        set_hip_error("could not select the GPU device", e);
        return nullptr;
    }
    if (!g_stream[device]) {
        // default priority on purpose: streams of different priorities are
        // never mixed in this library
        e = hipStreamCreateWithFlags(&g_stream[device], hipStreamNonBlocking);
        if (e != hipSuccess) {
            g_stream[device] = nullptr;
            set_hip_error("could not create a HIP stream", e);
            return nullptr;
        }
    }
    return g_stream[device];
}

} // namespace fa2

using namespace fa2;

extern "C" const char *fa2_gpu_error(void) { return g_err; }

extern "C" int fa2_gpu_ndevices(void) { return device_count(); }

extern "C" int fa2_gpu_device_info(int idx, fa2_gpu_devinfo *info)
{
    if (!info || idx < 0 || idx >= device_count()) return 1;
    memset(info, 0, sizeof(*info));
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, idx);
    if (e != hipSuccess) {
        set_hip_error("could not query the GPU device", e);
        return 1;
    }
    snprintf(info->name, sizeof(info->name), "%s", prop.name);
    snprintf(info->vendor, sizeof(info->vendor), "Advanced Micro Devices, Inc.");
    int rt = 0;
    (void) hipRuntimeGetVersion(&rt);
    snprintf(info->platform, sizeof(info->platform), "HIP %d.%d", rt / 10000000,
             (rt / 100000) % 100);
    snprintf(info->version, sizeof(info->version), "%s", prop.gcnArchName);
    info->is_gpu = 1;
    info->has_double = 1; // every CDNA device has full-rate FP64
    info->memory = (double) prop.totalGlobalMem;
    return 0;
}

extern "C" void fa2_gpu_shutdown(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (int d = 0; d < kMaxDevices; d++) {
        if (g_stream[d]) {
            (void) hipSetDevice(d);
            (void) hipStreamSynchronize(g_stream[d]);
            (void) hipStreamDestroy(g_stream[d]);
            g_stream[d] = nullptr;
        }
    }
    g_ndev = -1;
}

extern "C" double fa2_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}
