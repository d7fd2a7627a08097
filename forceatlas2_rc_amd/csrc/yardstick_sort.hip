// libfa2yard.so: rocPRIM's device radix sort of (key, value) pairs, timed.
// A yardstick for the engine's own sort of the Morton codes, loaded by the
// benchmark only: nothing in libfa2hip.so links or calls it, and the
// product's sort (engine.hip) uses no library sort.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <cstdint>
#include <vector>

namespace {

struct Buf {
    void *p = nullptr;
    ~Buf()
    {
        if (p) (void) hipFree(p);
    }
};

} // namespace

// Sort n (keys[i], i) pairs on all 32 key bits `reps` times on `device`
// (after two untimed sorts) and return the average device time of one sort
// in *ms (HIP events around the timed sorts, on one stream). *sorted is 1
// when the result is the stable order of the keys (checked on the host).
// 0 on success, 1 on a HIP error.
extern "C" int fa2yard_radix_sort_pairs(int device, int n, const uint32_t *keys, int reps,
                                        double *ms, int *sorted)
{
    if (n < 1 || reps < 1 || hipSetDevice(device) != hipSuccess) return 1;
    const size_t bytes = (size_t) n * sizeof(uint32_t);
    std::vector<uint32_t> vals(n);
    for (int i = 0; i < n; i++) vals[i] = (uint32_t) i;
    Buf ki, vi, ko, vo, tmp;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMalloc(&ki.p, bytes) || hipMalloc(&vi.p, bytes) || hipMalloc(&ko.p, bytes) ||
        hipMalloc(&vo.p, bytes))
        return 1;
    if (hipMemcpy(ki.p, keys, bytes, hipMemcpyHostToDevice) ||
        hipMemcpy(vi.p, vals.data(), bytes, hipMemcpyHostToDevice))
        return 1;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return 1;
    int rc = 1;
    size_t tmp_bytes = 0;
    auto sort = [&]() {
        return rocprim::radix_sort_pairs(tmp.p, tmp_bytes, (const uint32_t *) ki.p,
                                         (uint32_t *) ko.p, (const uint32_t *) vi.p,
                                         (uint32_t *) vo.p, (size_t) n, 0, 32, st);
    };
    if (sort() == hipSuccess && hipMalloc(&tmp.p, tmp_bytes) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        bool ok = sort() == hipSuccess && sort() == hipSuccess &&
                  hipEventRecord(e0, st) == hipSuccess;
        for (int r = 0; ok && r < reps; r++) ok = sort() == hipSuccess;
        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
        float t = 0.f;
        if (ok && hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
            *ms = (double) t / reps;
            std::vector<uint32_t> k(n), v(n);
            if (hipMemcpy(k.data(), ko.p, bytes, hipMemcpyDeviceToHost) == hipSuccess &&
                hipMemcpy(v.data(), vo.p, bytes, hipMemcpyDeviceToHost) == hipSuccess) {
                int good = 1;
                for (int i = 0; i < n && good; i++) {
                    good = k[i] == keys[v[i]];
                    if (i > 0 && good)
                        good = k[i - 1] < k[i] || (k[i - 1] == k[i] && v[i - 1] < v[i]);
                }
                *sorted = good;
                rc = 0;
            }
        }
    }
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    (void) hipStreamDestroy(st);
    return rc;
}
