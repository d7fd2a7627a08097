// Double-precision kernels of the repulsion offload (fa2_gpu_open with
// use_double). They follow the CPU code of the reference operation by
// operation (src/forceatlas2.c:265-368): same expressions, same order of the
// terms, no fused multiply-add (this file is compiled with
// -ffp-contract=off), IEEE division and square root. The double-precision
// GPU layout then stays within 1e-9 of the CPU one even where the dynamics
// amplify rounding differences (prevent_overlap), as the reference's own
// tests require (tests/testthat/test-gpu.R:34-58).
#include "common.h"
#include "kernels_common.h"
#include "partial_f64.h"

namespace fa2 {

namespace {

constexpr int kThreads = 256;

// rep_factor of src/forceatlas2.c:265-275
__device__ inline double ref_factor(bool adjust, double scaling, double d2, double mm, double ss)
{
    if (adjust) {
        const double d = __builtin_sqrt(d2) - ss;
        if (d > 0) return scaling * mm / (d * d);
        if (d < 0) return 100.0 * scaling * mm;
        return 0.0;
    }
    return d2 > 0 ? scaling * mm / d2 : 0.0;
}

// repulse_node_exact (src/forceatlas2.c:278-294): every other node, in
// increasing order, staged through LDS without changing that order.
template <bool ADJUST>
__global__ void __launch_bounds__(kThreads)
k_exact_f64(int n, double scaling, const double *__restrict__ x, const double *__restrict__ y,
            const double *__restrict__ m, const double *__restrict__ s, double2 *__restrict__ f)
{
    __shared__ double4 tile[kThreads];
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const int ii = i < n ? i : n - 1;
    const double xi = x[ii], yi = y[ii], mi = m[ii];
    const double si = ADJUST ? s[ii] : 0.0;
    double fx = 0.0, fy = 0.0;
    for (int j0 = 0; j0 < n; j0 += kThreads) {
        __syncthreads();
        const int jl = j0 + (int) threadIdx.x;
        if (jl < n) tile[threadIdx.x] = make_double4(x[jl], y[jl], m[jl], ADJUST ? s[jl] : 0.0);
        __syncthreads();
        const int cnt = min(kThreads, n - j0);
        for (int t = 0; t < cnt; t++) {
            const int j = j0 + t;
            if (j == ii) continue;
            const double4 o = tile[t];
            const double xd = xi - o.x, yd = yi - o.y;
            const double fc = ref_factor(ADJUST, scaling, xd * xd + yd * yd, mi * o.z,
                                         ADJUST ? si + o.w : 0.0);
            fx += xd * fc;
            fy += yd * fc;
        }
    }
    if (i < n) f[i] = make_double2(fx, fy);
}

// repulse_node_bh (src/forceatlas2.c:326-368) along the skip links of the
// host tree (src/gpu.c:82-123): rr = (cx, cy, mass, size^2), ri = (start,
// count, last child, skip), nodes in tree order.
template <bool ADJUST>
__global__ void __launch_bounds__(kThreads)
k_bh_f64(int n, double scaling, double theta2, const double *__restrict__ px,
         const double *__restrict__ py, const double *__restrict__ pm,
         const double *__restrict__ ps, const double4 *__restrict__ rr,
         const int4 *__restrict__ ri, double2 *__restrict__ f)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) return;
    const double xi = px[k], yi = py[k], mi = pm[k];
    const double si = ADJUST ? ps[k] : 0.0;
    double fx = 0.0, fy = 0.0;
    int cur = 0;
    for (int steps = 2 * n + 1; cur >= 0 && steps > 0; steps--) {
        const int4 info = ri[cur];
        if (info.y > 1) {
            const double4 c = rr[cur];
            const double xd = xi - c.x, yd = yi - c.y;
            const double d2 = xd * xd + yd * yd;
            if (d2 * theta2 > c.w) {
                const double fc = scaling * (mi * c.z) / d2;
                fx += xd * fc;
                fy += yd * fc;
                cur = info.w;
                continue;
            }
            if (info.z >= 0) {
                cur = info.z;
                continue;
            }
        }
        for (int j = info.x; j < info.x + info.y; j++) {
            if (j == k) continue;
            const double xd = xi - px[j], yd = yi - py[j];
            const double fc = ref_factor(ADJUST, scaling, xd * xd + yd * yd, mi * pm[j],
                                         ADJUST ? si + ps[j] : 0.0);
            fx += xd * fc;
            fy += yd * fc;
        }
        cur = info.w;
    }
    f[k] = make_double2(fx, fy);
}

} // namespace

hipError_t launch_exact_f64(hipStream_t st, int n, double scaling, bool adjust, const double *x,
                            const double *y, const double *m, const double *s, double2 *f)
{
    const unsigned blocks = ceil_div(n, kThreads);
    if (adjust)
        k_exact_f64<true><<<blocks, kThreads, 0, st>>>(n, scaling, x, y, m, s, f);
    else
        k_exact_f64<false><<<blocks, kThreads, 0, st>>>(n, scaling, x, y, m, s, f);
    return hipGetLastError();
}

hipError_t launch_bh_f64(hipStream_t st, int n, double scaling, double theta2, bool adjust,
                         const double *px, const double *py, const double *pm, const double *ps,
                         const double4 *rr, const int4 *ri, double2 *f)
{
    const unsigned blocks = ceil_div(n, kThreads);
    if (adjust)
        k_bh_f64<true><<<blocks, kThreads, 0, st>>>(n, scaling, theta2, px, py, pm, ps, rr, ri, f);
    else
        k_bh_f64<false><<<blocks, kThreads, 0, st>>>(n, scaling, theta2, px, py, pm, ps, rr, ri,
                                                     f);
    return hipGetLastError();
}

} // namespace fa2
