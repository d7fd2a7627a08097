// Device helpers shared by the kernels of libfa2hip.so (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fa2 {

template <typename R> struct Vec;
template <> struct Vec<float> {
    using v2 = float2;
    using v4 = float4;
};
template <> struct Vec<double> {
    using v2 = double2;
    using v4 = double4;
};

template <typename R> __host__ __device__ inline typename Vec<R>::v2 mk2(R a, R b)
{
    typename Vec<R>::v2 v;
    v.x = a;
    v.y = b;
    return v;
}

template <typename R> __host__ __device__ inline typename Vec<R>::v4 mk4(R a, R b, R c, R d)
{
    typename Vec<R>::v4 v;
    v.x = a;
    v.y = b;
    v.z = c;
    v.w = d;
    return v;
}

// Loads and stores of data that streams through a kernel once (the
// attraction's incidence lists, the forces the traversal leaves for the next
// phase, the forces and accumulators of the displacement), with the
// non-temporal hint: the lines do not displace what the
// kernels running beside them reuse from L2 (the traversal's records near the
// tree's top, the gathered positions). Only the cache policy differs.
template <typename T> __device__ inline T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }
__device__ inline double2 ld_stream(const double2 *p)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v v = __builtin_nontemporal_load((const d2v *) p);
    return make_double2(v.x, v.y);
}
__device__ inline void st_stream(double2 *p, double x, double y)
{
    typedef double d2v __attribute__((ext_vector_type(2)));
    const d2v v = {x, y};
    __builtin_nontemporal_store(v, (d2v *) p);
}

// Reciprocal for the bulk force arithmetic of a positive, finite x:
// v_rcp_f32 (1 ulp) in FP32; in FP64, v_rcp_f64 refined by two
// Newton-Raphson steps in fused multiply-adds (each step squares the
// relative error), within about an ulp of 1/x. The IEEE FP64 division would
// add a scaling, a third refinement and a fix-up: twice the instructions.
__device__ inline float frcp_r(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double frcp_r(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Square root for the bulk arithmetic: v_sqrt_f32 (1 ulp) in FP32, IEEE in
// FP64.
__device__ inline float fsqrt_r(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fsqrt_r(double x) { return __builtin_sqrt(x); }

// a * a + b * b with each operation rounded on its own (no fused
// multiply-add): the squared distance of the Barnes-Hut acceptance test and
// the squared diagonal of a region, as the reference computes them
// (pipeline.cl:38 turns contraction off). With the same region records, the
// device then takes exactly the reference's decisions.
template <typename R> __device__ inline R sumsq_rn(R a, R b)
{
#pragma clang fp contract(off)
    return a * a + b * b;
}

__device__ inline float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ inline float flog1p_r(float x) { return log1pf(x); }
__device__ inline double flog1p_r(double x) { return log1p(x); }

// Repulsion factor between two nodes without sizes (reference rep_factor,
// inst/opencl/pipeline.cl:46-55); smm = scaling * m_i * m_j.
template <typename R> __device__ inline R pair_factor(R d2, R smm)
{
    return d2 > (R) 0 ? smm * frcp_r(d2) : (R) 0;
}

// The same with node sizes (prevent_overlap): d = |x_i - x_j| - (s_i + s_j),
// scaling m_i m_j / d^2 apart, 100 scaling m_i m_j overlapping. The gap d is
// a difference of nearly equal numbers near contact, where the force grows
// as 1/d^2: it is computed in FP64 from the FP32 inputs, so that its relative
// error does not grow as the nodes touch.
template <typename R> __device__ inline R pair_factor_adjust(R xd, R yd, R smm, R si, R sj)
{
    const double dx = (double) xd, dy = (double) yd;
    const double d = __builtin_sqrt(dx * dx + dy * dy) - ((double) si + (double) sj);
    if (d > 0.0) return (R) ((double) smm / (d * d));
    if (d < 0.0) return (R) 100 * smm;
    return (R) 0;
}

// Compensated (Kahan) sum of a 2-vector in the working precision: long sums
// of terms of both signs keep about twice the precision of a plain sum.
template <typename R> struct KahanSum2 {
    R x = (R) 0, y = (R) 0, cx = (R) 0, cy = (R) 0;
    __device__ inline void add(R tx, R ty)
    {
        const R ux = tx - cx, uy = ty - cy;
        const R sx = x + ux, sy = y + uy;
        cx = (sx - x) - ux;
        cy = (sy - y) - uy;
        x = sx;
        y = sy;
    }
    __device__ inline R sum_x() const { return x - cx; }
    __device__ inline R sum_y() const { return y - cy; }
};

// Lane index inside the wave64.
__device__ inline int lane_id() { return (int) __lane_id(); }

// Wave-wide sums in a fixed order (xor butterfly): every lane ends with a
// value; lane 0's is the one used, so the result is reproducible.
template <typename T> __device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T> __device__ inline T wave_min(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

template <typename T> __device__ inline T wave_max(T v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const T w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

} // namespace fa2
