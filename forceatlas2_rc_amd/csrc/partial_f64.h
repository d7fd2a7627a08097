// Double-precision repulsion kernels that follow the reference CPU code
// operation by operation (partial_f64.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace fa2 {

hipError_t launch_exact_f64(hipStream_t st, int n, double scaling, bool adjust, const double *x,
                            const double *y, const double *m, const double *s, double2 *f);

hipError_t launch_bh_f64(hipStream_t st, int n, double scaling, double theta2, bool adjust,
                         const double *px, const double *py, const double *pm, const double *ps,
                         const double4 *rr, const int4 *ri, double2 *f);

} // namespace fa2
