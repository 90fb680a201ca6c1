// Device helper shared by the inversion kernels (FP64 and mixed paths).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "leap_internal.h"

namespace leap {
namespace dev {

// delta[d][n] = arg(exp(-2 pi i (cal1[a1_n] - cal1[a2_n])) * avg[d][n][pol 0]) for
// n < K - 1 (row n of avgData, CpuLeapCalibrator.cc:168-175; A * cal1 in gather
// form), 0 for the constraint row n = K - 1.  Written out without contraction so
// every kernel that uses it computes the same bits.
__device__ inline double delta_phase(const double* __restrict__ avg, uint32_t nbl, uint32_t P,
                                     const int2* __restrict__ pairs, int32_t K, int32_t N,
                                     const double* __restrict__ cal1, uint32_t d, int32_t n) {
#pragma clang fp contract(off)
    if (n >= K - 1) return 0.0;
    const int2 p = pairs[n];
    const double acal1 = cal1[size_t(d) * N + p.x] - cal1[size_t(d) * N + p.y];
    double sn, cs;
    sincos(-kTwoPi * acal1, &sn, &cs);
    const double* a = avg + (size_t(d) * nbl + n) * (2 * P);
    const double re = cs * a[0] - sn * a[1];
    const double im = cs * a[1] + sn * a[0];
    return atan2(im, re);
}

}  // namespace dev
}  // namespace leap
