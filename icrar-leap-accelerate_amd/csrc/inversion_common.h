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
// every kernel that uses it computes the same bits.  delta_phase_of is the
// arithmetic on gathered values (acal1 = cal1[a1] - cal1[a2], avgData pol 0),
// for kernels that gather several elements before computing them.
__device__ inline double delta_phase_of(double acal1, double re0, double im0) {
#pragma clang fp contract(off)
    double sn, cs;
    sincos(-kTwoPi * acal1, &sn, &cs);
    const double re = cs * re0 - sn * im0;
    const double im = cs * im0 + sn * re0;
    return atan2(im, re);
}

__device__ inline double delta_phase(const double* __restrict__ avg, uint32_t nbl, uint32_t P,
                                     const int2* __restrict__ pairs, int32_t K, int32_t N,
                                     const double* __restrict__ cal1, uint32_t d, int32_t n) {
#pragma clang fp contract(off)
    if (n >= K - 1) return 0.0;
    const int2 p = pairs[n];
    const double acal1 = cal1[size_t(d) * N + p.x] - cal1[size_t(d) * N + p.y];
    const double* a = avg + (size_t(d) * nbl + n) * (2 * P);
    return delta_phase_of(acal1, a[0], a[1]);
}

}  // namespace dev
}  // namespace leap
