// exp(i theta) in FP64 for the batched Horner kernels, host and device.
//
// Cody & Waite argument reduction by pi/2 with a two-part constant applied
// through fused multiply-adds (the product n * hi is exact inside the fma, and
// hi + lo carries pi/2 to about 2^-107 relative), then the Taylor series of sin
// and cos on |r| <= pi/4: sin to r^17 (next term < 1e-19), cos to r^16 (next
// term < 3e-18).  Quadrant q = n mod 4 maps (cos r, sin r) to (cos, sin) of
// theta.  Accurate to a few ulp for |theta| < 2^40; the kernels' arguments
// (2 pi w nu / c) stay below 1e6.
//
// Contraction is off so every fused operation is one written here: the host
// build and every kernel that uses these helpers evaluate the same sequence.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace leap {

struct Phasor64 {
    double c, s;  // cos theta, sin theta
};

__host__ __device__ inline Phasor64 phasor64(double theta) {
#pragma clang fp contract(off)
    constexpr double kTwoOverPi = 0.63661977236758134308;   // 2 / pi
    constexpr double kPiO2Hi = 1.5707963267948965580;       // double(pi / 2)
    constexpr double kPiO2Lo = 6.1232339957367658e-17;      // pi / 2 - double(pi / 2)
    const double n = rint(theta * kTwoOverPi);
    double r = fma(-n, kPiO2Hi, theta);
    r = fma(-n, kPiO2Lo, r);
    const double r2 = r * r;
    // sin r = r + r^3 * (-1/3! + r^2 (1/5! - ...)), cos r = 1 + r^2 (-1/2! + r^2 (1/4! - ...))
    double ps = 1.0 / 355687428096000.0;                // 1/17!
    ps = fma(ps, r2, -1.0 / 1307674368000.0);           // 1/15!
    ps = fma(ps, r2, 1.0 / 6227020800.0);               // 1/13!
    ps = fma(ps, r2, -1.0 / 39916800.0);                // 1/11!
    ps = fma(ps, r2, 1.0 / 362880.0);                   // 1/9!
    ps = fma(ps, r2, -1.0 / 5040.0);                    // 1/7!
    ps = fma(ps, r2, 1.0 / 120.0);                      // 1/5!
    ps = fma(ps, r2, -1.0 / 6.0);                       // 1/3!
    const double sn = fma(r * r2, ps, r);
    double pc = 1.0 / 20922789888000.0;                 // 1/16!
    pc = fma(pc, r2, -1.0 / 87178291200.0);             // 1/14!
    pc = fma(pc, r2, 1.0 / 479001600.0);                // 1/12!
    pc = fma(pc, r2, -1.0 / 3628800.0);                 // 1/10!
    pc = fma(pc, r2, 1.0 / 40320.0);                    // 1/8!
    pc = fma(pc, r2, -1.0 / 720.0);                     // 1/6!
    pc = fma(pc, r2, 1.0 / 24.0);                       // 1/4!
    pc = fma(pc, r2, -0.5);                             // 1/2!
    const double cs = fma(r2, pc, 1.0);
    const int q = int(static_cast<int64_t>(n) & 3);
    Phasor64 out;
    out.c = (q & 1) ? sn : cs;
    out.s = (q & 1) ? cs : sn;
    if (q == 1 || q == 2) out.c = -out.c;
    if (q >= 2) out.s = -out.s;
    return out;
}

// phasor64 of M arguments with the steps of the M evaluations interleaved
// (independent chains side by side); the same operations as phasor64 for
// each argument, so the same bits.
template <int M>
__device__ inline void phasor64_n(const double* theta, Phasor64* out) {
#pragma clang fp contract(off)
    constexpr double kTwoOverPi = 0.63661977236758134308;
    constexpr double kPiO2Hi = 1.5707963267948965580;
    constexpr double kPiO2Lo = 6.1232339957367658e-17;
    constexpr double ks[8] = {-1.0 / 1307674368000.0, 1.0 / 6227020800.0, -1.0 / 39916800.0, 1.0 / 362880.0,
                              -1.0 / 5040.0,          1.0 / 120.0,        -1.0 / 6.0,        0.0};
    constexpr double kc[8] = {-1.0 / 87178291200.0, 1.0 / 479001600.0, -1.0 / 3628800.0, 1.0 / 40320.0,
                              -1.0 / 720.0,         1.0 / 24.0,        -0.5,             0.0};
    double n[M], r[M], r2[M], ps[M], pc[M];
#pragma unroll
    for (int j = 0; j < M; ++j) n[j] = rint(theta[j] * kTwoOverPi);
#pragma unroll
    for (int j = 0; j < M; ++j) r[j] = fma(-n[j], kPiO2Hi, theta[j]);
#pragma unroll
    for (int j = 0; j < M; ++j) r[j] = fma(-n[j], kPiO2Lo, r[j]);
#pragma unroll
    for (int j = 0; j < M; ++j) {
        r2[j] = r[j] * r[j];
        ps[j] = 1.0 / 355687428096000.0;
        pc[j] = 1.0 / 20922789888000.0;
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
#pragma unroll
        for (int j = 0; j < M; ++j) {
            ps[j] = fma(ps[j], r2[j], ks[k]);
            pc[j] = fma(pc[j], r2[j], kc[k]);
        }
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const double sn = fma(r[j] * r2[j], ps[j], r[j]);
        const double cs = fma(r2[j], pc[j], 1.0);
        const int q = int(static_cast<int64_t>(n[j]) & 3);
        out[j].c = (q & 1) ? sn : cs;
        out[j].s = (q & 1) ? cs : sn;
        if (q == 1 || q == 2) out[j].c = -out[j].c;
        if (q >= 2) out[j].s = -out[j].s;
    }
}

}  // namespace leap
