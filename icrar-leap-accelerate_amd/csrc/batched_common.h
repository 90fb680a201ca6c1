// Device helpers shared by the batched kernels: complex pairs with the fused
// operations written out, the FP64 phasor and the reference's shift factor.
#pragma once

#include "leap_internal.h"
#include "phasor.h"

// Global-memory pointers: pointers read from a descriptor table are otherwise
// generic, and flat memory instructions share the LDS wait counter.
#define LEAP_GLOBAL __attribute__((address_space(1)))

namespace leap {
namespace dev {

template <typename T>
struct C2 {
    T x, y;
};

// Complex product with the fused operations written out (no contraction), so
// every kernel evaluates it the same way.
template <typename T>
__device__ inline C2<T> cmul(C2<T> a, C2<T> b) {
#pragma clang fp contract(off)
    return {fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)};
}

// a * z + v with fused multiply-adds.
template <typename T>
__device__ inline C2<T> horner_step(C2<T> a, C2<T> z, C2<T> v) {
    C2<T> r;
    r.x = fma(a.x, z.x, fma(-a.y, z.y, v.x));
    r.y = fma(a.x, z.y, fma(a.y, z.x, v.y));
    return r;
}

constexpr double kTwoPiHi = 6.283185307179586232;           // double(2 pi)
constexpr double kTwoPiLo = 2.4492935982947064e-16;         // 2 pi - double(2 pi)
constexpr double kInvTwoPi = 0.15915494309189533577;

// exp(i theta) with the argument reduced modulo 2 pi in FP64 and the sincos in FP32.
__device__ inline C2<float> phasor_mixed(double theta) {
    const double n = rint(theta * kInvTwoPi);
    double r = fma(-n, kTwoPiHi, theta);
    r = fma(-n, kTwoPiLo, r);
    float s, c;
    sincosf(float(r), &s, &c);
    return {c, s};
}

__device__ inline C2<double> phasor_f64(double theta) {
    const Phasor64 p = phasor64(theta);
    return {p.c, p.s};
}

// -2 pi ((dd . uvw).z - w) (CpuLeapCalibrator.cc:190) with the fused operations written out.
__device__ inline double shift_factor(const double dz[3], double u, double v, double w) {
#pragma clang fp contract(off)
    double t = dz[0] * u;
    t = fma(dz[1], v, t);
    t = fma(dz[2], w, t);
    return -kTwoPi * (t - w);
}

}  // namespace dev
}  // namespace leap
