// NUFFT channel sums of the mixed-precision batched path on gfx950: the
// kernel templates, instantiated for an FP64 transform (kernels_nufft.hip)
// and an FP32 transform (kernels_nufft32.hip).
//
// For row (t, baseline) and direction d the reference sums
//   sum_c vis[c][0] exp(i sf_d (f0 + c df) / c_light)     (CpuLeapCalibrator.cc:181-217)
// over the channels.  With b = sf_d df / c_light this is
// exp(i sf_d kappa_mid) * sum_{k=-h}^{C-1-h} v_{k+h} exp(i k b), h the plan's
// frequency offset (C / 2 rounded to a multiple of 64), kappa_mid =
// (f0 + h df) / c_light, and the trigonometric polynomial is evaluated by a
// type-2 non-uniform FFT (the plan and its error are described in
// nufft_plan.cpp): each channel's value is multiplied by its precorrection
// factor (here, from the cube, or by the pack kernel), one N-point inverse FFT
// per row puts the polynomial on a uniform grid of N points over one period,
// and each direction interpolates W grid points with the "exponential of
// semicircle" kernel.  Per row and direction this costs about 130 operations
// instead of the 4 C FMAs of the Horner kernel; the FFT's cost is shared by
// every direction of the row.
//
// Precision (template parameter F): the transform, the grid in LDS, the
// interpolation weights and the W-tap sums run in F; the shift factor, the
// grid coordinate, the phasor exp(i sf kappa_mid) and the sums over rows run
// in FP64 in both.  With F = float the grid and its exchanges take half the
// LDS bytes and the transform half the registers; the channel sum's error
// grows from about 1e-11 to about 2e-8 of sum |v| (nufft_plan.cpp).
//
// Work decomposition: a workgroup is one wave.  It owns one baseline and a
// block of 64 * NDL directions (NDL per lane) and walks the baseline's rows in
// order, prefetching the next row's channels while it works on the current
// one.  The FFT is a Stockham transform whose butterflies run in registers
// (radices 16, 16, 4 for N = 1024; 8, 8, 8 for 512; 4, 4, 4, 4 for 256), with
// the exchanges between passes through the wave's LDS (rows of 16 elements
// padded by one element against bank conflicts), except the second exchange
// of N = 1024, a transpose across 16-lane rows done with cross-lane swaps
// (transpose_rows_1024); the last pass writes the grid
// in natural order plus W - 1 wrapped elements, so a direction's W grid points
// are consecutive.  The transform's input is the channels in their own order
// (channel c at index c; lane l owns channels l + 64 i, which are its register
// slots i, the upper half being the zero padding); the frequency offset h of
// the plan then multiplies output m by exp(-2 pi i h m / N), which for
// h = 64 H depends on m mod (N / 64) = lane mod (N / 64) only and is folded into
// the last pass's twiddles.  Wave-scope fences order the LDS exchanges; there
// are no workgroup barriers and no atomics, and each direction's row sums are
// added in row order, so results are reproducible.
#pragma once

#include <algorithm>
#include <type_traits>

#include "batched_common.h"

// No contraction anywhere in these kernels: every fused operation is one
// written as fma(), so a kernel gives the same bits whether a row's values
// come from the cube (precorrection multiplied here) or from packed rows
// (multiplied by the pack kernel).
#pragma clang fp contract(off)

namespace leap {
namespace nufft {

using namespace dev;
using Z = C2<double>;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

constexpr int kWave = 64;

template <typename F>
using AS4 = __attribute__((address_space(4))) F;

// Which rows a workgroup sums: chunks == 1: workgroup x takes baseline
// bl_base + x, all T rows; chunks > 1: workgroup x takes baseline
// bl_base + x / chunks, rows [c T / chunks, (c + 1) T / chunks) with c = x mod
// chunks (output mode kNufftPartial).
struct NufftSpan {
    uint32_t bl_base;
    uint32_t chunks;
};

// Output modes (the kernel's mode argument): set avgData (pol 0, the other
// pols zero), add to avgData, or write the row sums to avg[x][direction]
// (complex; the partial-sum array of a chunked launch, summed by nufft_tail_sum).
constexpr int32_t kNufftSet = 0, kNufftAccumulate = 1, kNufftPartial = 2;

struct NufftArgs {
    const double* coef;   // device: [W / 2][kNufftMaxWidth] weight polynomials of taps j < W / 2, powers of s
    const float* coef32;  // the same coefficients rounded to FP32 (FP32 transforms)
    double kappa_mid;     // (f0 + h df) / c_light
    double dkappa;        // df / c_light
    int32_t h;            // frequency offset, a multiple of 64
};

template <typename F>
__device__ inline C2<F> zadd(C2<F> a, C2<F> b) { return {a.x + b.x, a.y + b.y}; }
template <typename F>
__device__ inline C2<F> zsub(C2<F> a, C2<F> b) { return {a.x - b.x, a.y - b.y}; }
template <typename F>
__device__ inline C2<F> zmuli(C2<F> a) { return {-a.y, a.x}; }  // i * a

constexpr double kCos4 = 0.70710678118654752440;   // cos(pi / 4)
constexpr double kCos8 = 0.92387953251128675613;   // cos(pi / 8)
constexpr double kSin8 = 0.38268343236508977173;   // sin(pi / 8)

// Inverse DFTs in place, natural order: X_k = sum_j x_j exp(+2 pi i j k / R).
template <typename F>
__device__ inline void idft4(C2<F>& a0, C2<F>& a1, C2<F>& a2, C2<F>& a3) {
    const C2<F> t0 = zadd(a0, a2), t1 = zsub(a0, a2), t2 = zadd(a1, a3), t3 = zmuli(zsub(a1, a3));
    a0 = zadd(t0, t2);
    a2 = zsub(t0, t2);
    a1 = zadd(t1, t3);
    a3 = zsub(t1, t3);
}

// idft4 of (a0, a1, 0, 0).
template <typename F>
__device__ inline void idft4_half(C2<F>& a0, C2<F>& a1, C2<F>& a2, C2<F>& a3) {
    const C2<F> x0 = a0, t1 = zmuli(a1);
    a0 = zadd(x0, a1);
    a2 = zsub(x0, a1);
    a3 = zsub(x0, t1);
    a1 = zadd(x0, t1);
}

// idft4 of (a0, .., a_{NZ-1}, 0, ..): NZ inputs that are not known to be zero.
template <int NZ, typename F>
__device__ inline void idft4_nz(C2<F>& a0, C2<F>& a1, C2<F>& a2, C2<F>& a3) {
    static_assert(NZ >= 1 && NZ <= 4, "1 to 4 inputs");
    if constexpr (NZ == 4) {
        idft4(a0, a1, a2, a3);
    } else if constexpr (NZ == 3) {
        const C2<F> t0 = zadd(a0, a2), t1 = zsub(a0, a2), t3 = zmuli(a1);
        a0 = zadd(t0, a1);
        a2 = zsub(t0, a1);
        a1 = zadd(t1, t3);
        a3 = zsub(t1, t3);
    } else if constexpr (NZ == 2) {
        idft4_half(a0, a1, a2, a3);
    } else {
        a1 = a0;
        a2 = a0;
        a3 = a0;
    }
}

// Number of the inputs x[first + stride k], k < 4, below NZ (at least 1).
constexpr int sub_nz(int NZ, int first, int stride) {
    int n = 0;
    for (int k = 0; k < 4; ++k) n += first + stride * k < NZ ? 1 : 0;
    return n < 1 ? 1 : n;
}

// NZ: the inputs from x[NZ] on are zero (the first pass's padding).
template <int NZ, typename F>
__device__ inline void idft8(C2<F>* x) {
    static_assert(NZ >= 2 && NZ <= 8, "2 to 8 inputs");
    // even / odd halves, then X_k = E_k + w8^k O_k, X_{k+4} = E_k - w8^k O_k
    C2<F> e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
    idft4_nz<sub_nz(NZ, 0, 2)>(e[0], e[1], e[2], e[3]);
    idft4_nz<sub_nz(NZ, 1, 2)>(o[0], o[1], o[2], o[3]);
    o[1] = cmul(o[1], C2<F>{F(kCos4), F(kCos4)});
    o[2] = zmuli(o[2]);
    o[3] = cmul(o[3], C2<F>{F(-kCos4), F(kCos4)});
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x[k] = zadd(e[k], o[k]);
        x[k + 4] = zsub(e[k], o[k]);
    }
}

template <int NZ, typename F>
__device__ inline void idft16(C2<F>* x) {
    static_assert(NZ >= 4 && NZ <= 16, "4 to 16 inputs");
    // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, twiddle w16^(n2 k1), DFT4 over n2
    C2<F> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        y[n2][0] = x[n2];
        y[n2][1] = x[4 + n2];
        y[n2][2] = x[8 + n2];
        y[n2][3] = x[12 + n2];
    }
    idft4_nz<sub_nz(NZ, 0, 4)>(y[0][0], y[0][1], y[0][2], y[0][3]);
    idft4_nz<sub_nz(NZ, 1, 4)>(y[1][0], y[1][1], y[1][2], y[1][3]);
    idft4_nz<sub_nz(NZ, 2, 4)>(y[2][0], y[2][1], y[2][2], y[2][3]);
    idft4_nz<sub_nz(NZ, 3, 4)>(y[3][0], y[3][1], y[3][2], y[3][3]);
    const C2<F> w1{F(kCos8), F(kSin8)}, w2{F(kCos4), F(kCos4)}, w3{F(kSin8), F(kCos8)}, w6{F(-kCos4), F(kCos4)},
        w9{F(-kCos8), F(-kSin8)};
    y[1][1] = cmul(y[1][1], w1);
    y[1][2] = cmul(y[1][2], w2);
    y[1][3] = cmul(y[1][3], w3);
    y[2][1] = cmul(y[2][1], w2);
    y[2][2] = zmuli(y[2][2]);
    y[2][3] = cmul(y[2][3], w6);
    y[3][1] = cmul(y[3][1], w3);
    y[3][2] = cmul(y[3][2], w6);
    y[3][3] = cmul(y[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        C2<F> z0 = y[0][k1], z1 = y[1][k1], z2 = y[2][k1], z3 = y[3][k1];
        idft4(z0, z1, z2, z3);
        x[k1] = z0;
        x[k1 + 4] = z1;
        x[k1 + 8] = z2;
        x[k1 + 12] = z3;
    }
}

// NZ: the inputs from x[NZ] on are zero (the first pass's padding); NZ = R for
// the other passes.
template <int R, int NZ, typename F>
__device__ inline void idft(C2<F>* x) {
    static_assert(R == 4 || R == 8 || R == 16, "radices 4, 8, 16");
    if constexpr (R == 4) {
        idft4_nz<NZ>(x[0], x[1], x[2], x[3]);
    } else if constexpr (R == 8) {
        idft8<NZ>(x);
    } else {
        idft16<NZ>(x);
    }
}

// Orders the wave's LDS accesses across lanes (the workgroup is one wave).
__device__ inline void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t padded(uint32_t e) { return e + (e >> 4); }

// Stockham passes of an N-point transform on one wave.  Register slot m of a
// lane holds element lane + 64 m on entry to every pass.  A pass of radix R
// and span NS has B = N / (64 R) butterflies per lane: butterfly j = lane + 64 q
// takes elements j + r N / R (slots q + B r), multiplies element r by
// w^r, w = exp(2 pi i (j mod NS) / (NS R)), transforms, and its output r goes
// to (j / NS) NS R + j mod NS + r NS.  NZ: inputs r >= NZ are zero (first pass:
// the lane's channel slots end at NZ; NZ = R otherwise).  PHI: every output of
// the lane is also multiplied by phi (folded into the input twiddles: element
// r by phi w^r).
template <int N, int R, int NS, int NZ, bool PHI, typename F>
__device__ inline void pass_compute(C2<F>* v, const C2<F>* base, C2<F> phi) {
    constexpr int B = N / (kWave * R);
#pragma unroll
    for (int q = 0; q < B; ++q) {
        C2<F> x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[q + B * r];
        if constexpr (PHI) {
            C2<F> t = phi;
            x[0] = cmul(x[0], t);
#pragma unroll
            for (int r = 1; r < R; ++r) {
                t = cmul(t, base[q]);
                x[r] = cmul(x[r], t);
            }
        } else if constexpr (NS > 1) {
            // Odd and even powers as two interleaved chains (step w^2): half
            // the dependent depth of one chain, three values live.
            const C2<F> w1 = base[q];
            const C2<F> w2 = cmul(w1, w1);
            C2<F> po = w1, pe = w2;
            x[1] = cmul(x[1], po);
            x[2] = cmul(x[2], pe);
#pragma unroll
            for (int r = 3; r + 1 < R; r += 2) {
                po = cmul(po, w2);
                pe = cmul(pe, w2);
                x[r] = cmul(x[r], po);
                x[r + 1] = cmul(x[r + 1], pe);
            }
            if constexpr (R % 2 == 0) {
                po = cmul(po, w2);
                x[R - 1] = cmul(x[R - 1], po);
            }
        }
        idft<R, NZ>(x);
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + B * r] = x[r];
    }
}

// For every pass here (o0 mod 16) + (r NS mod 16) < 16, so
// padded(o0 + r NS) = padded(o0) + r NS + (r NS) / 16: one address per lane
// plus compile-time offsets.
template <int N, int R, int NS, typename F>
__device__ inline void pass_store_padded(C2<F>* lds, const C2<F>* v, uint32_t lane) {
    constexpr int B = N / (kWave * R);
    static_assert(NS >= 16 || (NS * R <= 16 && 16 % (NS * R) == 0) || (NS == 8 && R == 8) || (NS == 4 && R == 4),
                  "padded store offsets");
#pragma unroll
    for (int q = 0; q < B; ++q) {
        const uint32_t j = lane + kWave * q;
        const uint32_t o0 = (j / NS) * NS * R + j % NS;
        C2<F>* base = lds + padded(o0);
#pragma unroll
        for (int r = 0; r < R; ++r) base[r * NS + (r * NS) / 16] = v[q + B * r];
    }
}

// Element lane + 64 m sits at padded(lane) + 68 m.
template <int N, typename F>
__device__ inline void load_padded(const C2<F>* lds, C2<F>* v, uint32_t lane) {
    const C2<F>* base = lds + padded(lane);
#pragma unroll
    for (int m = 0; m < N / kWave; ++m) v[m] = base[(kWave + kWave / 16) * m];
}

// Last pass (NS R = N): output r of butterfly q is element lane + 64 (q + B r),
// i.e. slot m goes to lane + 64 m; the first W - 1 elements are repeated after N.
template <int N, int W, typename F>
__device__ inline void store_grid(C2<F>* lds, const C2<F>* v, uint32_t lane) {
    C2<F>* base = lds + lane;
#pragma unroll
    for (int m = 0; m < N / kWave; ++m) base[kWave * m] = v[m];
    if (lane < uint32_t(W - 1)) base[N] = v[0];
}

// Hides a value from loop-invariant code motion, so powers derived from it are
// recomputed where they are used instead of being kept in registers across
// the row loop.
template <typename F>
__device__ inline void opaque(C2<F>& z) {
    asm volatile("" : "+v"(z.x), "+v"(z.y));
}

template <typename F>
__device__ inline C2<F> phasor_as(double theta) {
    const Phasor64 a = phasor64(theta);
    return {F(a.c), F(a.s)};
}

// Per-lane base twiddles of the passes after the first (row-independent; the
// powers each pass needs are recomputed per row from these, see opaque()).
template <int N, typename F>
struct Twiddles;

// phi = exp(-2 pi i H (lane mod V) / V), V = N / 64: the frequency offset
// h = 64 H applied to the lane's outputs (see the file comment).
template <int N, typename F>
__device__ inline C2<F> offset_phase(uint32_t lane, int32_t H) {
    constexpr int V = N / kWave;
    return phasor_as<F>(-kTwoPi * double((uint32_t(H) * (lane % V)) % V) / double(V));
}

template <typename F>
struct Twiddles<1024, F> {  // passes (16, 1), (16, 16), (4, 256)
    C2<F> p2[1], p3[4];
    __device__ void hide() {
        opaque(p2[0]);
#pragma unroll
        for (int q = 0; q < 4; ++q) opaque(p3[q]);
    }
    __device__ void init(uint32_t lane) {
        p2[0] = phasor_as<F>(kTwoPi * double(lane % 16) / 256.0);
#pragma unroll
        for (int q = 0; q < 4; ++q) p3[q] = phasor_as<F>(kTwoPi * double(lane + kWave * q) / 1024.0);
    }
};

template <typename F>
struct Twiddles<512, F> {  // passes (8, 1), (8, 8), (8, 64)
    C2<F> p2[1], p3[1];
    __device__ void hide() {
        opaque(p2[0]);
        opaque(p3[0]);
    }
    __device__ void init(uint32_t lane) {
        p2[0] = phasor_as<F>(kTwoPi * double(lane % 8) / 64.0);
        p3[0] = phasor_as<F>(kTwoPi * double(lane) / 512.0);
    }
};

template <typename F>
struct Twiddles<256, F> {  // passes (4, 1), (4, 4), (4, 16), (4, 64)
    C2<F> p2[1], p3[1], p4[1];
    __device__ void hide() {
        opaque(p2[0]);
        opaque(p3[0]);
        opaque(p4[0]);
    }
    __device__ void init(uint32_t lane) {
        p2[0] = phasor_as<F>(kTwoPi * double(lane % 4) / 16.0);
        p3[0] = phasor_as<F>(kTwoPi * double(lane % 16) / 64.0);
        p4[0] = phasor_as<F>(kTwoPi * double(lane) / 256.0);
    }
};

// Swaps 32-bit words of a and b across 16-lane rows: ROWS = 32 exchanges the
// upper 32 lanes of a with the lower 32 lanes of b (v_permlane32_swap), ROWS =
// 16 the odd rows of a with the even rows of b (v_permlane16_swap).
template <int ROWS, typename F>
__device__ inline void swap_rows(C2<F>& a, C2<F>& b) {
    constexpr int kWords = int(sizeof(C2<F>) / 4);
    uint32_t wa[kWords], wb[kWords];
    __builtin_memcpy(wa, &a, sizeof(a));
    __builtin_memcpy(wb, &b, sizeof(b));
#pragma unroll
    for (int i = 0; i < kWords; ++i) {
        if constexpr (ROWS == 32) {
            const auto r = __builtin_amdgcn_permlane32_swap(wa[i], wb[i], false, false);
            wa[i] = r[0];
            wb[i] = r[1];
        } else {
            static_assert(ROWS == 16, "16- or 32-lane swaps");
            const auto r = __builtin_amdgcn_permlane16_swap(wa[i], wb[i], false, false);
            wa[i] = r[0];
            wb[i] = r[1];
        }
    }
    __builtin_memcpy(&a, wa, sizeof(a));
    __builtin_memcpy(&b, wb, sizeof(b));
}

// The exchange between the second and third pass of the 1024-point transform,
// in registers.  After pass 2 lane 16 a + b holds output 4 c + d of its
// butterfly, element 256 a + 64 c + 16 d + b, in slot 4 c + d; pass 3 wants
// element L + 64 m in slot m of lane L, i.e. that value in lane 16 d + b, slot
// 4 a + c.  For each c this is a 4 x 4 transpose between the 16-lane row a and
// the slot index d: a 32-lane swap exchanges bit 1 of the two, a 16-lane swap
// bit 0.  No LDS traffic (it replaces a 16 KB store and load per row).
template <typename F>
__device__ inline void transpose_rows_1024(C2<F>* v) {
    C2<F> out[16];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        C2<F> t[4] = {v[4 * c], v[4 * c + 1], v[4 * c + 2], v[4 * c + 3]};
        swap_rows<32>(t[0], t[2]);
        swap_rows<32>(t[1], t[3]);
        swap_rows<16>(t[0], t[1]);
        swap_rows<16>(t[2], t[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k) out[4 * k + c] = t[k];
    }
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = out[m];
}

// g_m = sum_k f~_k exp(2 pi i k m / N) into lds[0, N + W - 1), from the lane's
// slots (slot i < CPL <= N / 128: channel lane + 64 i; the rest zero) and the
// lane's offset phase phi.
template <int N, int W, int CPL, typename F>
__device__ inline void row_fft(C2<F>* v, C2<F>* lds, const Twiddles<N, F>& tw, C2<F> phi, uint32_t lane) {
    static_assert(CPL >= 1 && 2 * kWave * CPL <= N, "channel slots within the lower half");
    if constexpr (N == 1024) {
        pass_compute<N, 16, 1, CPL, false>(v, (const C2<F>*)nullptr, phi);
        wave_fence();  // the previous row's interpolation has read the grid
        pass_store_padded<N, 16, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 16, 16, 16, false>(v, tw.p2, phi);
        transpose_rows_1024(v);
        pass_compute<N, 4, 256, 4, true>(v, tw.p3, phi);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    } else if constexpr (N == 512) {
        pass_compute<N, 8, 1, CPL, false>(v, (const C2<F>*)nullptr, phi);
        wave_fence();
        pass_store_padded<N, 8, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 8, 8, 8, false>(v, tw.p2, phi);
        wave_fence();
        pass_store_padded<N, 8, 8>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 8, 64, 8, true>(v, tw.p3, phi);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    } else {
        static_assert(N == 256, "grid sizes 256, 512, 1024");
        pass_compute<N, 4, 1, CPL, false>(v, (const C2<F>*)nullptr, phi);
        wave_fence();
        pass_store_padded<N, 4, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 4, 4, false>(v, tw.p2, phi);
        wave_fence();
        pass_store_padded<N, 4, 4>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 16, 4, false>(v, tw.p3, phi);
        wave_fence();
        pass_store_padded<N, 4, 16>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 64, 4, true>(v, tw.p4, phi);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    }
}

// A precorrected value in the transform's type: the FP64 product, rounded once
// for an FP32 transform (packed rows hold the same FP64 product, so both
// sources give the same bits).
template <typename F>
__device__ inline C2<F> to_transform(double re, double im) {
    if constexpr (std::is_same_v<F, double>)
        return {re, im};
    else
        return {float(re), float(im)};
}

// (dd . uvw).z - w in metres, the path difference whose -2 pi multiple is the
// reference's shift factor (CpuLeapCalibrator.cc:190).
__device__ inline double path_difference(const double dz[3], double u, double v, double w) {
    double t = dz[0] * u;
    t = fma(dz[1], v, t);
    t = fma(dz[2], w, t);
    return t - w;
}

// exp(2 pi i tr) for a phase tr in turns: k = rint(64 tr), r = 2 pi (tr - k / 64)
// with |r| <= pi / 64 (tr - k / 64 is exact for |tr| < 2^46), the table entry
// exp(2 pi i (k mod 64) / 64) times the series of exp(i r): cos to r^8 and sin
// to r^7, whose first omitted terms are below 2.5e-20 and 4.6e-18.
constexpr int kTurnTable = 64;

// One entry per lane; a wave fence orders it before the first read.
__device__ inline void turn_table_init(Z* table, uint32_t lane) {
    if (lane < uint32_t(kTurnTable)) {
        const Phasor64 p = phasor64(kTwoPi * double(lane) / double(kTurnTable));
        table[lane] = {p.c, p.s};
    }
}

// exp(2 pi i dl[j] scale) for the NDL directions of a lane, their steps
// interleaved; the table reads are issued before the series that hides them.
template <int NDL>
__device__ inline void turn_phasor_n(const Z* table, const double* dl, double scale, Z* out) {
    double tr[NDL], kf[NDL], r[NDL], r2[NDL], c[NDL], sn[NDL];
    Z e[NDL];
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        tr[j] = dl[j] * scale;
        kf[j] = rint(tr[j] * double(kTurnTable));
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) e[j] = table[int32_t(kf[j]) & (kTurnTable - 1)];
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        r[j] = fma(kf[j], -1.0 / kTurnTable, tr[j]) * kTwoPi;
        r2[j] = r[j] * r[j];
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        c[j] = fma(1.0 / 40320.0, r2[j], -1.0 / 720.0);
        sn[j] = fma(-1.0 / 5040.0, r2[j], 1.0 / 120.0);
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        c[j] = fma(c[j], r2[j], 1.0 / 24.0);
        sn[j] = fma(sn[j], r2[j], -1.0 / 6.0);
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        c[j] = fma(c[j], r2[j], -0.5);
        sn[j] = fma(r[j] * r2[j], sn[j], r[j]);
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) c[j] = fma(c[j], r2[j], 1.0);
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        out[j].x = fma(e[j].x, c[j], -(e[j].y * sn[j]));
        out[j].y = fma(e[j].x, sn[j], e[j].y * c[j]);
    }
}

// Row sources.  The lane owns channels c = lane + 64 i, i < CPL (C <= 64 CPL
// <= N / 2); load() fetches them for one row (channels past C read channel
// C - 1 and are zeroed by freq()), freq() returns the precorrected pol-0
// values f~.
//   P == 0: packed rows of FP64 complex, already precorrected (pack kernel).
//   P >= 1: the complex64 cube [row][channel][pol]: a channel whose 2P floats
//           are not all finite is zero (CpuLeapCalibrator.cc:201-214), pol 0 is
//           widened and multiplied by the channel's precorrection factor.
//   P = 16 + Q (Q = 1, 2, 4): the complex128 cube with Q pols, the same rule.
// FULL (complex128 only): every pol loaded whole (16-byte loads) instead of
// the upper words of pols 1 ..
template <int N, int P, int CPL, bool WIDE = (P > 16), bool FULL = false>
struct Source {
    float raw[CPL][2 * P];
    double sc[CPL];    // precorrection, 0 past C
    uint32_t off[CPL];  // float offset of the channel (clamped) in a row
    __device__ void init(const LEAP_GLOBAL double* scale, uint32_t lane, uint32_t C) {
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            const uint32_t c = lane + kWave * i;
            sc[i] = c < C ? scale[c] : 0.0;
            off[i] = min(c, C - 1) * (2 * P);
        }
    }
    // base advanced by `rows` rows of C channels
    __device__ static const void* advance(const void* base, size_t rows, uint32_t C) {
        return (const float*)base + rows * C * (2 * P);
    }
    __device__ void load(const void* base, size_t row, uint32_t, uint32_t C) {
        const LEAP_GLOBAL float* r = (const LEAP_GLOBAL float*)base + row * C * (2 * P);
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            const LEAP_GLOBAL float* q = r + off[i];
            if constexpr (P == 4 || P == 2) {
#pragma unroll
                for (int h = 0; h < P / 2; ++h) {
                    const f32x4 f = ((const LEAP_GLOBAL f32x4*)q)[h];
                    raw[i][4 * h] = f.x;
                    raw[i][4 * h + 1] = f.y;
                    raw[i][4 * h + 2] = f.z;
                    raw[i][4 * h + 3] = f.w;
                }
            } else {
                const f32x2 f = *(const LEAP_GLOBAL f32x2*)q;
                raw[i][0] = f.x;
                raw[i][1] = f.y;
            }
        }
    }
    template <typename F>
    __device__ void freq(C2<F>* x, uint32_t, uint32_t) const {
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            bool finite = true;
#pragma unroll
            for (int k = 0; k < 2 * P; ++k) finite = finite && __builtin_isfinite(raw[i][k]);
            const float re = finite ? raw[i][0] : 0.0f, im = finite ? raw[i][1] : 0.0f;
            x[i] = to_transform<F>(double(re) * sc[i], double(im) * sc[i]);
        }
    }
};

// The complex128 cube.  A lane's channels lie 16 Q bytes apart, so each load
// instruction of the wave touches a cache line per lane or two, and the
// vector-memory pipeline's work grows with the number of load instructions.
// FULL: every pol is loaded whole with one 16-byte load (Q loads and 4 Q
// registers per channel slot); otherwise pol 0 is loaded whole and the other
// pols only as the upper words of their real and imaginary parts (sign,
// exponent and the top of the mantissa: enough to tell whether the value is
// finite), 2 Q - 1 loads and 4 + 2 (Q - 1) registers per slot.  Each row's
// bytes cross HBM once either way (the words share the channel's lines).
template <int N, int SP, int CPL, bool FULL>
struct Source<N, SP, CPL, true, FULL> {
    static constexpr int Q = SP - 16;
    static_assert(Q == 1 || Q == 2 || Q == 4, "complex128 cube with 1, 2 or 4 pols");
    static constexpr int NH = Q > 1 && !FULL ? 2 * (Q - 1) : 1;
    static constexpr int NP = FULL ? Q : 1;
    double re[CPL][NP], im[CPL][NP];  // pol 0, then (FULL) pols 1 .. Q - 1
    uint32_t hi[CPL][NH];             // upper words of the real and imaginary parts of pols 1 .. Q - 1
    double sc[CPL];        // precorrection, 0 past C
    uint32_t off[CPL];     // double offset of the channel (clamped) in a row
    __device__ void init(const LEAP_GLOBAL double* scale, uint32_t lane, uint32_t C) {
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            const uint32_t c = lane + kWave * i;
            sc[i] = c < C ? scale[c] : 0.0;
            off[i] = min(c, C - 1) * (2 * Q);
        }
    }
    __device__ static const void* advance(const void* base, size_t rows, uint32_t C) {
        return (const double*)base + rows * C * (2 * Q);
    }
    __device__ void load(const void* base, size_t row, uint32_t, uint32_t C) {
        const LEAP_GLOBAL double* r = (const LEAP_GLOBAL double*)base + row * C * (2 * Q);
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            const LEAP_GLOBAL double* q = r + off[i];
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const f64x2 pk = ((const LEAP_GLOBAL f64x2*)q)[k];
                re[i][k] = pk.x;
                im[i][k] = pk.y;
            }
            if constexpr (Q > 1 && !FULL) {
                const LEAP_GLOBAL uint32_t* w = (const LEAP_GLOBAL uint32_t*)(q + 2);
#pragma unroll
                for (int k = 0; k < Q - 1; ++k) {
                    hi[i][2 * k] = w[4 * k + 1];
                    hi[i][2 * k + 1] = w[4 * k + 3];
                }
            }
        }
    }
    template <typename F>
    __device__ void freq(C2<F>* x, uint32_t, uint32_t) const {
        constexpr uint32_t kExp = 0x7ff00000u;  // an all-ones exponent: Inf or NaN
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            bool finite = true;
#pragma unroll
            for (int k = 0; k < NP; ++k) finite = finite && __builtin_isfinite(re[i][k]) && __builtin_isfinite(im[i][k]);
            if constexpr (Q > 1 && !FULL) {
                uint32_t e = 0;
#pragma unroll
                for (int k = 0; k < NH; ++k) e = max(e, hi[i][k] & kExp);
                finite = finite && e != kExp;
            }
            const double a = finite ? re[i][0] : 0.0, b = finite ? im[i][0] : 0.0;
            x[i] = to_transform<F>(a * sc[i], b * sc[i]);
        }
    }
};

template <int N, int CPL>
struct Source<N, 0, CPL, false, false> {
    Z raw[CPL];
    __device__ void init(const LEAP_GLOBAL double*, uint32_t, uint32_t) {}
    __device__ static const void* advance(const void* base, size_t rows, uint32_t C) {
        return (const Z*)base + rows * C;
    }
    __device__ void load(const void* base, size_t row, uint32_t lane, uint32_t C) {
        const LEAP_GLOBAL Z* r = (const LEAP_GLOBAL Z*)base + row * C;
#pragma unroll
        for (int i = 0; i < CPL; ++i) {
            const uint32_t c = min(lane + kWave * i, C - 1);
            raw[i].x = r[c].x;
            raw[i].y = r[c].y;
        }
    }
    template <typename F>
    __device__ void freq(C2<F>* x, uint32_t lane, uint32_t C) const {
#pragma unroll
        for (int i = 0; i < CPL; ++i)
            x[i] = lane + kWave * i < C ? to_transform<F>(raw[i].x, raw[i].y) : C2<F>{F(0), F(0)};
    }
};

// The rows [0, t_count) of baseline bl (rows t nbl + bl of the row source wr
// and of uvr) for the directions d0 + lane + 64 j, j < NDL, of D; the sums go
// to out + d dstride (mode kNufftSet: pol 0, pols 1 .. Pout - 1 zeroed;
// kNufftAccumulate: added; kNufftPartial: written).  s_g is the wave's grid,
// s_turn its phasor table (filled, and fenced, before the call).
// SP: the row source (0 packed rows, 1 / 2 / 4 a complex64 cube with that many pols).
// CPL: channel slots per lane (C <= 64 CPL <= N / 2).
// F: the type of the transform, the grid and the interpolation (double or float).
template <int N, int W, int NDL, int SP, int CPL, typename F>
__device__ __forceinline__ void nufft_rows(const void* wr, const LEAP_GLOBAL double* __restrict__ scale,
                                           const LEAP_GLOBAL double* __restrict__ uvr,
                                           const LEAP_GLOBAL double* __restrict__ dd, uint32_t D, uint32_t t_count,
                                           uint32_t nbl, uint32_t bl, uint32_t C, const NufftArgs& args, uint32_t d0,
                                           LEAP_GLOBAL double* __restrict__ out, uint32_t dstride, int32_t mode,
                                           uint32_t Pout, C2<F>* s_g, const Z* s_turn, uint32_t lane) {
#pragma clang fp contract(off)
    using ZF = C2<F>;
    constexpr int V = N / kWave;
    static_assert(W % 2 == 0 && W <= kNufftMaxWidth && N / 16 >= W - 1, "kernel width");
    double dz[NDL][3];
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        const uint32_t d = d0 + lane + kWave * j;
        const uint32_t dc = d < D ? d : D - 1;
        dz[j][0] = dd[9 * size_t(dc) + 6];
        dz[j][1] = dd[9 * size_t(dc) + 7];
        dz[j][2] = dd[9 * size_t(dc) + 8];
    }
    Twiddles<N, F> tw;
    tw.init(lane);
    Z sum[NDL];
#pragma unroll
    for (int j = 0; j < NDL; ++j) sum[j] = {0.0, 0.0};
    const ZF phi = offset_phase<N, F>(lane, args.h / kWave);
    const double neg_dkappa = -args.dkappa, neg_kappa_mid = -args.kappa_mid;
    const F* coef;
    if constexpr (std::is_same_v<F, double>)
        coef = args.coef;
    else
        coef = args.coef32;
    // The next row's loads are issued right after the current row is converted
    // and stay in flight while it is transformed and interpolated.  The load is
    // unconditional (the last row reloads itself), so the row buffer keeps its
    // registers from one iteration to the next.
    // complex128 cubes: whole-pol loads where the registers allow it at two
    // waves per SIMD (one or two directions per lane, the 1024-point grid, at
    // most 6 channel slots: 249 / 251 VGPRs with 4 pols)
    Source<N, SP, CPL, (SP > 16), (SP > 16 && NDL <= 2 && N == 1024 && CPL <= 6)> src;
    src.init(scale, lane, C);
    if (t_count > 0) src.load(wr, bl, lane, C);
    for (uint32_t t = 0; t < t_count; ++t) {
        ZF v[V];
        src.template freq<F>(v, lane, C);
#pragma unroll
        for (int m = CPL; m < V; ++m) v[m] = ZF{F(0), F(0)};  // padding (not read by the first pass)
        __builtin_amdgcn_sched_barrier(0);
        const size_t row = size_t(t) * nbl + bl;
        src.load(wr, size_t(t + 1 < t_count ? t + 1 : t) * nbl + bl, lane, C);
        const double u = uvr[3 * row], vv = uvr[3 * row + 1], ww = uvr[3 * row + 2];
        tw.hide();
        row_fft<N, W, CPL>(v, s_g, tw, phi, lane);
        // Interpolation.  Per direction: the path difference
        // dl = dd_z . uvw - w (the shift factor is sf = -2 pi dl), so b / 2 pi =
        // -dl dkappa turns; grid coordinate ug = N frac(b / 2 pi), first tap
        // m0 = ceil(ug - W/2), offset s = ug - m0 - (W - 1)/2 in (-1/2, 1/2]; tap
        // pairs (jp, W - 1 - jp) take weights P_jp(s), P_jp(-s),
        // P_jp(s) = E(s^2) + s O(s^2).  The pairs are the outer loop so that at
        // most two pairs' coefficients are live.  The NDL directions' steps are
        // interleaved (independent chains side by side) with each direction's
        // operations unchanged.  The grid coordinate is formed in FP64; s, the
        // weights and the tap sums in F.
        double dl[NDL], sv[NDL], x[NDL], m0f[NDL];
        F svf[NDL], s2[NDL];
        const ZF* gp[NDL];
#pragma unroll
        for (int j = 0; j < NDL; ++j) dl[j] = path_difference(dz[j], u, vv, ww);
#pragma unroll
        for (int j = 0; j < NDL; ++j) x[j] = dl[j] * neg_dkappa;  // turns of b
#pragma unroll
        for (int j = 0; j < NDL; ++j) {
            x[j] = __builtin_amdgcn_fract(x[j]) * double(N);  // grid coordinate ug in [0, N)
            m0f[j] = ceil(x[j] - 0.5 * W);
        }
#pragma unroll
        for (int j = 0; j < NDL; ++j) {
            sv[j] = (x[j] - m0f[j]) - 0.5 * (W - 1);
            svf[j] = F(sv[j]);
            s2[j] = svf[j] * svf[j];
            // m0 in [-W/2, N - W/2]; a negative m0 wraps to m0 + N, whose taps
            // end at N + W/2 - 1 <= N + W - 2 (the wrapped copy).  The mask also
            // keeps the taps of a non-finite uvw inside the grid.
            const int32_t m0 = int32_t(m0f[j]) & (N - 1);
            gp[j] = s_g + m0;
        }
        ZF acc[NDL];
#pragma unroll
        for (int j = 0; j < NDL; ++j) acc[j] = {F(0), F(0)};
        // A pair's coefficients are read with scalar loads one pair ahead and
        // not kept across the loop (all W/2 pairs do not fit in SGPRs).  Scalar
        // and LDS loads share one wait counter and scalar loads return out of
        // order, so every wait is for all of them: the pair's 2 NDL grid reads
        // are issued together with the next pair's coefficients, ahead of the
        // weight polynomials that hide their latency, and waited for once.
        F cc[W];
        {
            const F* cp = coef;
            asm volatile("" : "+s"(cp));
            const AS4<F>* cf = (const AS4<F>*)cp;
#pragma unroll
            for (int k = 0; k < W; ++k) cc[k] = cf[k];
        }
#pragma unroll
        for (int jp = 0; jp < W / 2; ++jp) {
            ZF ga[NDL], gb[NDL];
#pragma unroll
            for (int j = 0; j < NDL; ++j) {
                ga[j] = gp[j][jp];
                gb[j] = gp[j][W - 1 - jp];
            }
            F cn[W];
            if (jp + 1 < W / 2) {
                const F* cp = coef + (jp + 1) * kNufftMaxWidth;
                asm volatile("" : "+s"(cp));
                const AS4<F>* cf = (const AS4<F>*)cp;
#pragma unroll
                for (int k = 0; k < W; ++k) cn[k] = cf[k];
            }
            __builtin_amdgcn_sched_barrier(0);
            F wl[NDL], wh[NDL], e[NDL], o[NDL];
#pragma unroll
            for (int j = 0; j < NDL; ++j) {
                e[j] = cc[W - 2];
                o[j] = cc[W - 1];
            }
#pragma unroll
            for (int k = W - 4; k >= 0; k -= 2) {
#pragma unroll
                for (int j = 0; j < NDL; ++j) {
                    e[j] = fma(e[j], s2[j], cc[k]);
                    o[j] = fma(o[j], s2[j], cc[k + 1]);
                }
            }
#pragma unroll
            for (int j = 0; j < NDL; ++j) {
                wl[j] = fma(svf[j], o[j], e[j]);
                wh[j] = fma(-svf[j], o[j], e[j]);
            }
#pragma unroll
            for (int j = 0; j < NDL; ++j) {
                acc[j].x = fma(ga[j].x, wl[j], acc[j].x);
                acc[j].y = fma(ga[j].y, wl[j], acc[j].y);
                acc[j].x = fma(gb[j].x, wh[j], acc[j].x);
                acc[j].y = fma(gb[j].y, wh[j], acc[j].y);
            }
            if (jp + 1 < W / 2) {
#pragma unroll
                for (int k = 0; k < W; ++k) cc[k] = cn[k];
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        // exp(i sf kappa_mid) = exp(2 pi i tr), tr = -dl kappa_mid turns, from
        // the turn table and a short series; added to the row sums with FMAs.
        Z ph[NDL];
        turn_phasor_n<NDL>(s_turn, dl, neg_kappa_mid, ph);
#pragma unroll
        for (int j = 0; j < NDL; ++j) {
            const double ax = double(acc[j].x), ay = double(acc[j].y);
            sum[j].x = fma(ph[j].x, ax, sum[j].x);
            sum[j].x = fma(-ph[j].y, ay, sum[j].x);
            sum[j].y = fma(ph[j].x, ay, sum[j].y);
            sum[j].y = fma(ph[j].y, ax, sum[j].y);
        }
    }
    // The output indices are recomputed here from an opaque copy of the lane,
    // so they are not kept in registers across the row loop.
    uint32_t ln = lane;
    asm volatile("" : "+v"(ln));
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        const uint32_t d = d0 + ln + kWave * j;
        if (d >= D) continue;
        LEAP_GLOBAL double* o = out + size_t(d) * dstride;
        if (mode == kNufftPartial) {
            o[0] = sum[j].x;
            o[1] = sum[j].y;
        } else if (mode == kNufftAccumulate) {
            o[0] += sum[j].x;
            o[1] += sum[j].y;
        } else {
            o[0] = sum[j].x;
            o[1] = sum[j].y;
            for (uint32_t q = 2; q < 2 * Pout; ++q) o[q] = 0.0;  // pols not accumulated
        }
    }
}

// One launch over a set's baselines (blockIdx.x, resolved through the span)
// and direction blocks of 64 NDL (blockIdx.y).
template <int N, int W, int NDL, int SP, int CPL, typename F>
__global__ __launch_bounds__(kWave) void rotate_average_nufft(const void* __restrict__ w,
                                                              const LEAP_GLOBAL double* __restrict__ scale,
                                                              const LEAP_GLOBAL double* __restrict__ uvw,
                                                              const LEAP_GLOBAL double* __restrict__ dd, uint32_t D,
                                                              uint32_t T, uint32_t nbl, uint32_t C, uint32_t Pout,
                                                              NufftArgs args, LEAP_GLOBAL double* __restrict__ avg,
                                                              int32_t mode, NufftSpan span) {
    __shared__ C2<F> s_g[N + N / 16];
    __shared__ Z s_turn[kTurnTable];
    const uint32_t lane = threadIdx.x;
    turn_table_init(s_turn, lane);
    wave_fence();
    // The span is resolved here into the row source, uvw and output pointers and
    // a row count, so the row loop holds no more values than an unchunked one.
    const uint32_t item = blockIdx.x;
    const uint32_t chunk = span.chunks > 1 ? item % span.chunks : 0;
    const uint32_t bl = span.bl_base + (span.chunks > 1 ? item / span.chunks : item);
    const uint32_t t_begin = uint32_t(uint64_t(chunk) * T / max(span.chunks, 1u));
    const uint32_t t_count = uint32_t(uint64_t(chunk + 1) * T / max(span.chunks, 1u)) - t_begin;
    const void* wr = Source<N, SP, CPL>::advance(w, size_t(t_begin) * nbl, C);
    const LEAP_GLOBAL double* uvr = uvw + size_t(t_begin) * nbl * 3;
    LEAP_GLOBAL double* out = mode == kNufftPartial ? avg + size_t(item) * D * 2 : avg + size_t(bl) * Pout * 2;
    const uint32_t dstride = mode == kNufftPartial ? 2u : nbl * Pout * 2;  // doubles between directions
    nufft_rows<N, W, NDL, SP, CPL, F>(wr, scale, uvr, dd, D, t_count, nbl, bl, C, args, blockIdx.y * (kWave * NDL),
                                      out, dstride, mode, Pout, s_g, s_turn, lane);
}

// avgData of the chunked baselines: the chunks' partial sums added in chunk
// order (each chunk summed its rows in row order).
static __global__ void nufft_tail_sum(const double* __restrict__ partial, uint32_t rem, uint32_t chunks, uint32_t D,
                               uint32_t nbl, uint32_t bl_base, uint32_t Pout, double* __restrict__ avg) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(rem) * D) return;
    const uint32_t r = uint32_t(i / D), d = uint32_t(i % D);
    double x = 0.0, y = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) {
        const double* p = partial + ((size_t(r) * chunks + c) * D + d) * 2;
        x += p[0];
        y += p[1];
    }
    double* o = avg + (size_t(d) * nbl + bl_base + r) * Pout * 2;
    o[0] = x;
    o[1] = y;
    for (uint32_t q = 2; q < 2 * Pout; ++q) o[q] = 0.0;
}

// Workgroups of one NUFFT kernel resident on the whole device at once.
template <class K>
uint32_t resident_workgroups(K kernel) {
    int per_cu = 0, dev = 0, cus = 0;
    LEAP_HIP(hipGetDevice(&dev));
    LEAP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    LEAP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave, 0));
    return uint32_t(std::max(1, per_cu) * std::max(1, cus));
}

// One launch of rotate_average_nufft<N, W, NDL, SP, CPL, F> over nbl_grid baselines.
// When a single direction block covers D and the baselines do not fill whole
// rounds of resident workgroups, the whole rounds run as usual and the
// remaining baselines' rows are split into chunks so that the last round is
// full (and short): their partial sums go to the tail scratch and are added by
// nufft_tail_sum.  Not for accumulating launches.
template <int N, int W, int NDL, int SP, int CPL, typename F>
void launch_nufft_tile(hipStream_t s, uint32_t nbl_grid, const void* w, const LEAP_GLOBAL double* gs,
                       const LEAP_GLOBAL double* gu, const LEAP_GLOBAL double* gd, uint32_t D, uint32_t T, uint32_t nbl,
                       uint32_t C, uint32_t Pout, const NufftArgs& a, LEAP_GLOBAL double* ga, int32_t accumulate,
                       double* tail, size_t tail_doubles) {
    auto* kernel = rotate_average_nufft<N, W, NDL, SP, CPL, F>;
    const uint32_t ny = (D + kWave * NDL - 1) / (kWave * NDL);
    if (tail && !accumulate && ny == 1 && T >= 2) {
        static const uint32_t slots = resident_workgroups(kernel);
        const uint32_t full = nbl_grid / slots * slots, rem = nbl_grid - full;
        const uint32_t chunks = rem ? std::min(T, (slots + rem - 1) / rem) : 0;
        if (rem && chunks >= 2 && size_t(rem) * chunks * D * 2 <= tail_doubles) {
            // Each launch is checked on its own: a later successful launch may
            // reset the error state of an earlier failed one.
            if (full) {
                kernel<<<dim3(full, 1), kWave, 0, s>>>(w, gs, gu, gd, D, T, nbl, C, Pout, a, ga, kNufftSet,
                                                       NufftSpan{0, 1});
                LEAP_HIP(hipGetLastError());
            }
            kernel<<<dim3(rem * chunks, 1), kWave, 0, s>>>(w, gs, gu, gd, D, T, nbl, C, Pout, a,
                                                           (LEAP_GLOBAL double*)tail, kNufftPartial,
                                                           NufftSpan{full, chunks});
            LEAP_HIP(hipGetLastError());
            const size_t n = size_t(rem) * D;
            nufft_tail_sum<<<dim3(uint32_t((n + 255) / 256)), 256, 0, s>>>(tail, rem, chunks, D, nbl, full, Pout,
                                                                        (double*)ga);
            return;
        }
    }
    kernel<<<dim3(nbl_grid, ny), kWave, 0, s>>>(w, gs, gu, gd, D, T, nbl, C, Pout, a, ga,
                                               accumulate ? kNufftAccumulate : kNufftSet, NufftSpan{0, 1});
}

template <int V>
using IC = std::integral_constant<int, V>;

// Channel slots per lane: N / 128 in general; for the 1024-point grid with
// kernel width 10 and an FP64 transform (257..384 channels: 5 or 6 slots), a
// kernel with 6 slots, which skips the loads, checks and first-pass butterflies
// of the two slots that are padding.
template <int N, int W, typename F>
constexpr int nufft_slots(uint32_t C) {
    if constexpr (N == 1024 && (W == 10 || W == 16) && std::is_same_v<F, double>) return C <= 6 * kWave ? 6 : 8;
    return N / (2 * kWave);
}

// Calls fn(IC<NDL>, IC<SP>, IC<CPL>) for a runtime tile (64 NDL directions per
// workgroup: 641, 1282, 2564), row source (src_pols: 0 packed rows, 1 / 2 / 4
// complex64 pols, 17 / 18 / 20 complex128 pols, the last for the width-16
// kernel only) and channel count.
template <int N, int W, typename F, class Fn>
void dispatch_nufft(int32_t tile, int32_t src_pols, uint32_t C, Fn&& fn) {
    auto with_ndl = [&](auto sp, auto cpl) {
        switch (tile) {
            case 641: fn(IC<1>{}, sp, cpl); break;
            case 1282: fn(IC<2>{}, sp, cpl); break;
            case 2564: fn(IC<4>{}, sp, cpl); break;
            default: throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported NUFFT tile " + std::to_string(tile));
        }
    };
    auto with_cpl = [&](auto sp) {
        if constexpr (N == 1024 && (W == 10 || W == 16) && std::is_same_v<F, double>) {
            if (nufft_slots<N, W, F>(C) == 6) {
                with_ndl(sp, IC<6>{});
                return;
            }
        }
        with_ndl(sp, IC<N / (2 * kWave)>{});
    };
    switch (src_pols) {
        case 0: with_cpl(IC<0>{}); break;
        case 1: with_cpl(IC<1>{}); break;
        case 2: with_cpl(IC<2>{}); break;
        case 4: with_cpl(IC<4>{}); break;
        case 17:
        case 18:
        case 20:
            if constexpr (W == 16) {
                if (src_pols == 17)
                    with_cpl(IC<17>{});
                else if (src_pols == 18)
                    with_cpl(IC<18>{});
                else
                    with_cpl(IC<20>{});
                break;
            }
            [[fallthrough]];
        default: throw Error(LEAP_ERR_INTERNAL, "unsupported NUFFT row source");
    }
}

template <int N, int W, typename F>
void launch_nw(hipStream_t s, int32_t tile, uint32_t nbl_grid, const void* w, int32_t src_pols, const double* scale,
               const double* uvw, const double* dd, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t Pout,
               const NufftArgs& a, double* avg, int32_t accumulate, double* tail, size_t tail_doubles) {
    dispatch_nufft<N, W, F>(tile, src_pols, C, [&](auto ndl, auto sp, auto cpl) {
        launch_nufft_tile<N, W, decltype(ndl)::value, decltype(sp)::value, decltype(cpl)::value, F>(
            s, nbl_grid, w, (const LEAP_GLOBAL double*)scale, (const LEAP_GLOBAL double*)uvw,
            (const LEAP_GLOBAL double*)dd, D, T, nbl, C, Pout, a, (LEAP_GLOBAL double*)avg, accumulate, tail,
            tail_doubles);
    });
}

// Several sets in one launch: workgroup x takes item x = (set | direction block
// << 16, baseline) and sums all T rows of that baseline for the block's
// directions into pol 0 of the set's avgData (the items come in decreasing
// cost, so the dispatcher fills the GPU largest first).  Every set of a launch
// has the same grid, kernel width, channel slots and complex64 cube pols; each
// (baseline, directions) sum is the code and the order of the single-set
// kernel.
template <int N, int W, int NDL, int SP, int CPL, typename F>
__global__ __launch_bounds__(kWave) void rotate_average_nufft_sets(const LEAP_GLOBAL NufftSetDesc* __restrict__ sets,
                                                                   const LEAP_GLOBAL uint32_t* __restrict__ items) {
    __shared__ C2<F> s_g[N + N / 16];
    __shared__ Z s_turn[kTurnTable];
    const uint32_t lane = threadIdx.x;
    turn_table_init(s_turn, lane);
    wave_fence();
    const uint32_t key = items[2 * blockIdx.x], bl = items[2 * blockIdx.x + 1];
    const LEAP_GLOBAL NufftSetDesc& sd = sets[key & 0xffffu];
    const uint32_t nbl = sd.nbl;
    const NufftArgs a{sd.coef, sd.coef32, sd.kappa_mid, sd.dkappa, sd.h};
    nufft_rows<N, W, NDL, SP, CPL, F>(sd.vis, (const LEAP_GLOBAL double*)sd.scale, (const LEAP_GLOBAL double*)sd.uvw,
                                      (const LEAP_GLOBAL double*)sd.dd, sd.D, sd.T, nbl, bl, sd.C, a,
                                      (key >> 16) * (kWave * NDL), (LEAP_GLOBAL double*)sd.avg + size_t(bl) * 2,
                                      nbl * 2, kNufftSet, 1, s_g, s_turn, lane);
}

template <int N, int W, typename F>
void launch_nw_sets(hipStream_t s, int32_t tile, int32_t src_pols, uint32_t C, const NufftSetDesc* sets,
                    const uint32_t* items, uint32_t n_items) {
    dispatch_nufft<N, W, F>(tile, src_pols, C, [&](auto ndl, auto sp, auto cpl) {
        if constexpr (decltype(sp)::value == 0) {
            throw Error(LEAP_ERR_INTERNAL, "the multi-set NUFFT kernel reads complex64 cubes");
        } else {
            rotate_average_nufft_sets<N, W, decltype(ndl)::value, decltype(sp)::value, decltype(cpl)::value, F>
                <<<dim3(n_items), kWave, 0, s>>>((const LEAP_GLOBAL NufftSetDesc*)sets,
                                                 (const LEAP_GLOBAL uint32_t*)items);
        }
    });
}

// Launch arguments shared by both transform precisions.
struct NufftLaunch {
    int32_t tile;
    const NufftPlan* plan;
    const void* src;
    int32_t src_pols;  // 0: packed rows; 1 / 2 / 4: complex64 cube; 17 / 18 / 20: complex128 cube
    const double* scale;
    const double* uvw;
    const double* dd;
    uint32_t D, T, nbl, nbl_grid, C, P_out;
    NufftArgs args;
    double* avg;
    int32_t accumulate;
    double* tail;
    size_t tail_doubles;
};

}  // namespace nufft

// The instantiations of one transform precision (kernels_nufft.hip: double,
// kernels_nufft32.hip: float).
void launch_rotate_nufft_f64(hipStream_t s, const nufft::NufftLaunch& l);
void launch_rotate_nufft_f32(hipStream_t s, const nufft::NufftLaunch& l);
// The width-16 kernel (kernels_nufft_w16.hip) and its multi-set form
// (kernels_nufft_w16_sets.hip; src: the cube's pols, plus 16 for complex128).
void launch_rotate_nufft_w16(hipStream_t s, const nufft::NufftLaunch& l);
void launch_rotate_nufft_sets_w16(hipStream_t s, int32_t tile, const NufftPlan& plan, uint32_t C, uint32_t src,
                                  const NufftSetDesc* sets, const uint32_t* items, uint32_t n_items);

}  // namespace leap
