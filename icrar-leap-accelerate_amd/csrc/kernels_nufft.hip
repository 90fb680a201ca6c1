// NUFFT channel sums of the mixed-precision batched path on gfx950.
//
// For row (t, baseline) and direction d the reference sums
//   sum_c vis[c][0] exp(i sf_d (f0 + c df) / c_light)     (CpuLeapCalibrator.cc:181-217)
// over the channels.  With b = sf_d df / c_light this is
// exp(i sf_d kappa_mid) * sum_{k=-h}^{C-1-h} v_{k+h} exp(i k b), h = C / 2,
// kappa_mid = (f0 + h df) / c_light, and the trigonometric polynomial is
// evaluated by a type-2 non-uniform FFT (the plan and its error are described
// in nufft_plan.cpp): the packed rows hold v_c already multiplied by the
// precorrection factor of channel c, one N-point inverse FFT per row puts the
// polynomial on a uniform grid of N points over one period, and each direction
// interpolates W grid points with the "exponential of semicircle" kernel.  Per
// row and direction this costs about 130 FP64 operations instead of the 4 C
// FMAs of the Horner kernel; the FFT's cost is shared by every direction of
// the row.
//
// Work decomposition: a workgroup is one wave.  It owns one baseline and a
// block of 64 * NDL directions (NDL per lane) and walks the baseline's rows in
// order, prefetching the next row's frequencies while it works on the current
// one.  The FFT is a Stockham transform whose butterflies run in registers
// (radices 16, 16, 4 for N = 1024; 8, 8, 8 for 512; 4, 4, 4, 4 for 256), with
// the exchanges between passes through the wave's LDS (rows of 16 elements
// padded by one element against bank conflicts); the last pass writes the grid
// in natural order plus W - 1 wrapped elements, so a direction's W grid points
// are consecutive.  Wave-scope fences order the LDS exchanges; there are no
// workgroup barriers and no atomics, and each direction's row sums are added
// in row order, so results are reproducible.
#include "batched_common.h"

namespace leap {
namespace {

using namespace dev;
using Z = C2<double>;

constexpr int kWave = 64;

struct NufftArgs {
    double coef[kNufftMaxWidth / 2][kNufftMaxWidth];  // weight polynomial of taps j < W / 2, powers of s
    double kappa_mid;                                 // (f0 + h df) / c_light
    double dkappa;                                    // df / c_light
    int32_t h;                                        // frequency offset, a multiple of 64
};

__device__ inline Z zadd(Z a, Z b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline Z zsub(Z a, Z b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline Z zmuli(Z a) { return {-a.y, a.x}; }  // i * a

constexpr double kCos4 = 0.70710678118654752440;   // cos(pi / 4)
constexpr double kCos8 = 0.92387953251128675613;   // cos(pi / 8)
constexpr double kSin8 = 0.38268343236508977173;   // sin(pi / 8)

// Inverse DFTs in place, natural order: X_k = sum_j x_j exp(+2 pi i j k / R).
__device__ inline void idft4(Z& a0, Z& a1, Z& a2, Z& a3) {
    const Z t0 = zadd(a0, a2), t1 = zsub(a0, a2), t2 = zadd(a1, a3), t3 = zmuli(zsub(a1, a3));
    a0 = zadd(t0, t2);
    a2 = zsub(t0, t2);
    a1 = zadd(t1, t3);
    a3 = zsub(t1, t3);
}

template <int R>
__device__ inline void idft(Z* x);

template <>
__device__ inline void idft<4>(Z* x) {
    idft4(x[0], x[1], x[2], x[3]);
}

template <>
__device__ inline void idft<8>(Z* x) {
    // even / odd halves, then X_k = E_k + w8^k O_k, X_{k+4} = E_k - w8^k O_k
    Z e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
    idft4(e[0], e[1], e[2], e[3]);
    idft4(o[0], o[1], o[2], o[3]);
    o[1] = cmul(o[1], Z{kCos4, kCos4});
    o[2] = zmuli(o[2]);
    o[3] = cmul(o[3], Z{-kCos4, kCos4});
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        x[k] = zadd(e[k], o[k]);
        x[k + 4] = zsub(e[k], o[k]);
    }
}

template <>
__device__ inline void idft<16>(Z* x) {
    // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, twiddle w16^(n2 k1), DFT4 over n2
    Z y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        y[n2][0] = x[n2];
        y[n2][1] = x[4 + n2];
        y[n2][2] = x[8 + n2];
        y[n2][3] = x[12 + n2];
        idft4(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    const Z w1{kCos8, kSin8}, w2{kCos4, kCos4}, w3{kSin8, kCos8}, w6{-kCos4, kCos4}, w9{-kCos8, -kSin8};
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
        Z z0 = y[0][k1], z1 = y[1][k1], z2 = y[2][k1], z3 = y[3][k1];
        idft4(z0, z1, z2, z3);
        x[k1] = z0;
        x[k1 + 4] = z1;
        x[k1 + 8] = z2;
        x[k1 + 12] = z3;
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
// to (j / NS) NS R + j mod NS + r NS.
template <int N, int R, int NS>
__device__ inline void pass_compute(Z* v, const Z* base) {
    constexpr int B = N / (kWave * R);
#pragma unroll
    for (int q = 0; q < B; ++q) {
        Z x[R];
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = v[q + B * r];
        if constexpr (NS > 1) {
            Z p = base[q];
            x[1] = cmul(x[1], p);
#pragma unroll
            for (int r = 2; r < R; ++r) {
                p = cmul(p, base[q]);
                x[r] = cmul(x[r], p);
            }
        }
        idft<R>(x);
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + B * r] = x[r];
    }
}

// For every pass here (o0 mod 16) + (r NS mod 16) < 16, so
// padded(o0 + r NS) = padded(o0) + r NS + (r NS) / 16: one address per lane
// plus compile-time offsets.
template <int N, int R, int NS>
__device__ inline void pass_store_padded(Z* lds, const Z* v, uint32_t lane) {
    constexpr int B = N / (kWave * R);
    static_assert(NS >= 16 || (NS * R <= 16 && 16 % (NS * R) == 0) || (NS == 8 && R == 8) || (NS == 4 && R == 4),
                  "padded store offsets");
#pragma unroll
    for (int q = 0; q < B; ++q) {
        const uint32_t j = lane + kWave * q;
        const uint32_t o0 = (j / NS) * NS * R + j % NS;
        Z* base = lds + padded(o0);
#pragma unroll
        for (int r = 0; r < R; ++r) base[r * NS + (r * NS) / 16] = v[q + B * r];
    }
}

// Element lane + 64 m sits at padded(lane) + 68 m.
template <int N>
__device__ inline void load_padded(const Z* lds, Z* v, uint32_t lane) {
    const Z* base = lds + padded(lane);
#pragma unroll
    for (int m = 0; m < N / kWave; ++m) v[m] = base[(kWave + kWave / 16) * m];
}

// Last pass (NS R = N): output r of butterfly q is element lane + 64 (q + B r),
// i.e. slot m goes to lane + 64 m; the first W - 1 elements are repeated after N.
template <int N, int W>
__device__ inline void store_grid(Z* lds, const Z* v, uint32_t lane) {
    Z* base = lds + lane;
#pragma unroll
    for (int m = 0; m < N / kWave; ++m) base[kWave * m] = v[m];
    if (lane < uint32_t(W - 1)) base[N] = v[0];
}

// Hides a value from loop-invariant code motion, so powers derived from it are
// recomputed where they are used instead of being kept in registers across
// the row loop.
__device__ inline void opaque(Z& z) {
    asm volatile("" : "+v"(z.x), "+v"(z.y));
}

// Per-lane base twiddles of the passes after the first (row-independent; the
// powers each pass needs are recomputed per row from these, see opaque()).
template <int N>
struct Twiddles;

template <>
struct Twiddles<1024> {  // passes (16, 1), (16, 16), (4, 256)
    Z p2[1], p3[4];
    __device__ void hide() {
        opaque(p2[0]);
#pragma unroll
        for (int q = 0; q < 4; ++q) opaque(p3[q]);
    }
    __device__ void init(uint32_t lane) {
        const Phasor64 a = phasor64(kTwoPi * double(lane % 16) / 256.0);
        p2[0] = {a.c, a.s};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const Phasor64 b = phasor64(kTwoPi * double(lane + kWave * q) / 1024.0);
            p3[q] = {b.c, b.s};
        }
    }
};

template <>
struct Twiddles<512> {  // passes (8, 1), (8, 8), (8, 64)
    Z p2[1], p3[1];
    __device__ void hide() {
        opaque(p2[0]);
        opaque(p3[0]);
    }
    __device__ void init(uint32_t lane) {
        const Phasor64 a = phasor64(kTwoPi * double(lane % 8) / 64.0);
        const Phasor64 b = phasor64(kTwoPi * double(lane) / 512.0);
        p2[0] = {a.c, a.s};
        p3[0] = {b.c, b.s};
    }
};

template <>
struct Twiddles<256> {  // passes (4, 1), (4, 4), (4, 16), (4, 64)
    Z p2[1], p3[1], p4[1];
    __device__ void hide() {
        opaque(p2[0]);
        opaque(p3[0]);
        opaque(p4[0]);
    }
    __device__ void init(uint32_t lane) {
        const Phasor64 a = phasor64(kTwoPi * double(lane % 4) / 16.0);
        const Phasor64 b = phasor64(kTwoPi * double(lane % 16) / 64.0);
        const Phasor64 c = phasor64(kTwoPi * double(lane) / 256.0);
        p2[0] = {a.c, a.s};
        p3[0] = {b.c, b.s};
        p4[0] = {c.c, c.s};
    }
};

// g_m = sum_k f~_k exp(2 pi i k m / N) from the lane's slots (element lane + 64 m
// = frequency index mod N) into lds[0, N + W - 1).
template <int N, int W>
__device__ inline void row_fft(Z* v, Z* lds, const Twiddles<N>& tw, uint32_t lane) {
    if constexpr (N == 1024) {
        pass_compute<N, 16, 1>(v, nullptr);
        wave_fence();  // the previous row's interpolation has read the grid
        pass_store_padded<N, 16, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 16, 16>(v, tw.p2);
        wave_fence();
        pass_store_padded<N, 16, 16>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 256>(v, tw.p3);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    } else if constexpr (N == 512) {
        pass_compute<N, 8, 1>(v, nullptr);
        wave_fence();
        pass_store_padded<N, 8, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 8, 8>(v, tw.p2);
        wave_fence();
        pass_store_padded<N, 8, 8>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 8, 64>(v, tw.p3);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    } else {
        static_assert(N == 256, "grid sizes 256, 512, 1024");
        pass_compute<N, 4, 1>(v, nullptr);
        wave_fence();
        pass_store_padded<N, 4, 1>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 4>(v, tw.p2);
        wave_fence();
        pass_store_padded<N, 4, 4>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 16>(v, tw.p3);
        wave_fence();
        pass_store_padded<N, 4, 16>(lds, v, lane);
        wave_fence();
        load_padded<N>(lds, v, lane);
        pass_compute<N, 4, 64>(v, tw.p4);
        wave_fence();
        store_grid<N, W>(lds, v, lane);
        wave_fence();
    }
}

// Channels of one row owned by the lane: c = lane + 64 i, i < N / 128 (C <= N / 2),
// zero beyond C.
template <int N>
__device__ inline void load_channels(Z* x, const LEAP_GLOBAL Z* __restrict__ wr, uint32_t lane, uint32_t C) {
#pragma unroll
    for (int i = 0; i < N / (2 * kWave); ++i) {
        const uint32_t c = lane + kWave * i;
        Z z{0.0, 0.0};
        if (c < C) {
            z.x = wr[c].x;
            z.y = wr[c].y;
        }
        x[i] = z;
    }
}

// Slot m holds grid index n = lane + 64 m, frequency k = n (n < N / 2) or
// n - N, channel c = k + h = lane + 64 i with i = m + H or m + H - N / 64
// (h = 64 H): the lane's own channels, placed by a compile-time rotation.
template <int N, int H>
__device__ inline void place_channels_h(Z* v, const Z* x) {
    constexpr int V = N / kWave;
#pragma unroll
    for (int m = 0; m < V; ++m) {
        const int i = (m < V / 2) ? m + H : m + H - V;
        v[m] = (i >= 0 && i < V / 2) ? x[i] : Z{0.0, 0.0};
    }
}

template <int N>
__device__ inline void place_channels(Z* v, const Z* x, int32_t H) {
    // h <= N / 4, so H <= N / 256
    switch (H) {
        case 0: place_channels_h<N, 0>(v, x); break;
        case 1: place_channels_h<N, 1>(v, x); break;
        case 2: if constexpr (N >= 512) place_channels_h<N, 2>(v, x); break;
        case 3: if constexpr (N >= 1024) place_channels_h<N, 3>(v, x); break;
        default: if constexpr (N >= 1024) place_channels_h<N, 4>(v, x); break;
    }
}

template <int N, int W, int NDL>
__global__ __launch_bounds__(kWave) void rotate_average_nufft(const LEAP_GLOBAL Z* __restrict__ w,
                                                              const LEAP_GLOBAL double* __restrict__ uvw,
                                                              const LEAP_GLOBAL double* __restrict__ dd, uint32_t D,
                                                              uint32_t T, uint32_t nbl, uint32_t C, uint32_t P,
                                                              NufftArgs args, LEAP_GLOBAL double* __restrict__ avg,
                                                              int32_t accumulate) {
#pragma clang fp contract(off)
    constexpr int V = N / kWave;
    static_assert(W % 2 == 0 && W <= kNufftMaxWidth && N / 16 >= W - 1, "kernel width");
    __shared__ Z s_g[N + N / 16];
    const uint32_t lane = threadIdx.x;
    const uint32_t bl = blockIdx.x;
    const uint32_t d0 = blockIdx.y * (kWave * NDL);
    double dz[NDL][3];
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        const uint32_t d = d0 + lane + kWave * j;
        const uint32_t dc = d < D ? d : D - 1;
        dz[j][0] = dd[9 * size_t(dc) + 6];
        dz[j][1] = dd[9 * size_t(dc) + 7];
        dz[j][2] = dd[9 * size_t(dc) + 8];
    }
    Twiddles<N> tw;
    tw.init(lane);
    Z sum[NDL];
#pragma unroll
    for (int j = 0; j < NDL; ++j) sum[j] = {0.0, 0.0};
    const int32_t H = args.h / kWave;
    Z next[V / 2];
    if (T > 0) load_channels<N>(next, w + size_t(bl) * C, lane, C);
    for (uint32_t t = 0; t < T; ++t) {
        Z v[V];
        place_channels<N>(v, next, H);
        const size_t row = size_t(t) * nbl + bl;
        if (t + 1 < T) load_channels<N>(next, w + (row + nbl) * C, lane, C);
        const double u = uvw[3 * row], vv = uvw[3 * row + 1], ww = uvw[3 * row + 2];
        tw.hide();
        row_fft<N, W>(v, s_g, tw, lane);
#pragma unroll
        for (int j = 0; j < NDL; ++j) {
            const double sf = shift_factor(dz[j], u, vv, ww);
            double x = sf * args.dkappa * kInvTwoPi;  // turns of b
            x -= floor(x);                           // [0, 1]
            const double ug = x * double(N);          // grid coordinate in [0, N]
            const double m0f = ceil(ug - 0.5 * W);
            const double s = (ug - m0f) - 0.5 * (W - 1);  // (-1/2, 1/2]
            int32_t m0 = int32_t(m0f);
            if (m0 < 0) m0 += N;  // taps m0 .. m0 + W - 1 <= N + W - 2 (wrapped copy)
            m0 = min(max(m0, 0), N - 1);  // a non-finite uvw stays inside the grid
            const Z* gp = s_g + m0;
            const double s2 = s * s;
            Z acc{0.0, 0.0};
#pragma unroll
            for (int jp = 0; jp < W / 2; ++jp) {
                // P_jp(s) = E(s^2) + s O(s^2); tap W - 1 - jp has P_jp(-s)
                double e = args.coef[jp][W - 2];
                double o = args.coef[jp][W - 1];
#pragma unroll
                for (int k = W - 4; k >= 0; k -= 2) {
                    e = fma(e, s2, args.coef[jp][k]);
                    o = fma(o, s2, args.coef[jp][k + 1]);
                }
                const double wl = fma(s, o, e), wh = fma(-s, o, e);
                const Z g0 = gp[jp], g1 = gp[W - 1 - jp];
                acc.x = fma(g0.x, wl, acc.x);
                acc.y = fma(g0.y, wl, acc.y);
                acc.x = fma(g1.x, wh, acc.x);
                acc.y = fma(g1.y, wh, acc.y);
            }
            const Z r = cmul(phasor_f64(sf * args.kappa_mid), acc);
            sum[j].x += r.x;
            sum[j].y += r.y;
            // one direction's grid reads in flight at a time (registers)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int j = 0; j < NDL; ++j) {
        const uint32_t d = d0 + lane + kWave * j;
        if (d >= D) continue;
        LEAP_GLOBAL double* o = avg + (size_t(d) * nbl + bl) * P * 2;
        if (accumulate) {
            o[0] += sum[j].x;
            o[1] += sum[j].y;
        } else {
            o[0] = sum[j].x;
            o[1] = sum[j].y;
            for (uint32_t q = 2; q < 2 * P; ++q) o[q] = 0.0;  // pols not accumulated
        }
    }
}

template <int N, int W>
void launch_nw(hipStream_t s, int32_t tile, const dim3& grid_x, const Z* w, const double* uvw, const double* dd,
               uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, const NufftArgs& a, double* avg,
               int32_t accumulate) {
    const auto* gw = (const LEAP_GLOBAL Z*)w;
    const auto* gu = (const LEAP_GLOBAL double*)uvw;
    const auto* gd = (const LEAP_GLOBAL double*)dd;
    auto* ga = (LEAP_GLOBAL double*)avg;
    switch (tile) {
        case 641:
            rotate_average_nufft<N, W, 1><<<dim3(grid_x.x, (D + 63) / 64), kWave, 0, s>>>(gw, gu, gd, D, T, nbl, C, P,
                                                                                        a, ga, accumulate);
            break;
        case 1282:
            rotate_average_nufft<N, W, 2><<<dim3(grid_x.x, (D + 127) / 128), kWave, 0, s>>>(gw, gu, gd, D, T, nbl, C,
                                                                                          P, a, ga, accumulate);
            break;
        case 2564:
            rotate_average_nufft<N, W, 4><<<dim3(grid_x.x, (D + 255) / 256), kWave, 0, s>>>(gw, gu, gd, D, T, nbl, C,
                                                                                          P, a, ga, accumulate);
            break;
        default: throw Error(LEAP_ERR_INVALID_ARGUMENT, "unsupported NUFFT tile " + std::to_string(tile));
    }
}

}  // namespace

void launch_rotate_nufft(hipStream_t s, int32_t tile, const NufftPlan& plan, const double* w, const double* uvw,
                         const double* dd, uint32_t D, uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C,
                         uint32_t P, double f0, double df, double* avg, int32_t accumulate) {
    if (nbl_grid == 0 || D == 0) return;
    require(plan.N != 0 && plan.scale.size() == C, LEAP_ERR_INTERNAL, "NUFFT plan does not match the channel count");
    NufftArgs a{};
    for (int j = 0; j < kNufftMaxWidth / 2; ++j)
        for (int m = 0; m < kNufftMaxWidth; ++m) a.coef[j][m] = plan.coef[j][m];
    a.kappa_mid = (f0 + double(plan.h) * df) / kSpeedOfLight;
    a.dkappa = df / kSpeedOfLight;
    a.h = int32_t(plan.h);
    const Z* wp = reinterpret_cast<const Z*>(w);
    const dim3 gx(nbl_grid);
#define LEAP_NUFFT(N_, W_)                                                                          \
    if (plan.N == N_ && plan.W == W_) {                                                             \
        launch_nw<N_, W_>(s, tile, gx, wp, uvw, dd, D, T, nbl, C, P, a, avg, accumulate);            \
        LEAP_HIP(hipGetLastError());                                                                \
        return;                                                                                     \
    }
    LEAP_NUFFT(256, 10)
    LEAP_NUFFT(256, 12)
    LEAP_NUFFT(512, 10)
    LEAP_NUFFT(512, 12)
    LEAP_NUFFT(1024, 10)
    LEAP_NUFFT(1024, 12)
#undef LEAP_NUFFT
    throw Error(LEAP_ERR_INTERNAL, "no NUFFT kernel for grid " + std::to_string(plan.N) + ", width " +
                                       std::to_string(plan.W));
}

}  // namespace leap
