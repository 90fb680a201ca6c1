// FP64 kernels of the per-direction ("plain") calibration path for gfx950.
//
//   dd_kernel               MetaData::GenerateDDMatrix (model/cpu/MetaData.cc:166-203)
//   rotate_average_f64      CpuLeapCalibrator::RotateVisibilities (CpuLeapCalibrator.cc:181-217)
//   assemble_ad_kernel      pinv(A) = G+ A^T from the Gram pseudo-inverse
//   phase_i1 / cal1 / delta / gemv kernels
//                           the inversion of CpuLeapCalibrator::PhaseRotate (:159-178)
//
// Layouts in HBM: visibilities MS-native [row][channel][pol] complex (c128 or
// c64); uvw [row][3] f64; avgData [direction][baseline][pol] complex128;
// Ad stored transposed as AdT [K][N] (row k = column k of Ad) so the GEMV
// reads whole rows; cal [direction][N].
#include "inversion_common.h"
#include "leap_internal.h"

namespace leap {
namespace {

constexpr int kBlock = 256;       // 4 wave64 per workgroup
constexpr int kWaves = kBlock / 64;
constexpr uint32_t kTChunk = 512;  // timesteps whose phase factors sit in LDS at once

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---------------------------------------------------------------------------
// dd matrices, one thread per direction; dd row-major [d][3][3].
// ---------------------------------------------------------------------------
__global__ void dd_kernel(double ra0, double dec0, const double* __restrict__ dirs, uint32_t D,
                          double* __restrict__ dd) {
#pragma clang fp contract(off)
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    const double ang1 = kPi / 2.0 - dec0;
    const double ang2 = dirs[2 * d] - ra0;
    const double ang3 = -kPi / 2.0 + dirs[2 * d + 1];
    double s1, c1, s2, c2, s3, c3;
    sincos(ang1, &s1, &c1);
    sincos(ang2, &s2, &c2);
    sincos(ang3, &s3, &c3);
    const double m1[9] = {1, 0, 0, 0, c1, -s1, 0, s1, c1};
    const double m2[9] = {c2, s2, 0, -s2, c2, 0, 0, 0, 1};
    const double m3[9] = {1, 0, 0, 0, c3, -s3, 0, s3, c3};
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            t[i * 3 + j] = m3[i * 3 + 0] * m2[j] + m3[i * 3 + 1] * m2[3 + j] + m3[i * 3 + 2] * m2[6 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            dd[9 * d + i * 3 + j] = t[i * 3 + 0] * m1[j] + t[i * 3 + 1] * m1[3 + j] + t[i * 3 + 2] * m1[6 + j];
}

// ---------------------------------------------------------------------------
// rotate_average_f64: one workgroup per baseline.  It streams the baseline's
// T_interval rows (C x P complex each, contiguous) once, rotates every
// (row, channel) by exp(i * shift) with shift = sf(row) / lambda(channel),
// skips a (row, channel) whose rotated polarisations hold a NaN, and sums the
// rest in registers; a wave64 butterfly plus an LDS step then reduce the
// workgroup to avgData[baseline][pol].  No atomics and no write-back to the
// cube.  The sum order is fixed, so results are reproducible bit for bit.
// ---------------------------------------------------------------------------
template <typename T, int P>
struct VisLoad;

template <int P>
struct VisLoad<double, P> {
    double v[2 * P];
    __device__ inline void load(const double* __restrict__ p) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const double2 x = reinterpret_cast<const double2*>(p)[q];
            v[2 * q] = x.x;
            v[2 * q + 1] = x.y;
        }
    }
};

template <int P>
struct VisLoad<float, P> {
    double v[2 * P];
    __device__ inline void load(const float* __restrict__ p) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float2 x = reinterpret_cast<const float2*>(p)[q];
            v[2 * q] = double(x.x);  // the reference widens complex<float> (ms/utils.h:141-142)
            v[2 * q + 1] = double(x.y);
        }
    }
};

template <typename T, int P>
__global__ __launch_bounds__(kBlock) void rotate_average_f64(
    const T* __restrict__ vis, const double* __restrict__ uvw, const double* __restrict__ dd,
    uint32_t row_start, uint32_t T_steps, uint32_t nbl, uint32_t C,
    const double* __restrict__ lambda, double* __restrict__ avg) {
#pragma clang fp contract(off)
    __shared__ double s_sf[kTChunk];
    __shared__ double s_red[kWaves][2 * P];
    const uint32_t bl = blockIdx.x;
    const double dz0 = dd[6], dz1 = dd[7], dz2 = dd[8];
    double acc[2 * P];
#pragma unroll
    for (int q = 0; q < 2 * P; ++q) acc[q] = 0.0;

    for (uint32_t t0 = 0; t0 < T_steps; t0 += kTChunk) {
        const uint32_t tn = min(kTChunk, T_steps - t0);
        __syncthreads();
        for (uint32_t t = threadIdx.x; t < tn; t += kBlock) {
            const size_t row = size_t(row_start) + size_t(t0 + t) * nbl + bl;
            const double u = uvw[3 * row], v = uvw[3 * row + 1], w = uvw[3 * row + 2];
            const double wrot = dz0 * u + dz1 * v + dz2 * w;  // (dd * uvw).z, MetaData.cc:227
            s_sf[t] = -kTwoPi * (wrot - w);                     // CpuLeapCalibrator.cc:190
        }
        __syncthreads();
        // Pair (t, c) of this chunk handled by this thread: i = threadIdx.x + k * kBlock.
        uint32_t t = 0, c = threadIdx.x;
        while (c >= C) { c -= C; ++t; }
        for (; t < tn;) {
            const size_t row = size_t(row_start) + size_t(t0 + t) * nbl + bl;
            VisLoad<T, P> x;
            x.load(vis + (row * C + c) * (2 * P));
            const double shift = s_sf[t] / lambda[c];  // CpuLeapCalibrator.cc:195
            double sn, cs;
            sincos(shift, &sn, &cs);
            double r[2 * P];
            bool nan = false;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const double a = x.v[2 * q], b = x.v[2 * q + 1];
                r[2 * q] = a * cs - b * sn;  // (a + ib)(cs + i sn), :198
                r[2 * q + 1] = a * sn + b * cs;
                nan |= __builtin_isnan(r[2 * q]) || __builtin_isnan(r[2 * q + 1]);  // :201-206
            }
            if (!nan) {
#pragma unroll
                for (int q = 0; q < 2 * P; ++q) acc[q] += r[q];  // :208-214
            }
            c += kBlock;
            while (c >= C) { c -= C; ++t; }
        }
    }
    // Workgroup reduction: wave butterfly, then the kWaves partials in order.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 2 * P; ++q) {
        const double s = wave_sum(acc[q]);
        if (lane == 0) s_red[wave][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < 2 * P) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += s_red[w][threadIdx.x];
        avg[size_t(bl) * (2 * P) + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Ad assembly: AdT[k][i] = G+(i, a1_k) - G+(i, a2_k) for k < K-1 and
// AdT[K-1][i] = G+(i, c); G+ is symmetric, so G+(i, a) = Gp[a * N + i].
// ---------------------------------------------------------------------------
__global__ void assemble_ad_kernel(const double* __restrict__ Gp, int32_t N,
                                   const int2* __restrict__ pairs, int32_t n_pairs,
                                   int32_t constraint, double* __restrict__ AdT) {
    const int32_t k = blockIdx.x;
    if (k < n_pairs) {
        const int2 p = pairs[k];
        for (int32_t i = threadIdx.x; i < N; i += blockDim.x)
            AdT[size_t(k) * N + i] = Gp[size_t(p.x) * N + i] - Gp[size_t(p.y) * N + i];
    } else {
        for (int32_t i = threadIdx.x; i < N; i += blockDim.x)
            AdT[size_t(k) * N + i] = Gp[size_t(constraint) * N + i];
    }
}

// ---------------------------------------------------------------------------
// Inversion, per direction d (grid.y):
//   phase[d][k] = arg(avg[d][I1[k]][pol 0]) for k < K1 - 1, 0 at K1 - 1
//   cal1[d]     = Ad1 * phase[d]
//   delta[d][n] = arg(exp(-2 pi i (cal1[a1_n] - cal1[a2_n])) * avg[d][n][pol 0]), n < K - 1
//                 (row n of avgData, as the reference indexes it; A * cal1 in
//                 gather form, exact because each A row is e_a1 - e_a2)
//   cal[d]      = Ad * delta[d] + cal1[d]
// ---------------------------------------------------------------------------
__global__ void phase_i1_kernel(const double* __restrict__ avg, uint32_t nbl, uint32_t P,
                                const int32_t* __restrict__ I1, int32_t K1, double* __restrict__ phase) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    if (k >= K1) return;
    double v = 0.0;
    if (k < K1 - 1) {
        int32_t idx = I1[k];
        if (idx < 0) idx += int32_t(nbl);  // VectorRangeSelect wrap (math/cpu/vector.h:97-104)
        const double* a = avg + (size_t(d) * nbl + idx) * (2 * P);
        v = atan2(a[1], a[0]);
    }
    phase[size_t(d) * K1 + k] = v;
}

// cal1[d][i] = sum_k Ad1T[k][i] phase[d][k], summed in k order.  One wave per
// (kCal1Dirs directions, 64 antennas): lane = antenna, the wave's phases are
// wave-uniform (scalar loads), every Ad1T element read feeds kCal1Dirs chains,
// and the loop is unrolled so several rows of Ad1T are in flight.
constexpr int kCal1Dirs = 4;

__global__ __launch_bounds__(64) void cal1_kernel(const double* __restrict__ Ad1T, int32_t K1, int32_t N,
                                                  const double* __restrict__ phase, uint32_t D,
                                                  double* __restrict__ cal1) {
    const int32_t i = blockIdx.x * 64 + threadIdx.x;
    const uint32_t d0 = blockIdx.y * kCal1Dirs;
    const int32_t ic = i < N ? i : N - 1;
    const double* ph[kCal1Dirs];
#pragma unroll
    for (int j = 0; j < kCal1Dirs; ++j) ph[j] = phase + size_t(d0 + j < D ? d0 + j : D - 1) * K1;
    double acc[kCal1Dirs];
#pragma unroll
    for (int j = 0; j < kCal1Dirs; ++j) acc[j] = 0.0;
#pragma unroll 8
    for (int32_t k = 0; k < K1; ++k) {
        const double b = Ad1T[size_t(k) * N + ic];
#pragma unroll
        for (int j = 0; j < kCal1Dirs; ++j) acc[j] = fma(b, ph[j][k], acc[j]);
    }
    if (i < N) {
#pragma unroll
        for (int j = 0; j < kCal1Dirs; ++j)
            if (d0 + j < D) cal1[size_t(d0 + j) * N + i] = acc[j];
    }
}

__global__ void delta_kernel(const double* __restrict__ avg, uint32_t nbl, uint32_t P,
                             const int2* __restrict__ pairs, int32_t K, int32_t N,
                             const double* __restrict__ cal1, double* __restrict__ delta) {
    const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    if (n >= K) return;
    delta[size_t(d) * K + n] = dev::delta_phase(avg, nbl, P, pairs, K, N, cal1, d, n);
}

// Split-K GEMV: partial[d][chunk][i] = sum over the chunk's k of AdT[k][i] * delta[d][k].
// 64 rows per chunk: about two workgroups per CU at 32,896 baselines, and a short dependent
// chain per thread.
constexpr int32_t kGemvChunk = 64;

__global__ __launch_bounds__(kBlock) void gemv_partial_kernel(
    const double* __restrict__ AdT, int32_t K, int32_t N, const double* __restrict__ delta,
    double* __restrict__ partial) {
    const int32_t chunk = blockIdx.x;
    const uint32_t d = blockIdx.y;
    const int32_t k0 = chunk * kGemvChunk;
    const int32_t k1 = min(K, k0 + kGemvChunk);
    __shared__ double s_x[kGemvChunk];
    for (int32_t k = k0 + threadIdx.x; k < k1; k += kBlock) s_x[k - k0] = delta[size_t(d) * K + k];
    __syncthreads();
    for (int32_t i = threadIdx.x; i < N; i += kBlock) {
        double s = 0.0;
        for (int32_t k = k0; k < k1; ++k) s += AdT[size_t(k) * N + i] * s_x[k - k0];
        partial[(size_t(d) * gridDim.x + chunk) * N + i] = s;
    }
}

// cal[d][i] = sum over chunks of partial[d][chunk][i] + cal1[d][i]: one workgroup per output,
// each lane a strided slice of the chunks, then a fixed-order tree in LDS.
__global__ __launch_bounds__(kBlock) void gemv_reduce_kernel(const double* __restrict__ partial, int32_t n_chunks,
                                                             int32_t N, const double* __restrict__ cal1,
                                                             double* __restrict__ cal) {
    __shared__ double s_red[kBlock];
    const int32_t i = blockIdx.x;
    const uint32_t d = blockIdx.y;
    double acc = 0.0;
    for (int32_t c = threadIdx.x; c < n_chunks; c += kBlock) acc += partial[(size_t(d) * n_chunks + c) * N + i];
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t stride = kBlock / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s_red[threadIdx.x] += s_red[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) cal[size_t(d) * N + i] = s_red[0] + cal1[size_t(d) * N + i];
}

// ---------------------------------------------------------------------------
// Moore-Penrose inverse of the Gram matrix G = A^T A (N x N, column-major), for
// A the baseline rows e_a1 - e_a2 plus the constraint row e_c.  G is a graph
// Laplacian plus e_c e_c^T, formed exactly (small integers).  Its null space is
// spanned by the indicators of the baseline graph's components without c, so
// per component S (members sorted, from the host's union-find):
//   c in S:      G+_SS = G_SS^-1                      (positive definite)
//   c not in S:  G+_SS = (G_SS + J / |S|)^-1 - J / |S|  (L_S 1 = 0)
// and G+ is zero between components and on unused antennas.  Each block is
// inverted as M = L L^T (left-looking Cholesky), W = L^-1 (forward
// substitution), M^-1 = W^T W, every sum in the order of the host routine.
// ---------------------------------------------------------------------------
__global__ void fill_zero_kernel(double* __restrict__ x, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// Column a of G: the degree of antenna a (+1 at the constraint) on the diagonal,
// -1 per row joining a and b at row b.  csr_ent holds 2 k + (a is antenna2 of row k).
__global__ void gram_columns_kernel(const int32_t* __restrict__ csr_off, const int32_t* __restrict__ csr_ent,
                                    const int2* __restrict__ pairs, int32_t N, int32_t constraint,
                                    double* __restrict__ G) {
    const int32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    double* col = G + size_t(a) * N;
    col[a] = double(csr_off[a + 1] - csr_off[a]) + (a == constraint ? 1.0 : 0.0);
    for (int32_t e = csr_off[a]; e < csr_off[a + 1]; ++e) {
        const int32_t v = csr_ent[e];
        const int2 p = pairs[v >> 1];
        col[(v & 1) ? p.x : p.y] -= 1.0;
    }
}

// One workgroup per component; ws holds 2 n^2 doubles per component (M, then W).
// *status becomes 1 when a pivot is not positive (the block is not positive definite).
__global__ __launch_bounds__(kBlock) void gram_block_pinv_kernel(const double* __restrict__ G, int32_t N,
                                                                 const int32_t* __restrict__ comp_off,
                                                                 const int32_t* __restrict__ members,
                                                                 const uint8_t* __restrict__ has_c,
                                                                 const int64_t* __restrict__ ws_off,
                                                                 double* __restrict__ ws, double* __restrict__ Gp,
                                                                 int32_t* __restrict__ status) {
#pragma clang fp contract(off)
    const int32_t c = blockIdx.x;
    const int32_t n = comp_off[c + 1] - comp_off[c];
    const int32_t* S = members + comp_off[c];
    double* M = ws + ws_off[c];
    double* Wm = M + size_t(n) * n;
    auto at = [&](int32_t i, int32_t j) -> double& { return M[size_t(j) * n + i]; };
    auto w = [&](int32_t i, int32_t j) -> double& { return Wm[size_t(j) * n + i]; };
    const double inv_n = 1.0 / double(n);
    const bool shifted = !has_c[c];
    for (int32_t idx = threadIdx.x; idx < n * n; idx += kBlock) {
        const int32_t i = idx % n, j = idx / n;
        double x = G[size_t(S[j]) * N + S[i]];
        if (shifted) x += inv_n;
        M[idx] = x;
        Wm[idx] = 0.0;
    }
    __syncthreads();
    __shared__ double s_ljj;
    for (int32_t j = 0; j < n; ++j) {  // L in the lower triangle
        if (threadIdx.x == 0) {
            double d = at(j, j);
            for (int32_t k = 0; k < j; ++k) d -= at(j, k) * at(j, k);
            if (!(d > 0.0)) {
                *status = 1;
                d = 1.0;
            }
            s_ljj = sqrt(d);
            at(j, j) = s_ljj;
        }
        __syncthreads();
        const double ljj = s_ljj;
        for (int32_t i = j + 1 + int32_t(threadIdx.x); i < n; i += kBlock) {
            double sum = at(i, j);
            for (int32_t k = 0; k < j; ++k) sum -= at(i, k) * at(j, k);
            at(i, j) = sum / ljj;
        }
        __syncthreads();
    }
    for (int32_t j = threadIdx.x; j < n; j += kBlock) {  // W = L^-1, one column per thread
        w(j, j) = 1.0 / at(j, j);
        for (int32_t i = j + 1; i < n; ++i) {
            double sum = 0.0;
            for (int32_t k = j; k < i; ++k) sum -= at(i, k) * w(k, j);
            w(i, j) = sum / at(i, i);
        }
    }
    __syncthreads();
    for (int32_t idx = threadIdx.x; idx < n * n; idx += kBlock) {  // M^-1 = W^T W, scattered into G+
        const int32_t i = idx % n, j = idx / n;
        const int32_t lo = i > j ? i : j, a = i > j ? i : j, b = i > j ? j : i;
        double sum = 0.0;
        for (int32_t k = lo; k < n; ++k) sum += w(k, a) * w(k, b);
        if (shifted) sum -= inv_n;
        Gp[size_t(S[j]) * N + S[i]] = sum;
    }
}

template <typename T>
void launch_rotate_typed(hipStream_t s, const T* vis, const double* uvw, const double* dd,
                         uint32_t row_start, uint32_t T_steps, uint32_t nbl, uint32_t C, uint32_t P,
                         const double* lambda, double* avg) {
    const dim3 grid(nbl), block(kBlock);
    switch (P) {
        case 1: rotate_average_f64<T, 1><<<grid, block, 0, s>>>(vis, uvw, dd, row_start, T_steps, nbl, C, lambda, avg); break;
        case 2: rotate_average_f64<T, 2><<<grid, block, 0, s>>>(vis, uvw, dd, row_start, T_steps, nbl, C, lambda, avg); break;
        case 3: rotate_average_f64<T, 3><<<grid, block, 0, s>>>(vis, uvw, dd, row_start, T_steps, nbl, C, lambda, avg); break;
        case 4: rotate_average_f64<T, 4><<<grid, block, 0, s>>>(vis, uvw, dd, row_start, T_steps, nbl, C, lambda, avg); break;
        default: throw Error(LEAP_ERR_NOT_IMPLEMENTED, "unsupported polarisation count");
    }
    LEAP_HIP(hipGetLastError());
}

}  // namespace

void launch_dd(hipStream_t s, double ra0, double dec0, const double* d_dirs, uint32_t D, double* d_dd) {
    if (D == 0) return;
    dd_kernel<<<(D + 63) / 64, 64, 0, s>>>(ra0, dec0, d_dirs, D, d_dd);
    LEAP_HIP(hipGetLastError());
}

void launch_rotate_f64(hipStream_t s, const void* vis, int32_t vis_dtype, const double* uvw,
                       const double* dd, uint32_t row_start, uint32_t T_steps, uint32_t nbl,
                       uint32_t C, uint32_t P, const double* lambda, double* avg) {
    if (nbl == 0) return;
    if (vis_dtype == LEAP_VIS_C64)
        launch_rotate_typed(s, static_cast<const float*>(vis), uvw, dd, row_start, T_steps, nbl, C, P, lambda, avg);
    else
        launch_rotate_typed(s, static_cast<const double*>(vis), uvw, dd, row_start, T_steps, nbl, C, P, lambda, avg);
}

void launch_gram_pinv(hipStream_t s, const GramPinv& g) {
    const size_t nn = size_t(g.N) * g.N;
    fill_zero_kernel<<<uint32_t((nn + 255) / 256), 256, 0, s>>>(g.G, nn);
    LEAP_HIP(hipGetLastError());
    fill_zero_kernel<<<uint32_t((nn + 255) / 256), 256, 0, s>>>(g.Gp, nn);
    LEAP_HIP(hipGetLastError());
    gram_columns_kernel<<<(g.N + 63) / 64, 64, 0, s>>>(g.csr_off, g.csr_ent, g.pairs, g.N, g.constraint, g.G);
    LEAP_HIP(hipGetLastError());
    if (g.n_components > 0) {
        gram_block_pinv_kernel<<<g.n_components, kBlock, 0, s>>>(g.G, g.N, g.comp_off, g.members, g.has_c, g.ws_off,
                                                                 g.ws, g.Gp, g.status);
        LEAP_HIP(hipGetLastError());
    }
}

void launch_assemble_ad(hipStream_t s, const double* Gp, int32_t N, const int2* pairs, int32_t n_pairs,
                        int32_t constraint, double* AdT) {
    assemble_ad_kernel<<<n_pairs + 1, kBlock, 0, s>>>(Gp, N, pairs, n_pairs, constraint, AdT);
    LEAP_HIP(hipGetLastError());
}

void launch_inversion_f64(hipStream_t s, const leap_metadata& md, const double* avg, uint32_t D,
                          double* phase, double* cal1, double* delta, double* partial, double* cal) {
    if (D == 0) return;
    const int32_t K = md.A.rows, N = md.A.cols, K1 = md.A1.rows;
    const uint32_t nbl = md.nbl, P = md.pols;
    phase_i1_kernel<<<dim3((K1 + 255) / 256, D), 256, 0, s>>>(avg, nbl, P, md.d_I1.as<int32_t>(), K1, phase);
    LEAP_HIP(hipGetLastError());
    cal1_kernel<<<dim3((N + 63) / 64, (D + kCal1Dirs - 1) / kCal1Dirs), 64, 0, s>>>(md.d_Ad1T.as<double>(), K1, N,
                                                                                phase, D, cal1);
    LEAP_HIP(hipGetLastError());
    delta_kernel<<<dim3((K + 255) / 256, D), 256, 0, s>>>(avg, nbl, P, md.d_pairs.as<int2>(), K, N, cal1, delta);
    LEAP_HIP(hipGetLastError());
    const int32_t chunks = (K + kGemvChunk - 1) / kGemvChunk;
    gemv_partial_kernel<<<dim3(chunks, D), kBlock, 0, s>>>(md.d_AdT.as<double>(), K, N, delta, partial);
    LEAP_HIP(hipGetLastError());
    gemv_reduce_kernel<<<dim3(N, D), kBlock, 0, s>>>(partial, chunks, N, cal1, cal);
    LEAP_HIP(hipGetLastError());
}

void launch_inversion_head(hipStream_t s, const leap_metadata& md, const double* avg, uint32_t avg_pols, uint32_t D,
                           double* phase, double* cal1, double* delta) {
    if (D == 0) return;
    const int32_t K = md.A.rows, N = md.A.cols, K1 = md.A1.rows;
    const uint32_t nbl = md.nbl, P = avg_pols;
    phase_i1_kernel<<<dim3((K1 + 255) / 256, D), 256, 0, s>>>(avg, nbl, P, md.d_I1.as<int32_t>(), K1, phase);
    LEAP_HIP(hipGetLastError());
    cal1_kernel<<<dim3((N + 63) / 64, (D + kCal1Dirs - 1) / kCal1Dirs), 64, 0, s>>>(md.d_Ad1T.as<double>(), K1, N,
                                                                                phase, D, cal1);
    LEAP_HIP(hipGetLastError());
    if (delta) {
        delta_kernel<<<dim3((K + 255) / 256, D), 256, 0, s>>>(avg, nbl, P, md.d_pairs.as<int2>(), K, N, cal1, delta);
        LEAP_HIP(hipGetLastError());
    }
}

size_t inversion_partial_doubles(int32_t K, int32_t N, uint32_t D) {
    return size_t((K + kGemvChunk - 1) / kGemvChunk) * size_t(N) * D;
}

}  // namespace leap
