// Adjoint of rotate -> average for the calibration-residual gradient.
//
// Forward (per row r = (t, baseline), direction d, pol 0):
//   sf(r, d)  = -2 pi ((dz_d . uvw(r)) - w(r))
//   avg(b, d) = sum_{t, c} v(r, c) exp(i sf kappa_c),   kappa_c = kappa0 + c dkappa
// Given abar(b, d) = dL/dRe avg + i dL/dIm avg (so dL = Re(conj(abar) d avg)):
//   dL/dsf(r, d) = Re(conj(abar) * i * sum_c kappa_c v_c exp(i sf kappa_c))
//                = Re(conj(abar) * i * e0 * (kappa0 P(z) + dkappa z P'(z)))
// with e0 = exp(i sf kappa0), z = exp(i sf dkappa), P(z) = sum_c v_c z^c.  When
// the forward pass has kept each row's sums e0 P(z) and e0 z P'(z) (32 bytes per
// row and direction, rotate_average_horner64_q), adjoint_q_kernel reads them
// and the cube is not read again; otherwise rotate_adjoint_kernel evaluates P
// and P' together by Horner's rule (p' <- p' z + p; p <- p z + v) from the
// cube.  Then
//   dL/d dz_d   += -2 pi dL/dsf(r, d) uvw(r)            (per baseline, reduced later)
//   dL/d uvw(r)  = sum_d -2 pi dL/dsf(r, d) (dz_d - e_z)
// and the antenna gradient is dL/d dp_a = sum_t M_t^T (sum_{b: a2 = a} g_uvw - sum_{b: a1 = a} g_uvw).
#include <algorithm>

#include "cube_ring.h"
#include "leap_internal.h"

namespace leap {
namespace {

constexpr int kThreads = 256;

struct Cx {
    double x, y;
};

__device__ inline Cx cmulx(Cx a, Cx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline Cx hstep(Cx a, Cx z, Cx v) {
    return {fma(a.x, z.x, fma(-a.y, z.y, v.x)), fma(a.x, z.y, fma(a.y, z.x, v.y))};
}
__device__ inline Cx expi(double t) {
    double s, c;
    sincos(t, &s, &c);
    return {c, s};
}

// uvw'(row) = uvw0(row) + M_t (dp[a2] - dp[a1])
__global__ void perturb_uvw_kernel(const double* __restrict__ uvw0, const double* __restrict__ M,
                                   const int2* __restrict__ bl_ants, const double* __restrict__ dp, uint32_t nbl,
                                   uint32_t rows, double* __restrict__ uvw) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t t = r / nbl, b = r - t * nbl;
    const int2 a = bl_ants[b];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = dp[3 * a.y + k] - dp[3 * a.x + k];
    for (int i = 0; i < 3; ++i) {
        double s = uvw0[3 * size_t(r) + i];
        for (int k = 0; k < 3; ++k) s += M[9 * size_t(t) + 3 * i + k] * d[k];
        uvw[3 * size_t(r) + i] = s;
    }
}

// One workgroup per baseline; DL lanes per row with ND directions each (direction lane + DL j,
// D <= DL ND), RP = 256/DL rows side by side, channels in chunks of CC = 256/RP (at least 16).
// A cube the ring accepts (cube_ring.h: complex128, or complex64 with 2 or 4 pols, 16-byte
// aligned) streams into LDS with global_load_lds, two chunks ahead for complex64 and one for
// complex128; others are staged element by element.  The Horner recurrences run over the chunks
// in order, so the chunk size does not change the result.
template <typename VT, int DL, int ND>
__global__ __launch_bounds__(kThreads) void rotate_adjoint_kernel(
    const VT* __restrict__ vis, const double* __restrict__ uvw, const double* __restrict__ dd,
    const double* __restrict__ abar, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double kappa0,
    double dkappa, double* __restrict__ g_uvw, double* __restrict__ g_dz_partial) {
    constexpr int RP = kThreads / DL;
    constexpr int kCC = 256 / RP > 16 ? 256 / RP : 16;  // channels per chunk
    constexpr int DW = DL * ND;                          // directions of the workgroup
    using Ring = dev::CubeRing<VT, RP, kCC, sizeof(VT) == 4 ? 2 : 1>;
    // one padding element per row for fewer than 16 lanes per row (rows kCC * 16 bytes apart
    // would share LDS banks within a 16-lane group)
    __shared__ Cx s_v[RP][kCC + (DL < 16 ? 1 : 0)];
    __shared__ dev::f32x4 s_raw[Ring::kBytes / 16];
    __shared__ double s_g[RP][DW][3];
    const uint32_t bl = blockIdx.x;
    const uint32_t lane_d = threadIdx.x % DL;
    const uint32_t rsub = threadIdx.x / DL;
    bool dir_ok[ND];
    bool any_dir = false;
    double dz[ND][3], gdz[ND][3];
    Cx ab[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const uint32_t dj = lane_d + DL * j;
        dir_ok[j] = dj < D;
        any_dir = any_dir || dir_ok[j];
        const uint32_t dc = dir_ok[j] ? dj : 0;
        dz[j][0] = dd[9 * dc + 6];
        dz[j][1] = dd[9 * dc + 7];
        dz[j][2] = dd[9 * dc + 8];
        ab[j] = Cx{0.0, 0.0};
        if (dir_ok[j]) ab[j] = {abar[2 * (size_t(dc) * nbl + bl)], abar[2 * (size_t(dc) * nbl + bl) + 1]};
        gdz[j][0] = gdz[j][1] = gdz[j][2] = 0.0;
    }
    const uint32_t n_chunks = (C + kCC - 1) / kCC;
    Ring ring;
    const bool streamed = Ring::accepts((const void*)vis, P);
    // address-space casts: the copies take global and LDS pointers
    if (streamed) ring.start((LEAP_LDS dev::f32x4*)(&s_raw[0]), (const LEAP_GLOBAL VT*)vis, P, T, nbl, bl, C);

    for (uint32_t tb = 0; tb < T; tb += RP) {
        const uint32_t t = tb + rsub;
        const bool row_ok = t < T;
        double u = 0.0, v = 0.0, w = 0.0;
        if (row_ok) {
            const size_t row = size_t(t) * nbl + bl;
            u = uvw[3 * row];
            v = uvw[3 * row + 1];
            w = uvw[3 * row + 2];
        }
        double sf[ND];
        Cx z[ND], p[ND], dp[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            sf[j] = -kTwoPi * ((dz[j][0] * u + dz[j][1] * v + dz[j][2] * w) - w);
            z[j] = expi(sf[j] * dkappa);
            p[j] = {0.0, 0.0};
            dp[j] = {0.0, 0.0};
        }
        for (int ch = int(n_chunks) - 1; ch >= 0; --ch) {
            const uint32_t c0 = uint32_t(ch) * kCC;
            const uint32_t cn = min(uint32_t(kCC), C - c0);
            if (streamed) {
                // this chunk's copy has landed and the previous chunk's recurrences are done
                const LEAP_LDS dev::f32x4* slot = ring.next_slot();
                __syncthreads();
                ring.convert(slot, tb, cn, [&](uint32_t r, uint32_t c, double re, double im) { s_v[r][c] = Cx{re, im}; });
                __syncthreads();
                ring.issue();  // into the slot just converted
            } else {
                __syncthreads();
                for (uint32_t i = threadIdx.x; i < RP * uint32_t(kCC); i += kThreads) {
                    const uint32_t r = i / kCC, c = i - r * kCC;
                    const uint32_t tt = tb + r;
                    Cx val = {0.0, 0.0};
                    if (tt < T && c < cn) {
                        const size_t e = ((size_t(tt) * nbl + bl) * C + c0 + c) * P;
                        bool finite = true;
                        double x0 = 0.0, y0 = 0.0;
                        for (uint32_t q = 0; q < P; ++q) {
                            const double xr = vis[2 * (e + q)], xi = vis[2 * (e + q) + 1];
                            finite = finite && __builtin_isfinite(xr) && __builtin_isfinite(xi);
                            if (q == 0) {
                                x0 = xr;
                                y0 = xi;
                            }
                        }
                        if (finite) val = {x0, y0};
                    }
                    s_v[r][c] = val;
                }
                __syncthreads();
            }
            if (!row_ok || !any_dir) continue;
            // Whole chunks: only a row's first chunk can be partial; its entries past cn are
            // zero and its recurrences start at +0, which zero entries leave at +0.
            for (int k = kCC - 1; k >= 0; --k) {
                const Cx x = s_v[rsub][k];
#pragma unroll
                for (int j = 0; j < ND; ++j) {
                    dp[j] = hstep(dp[j], z[j], p[j]);  // P' <- P' z + P (uses P before its update)
                    p[j] = hstep(p[j], z[j], x);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double gsf = 0.0;
            if (row_ok && dir_ok[j]) {
                const Cx e0 = expi(sf[j] * kappa0);
                const Cx zdp = cmulx(z[j], dp[j]);
                const Cx s1 = cmulx(e0, Cx{kappa0 * p[j].x + dkappa * zdp.x, kappa0 * p[j].y + dkappa * zdp.y});
                // Re(conj(ab) * i * s1) with conj(ab) = (ab.x, -ab.y) and i * s1 = (-s1.y, s1.x)
                gsf = -ab[j].x * s1.y + ab[j].y * s1.x;
                gdz[j][0] += -kTwoPi * gsf * u;
                gdz[j][1] += -kTwoPi * gsf * v;
                gdz[j][2] += -kTwoPi * gsf * w;
            }
            // g_uvw(row) = sum over directions of -2 pi gsf (dz - e_z), reduced in LDS.
            s_g[rsub][lane_d + DL * j][0] = -kTwoPi * gsf * dz[j][0];
            s_g[rsub][lane_d + DL * j][1] = -kTwoPi * gsf * dz[j][1];
            s_g[rsub][lane_d + DL * j][2] = -kTwoPi * gsf * (dz[j][2] - 1.0);
        }
        __syncthreads();
        if (lane_d < 3 && row_ok) {
            double s = 0.0;
            for (int j = 0; j < DW; ++j) s += s_g[rsub][j][lane_d];  // direction order
            g_uvw[3 * (size_t(t) * nbl + bl) + lane_d] = s;
        }
    }
    // Partials of this baseline's rows per rsub lane; the RP lanes are summed in order below.
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ND; ++j)
        for (int k = 0; k < 3; ++k) s_g[rsub][lane_d + DL * j][k] = dir_ok[j] ? gdz[j][k] : 0.0;
    __syncthreads();
    if (rsub == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            if (!dir_ok[j]) continue;
            const uint32_t dj = lane_d + DL * j;
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int r = 0; r < RP; ++r) s += s_g[r][dj][k];
                g_dz_partial[(size_t(bl) * D + dj) * 3 + k] = s;
            }
        }
    }
}

// The same outputs from the forward pass's per-row sums q[row][d] = (e0 P(z),
// e0 z P'(z)) (rotate_average_horner64 with row sums), so the cube is not read
// again: s1 = kappa0 q0 + dkappa q1 is the sum above.  Lanes, rows and the
// reductions as in rotate_adjoint_kernel.
template <int DL, int ND>
__global__ __launch_bounds__(kThreads) void adjoint_q_kernel(const double* __restrict__ q,
                                                             const double* __restrict__ uvw,
                                                             const double* __restrict__ dd,
                                                             const double* __restrict__ abar, uint32_t D, uint32_t T,
                                                             uint32_t nbl, double kappa0, double dkappa,
                                                             double* __restrict__ g_uvw,
                                                             double* __restrict__ g_dz_partial) {
    constexpr int RP = kThreads / DL;
    constexpr int DW = DL * ND;
    __shared__ double s_g[RP][DW][3];
    const uint32_t bl = blockIdx.x;
    const uint32_t lane_d = threadIdx.x % DL;
    const uint32_t rsub = threadIdx.x / DL;
    bool dir_ok[ND];
    double dz[ND][3], gdz[ND][3];
    Cx ab[ND];
#pragma unroll
    for (int j = 0; j < ND; ++j) {
        const uint32_t dj = lane_d + DL * j;
        dir_ok[j] = dj < D;
        const uint32_t dc = dir_ok[j] ? dj : 0;
        dz[j][0] = dd[9 * dc + 6];
        dz[j][1] = dd[9 * dc + 7];
        dz[j][2] = dd[9 * dc + 8];
        ab[j] = Cx{0.0, 0.0};
        if (dir_ok[j]) ab[j] = {abar[2 * (size_t(dc) * nbl + bl)], abar[2 * (size_t(dc) * nbl + bl) + 1]};
        gdz[j][0] = gdz[j][1] = gdz[j][2] = 0.0;
    }
    for (uint32_t tb = 0; tb < T; tb += RP) {
        const uint32_t t = tb + rsub;
        const bool row_ok = t < T;
        const size_t row = size_t(t) * nbl + bl;
        double u = 0.0, v = 0.0, w = 0.0;
        if (row_ok) {
            u = uvw[3 * row];
            v = uvw[3 * row + 1];
            w = uvw[3 * row + 2];
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double gsf = 0.0;
            if (row_ok && dir_ok[j]) {
                const double* qp = q + (row * D + lane_d + DL * j) * 4;
                const Cx s1 = {kappa0 * qp[0] + dkappa * qp[2], kappa0 * qp[1] + dkappa * qp[3]};
                gsf = -ab[j].x * s1.y + ab[j].y * s1.x;
                gdz[j][0] += -kTwoPi * gsf * u;
                gdz[j][1] += -kTwoPi * gsf * v;
                gdz[j][2] += -kTwoPi * gsf * w;
            }
            s_g[rsub][lane_d + DL * j][0] = -kTwoPi * gsf * dz[j][0];
            s_g[rsub][lane_d + DL * j][1] = -kTwoPi * gsf * dz[j][1];
            s_g[rsub][lane_d + DL * j][2] = -kTwoPi * gsf * (dz[j][2] - 1.0);
        }
        __syncthreads();
        if (lane_d < 3 && row_ok) {
            double s = 0.0;
            for (int j = 0; j < DW; ++j) s += s_g[rsub][j][lane_d];  // direction order
            g_uvw[3 * row + lane_d] = s;
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < ND; ++j)
        for (int k = 0; k < 3; ++k) s_g[rsub][lane_d + DL * j][k] = dir_ok[j] ? gdz[j][k] : 0.0;
    __syncthreads();
    if (rsub == 0) {
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            if (!dir_ok[j]) continue;
            const uint32_t dj = lane_d + DL * j;
            for (int k = 0; k < 3; ++k) {
                double s = 0.0;
                for (int r = 0; r < RP; ++r) s += s_g[r][dj][k];
                g_dz_partial[(size_t(bl) * D + dj) * 3 + k] = s;
            }
        }
    }
}

// g_dz[d][k] = sum over baselines of g_dz_partial[b][d][k]: one workgroup per output
// i = 3 d + k, each lane a strided slice of the baselines, then a fixed-order tree in LDS.
__global__ __launch_bounds__(256) void reduce_gdz_kernel(const double* __restrict__ part, uint32_t nbl, uint32_t D,
                                                         double* __restrict__ out) {
    __shared__ double s_red[256];
    const uint32_t i = blockIdx.x, n_out = 3 * D;
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < nbl; b += 256) acc += part[size_t(b) * n_out + i];
    s_red[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s_red[threadIdx.x] += s_red[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = s_red[0];
}

// dL/d(ra, dec) of direction d from g_dz[d] = dL/d(row 2 of its DD matrix): the
// derivatives of dd row 2 = R_x(-pi/2 + dec) R_z(ra - ra0) R_x(pi/2 - dec0) row 2
// (MetaData.cc:166-203) in ra and dec; dirs[d] = (ra, dec).  One thread per direction.
__global__ void gdz_offsets_kernel(const double* __restrict__ g_dz, uint32_t D, double ra0, double dec0,
                                   const double* __restrict__ dirs, double* __restrict__ out) {
#pragma clang fp contract(off)
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    double s1, c1, s2, c2, s3, c3;
    sincos(kPi / 2.0 - dec0, &s1, &c1);
    sincos(dirs[2 * d] - ra0, &s2, &c2);
    sincos(-kPi / 2.0 + dirs[2 * d + 1], &s3, &c3);
    const double dra[3] = {-s3 * c2, -s3 * s2 * c1, s3 * s2 * s1};
    const double ddec[3] = {-c3 * s2, c3 * c2 * c1 - s3 * s1, -c3 * c2 * s1 - s3 * c1};
    const double* g = g_dz + 3 * size_t(d);
    out[2 * size_t(d)] = g[0] * dra[0] + g[1] * dra[1] + g[2] * dra[2];
    out[2 * size_t(d) + 1] = g[0] * ddec[0] + g[1] * ddec[1] + g[2] * ddec[2];
}

// For the n = D nbl (direction, baseline) entries of P pols each: out[2 i], out[2 i + 1] = pol 0
// of avg[i], and out[2 n + i] = its argument atan2(im, re).
__global__ void gather_pol0_kernel(const double* __restrict__ avg, uint32_t n, uint32_t P, double* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double re = avg[size_t(i) * P * 2], im = avg[size_t(i) * P * 2 + 1];
    out[2 * size_t(i)] = re;
    out[2 * size_t(i) + 1] = im;
    out[2 * size_t(n) + i] = atan2(im, re);
}

// One workgroup per antenna: sum over its baselines (sign +1 as antenna2, -1 as
// antenna1) and timesteps of M_t^T g_uvw(t, b).
__global__ __launch_bounds__(256) void antenna_grad_kernel(const double* __restrict__ g_uvw,
                                                           const double* __restrict__ M,
                                                           const int32_t* __restrict__ csr_off,
                                                           const int32_t* __restrict__ csr_bl,
                                                           const int32_t* __restrict__ csr_sign, uint32_t nbl,
                                                           uint32_t T, double* __restrict__ out) {
    __shared__ double s_red[256][3];
    const int32_t a = blockIdx.x;
    const int32_t j0 = csr_off[a], j1 = csr_off[a + 1];
    const uint32_t nj = uint32_t(j1 - j0);
    double acc[3] = {0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < nj * T; i += 256) {
        const uint32_t t = i / nj, j = i - t * nj;
        const int32_t b = csr_bl[j0 + j];
        const double sgn = double(csr_sign[j0 + j]);
        const double* g = g_uvw + 3 * (size_t(t) * nbl + b);
        const double* m = M + 9 * size_t(t);
        for (int k = 0; k < 3; ++k) acc[k] += sgn * (m[k] * g[0] + m[3 + k] * g[1] + m[6 + k] * g[2]);
    }
    for (int k = 0; k < 3; ++k) s_red[threadIdx.x][k] = acc[k];
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (int(threadIdx.x) < stride)
            for (int k = 0; k < 3; ++k) s_red[threadIdx.x][k] += s_red[threadIdx.x + stride][k];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[3 * a + threadIdx.x] = s_red[0][threadIdx.x];
}

// ---------------------------------------------------------------------------
// The inversion, loss and their adjoint for D directions (layouts [d][...]).
//
// Forward (CpuLeapCalibrator.cc:159-178, then the loss): with ph[b] = arg avg[b]
// (pol 0), cal1 = Ad1 phi1, dphi, cal = Ad [dphi; 0] + cal1 from the inversion
// kernels, the residual of row n < K - 1 is
//   res[n] = wrap(cal[a1(n)] - cal[a2(n)] - ph[I(n)]),   loss = sum_n res[n]^2.
// Backward, with Ad = G+ A^T (G+ the pseudo-inverse of the Gram matrix A^T A):
//   g_arg[I(n)] = -2 res[n];  g_cal = A_top^T (2 res)
//   y = G+ g_cal;  g_arg[n] += (A y)[n] = y[a1(n)] - y[a2(n)]       (row n of avgData)
//   g_cal1 = g_cal + A_top^T (-2 pi (A y))
//   y1 = G1+ g_cal1;  g_arg[I1(k)] += y1[a1'(k)] - y1[a2'(k)]
//   abar[b] = i g_arg[b] / conj(avg[b])   (d arg z = Im(dz / z))
// The antenna sums A^T v are gathers over each antenna's rows (CSR, in row
// order), so every sum has a fixed order and no atomics are used.
// ---------------------------------------------------------------------------
__device__ inline double wrap_pi(double x) {
#pragma clang fp contract(off)
    return x - kTwoPi * floor((x + kPi) / kTwoPi);
}

// res[d][n] and per-workgroup sums of res^2 (fixed-order tree) into part[d][block].
__global__ __launch_bounds__(256) void inv_residual_kernel(const double* __restrict__ cal, int32_t N,
                                                           const int2* __restrict__ pairs, const int32_t* __restrict__ I,
                                                           int32_t nI, const double* __restrict__ ph, uint32_t nbl,
                                                           double* __restrict__ res, double* __restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double s_red[256];
    const int32_t n = blockIdx.x * 256 + threadIdx.x;
    const uint32_t d = blockIdx.y;
    double r2 = 0.0;
    if (n < nI) {
        const int2 p = pairs[n];
        const double* c = cal + size_t(d) * N;
        const double r = wrap_pi(c[p.x] - c[p.y] - ph[size_t(d) * nbl + I[n]]);
        res[size_t(d) * nI + n] = r;
        r2 = r * r;
    }
    s_red[threadIdx.x] = r2;
    __syncthreads();
    for (uint32_t stride = 128; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) s_red[threadIdx.x] += s_red[threadIdx.x + stride];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[size_t(d) * gridDim.x + blockIdx.x] = s_red[0];
}

// loss[d] = sum of the workgroup sums in block order.
__global__ void inv_loss_kernel(const double* __restrict__ part, uint32_t blocks, uint32_t D,
                                double* __restrict__ loss) {
    const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    double s = 0.0;
    for (uint32_t k = 0; k < blocks; ++k) s += part[size_t(d) * blocks + k];
    loss[d] = s;
}

// g_arg[d][b] = 0 for every baseline, then -2 res[n] at b = I(n) (I is one-to-one).
__global__ void inv_garg_init_kernel(double* __restrict__ g_arg, uint32_t nbl) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nbl) g_arg[size_t(blockIdx.y) * nbl + b] = 0.0;
}

__global__ void inv_garg_res_kernel(const double* __restrict__ res, const int32_t* __restrict__ I, int32_t nI,
                                    uint32_t nbl, double* __restrict__ g_arg) {
#pragma clang fp contract(off)
    const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    if (n >= nI) return;
    double& g = g_arg[size_t(d) * nbl + I[n]];
    g -= 2.0 * res[size_t(d) * nI + n];
}

// Sum of the 64 lanes' values of a wave by a butterfly of xor shuffles: a fixed
// order, so every lane gets the same bits on every run.
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// out[d][a] = (base ? base[d][a] : 0) + sum over antenna a's rows e of
// sign_e * scale * src[d][row_e]; csr_ent[e] = 2 row + (sign < 0).  One wave
// per (a, d) (four per workgroup): lane l sums entries l, l + 64, ... in order,
// then the lanes are summed by wave_sum.
__global__ __launch_bounds__(256) void inv_antenna_sum_kernel(const double* __restrict__ src, int32_t nI,
                                                              double scale, const int32_t* __restrict__ csr_off,
                                                              const int32_t* __restrict__ csr_ent, int32_t N,
                                                              const double* __restrict__ base,
                                                              double* __restrict__ out) {
#pragma clang fp contract(off)
    const int32_t a = int32_t(blockIdx.x * 4 + threadIdx.x / 64);
    const uint32_t lane = threadIdx.x % 64, d = blockIdx.y;
    if (a >= N) return;  // whole waves
    const double* x = src + size_t(d) * nI;
    double acc = 0.0;
    for (int32_t e = csr_off[a] + int32_t(lane); e < csr_off[a + 1]; e += 64) {
        const int32_t v = csr_ent[e];
        const double t = scale * x[v >> 1];
        if (v & 1)
            acc -= t;
        else
            acc += t;
    }
    acc = wave_sum(acc);
    if (lane == 0) out[size_t(d) * N + a] = (base ? base[size_t(d) * N + a] : 0.0) + acc;
}

// y[d][a] = sum_i Gp[a N + i] g[d][i] (Gp is symmetric, so column a is row a):
// one wave per (a, d), lane l takes i = l, l + 64, ..., then wave_sum.
__global__ __launch_bounds__(256) void inv_gram_gemv_kernel(const double* __restrict__ Gp, int32_t N,
                                                            const double* __restrict__ g, double* __restrict__ y) {
#pragma clang fp contract(off)
    const int32_t a = int32_t(blockIdx.x * 4 + threadIdx.x / 64);
    const uint32_t lane = threadIdx.x % 64, d = blockIdx.y;
    if (a >= N) return;  // whole waves
    const double* col = Gp + size_t(a) * N;
    const double* x = g + size_t(d) * N;
    double t = 0.0;
    for (int32_t i = int32_t(lane); i < N; i += 64) t += col[i] * x[i];
    t = wave_sum(t);
    if (lane == 0) y[size_t(d) * N + a] = t;
}

// g = y[a1(n)] - y[a2(n)]: g_arg[d][n] += g (row n of avgData), gacal1[d][n] = -2 pi g.
__global__ void inv_rows_kernel(const double* __restrict__ y, int32_t N, const int2* __restrict__ pairs, int32_t nI,
                                uint32_t nbl, double* __restrict__ g_arg, double* __restrict__ gacal1) {
#pragma clang fp contract(off)
    const int32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    if (n >= nI) return;
    const int2 p = pairs[n];
    const double g = y[size_t(d) * N + p.x] - y[size_t(d) * N + p.y];
    g_arg[size_t(d) * nbl + n] += g;
    gacal1[size_t(d) * nI + n] = -kTwoPi * g;
}

// g_arg[d][I1(k)] += y1[a1'(k)] - y1[a2'(k)] (I1 is one-to-one).
__global__ void inv_rows1_kernel(const double* __restrict__ y1, int32_t N, const int2* __restrict__ pairs1,
                                 const int32_t* __restrict__ I1, int32_t nI1, uint32_t nbl,
                                 double* __restrict__ g_arg) {
#pragma clang fp contract(off)
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t d = blockIdx.y;
    if (k >= nI1) return;
    const int2 p = pairs1[k];
    int32_t b = I1[k];
    if (b < 0) b += int32_t(nbl);
    g_arg[size_t(d) * nbl + b] += y1[size_t(d) * N + p.x] - y1[size_t(d) * N + p.y];
}

// abar[d][b] = i g / conj(z) = (-g y / |z|^2, g x / |z|^2), 0 where z = 0; z = avg0[d][b].
__global__ void inv_abar_kernel(const double* __restrict__ avg0, const double* __restrict__ g_arg, uint32_t n,
                                double* __restrict__ abar) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = avg0[2 * size_t(i)], y = avg0[2 * size_t(i) + 1], m = x * x + y * y;
    const double g = g_arg[i];
    abar[2 * size_t(i)] = m == 0.0 ? 0.0 : -g * y / m;
    abar[2 * size_t(i) + 1] = m == 0.0 ? 0.0 : g * x / m;
}

}  // namespace

void launch_inversion_adjoint(hipStream_t s, const InversionAdjoint& a) {
    const int32_t nI = a.K - 1, nI1 = a.K1 - 1;
    const uint32_t D = a.D, nbl = a.nbl;
    const uint32_t rb = uint32_t(std::max(1, (nI + 255) / 256));
    inv_residual_kernel<<<dim3(rb, D), 256, 0, s>>>(a.cal, a.N, a.pairs, a.I, nI, a.ph, nbl, a.res, a.part);
    LEAP_HIP(hipGetLastError());
    inv_loss_kernel<<<(D + 63) / 64, 64, 0, s>>>(a.part, rb, D, a.loss);
    LEAP_HIP(hipGetLastError());
    if (!a.abar) return;
    inv_garg_init_kernel<<<dim3((nbl + 255) / 256, D), 256, 0, s>>>(a.g_arg, nbl);
    LEAP_HIP(hipGetLastError());
    const dim3 gn((std::max(1, nI) + 255) / 256, D), ga((a.N + 3) / 4, D);
    if (nI > 0) {
        inv_garg_res_kernel<<<gn, 256, 0, s>>>(a.res, a.I, nI, nbl, a.g_arg);
        LEAP_HIP(hipGetLastError());
    }
    inv_antenna_sum_kernel<<<ga, 256, 0, s>>>(a.res, nI, 2.0, a.csr_off, a.csr_ent, a.N, nullptr, a.g_cal);
    LEAP_HIP(hipGetLastError());
    inv_gram_gemv_kernel<<<ga, 256, 0, s>>>(a.Gp, a.N, a.g_cal, a.y);
    LEAP_HIP(hipGetLastError());
    if (nI > 0) {
        inv_rows_kernel<<<gn, 256, 0, s>>>(a.y, a.N, a.pairs, nI, nbl, a.g_arg, a.gacal1);
        LEAP_HIP(hipGetLastError());
    }
    inv_antenna_sum_kernel<<<ga, 256, 0, s>>>(a.gacal1, nI, 1.0, a.csr_off, a.csr_ent, a.N, a.g_cal, a.g_cal1);
    LEAP_HIP(hipGetLastError());
    inv_gram_gemv_kernel<<<ga, 256, 0, s>>>(a.G1p, a.N, a.g_cal1, a.y1);
    LEAP_HIP(hipGetLastError());
    if (nI1 > 0) {
        inv_rows1_kernel<<<dim3((nI1 + 255) / 256, D), 256, 0, s>>>(a.y1, a.N, a.pairs1, a.I1, nI1, nbl, a.g_arg);
        LEAP_HIP(hipGetLastError());
    }
    const uint32_t n = D * nbl;
    inv_abar_kernel<<<(n + 255) / 256, 256, 0, s>>>(a.avg0, a.g_arg, n, a.abar);
    LEAP_HIP(hipGetLastError());
}

void launch_perturb_uvw(hipStream_t s, const double* uvw0, const double* M, const int2* bl_ants, const double* dp,
                        uint32_t nbl, uint32_t rows, double* uvw) {
    perturb_uvw_kernel<<<(rows + 255) / 256, 256, 0, s>>>(uvw0, M, bl_ants, dp, nbl, rows, uvw);
    LEAP_HIP(hipGetLastError());
}

void launch_rotate_adjoint(hipStream_t s, const void* vis, int32_t vis_dtype, const double* uvw, const double* dd,
                           const double* abar, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double f0,
                           double df, double* g_uvw, double* g_dz_partial) {
    const double kappa0 = f0 / kSpeedOfLight, dkappa = df / kSpeedOfLight;
    const bool c64 = vis_dtype == LEAP_VIS_C64;
    require(D >= 1 && D <= 256, LEAP_ERR_NOT_IMPLEMENTED, "the adjoint handles 1 to 256 directions per call");
    const dim3 grid(nbl);
#define LEAP_ADJ(DL_, ND_)                                                                                     \
    if (c64)                                                                                                   \
        rotate_adjoint_kernel<float, DL_, ND_><<<grid, kThreads, 0, s>>>(static_cast<const float*>(vis), uvw, dd, \
                                                                         abar, D, T, nbl, C, P, kappa0, dkappa,   \
                                                                         g_uvw, g_dz_partial);                    \
    else                                                                                                       \
        rotate_adjoint_kernel<double, DL_, ND_><<<grid, kThreads, 0, s>>>(static_cast<const double*>(vis), uvw,   \
                                                                          dd, abar, D, T, nbl, C, P, kappa0,       \
                                                                          dkappa, g_uvw, g_dz_partial)
    if (D <= 16)
        LEAP_ADJ(8, 2);
    else if (D <= 32)
        LEAP_ADJ(32, 1);
    else if (D <= 64)
        LEAP_ADJ(64, 1);
    else if (D <= 128)
        LEAP_ADJ(128, 1);
    else
        LEAP_ADJ(256, 1);
#undef LEAP_ADJ
    LEAP_HIP(hipGetLastError());
}

void launch_adjoint_q(hipStream_t s, const double* q, const double* uvw, const double* dd, const double* abar,
                      uint32_t D, uint32_t T, uint32_t nbl, double f0, double df, double* g_uvw,
                      double* g_dz_partial) {
    const double kappa0 = f0 / kSpeedOfLight, dkappa = df / kSpeedOfLight;
    require(D >= 1 && D <= 256, LEAP_ERR_NOT_IMPLEMENTED, "the adjoint handles 1 to 256 directions per call");
    const dim3 grid(nbl);
#define LEAP_ADJ(DL_, ND_)                                                                                         \
    adjoint_q_kernel<DL_, ND_><<<grid, kThreads, 0, s>>>(q, uvw, dd, abar, D, T, nbl, kappa0, dkappa, g_uvw, \
                                                         g_dz_partial)
    if (D <= 16)
        LEAP_ADJ(16, 1);
    else if (D <= 32)
        LEAP_ADJ(32, 1);
    else if (D <= 64)
        LEAP_ADJ(64, 1);
    else if (D <= 128)
        LEAP_ADJ(128, 1);
    else
        LEAP_ADJ(256, 1);
#undef LEAP_ADJ
    LEAP_HIP(hipGetLastError());
}

void launch_gather_pol0(hipStream_t s, const double* avg, uint32_t n, uint32_t P, double* out) {
    if (n == 0) return;
    gather_pol0_kernel<<<(n + 255) / 256, 256, 0, s>>>(avg, n, P, out);
    LEAP_HIP(hipGetLastError());
}

void launch_reduce_gdz_offsets(hipStream_t s, const double* part, uint32_t nbl, uint32_t D, double ra0, double dec0,
                               const double* dirs, double* out) {
    reduce_gdz_kernel<<<D * 3, 256, 0, s>>>(part, nbl, D, out);
    LEAP_HIP(hipGetLastError());
    gdz_offsets_kernel<<<(D + 63) / 64, 64, 0, s>>>(out, D, ra0, dec0, dirs, out + 3 * size_t(D));
    LEAP_HIP(hipGetLastError());
}

void launch_antenna_grad(hipStream_t s, const double* g_uvw, const double* M, const int32_t* csr_off,
                         const int32_t* csr_bl, const int32_t* csr_sign, int32_t n_ant, uint32_t nbl, uint32_t T,
                         double* out) {
    antenna_grad_kernel<<<n_ant, 256, 0, s>>>(g_uvw, M, csr_off, csr_bl, csr_sign, nbl, T, out);
    LEAP_HIP(hipGetLastError());
}

}  // namespace leap
