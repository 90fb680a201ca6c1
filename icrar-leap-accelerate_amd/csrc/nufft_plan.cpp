// Plan of the NUFFT channel sums (mixed precision).
//
// For one row the channel sum of a direction is a trigonometric polynomial in
// b = sf * dkappa:
//   sum_c v_c exp(i sf (kappa0 + c dkappa)) = exp(i sf kappa0) F(b),
//   F(b) = sum_{c<C} v_c exp(i c b) = exp(i h b) sum_{k=-h}^{C-1-h} v_{k+h} exp(i k b),
// and F is evaluated at one b per direction by a type-2 non-uniform FFT
// (Dutt & Rokhlin 1993; the "exponential of semicircle" kernel of Barnett,
// Magland & af Klinteberg, SIAM J. Sci. Comput. 41 (2019) C479-C504):
//   1. f~_k = v_{k+h} * (2 pi / N) / psi^(k)      (precorrection, per channel)
//   2. g_m = sum_k f~_k exp(2 pi i k m / N), m < N  (one N-point inverse FFT per row)
//   3. F(b) ~= exp(i h b) sum_{j<W} g_{(m0+j) mod N} phi((u - m0 - j) / (W/2)),
//      u = N * frac(b / 2 pi), m0 = ceil(u - W/2),
// with phi(z) = exp(beta (sqrt(1 - z^2) - 1)) on |z| < 1, psi(x) = phi(x N / (W pi))
// and psi^ its Fourier transform.  The offset h is C / 2 rounded to a multiple
// of 64, so channel c's grid index is congruent to c modulo the wavefront
// width (each lane's channels land in its own FFT slots); N is the power of two
// >= 2C (at least 256), which keeps |k| <= N / 4; with sigma = N / (2 max|k|)
// >= 2 and beta = 0.97 pi W (1 - 1 / (2 sigma)):
//   FP64 transform: W = 10 for sigma >= 2.5, else 12.  The relative error of F
//     against sum |v_c| stays below 1e-10 (about 1e-11 with W = 10, 3e-12 with
//     W = 12) for every C up to 512.
//   FP32 transform: W = 8 for sigma >= 2.5, else 10 (the kernel's own error,
//     about 3e-9 and 2e-10, then stays below the FP32 transform's rounding,
//     about 2e-8 of sum |v_c|).
//   Width-16 plan (FP64 transform): W = 16, exp(-beta) from 1.3e-16 (sigma =
//     2) down to 2e-18, below FP64 rounding: with the evaluation point given
//     in turns the error is about 1e-15 of sum |v_c| for every C up to 512,
//     where FP64 Horner chains reach n eps (about 2e-14 at C = 384).
//
// Step 3's weights come from polynomials in s = u - m0 - (W - 1) / 2 in
// (-1/2, 1/2]: tap j's weight is phi((s + (W-1)/2 - j) / (W/2)) = P_j(s), and
// tap W-1-j's is P_j(-s) (phi is even), so only taps j < W/2 are fitted.  Each
// P_j is the degree-(W-1) Chebyshev interpolant on W Chebyshev points; its
// error is at the level of phi's truncation jump exp(-beta).
#include <algorithm>
#include <cmath>
#include <vector>

#include "leap_internal.h"

namespace leap {

namespace {

double es_kernel(double z, double beta) {
    if (!(std::fabs(z) < 1.0)) return 0.0;
    return std::exp(beta * (std::sqrt((1.0 - z) * (1.0 + z)) - 1.0));
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton's method on P_n from
// the Chebyshev-like first guesses; Golub & Welsch give the same rule).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
    x.assign(size_t(n), 0.0);
    w.assign(size_t(n), 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[size_t(i)] = -z;
        x[size_t(n - 1 - i)] = z;
        w[size_t(i)] = w[size_t(n - 1 - i)] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

}  // namespace

bool nufft_supported(uint32_t C) { return C >= 1 && C <= kNufftMaxChannels; }

std::vector<double> nufft_tables(const NufftPlan& p) {
    const size_t C = p.scale.size();
    std::vector<double> t(&p.coef[0][0], &p.coef[0][0] + kNufftCoefDoubles);
    t.insert(t.end(), p.scale.begin(), p.scale.end());
    t.resize(nufft_coef32_offset(uint32_t(C)) + (kNufftCoefDoubles + 1) / 2, 0.0);
    float* f = reinterpret_cast<float*>(t.data() + nufft_coef32_offset(uint32_t(C)));
    for (int i = 0; i < kNufftCoefDoubles; ++i) f[i] = float((&p.coef[0][0])[i]);
    return t;
}

NufftPlan make_nufft_plan(uint32_t C, bool fp32, bool w16) {
    require(nufft_supported(C), LEAP_ERR_INVALID_ARGUMENT,
            "the NUFFT channel sums support 1.." + std::to_string(kNufftMaxChannels) + " channels, not " +
                std::to_string(C));
    require(!(fp32 && w16), LEAP_ERR_INTERNAL, "the width-16 NUFFT plan has an FP64 transform");
    NufftPlan p;
    uint32_t N = 256;
    while (N < 2 * C) N *= 2;
    p.N = N;
    p.h = 64 * ((C + 64) / 128);  // 64 * round(C / 128)
    const uint32_t kmax = std::max<uint32_t>(1, std::max(p.h, C - 1 - std::min(C - 1, p.h)));
    const double sigma = double(N) / (2.0 * kmax);
    p.fp32 = fp32;
    p.w16 = w16;
    if (w16)
        p.W = 16;
    else if (fp32)
        p.W = sigma >= 2.5 ? 8 : 10;
    else
        p.W = sigma >= 2.5 ? 10 : 12;
    p.beta = 0.97 * M_PI * p.W * (1.0 - 1.0 / (2.0 * sigma));
    // psi^(k) = (W pi / N) int_{-1}^{1} phi(z) cos(k z W pi / N) dz
    std::vector<double> gx, gw;
    gauss_legendre(kNufftQuadratureNodes, gx, gw);
    const double half = p.W * M_PI / N;
    p.scale.resize(C);
    for (uint32_t c = 0; c < C; ++c) {
        const double k = double(int64_t(c) - int64_t(p.h));
        double acc = 0.0;
        for (size_t q = 0; q < gx.size(); ++q) acc += gw[q] * es_kernel(gx[q], p.beta) * std::cos(k * gx[q] * half);
        p.scale[c] = (2.0 * M_PI / N) / (half * acc);
    }
    // P_j(s) for j < W/2: Chebyshev interpolation in x = 2 s on [-1, 1], then monomials in s.
    const int W = int(p.W), n = W;  // n nodes, degree n - 1
    for (int j = 0; j < W / 2; ++j) {
        const size_t ns = size_t(n);
        std::vector<double> fx(ns), a(ns, 0.0);
        for (int i = 0; i < n; ++i) {
            const double x = std::cos(M_PI * (i + 0.5) / n);
            fx[size_t(i)] = es_kernel((x / 2 + (W - 1) / 2.0 - j) / (W / 2.0), p.beta);
        }
        for (int k = 0; k < n; ++k) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += fx[size_t(i)] * std::cos(M_PI * k * (i + 0.5) / n);
            a[size_t(k)] = (k == 0 ? 1.0 : 2.0) * s / n;
        }
        // sum_k a_k T_k(x) in monomials of x: T_0 = 1, T_1 = x, T_{k+1} = 2 x T_k - T_{k-1}
        std::vector<double> mono(size_t(n), 0.0), tkm1(size_t(n), 0.0), tk(size_t(n), 0.0);
        tkm1[0] = 1.0;
        mono[0] += a[0];
        if (n > 1) {
            tk[1] = 1.0;
            mono[1] += a[1];
        }
        for (int k = 2; k < n; ++k) {
            std::vector<double> t(size_t(n), 0.0);
            for (int m = 0; m < n; ++m) {
                if (m > 0) t[size_t(m)] += 2.0 * tk[size_t(m - 1)];
                t[size_t(m)] -= tkm1[size_t(m)];
            }
            for (int m = 0; m < n; ++m) mono[size_t(m)] += a[size_t(k)] * t[size_t(m)];
            tkm1 = tk;
            tk = t;
        }
        for (int m = 0; m < n; ++m) p.coef[j][m] = mono[size_t(m)] * std::ldexp(1.0, m);  // x^m = 2^m s^m
        for (int m = n; m < kNufftMaxWidth; ++m) p.coef[j][m] = 0.0;
    }
    return p;
}

}  // namespace leap
