/*
 * leap_oracle.cpp — CPU restatement of LEAP-Accelerate's calibration path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP calibrator.  Only
 * tests/, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product never calls it and it is never a fallback.
 *
 * Written from the reference's behaviour (ska-telescope/icrar-leap-accelerate
 * v0.8.0); every function cites the file:line it restates, relative to
 * src/icrar/leap-accelerate/.  Eigen, casacore and Boost are replaced by
 * plain loops; arithmetic stays FP64 and follows the reference's operation
 * order wherever that order is visible in its source.  Build without FMA
 * contraction (-ffp-contract=off), as the reference's x86-64 build is.
 *
 * Pins (see tests/test_oracle.py): the DD-matrix golden of MetaDataTests.cc:155-165,
 * the wavelength golden of MetaDataTests.cc:186, the pseudo-inverse known answers
 * of MatrixTests.cc:154-250, LAPACK (numpy) pseudo-inverses of incidence
 * matrices, brute-force PhaseMatrixFunction cases, the empty-input throw of
 * PhaseRotateTests.cc:248-274, closed forms of RotateVisibilities, and an
 * independent numpy restatement of the inversion.
 */
#include "leap_oracle.h"

#include <algorithm>
#include <limits>
#include <memory>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

struct oracle_error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;
constexpr double kSpeedOfLight = 299792458.0;  // common/constants.h:35

using cd = std::complex<double>;

// Column-major dense matrix.
struct Mat {
    int32_t rows = 0, cols = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(int32_t r, int32_t c) : rows(r), cols(c), v(size_t(r) * size_t(c), 0.0) {}
    double& operator()(int32_t i, int32_t j) { return v[size_t(j) * rows + i]; }
    double operator()(int32_t i, int32_t j) const { return v[size_t(j) * rows + i]; }
};

}  // namespace

struct oracle_metadata {
    // Constants (model/cpu/MetaData.h:76-101)
    uint32_t nbaselines = 0;
    uint32_t reference_antenna = 0;
    uint32_t channels = 0;
    uint32_t num_pols = 0;
    uint32_t stations = 0;
    uint32_t rows = 0;
    double freq_start_hz = 0, freq_inc_hz = 0;
    double phase_centre_ra_rad = 0, phase_centre_dec_rad = 0;
    Mat A, Ad, A1, Ad1;
    std::vector<int32_t> I, I1;
};

namespace {

// ---------------------------------------------------------------------------
// PhaseMatrixFunction (algorithm/cpu/PhaseMatrixFunction.cc:35-93)
// ---------------------------------------------------------------------------
void phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n,
                  int32_t ref_ant, Mat& A, std::vector<int32_t>& I) {
    // :41-44 — the reference joins the size checks with &&; all three arrays
    // share n here, so the check cannot fire.
    // :46-51 — refAnt is compared with the number of baselines.
    if (ref_ant >= 0 && ref_ant >= n) {
        throw oracle_error("refAnt " + std::to_string(ref_ant) + " is out of bounds");
    }
    // :52-57 — fg is indexed by refAnt as if it were a baseline index.
    if (ref_ant >= 0 && fg[ref_ant]) {
        throw oracle_error("refAnt " + std::to_string(ref_ant) + " is flagged");
    }
    if (n <= 0) {
        // Eigen's maxCoeff() of an empty vector is undefined; refuse instead.
        throw oracle_error("a1 and a2 must not be empty");
    }
    // :59 — columns = max(a1, a2) + 1
    int32_t maxa = a1[0];
    for (int32_t i = 0; i < n; ++i) maxa = std::max(maxa, std::max(a1[i], a2[i]));
    A = Mat(n + 1, maxa + 1);
    std::vector<int32_t> Ifull(size_t(n), -1);  // :60-61
    int32_t k = 0;
    for (int32_t i = 0; i < n; ++i) {  // :65-79
        if (a1[i] != a2[i]) {
            const bool uses_ref = ref_ant < 0 || a1[i] == ref_ant || a2[i] == ref_ant;
            if (!fg[i] && uses_ref) {
                A(k, a1[i]) = 1;
                A(k, a2[i]) = -1;
                Ifull[size_t(k)] = i;
                ++k;
            }
        }
    }
    const int32_t constraint = ref_ant < 0 ? 0 : ref_ant;  // :81-84
    A(k, constraint) = 1;                                   // :86
    ++k;
    // :89-90 — conservativeResize keeps the top-left block.
    Mat trimmed(k, A.cols);
    for (int32_t j = 0; j < A.cols; ++j)
        for (int32_t i = 0; i < k; ++i) trimmed(i, j) = A(i, j);
    A = std::move(trimmed);
    I.assign(Ifull.begin(), Ifull.begin() + (k - 1));
}

// ---------------------------------------------------------------------------
// SVDPseudoInverse (math/cpu/matrix_invert.h:49-56)
//
// The reference takes a thin BDCSVD and keeps singular values above
// max(m, n) * eps * sigma_max.  Restated with a textbook thin SVD: Householder
// QR (A = QR) followed by one-sided (Hestenes) Jacobi on R (Golub & Van Loan,
// Matrix Computations, 4th ed., 5.2 and 8.6.4).  For m < n the SVD of A^T is
// taken and transposed back.
// ---------------------------------------------------------------------------
struct SVD {
    Mat U;                   // m x r
    std::vector<double> s;   // r, descending
    Mat V;                   // n x r
};

// Thin SVD for m >= n.
SVD thin_svd_tall(const Mat& Ain) {
    const int32_t m = Ain.rows, n = Ain.cols;
    Mat R = Ain;  // overwritten by Householder vectors below the diagonal
    std::vector<double> tau(size_t(n), 0.0);
    for (int32_t j = 0; j < n; ++j) {
        double norm2 = 0;
        for (int32_t i = j; i < m; ++i) norm2 += R(i, j) * R(i, j);
        const double norm = std::sqrt(norm2);
        if (norm == 0.0) { tau[size_t(j)] = 0.0; continue; }
        const double alpha = R(j, j) > 0 ? -norm : norm;
        // v = x - alpha e1, stored with v0 = 1 (scaled)
        const double v0 = R(j, j) - alpha;
        for (int32_t i = j + 1; i < m; ++i) R(i, j) /= v0;
        tau[size_t(j)] = -v0 / alpha;  // H = I - tau v v^T with v(0) = 1
        R(j, j) = alpha;
        for (int32_t c = j + 1; c < n; ++c) {
            double dot = R(j, c);
            for (int32_t i = j + 1; i < m; ++i) dot += R(i, j) * R(i, c);
            dot *= tau[size_t(j)];
            R(j, c) -= dot;
            for (int32_t i = j + 1; i < m; ++i) R(i, c) -= dot * R(i, j);
        }
    }
    // W = upper triangle of R (n x n)
    Mat W(n, n), V(n, n);
    for (int32_t j = 0; j < n; ++j) {
        for (int32_t i = 0; i <= j; ++i) W(i, j) = R(i, j);
        V(j, j) = 1.0;
    }
    // One-sided Jacobi: orthogonalise the columns of W, accumulating V.
    const double eps = std::numeric_limits<double>::epsilon();
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool rotated = false;
        for (int32_t p = 0; p < n - 1; ++p) {
            for (int32_t q = p + 1; q < n; ++q) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int32_t i = 0; i < n; ++i) {
                    const double wp = W(i, p), wq = W(i, q);
                    alpha += wp * wp;
                    beta += wq * wq;
                    gamma += wp * wq;
                }
                if (gamma == 0.0 || std::fabs(gamma) <= eps * std::sqrt(alpha * beta)) continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = c * t;
                for (int32_t i = 0; i < n; ++i) {
                    const double wp = W(i, p), wq = W(i, q);
                    W(i, p) = c * wp - s * wq;
                    W(i, q) = s * wp + c * wq;
                    const double vp = V(i, p), vq = V(i, q);
                    V(i, p) = c * vp - s * vq;
                    V(i, q) = s * vp + c * vq;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> sig(static_cast<size_t>(n));
    for (int32_t j = 0; j < n; ++j) {
        double s2 = 0;
        for (int32_t i = 0; i < n; ++i) s2 += W(i, j) * W(i, j);
        sig[size_t(j)] = std::sqrt(s2);
    }
    std::vector<int32_t> order(static_cast<size_t>(n));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return sig[size_t(a)] > sig[size_t(b)]; });
    SVD out;
    out.s.resize(size_t(n));
    out.V = Mat(n, n);
    Mat Ur(m, n);  // [U_R ; 0] then Q applied
    for (int32_t jj = 0; jj < n; ++jj) {
        const int32_t j = order[size_t(jj)];
        out.s[size_t(jj)] = sig[size_t(j)];
        for (int32_t i = 0; i < n; ++i) out.V(i, jj) = V(i, j);
        if (sig[size_t(j)] > 0) {
            for (int32_t i = 0; i < n; ++i) Ur(i, jj) = W(i, j) / sig[size_t(j)];
        }
    }
    // U = Q [U_R; 0]: apply H_0 ... H_{n-1} in reverse order.
    for (int32_t j = n - 1; j >= 0; --j) {
        const double tj = tau[size_t(j)];
        if (tj == 0.0) continue;
        for (int32_t c = 0; c < n; ++c) {
            double dot = Ur(j, c);
            for (int32_t i = j + 1; i < m; ++i) dot += R(i, j) * Ur(i, c);
            dot *= tj;
            Ur(j, c) -= dot;
            for (int32_t i = j + 1; i < m; ++i) Ur(i, c) -= dot * R(i, j);
        }
    }
    out.U = std::move(Ur);
    return out;
}

Mat transpose(const Mat& a) {
    Mat t(a.cols, a.rows);
    for (int32_t j = 0; j < a.cols; ++j)
        for (int32_t i = 0; i < a.rows; ++i) t(j, i) = a(i, j);
    return t;
}

SVD thin_svd(const Mat& A) {
    if (A.rows >= A.cols) return thin_svd_tall(A);
    SVD t = thin_svd_tall(transpose(A));  // A^T = U' S V'^T  =>  A = V' S U'^T
    SVD out;
    out.U = std::move(t.V);
    out.V = std::move(t.U);
    out.s = std::move(t.s);
    return out;
}

Mat pseudo_inverse(const Mat& A) {
    const int32_t m = A.rows, n = A.cols;
    Mat Ad(n, m);
    if (m == 0 || n == 0) return Ad;
    SVD svd = thin_svd(A);
    const int32_t r = int32_t(svd.s.size());
    // matrix_invert.h:54 — tolerance from the largest singular value.
    const double tol = double(std::max(m, n)) * std::numeric_limits<double>::epsilon() * std::fabs(svd.s[0]);
    std::vector<double> sinv(size_t(r), 0.0);
    for (int32_t j = 0; j < r; ++j) sinv[size_t(j)] = std::fabs(svd.s[size_t(j)]) > tol ? 1.0 / svd.s[size_t(j)] : 0.0;
    // :55 — V * diag(sinv) * U^T
    for (int32_t k = 0; k < m; ++k) {
        for (int32_t j = 0; j < r; ++j) {
            const double u = svd.U(k, j) * sinv[size_t(j)];
            if (u == 0.0) continue;
            for (int32_t i = 0; i < n; ++i) Ad(i, k) += svd.V(i, j) * u;
        }
    }
    return Ad;
}

// ---------------------------------------------------------------------------
// MetaData::GenerateDDMatrix (model/cpu/MetaData.cc:166-203)
// ---------------------------------------------------------------------------
void dd_matrix(double ra0, double dec0, double ra, double dec, double dd[9]) {
    const double ang1 = kPi / 2.0 - dec0;  // :169
    const double ang2 = ra - ra0;          // :170
    const double ang3 = -kPi / 2.0 + dec;  // :171
    const double dd1[9] = {1, 0, 0, 0, std::cos(ang1), -std::sin(ang1), 0, std::sin(ang1), std::cos(ang1)};
    const double dd2[9] = {std::cos(ang2), std::sin(ang2), 0, -std::sin(ang2), std::cos(ang2), 0, 0, 0, 1};
    const double dd3[9] = {1, 0, 0, 0, std::cos(ang3), -std::sin(ang3), 0, std::sin(ang3), std::cos(ang3)};
    double t[9];
    for (int i = 0; i < 3; ++i)  // :202 — (dd3 * dd2) * dd1
        for (int j = 0; j < 3; ++j)
            t[i * 3 + j] = dd3[i * 3 + 0] * dd2[0 * 3 + j] + dd3[i * 3 + 1] * dd2[1 * 3 + j] + dd3[i * 3 + 2] * dd2[2 * 3 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            dd[i * 3 + j] = t[i * 3 + 0] * dd1[0 * 3 + j] + t[i * 3 + 1] * dd1[1 * 3 + j] + t[i * 3 + 2] * dd1[2 * 3 + j];
}

// Constants::GetChannelWavelength (model/cpu/MetaData.h:95-98)
inline double channel_wavelength(double f0, double df, int32_t i) {
    return kSpeedOfLight / (f0 + i * df);
}

// ---------------------------------------------------------------------------
// MeasurementSet flag helpers (ms/MeasurementSet.cc:181-240)
// ---------------------------------------------------------------------------
void filtered_baselines(const oracle_ms_view& ms, double thr, std::vector<uint8_t>& out) {
    const uint32_t nbl = ms.num_baselines;
    out.assign(nbl, 0);
    // :181-197 — FLAG of the first epoch, channel 0, polarisation 0.
    for (uint32_t i = 0; i < nbl; ++i) out[i] = ms.flag[i] ? 1 : 0;
    // :205-228 — short baselines.  The reference reads UVW through a slicer
    // that keeps only u, into a matrix of shape (1, rows), and then indexes it
    // as uv(i, 0), uv(i, 1); with casacore's unchecked release indexing those
    // are the u of rows i and i + 1.  Restated as such; the read past the last
    // row (only possible with one timestep) is taken as 0.
    if (thr > 0.0) {
        for (uint32_t i = 0; i < nbl; ++i) {
            const double x = ms.uvw[size_t(i) * 3 + 0];
            const double y = (i + 1 < ms.num_rows) ? ms.uvw[size_t(i + 1) * 3 + 0] : 0.0;
            if (std::sqrt(x * x + y * y) < thr) out[i] = 1;
        }
    }
}

void check_view(const oracle_ms_view* ms) {
    if (!ms) throw oracle_error("null measurement set");
    if (ms->vis_on_device) throw oracle_error("oracle reads host memory only");
    if (ms->num_baselines == 0 || ms->num_rows % ms->num_baselines != 0)
        throw oracle_error("number of rows not an integer multiple of baselines");  // MeasurementSet.cc:95-102
    if (ms->num_rows < ms->num_baselines) throw oracle_error("invalid number of rows");
}

// ---------------------------------------------------------------------------
// MetaData::MetaData (model/cpu/MetaData.cc:42-137)
// ---------------------------------------------------------------------------
oracle_metadata* metadata_create(const oracle_ms_view& ms, int32_t ref_ant, double thr) {
    auto md = std::make_unique<oracle_metadata>();
    md->nbaselines = ms.num_baselines;                                                  // :50
    md->reference_antenna = ref_ant >= 0 ? uint32_t(ref_ant) : ms.total_antennas - 1;   // :51
    md->channels = ms.num_channels;                                                     // :58
    md->freq_start_hz = ms.freq_start_hz;                                               // :59
    md->freq_inc_hz = ms.freq_inc_hz;                                                   // :60
    md->rows = ms.num_rows;
    md->num_pols = ms.num_pols;
    md->stations = ms.num_stations;
    md->phase_centre_ra_rad = ms.phase_centre_ra_rad;                                   // :76
    md->phase_centre_dec_rad = ms.phase_centre_dec_rad;                                 // :77
    std::vector<uint8_t> filtered;
    filtered_baselines(ms, thr, filtered);                                              // :84
    const int32_t n = int32_t(ms.num_baselines);                                        // :87-89
    phase_matrix(ms.antenna1, ms.antenna2, filtered.data(), n, int32_t(md->reference_antenna), md->A1, md->I1);  // :92
    phase_matrix(ms.antenna1, ms.antenna2, filtered.data(), n, -1, md->A, md->I);        // :97
    md->Ad1 = pseudo_inverse(md->A1);                                                    // :111
    md->Ad = pseudo_inverse(md->A);                                                      // :123
    // :129-136 only log a warning when Ad*A is not the identity.
    return md.release();
}

template <typename T>
inline cd load_vis(const void* base, size_t idx) {
    const T* p = static_cast<const T*>(base) + 2 * idx;
    return cd(double(p[0]), double(p[1]));  // ms/utils.h:141-142 widens to double
}

// ---------------------------------------------------------------------------
// CpuLeapCalibrator::RotateVisibilities (algorithm/cpu/CpuLeapCalibrator.cc:181-217)
// ---------------------------------------------------------------------------
template <typename T>
void rotate_visibilities(const oracle_metadata& md, const oracle_ms_view& ms, uint32_t row_start,
                         uint32_t n_rows, const double dd[9], std::vector<cd>& avg) {
    const uint32_t nbl = md.nbaselines, C = md.channels, P = md.num_pols;
    avg.assign(size_t(nbl) * P, cd(0.0, 0.0));  // CpuLeapCalibrator.cc:137
    std::vector<cd> pol(P);
    for (uint32_t b = 0; b < n_rows; ++b) {  // :187 — b indexes the integration's rows
        const size_t row = size_t(row_start) + b;
        const uint32_t md_baseline = b % nbl;  // :189
        const double* uvw = ms.uvw + row * 3;
        // MetaData::CalcUVW (MetaData.cc:220-229): rotated = dd * uvw; only z is used.
        const double wrot = dd[6] * uvw[0] + dd[7] * uvw[1] + dd[8] * uvw[2];
        const double shift_factor = -kTwoPi * (wrot - uvw[2]);  // :190
        for (uint32_t c = 0; c < C; ++c) {
            const double shift_rad = shift_factor / channel_wavelength(md.freq_start_hz, md.freq_inc_hz, int32_t(c));  // :195
            const cd ph = std::exp(cd(0.0, shift_rad));
            bool has_nan = false;
            for (uint32_t p = 0; p < P; ++p) {  // :196-199 (the product replaces the in-place write)
                pol[p] = load_vis<T>(ms.vis, (row * C + c) * P + p) * ph;
                has_nan |= std::isnan(pol[p].real()) || std::isnan(pol[p].imag());  // :201-206
            }
            if (!has_nan) {  // :208-214 — a sum, not a mean
                for (uint32_t p = 0; p < P; ++p) avg[size_t(md_baseline) * P + p] += pol[p];
            }
        }
    }
}

void rotate_dispatch(const oracle_metadata& md, const oracle_ms_view& ms, uint32_t row_start,
                     uint32_t n_rows, const double* direction, std::vector<cd>& avg) {
    if (size_t(row_start) + n_rows > ms.num_rows) throw oracle_error("row selection exceeds total range");  // utils.h:111-116
    double dd[9];
    dd_matrix(md.phase_centre_ra_rad, md.phase_centre_dec_rad, direction[0], direction[1], dd);  // MetaData.cc:205-213
    if (ms.vis_dtype == 1)
        rotate_visibilities<float>(md, ms, row_start, n_rows, dd, avg);
    else
        rotate_visibilities<double>(md, ms, row_start, n_rows, dd, avg);
}

// ---------------------------------------------------------------------------
// PhaseRotate, inversion half (algorithm/cpu/CpuLeapCalibrator.cc:159-178)
// ---------------------------------------------------------------------------
std::vector<double> apply_inversion(const oracle_metadata& md, const std::vector<cd>& avg) {
    const uint32_t P = md.num_pols;
    const int32_t nbl = int32_t(md.nbaselines);
    const int32_t N = md.A.cols;
    // :161-163 — arg of avgData(I1, pol 0), negative indices wrap (math/cpu/vector.h:88-107),
    // then a trailing 0 for the reference antenna.
    const int32_t K1 = md.A1.rows;
    std::vector<double> phase_i1(size_t(K1), 0.0);
    for (int32_t k = 0; k < K1 - 1; ++k) {
        int32_t idx = md.I1[size_t(k)];
        if (idx < 0) idx += nbl;
        phase_i1[size_t(k)] = std::arg(avg[size_t(idx) * P + 0]);  // math/vector_extensions.cc:28-31
    }
    // :165 — cal1 = Ad1 * phaseAnglesI1
    std::vector<double> cal1(size_t(N), 0.0);
    for (int32_t k = 0; k < K1; ++k)
        for (int32_t i = 0; i < N; ++i) cal1[size_t(i)] += md.Ad1(i, k) * phase_i1[size_t(k)];
    // :166 — ACal1 = A * cal1 (dense)
    const int32_t K = md.A.rows;
    std::vector<double> acal1(size_t(K), 0.0);
    for (int32_t j = 0; j < N; ++j)
        for (int32_t n = 0; n < K; ++n) acal1[size_t(n)] += md.A(n, j) * cal1[size_t(j)];
    // :168-176 — dInt row n uses avgData row n (not row I(n)); only pol 0 is kept.
    const int32_t nI = int32_t(md.I.size());
    std::vector<double> delta(size_t(nI) + 1, 0.0);
    for (int32_t n = 0; n < nI; ++n) {
        const cd rot = std::exp(cd(0.0, -kTwoPi * acal1[size_t(n)]));
        delta[size_t(n)] = std::arg(rot * avg[size_t(n) * P + 0]);
    }
    // :178 — Ad * deltaPhaseColumn + cal1
    std::vector<double> cal(size_t(N), 0.0);
    for (int32_t k = 0; k <= nI; ++k)
        for (int32_t i = 0; i < N; ++i) cal[size_t(i)] += md.Ad(i, k) * delta[size_t(k)];
    for (int32_t i = 0; i < N; ++i) cal[size_t(i)] += cal1[size_t(i)];
    return cal;
}

// Slice(start, interval, end) (common/Slice.cc:36-63), Evaluate (Slice.h:47-54),
// Range (Range.h:40-63).  Returns {start, interval, end} of the evaluated range.
void evaluate_slice(int32_t start, int32_t interval, int32_t end, int32_t T, int32_t r[3]) {
    if (start < 0) throw oracle_error("expected a positive integer");
    if (interval < -1) throw oracle_error("expected a positive integer or -1");
    if (interval == 0) throw oracle_error("expected a non zero integer");
    if (end < -1) throw oracle_error("expected a positive integer or -1");
    if (end != -1 && start >= end) throw oracle_error("range start must be greater than end");
    if (end != -1) {
        if (interval == -1) interval = end - start;
        if (interval > end - start) throw oracle_error("range increment out of bounds");
    }
    const int32_t rs = start == -1 ? T : start;
    const int32_t ri = interval == -1 ? T : interval;
    const int32_t re = end == -1 ? T : end;
    if (rs < 0 || ri < 1 || re < 0) throw oracle_error("expected a positive integer");
    r[0] = rs;
    r[1] = ri;
    r[2] = re;
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }

int oracle_phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n,
                        int32_t ref_ant, double* A_out, int32_t* I_out, int32_t* K, int32_t* N) {
    return guarded([&] {
        Mat A;
        std::vector<int32_t> I;
        phase_matrix(a1, a2, fg, n, ref_ant, A, I);
        *K = A.rows;
        *N = A.cols;
        const int32_t ld = n + 1;
        for (int32_t j = 0; j < A.cols; ++j)
            for (int32_t i = 0; i < A.rows; ++i) A_out[size_t(j) * ld + i] = A(i, j);
        std::copy(I.begin(), I.end(), I_out);
    });
}

int oracle_pinv(const double* A, int32_t m, int32_t n, double* Ad_out) {
    return guarded([&] {
        Mat a(m, n);
        std::copy(A, A + size_t(m) * n, a.v.begin());
        Mat ad = pseudo_inverse(a);
        std::copy(ad.v.begin(), ad.v.end(), Ad_out);
    });
}

int oracle_singular_values(const double* A, int32_t m, int32_t n, double* s_out) {
    return guarded([&] {
        Mat a(m, n);
        std::copy(A, A + size_t(m) * n, a.v.begin());
        SVD s = thin_svd(a);
        std::copy(s.s.begin(), s.s.end(), s_out);
    });
}

void oracle_dd_matrix(double ra0, double dec0, double ra, double dec, double* dd_out) {
    dd_matrix(ra0, dec0, ra, dec, dd_out);
}

double oracle_channel_wavelength(double f0, double df, int32_t channel) {
    return channel_wavelength(f0, df, channel);
}

int oracle_filtered_baselines(const oracle_ms_view* ms, double thr, uint8_t* out) {
    return guarded([&] {
        check_view(ms);
        std::vector<uint8_t> f;
        filtered_baselines(*ms, thr, f);
        std::copy(f.begin(), f.end(), out);
    });
}

int oracle_metadata_create(const oracle_ms_view* ms, int32_t ref_ant, double thr, oracle_metadata** out) {
    return guarded([&] {
        check_view(ms);
        *out = metadata_create(*ms, ref_ant, thr);
    });
}

void oracle_metadata_destroy(oracle_metadata* md) { delete md; }

void oracle_metadata_dims(const oracle_metadata* md, int32_t* dims) {
    dims[0] = md->A.rows;
    dims[1] = md->A.cols;
    dims[2] = md->A1.rows;
    dims[3] = int32_t(md->nbaselines);
    dims[4] = int32_t(md->reference_antenna);
}

// (Ad A).isApprox(I, 0.001) of MetaData.cc:129-136 with Eigen's definition of
// isApprox: ||X - I||_F^2 <= p^2 min(||X||_F^2, ||I||_F^2).  Returns 1 when the
// reference would log "Ad is degenerate".
namespace {
int32_t degenerate(const Mat& Ad, const Mat& A) {
    const int32_t n = A.cols, m = A.rows;
    double diff2 = 0.0, x2 = 0.0;
    for (int32_t j = 0; j < n; ++j)
        for (int32_t i = 0; i < n; ++i) {
            double x = 0.0;
            for (int32_t k = 0; k < m; ++k) x += Ad(i, k) * A(k, j);
            const double d = x - (i == j ? 1.0 : 0.0);
            diff2 += d * d;
            x2 += x * x;
        }
    const double p = 0.001;
    return diff2 <= p * p * std::min(x2, double(n)) ? 0 : 1;
}
}  // namespace

void oracle_metadata_degenerate(const oracle_metadata* md, int32_t* ad, int32_t* ad1) {
    if (ad) *ad = degenerate(md->Ad, md->A);
    if (ad1) *ad1 = degenerate(md->Ad1, md->A1);
}

void oracle_metadata_copy(const oracle_metadata* md, double* A, int32_t* I, double* Ad,
                          double* A1, int32_t* I1, double* Ad1) {
    if (A) std::copy(md->A.v.begin(), md->A.v.end(), A);
    if (I) std::copy(md->I.begin(), md->I.end(), I);
    if (Ad) std::copy(md->Ad.v.begin(), md->Ad.v.end(), Ad);
    if (A1) std::copy(md->A1.v.begin(), md->A1.v.end(), A1);
    if (I1) std::copy(md->I1.begin(), md->I1.end(), I1);
    if (Ad1) std::copy(md->Ad1.v.begin(), md->Ad1.v.end(), Ad1);
}

int oracle_rotate_visibilities(const oracle_metadata* md, const oracle_ms_view* ms, uint32_t row_start,
                               uint32_t n_rows, const double* direction, double* avg_out) {
    return guarded([&] {
        check_view(ms);
        std::vector<cd> avg;
        rotate_dispatch(*md, *ms, row_start, n_rows, direction, avg);
        std::memcpy(avg_out, avg.data(), avg.size() * sizeof(cd));
    });
}

int oracle_apply_inversion(const oracle_metadata* md, const double* avg_in, double* cal_out) {
    return guarded([&] {
        std::vector<cd> avg(size_t(md->nbaselines) * md->num_pols);
        for (size_t i = 0; i < avg.size(); ++i) avg[i] = cd(avg_in[2 * i], avg_in[2 * i + 1]);
        std::vector<double> cal = apply_inversion(*md, avg);
        std::copy(cal.begin(), cal.end(), cal_out);
    });
}

int oracle_slice_solutions(int32_t start, int32_t interval, int32_t end, int32_t timesteps) {
    int32_t r[3];
    int rc = guarded([&] { evaluate_slice(start, interval, end, timesteps, r); });
    if (rc != 0) return -1;
    return (r[2] - r[0]) / r[1];  // Range::GetSize (Range.h:56-63)
}

int oracle_calibrate(const oracle_ms_view* ms, const double* directions, int32_t n_directions,
                     int32_t slice_start, int32_t slice_interval, int32_t slice_end, double thr,
                     int32_t ref_ant, double* cal_out, double* epochs_out, int32_t* n_solutions,
                     int32_t* n_antennas) {
    return guarded([&] {
        check_view(ms);
        const int32_t T = int32_t(ms->num_rows / ms->num_baselines);  // MeasurementSet.cc:115-118
        int32_t r[3];
        evaluate_slice(slice_start, slice_interval, slice_end, T, r);  // CpuLeapCalibrator.cc:90
        const int32_t solutions = (r[2] - r[0]) / r[1];                 // :102
        const int32_t interval = r[1];
        *n_solutions = solutions;
        int32_t maxa = 0;
        for (uint32_t i = 0; i < ms->num_baselines; ++i)
            maxa = std::max(maxa, std::max(ms->antenna1[i], ms->antenna2[i]));
        *n_antennas = maxa + 1;
        if (!cal_out) return;
        std::unique_ptr<oracle_metadata> md(metadata_create(*ms, ref_ant, thr));  // :95-99
        const int32_t N = md->A.cols;
        const uint32_t nbl = ms->num_baselines;
        std::vector<cd> avg;
        for (int32_t s = 0; s < solutions; ++s) {  // :104
            // :107-109 — epochs[t] = TIME[t * nbl] (MeasurementSet.cc:120-130)
            if (epochs_out) {
                epochs_out[2 * s + 0] = ms->time[size_t(s) * interval * nbl];
                epochs_out[2 * s + 1] = ms->time[(size_t(s + 1) * interval - 1) * nbl];
            }
            // :113-119 — rows [s * interval * nbl, +interval * nbl); the slice start is not added.
            const uint32_t row_start = uint32_t(size_t(s) * interval * nbl);
            const uint32_t n_rows = uint32_t(size_t(interval) * nbl);
            for (int32_t d = 0; d < n_directions; ++d) {  // :132-139
                rotate_dispatch(*md, *ms, row_start, n_rows, directions + 2 * d, avg);
                std::vector<double> cal = apply_inversion(*md, avg);
                std::copy(cal.begin(), cal.end(), cal_out + (size_t(s) * n_directions + d) * N);
            }
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Residual loss and its gradient, written out with explicit sums.
// ---------------------------------------------------------------------------
namespace {

inline double wrap_pi(double x) {
    double r = std::fmod(x + kPi, 2.0 * kPi);
    if (r < 0) r += 2.0 * kPi;
    return r - kPi;
}

// dd row 2 and its derivatives in ra and dec (MetaData.cc:166-203 differentiated).
void dz_and_derivs(double ra0, double dec0, double ra, double dec, double dz[3], double dra[3], double ddec[3]) {
    const double s1 = std::sin(kPi / 2.0 - dec0), c1 = std::cos(kPi / 2.0 - dec0);
    const double s2 = std::sin(ra - ra0), c2 = std::cos(ra - ra0);
    const double s3 = std::sin(-kPi / 2.0 + dec), c3 = std::cos(-kPi / 2.0 + dec);
    dz[0] = -s3 * s2;
    dz[1] = s3 * c2 * c1 + c3 * s1;
    dz[2] = -s3 * c2 * s1 + c3 * c1;
    dra[0] = -s3 * c2;
    dra[1] = -s3 * s2 * c1;
    dra[2] = s3 * s2 * s1;
    ddec[0] = -c3 * s2;
    ddec[1] = c3 * c2 * c1 - s3 * s1;
    ddec[2] = -c3 * c2 * s1 - s3 * c1;
}

}  // namespace

extern "C" int oracle_residual_loss_grad(const oracle_ms_view* ms, const double* directions, const double* offsets,
                                         int32_t D, const double* antenna_delta, const double* enu_to_uvw,
                                         int32_t ref_ant, double* loss_out, double* grad_offsets,
                                         double* grad_antenna) {
    return guarded([&] {
        check_view(ms);
        std::unique_ptr<oracle_metadata> md(metadata_create(*ms, ref_ant, 0.0));
        const oracle_metadata& m = *md;
        const uint32_t nbl = ms->num_baselines, C = ms->num_channels, P = ms->num_pols, R = ms->num_rows;
        const int32_t N = m.A.cols, K = m.A.rows, K1 = m.A1.rows;
        const int32_t nI = K - 1;
        // perturbed uvw
        std::vector<double> uvw(size_t(R) * 3);
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t t = r / nbl, b = r % nbl;
            double dp[3] = {0, 0, 0};
            if (antenna_delta) {
                for (int k = 0; k < 3; ++k)
                    dp[k] = antenna_delta[3 * ms->antenna2[b] + k] - antenna_delta[3 * ms->antenna1[b] + k];
            }
            for (int i = 0; i < 3; ++i) {
                double s = ms->uvw[3 * size_t(r) + i];
                for (int k = 0; k < 3; ++k) s += enu_to_uvw[9 * size_t(t) + 3 * i + k] * dp[k];
                uvw[3 * size_t(r) + i] = s;
            }
        }
        std::vector<double> kappa(C);
        for (uint32_t c = 0; c < C; ++c) kappa[c] = (m.freq_start_hz + double(c) * m.freq_inc_hz) / kSpeedOfLight;
        double loss = 0.0;
        std::vector<double> g_uvw(size_t(R) * 3, 0.0);
        if (grad_offsets) std::fill(grad_offsets, grad_offsets + 2 * size_t(D), 0.0);
        for (int32_t d = 0; d < D; ++d) {
            const double ra = directions[2 * d] + (offsets ? offsets[2 * d] : 0.0);
            const double dec = directions[2 * d + 1] + (offsets ? offsets[2 * d + 1] : 0.0);
            double dz[3], dra[3], ddec[3];
            dz_and_derivs(m.phase_centre_ra_rad, m.phase_centre_dec_rad, ra, dec, dz, dra, ddec);
            // forward: avgData pol 0 with the NaN rule over all pols
            std::vector<cd> avg(nbl, cd(0, 0));
            std::vector<double> sf(R);
            for (uint32_t r = 0; r < R; ++r) {
                const double* u = &uvw[3 * size_t(r)];
                sf[r] = -kTwoPi * ((dz[0] * u[0] + dz[1] * u[1] + dz[2] * u[2]) - u[2]);
                for (uint32_t c = 0; c < C; ++c) {
                    bool finite = true;
                    for (uint32_t p = 0; p < P; ++p) {
                        const cd v = ms->vis_dtype == 1 ? load_vis<float>(ms->vis, (size_t(r) * C + c) * P + p)
                                                        : load_vis<double>(ms->vis, (size_t(r) * C + c) * P + p);
                        finite = finite && std::isfinite(v.real()) && std::isfinite(v.imag());
                    }
                    if (!finite) continue;
                    const cd v0 = ms->vis_dtype == 1 ? load_vis<float>(ms->vis, (size_t(r) * C + c) * P)
                                                     : load_vis<double>(ms->vis, (size_t(r) * C + c) * P);
                    avg[r % nbl] += v0 * std::exp(cd(0.0, sf[r] * kappa[c]));
                }
            }
            // forward inversion
            std::vector<double> phi1(size_t(K1), 0.0);
            for (int32_t k = 0; k + 1 < K1; ++k) phi1[k] = std::arg(avg[m.I1[k]]);
            std::vector<double> cal1(size_t(N), 0.0);
            for (int32_t k = 0; k < K1; ++k)
                for (int32_t i = 0; i < N; ++i) cal1[i] += m.Ad1(i, k) * phi1[k];
            std::vector<double> acal1(size_t(K), 0.0);
            for (int32_t j = 0; j < N; ++j)
                for (int32_t n = 0; n < K; ++n) acal1[n] += m.A(n, j) * cal1[j];
            std::vector<double> dphi(size_t(K), 0.0);
            for (int32_t n = 0; n < nI; ++n) dphi[n] = std::arg(std::exp(cd(0.0, -kTwoPi * acal1[n])) * avg[n]);
            std::vector<double> cal(cal1);
            for (int32_t k = 0; k < K; ++k)
                for (int32_t i = 0; i < N; ++i) cal[i] += m.Ad(i, k) * dphi[k];
            // residuals
            std::vector<double> res(static_cast<size_t>(nI));
            for (int32_t n = 0; n < nI; ++n) {
                double ac = 0.0;
                for (int32_t j = 0; j < N; ++j) ac += m.A(n, j) * cal[j];
                res[n] = wrap_pi(ac - std::arg(avg[m.I[n]]));
                loss += res[n] * res[n];
            }
            if (!grad_offsets && !grad_antenna) continue;
            // backward through the inversion: g_arg[b] = dL/d arg(avg[b])
            std::vector<double> g_arg(nbl, 0.0), g_cal(size_t(N), 0.0);
            for (int32_t n = 0; n < nI; ++n) {
                g_arg[m.I[n]] -= 2.0 * res[n];
                for (int32_t j = 0; j < N; ++j) g_cal[j] += 2.0 * res[n] * m.A(n, j);
            }
            std::vector<double> g_dphi(size_t(K), 0.0), g_cal1(g_cal);
            for (int32_t k = 0; k < K; ++k)
                for (int32_t i = 0; i < N; ++i) g_dphi[k] += m.Ad(i, k) * g_cal[i];
            std::vector<double> g_acal1(size_t(K), 0.0);
            for (int32_t n = 0; n < nI; ++n) {
                g_arg[n] += g_dphi[n];  // dInt row n uses avgData row n
                g_acal1[n] = -kTwoPi * g_dphi[n];
            }
            for (int32_t n = 0; n < K; ++n)
                for (int32_t j = 0; j < N; ++j) g_cal1[j] += m.A(n, j) * g_acal1[n];
            for (int32_t k = 0; k + 1 < K1; ++k) {
                double g = 0.0;
                for (int32_t i = 0; i < N; ++i) g += m.Ad1(i, k) * g_cal1[i];
                g_arg[m.I1[k]] += g;
            }
            // d arg(z) = Im(conj(z) dz) / |z|^2  ->  adjoint abar = i g / conj(z)
            std::vector<cd> abar(nbl);
            for (uint32_t b = 0; b < nbl; ++b)
                abar[b] = (avg[b] == cd(0, 0)) ? cd(0, 0) : cd(0.0, g_arg[b]) / std::conj(avg[b]);
            // backward through the rotation: dL/dsf(row) = Re(conj(abar) * sum_c i kappa_c v_c e^{i sf kappa_c})
            double g_dz[3] = {0, 0, 0};
            for (uint32_t r = 0; r < R; ++r) {
                cd s1(0, 0);
                for (uint32_t c = 0; c < C; ++c) {
                    bool finite = true;
                    for (uint32_t p = 0; p < P; ++p) {
                        const cd v = ms->vis_dtype == 1 ? load_vis<float>(ms->vis, (size_t(r) * C + c) * P + p)
                                                        : load_vis<double>(ms->vis, (size_t(r) * C + c) * P + p);
                        finite = finite && std::isfinite(v.real()) && std::isfinite(v.imag());
                    }
                    if (!finite) continue;
                    const cd v0 = ms->vis_dtype == 1 ? load_vis<float>(ms->vis, (size_t(r) * C + c) * P)
                                                     : load_vis<double>(ms->vis, (size_t(r) * C + c) * P);
                    s1 += kappa[c] * v0 * std::exp(cd(0.0, sf[r] * kappa[c]));
                }
                const double g_sf = (std::conj(abar[r % nbl]) * cd(0.0, 1.0) * s1).real();
                const double* u = &uvw[3 * size_t(r)];
                for (int i = 0; i < 3; ++i) g_dz[i] += -kTwoPi * g_sf * u[i];
                g_uvw[3 * size_t(r) + 0] += -kTwoPi * g_sf * dz[0];
                g_uvw[3 * size_t(r) + 1] += -kTwoPi * g_sf * dz[1];
                g_uvw[3 * size_t(r) + 2] += -kTwoPi * g_sf * (dz[2] - 1.0);
            }
            if (grad_offsets) {
                grad_offsets[2 * d] = g_dz[0] * dra[0] + g_dz[1] * dra[1] + g_dz[2] * dra[2];
                grad_offsets[2 * d + 1] = g_dz[0] * ddec[0] + g_dz[1] * ddec[1] + g_dz[2] * ddec[2];
            }
        }
        *loss_out = loss;
        if (grad_antenna) {
            int32_t nant = 0;
            for (uint32_t b = 0; b < nbl; ++b) nant = std::max(nant, std::max(ms->antenna1[b], ms->antenna2[b]) + 1);
            std::fill(grad_antenna, grad_antenna + 3 * size_t(nant), 0.0);
            for (uint32_t r = 0; r < R; ++r) {
                const uint32_t t = r / nbl, b = r % nbl;
                for (int k = 0; k < 3; ++k) {
                    double g = 0.0;  // (M_t^T g_uvw)_k
                    for (int i = 0; i < 3; ++i) g += enu_to_uvw[9 * size_t(t) + 3 * i + k] * g_uvw[3 * size_t(r) + i];
                    grad_antenna[3 * size_t(ms->antenna2[b]) + k] += g;
                    grad_antenna[3 * size_t(ms->antenna1[b]) + k] -= g;
                }
            }
        }
    });
}
