// Synthetic measurement sets (host columns + host or device visibilities).
//
// Rows are time-major with baselines in upper-triangle order (antenna1 <=
// antenna2, autocorrelations optional), the MS convention the reference's
// row % nbaselines mapping relies on (CpuLeapCalibrator.cc:189).  UVW follow
// Earth rotation of an east-north-up array at the given latitude, tracking
// the phase centre; uvw(baseline) = uvw(antenna2) - uvw(antenna1).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "leap_internal.h"
#include "synth.h"

namespace leap {
namespace {

constexpr double kEarthRotation = 7.2921150e-5;  // rad / s
constexpr double kEpoch0 = 5.0203e9;              // TIME of the first row, MJD seconds

// Direction-dependent matrix of the calibrator (same definition as the dd kernel).
void host_dd(double ra0, double dec0, double ra, double dec, double dd[9]) {
    const double ang1 = kPi / 2.0 - dec0, ang2 = ra - ra0, ang3 = -kPi / 2.0 + dec;
    const double m1[9] = {1, 0, 0, 0, std::cos(ang1), -std::sin(ang1), 0, std::sin(ang1), std::cos(ang1)};
    const double m2[9] = {std::cos(ang2), std::sin(ang2), 0, -std::sin(ang2), std::cos(ang2), 0, 0, 0, 1};
    const double m3[9] = {1, 0, 0, 0, std::cos(ang3), -std::sin(ang3), 0, std::sin(ang3), std::cos(ang3)};
    double t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            t[i * 3 + j] = m3[i * 3] * m2[j] + m3[i * 3 + 1] * m2[3 + j] + m3[i * 3 + 2] * m2[6 + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            dd[i * 3 + j] = t[i * 3] * m1[j] + t[i * 3 + 1] * m1[3 + j] + t[i * 3 + 2] * m1[6 + j];
}

__global__ __launch_bounds__(256) void synth_vis_kernel(SynthParams sp, const double* __restrict__ uvw,
                                                        uint64_t n_rows, void* __restrict__ out, int c64) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_rows * sp.C) return;
    const uint64_t row = i / sp.C;
    const uint32_t c = uint32_t(i - row * sp.C);
    const double u[3] = {uvw[3 * row], uvw[3 * row + 1], uvw[3 * row + 2]};
    double v[8];
    synth_visibility(sp, row, c, u, v);
    if (c64) {
        float* o = static_cast<float*>(out) + i * 2 * sp.P;
        for (uint32_t q = 0; q < 2 * sp.P; ++q) o[q] = float(v[q]);
    } else {
        double* o = static_cast<double*>(out) + i * 2 * sp.P;
        for (uint32_t q = 0; q < 2 * sp.P; ++q) o[q] = v[q];
    }
}

}  // namespace

void launch_synth_vis(hipStream_t s, const leap_synth_ms* ms, int32_t vis_dtype, void* out, const double* d_uvw) {
    const uint64_t total = uint64_t(ms->n_rows) * ms->cfg.n_channels;
    const uint64_t blocks = (total + 255) / 256;
    require(blocks < (1ull << 31), LEAP_ERR_RANGE, "synthetic cube too large for one launch");
    synth_vis_kernel<<<dim3(uint32_t(blocks)), 256, 0, s>>>(ms->params, d_uvw, ms->n_rows, out,
                                                            vis_dtype == LEAP_VIS_C64 ? 1 : 0);
    LEAP_HIP(hipGetLastError());
}

}  // namespace leap

using namespace leap;

namespace {
thread_local std::string g_synth_error;
}

extern "C" {

int leap_synth_create(const leap_synth_config* cfg_in, leap_synth_ms** out) {
    try {
        require(cfg_in && out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const leap_synth_config& cfg = *cfg_in;
        require(cfg.n_antennas >= 2, LEAP_ERR_INVALID_ARGUMENT, "need at least 2 antennas");
        require(cfg.n_pols >= 1 && cfg.n_pols <= 4, LEAP_ERR_INVALID_ARGUMENT, "n_pols must be 1..4");
        require(cfg.n_sources <= uint32_t(kMaxSynthSources), LEAP_ERR_INVALID_ARGUMENT, "too many sources");
        require(cfg.n_timesteps >= 1 && cfg.n_channels >= 1, LEAP_ERR_INVALID_ARGUMENT, "empty cube");
        auto ms = std::make_unique<leap_synth_ms>();
        ms->cfg = cfg;
        const uint32_t N = cfg.n_antennas;
        const uint64_t seed = cfg.seed;
        // Antenna positions (east, north, up) in metres.
        ms->enu.assign(size_t(N) * 3, 0.0);
        for (uint32_t a = 0; a < N; ++a) {
            double e, n;
            if (cfg.layout == 1) {
                // Three-arm logarithmic spiral from 20 m to array_radius_m.
                const uint32_t arm = a % 3, j = a / 3, per_arm = (N + 2) / 3;
                const double rmin = 20.0;
                const double frac = per_arm > 1 ? double(j) / double(per_arm - 1) : 0.0;
                const double r = rmin * std::pow(cfg.array_radius_m / rmin, frac);
                const double th = arm * (2.0 * kPi / 3.0) + 1.2 * std::log(r / rmin) +
                                  0.05 * (uniform01(seed, 10, a) - 0.5);
                e = r * std::cos(th);
                n = r * std::sin(th);
            } else {
                const double r = cfg.array_radius_m * std::sqrt(uniform01(seed, 11, a));
                const double th = 2.0 * kPi * uniform01(seed, 12, a);
                e = r * std::cos(th);
                n = r * std::sin(th);
            }
            ms->enu[3 * a] = e;
            ms->enu[3 * a + 1] = n;
            ms->enu[3 * a + 2] = 0.0;
        }
        // Baselines in upper-triangle order.
        for (uint32_t i = 0; i < N; ++i)
            for (uint32_t j = cfg.autocorrelations ? i : i + 1; j < N; ++j) {
                ms->a1.push_back(int32_t(i));
                ms->a2.push_back(int32_t(j));
            }
        ms->nbl = uint32_t(ms->a1.size());
        const uint64_t rows = uint64_t(ms->nbl) * cfg.n_timesteps;
        require(rows < (1ull << 32), LEAP_ERR_RANGE, "too many rows");
        ms->n_rows = uint32_t(rows);
        // Flags: whole baselines; baseline indices below n_antennas stay
        // unflagged so that any reference antenna index passes the reference's
        // flag check at that index (PhaseMatrixFunction.cc:52-57).
        ms->flag.assign(ms->nbl, 0);
        for (uint32_t b = N; b < ms->nbl; ++b)
            if (ms->a1[b] != ms->a2[b] && uniform01(seed, 13, b) <= cfg.flag_fraction) ms->flag[b] = 1;
        // Per-antenna phase screen U(-pi, pi).
        ms->antenna_phase.resize(N);
        for (uint32_t a = 0; a < N; ++a) ms->antenna_phase[a] = kPi * (2.0 * uniform01(seed, 14, a) - 1.0);
        ms->bl_phase.resize(ms->nbl);
        for (uint32_t b = 0; b < ms->nbl; ++b)
            ms->bl_phase[b] = ms->antenna_phase[ms->a1[b]] - ms->antenna_phase[ms->a2[b]];
        // UVW and TIME.
        const double lat = cfg.latitude_rad, dec = cfg.phase_centre_dec_rad;
        std::vector<double> xyz(size_t(N) * 3);
        for (uint32_t a = 0; a < N; ++a) {
            const double e = ms->enu[3 * a], n = ms->enu[3 * a + 1], up = ms->enu[3 * a + 2];
            xyz[3 * a] = -std::sin(lat) * n + std::cos(lat) * up;
            xyz[3 * a + 1] = e;
            xyz[3 * a + 2] = std::cos(lat) * n + std::sin(lat) * up;
        }
        ms->uvw.resize(size_t(rows) * 3);
        ms->time.resize(size_t(rows));
        std::vector<double> auvw(size_t(N) * 3);
        ms->enu_to_uvw.assign(size_t(cfg.n_timesteps) * 9, 0.0);
        const double h0 = -0.5 * kEarthRotation * cfg.integration_s * cfg.n_timesteps;
        for (uint32_t t = 0; t < cfg.n_timesteps; ++t) {
            const double H = h0 + kEarthRotation * cfg.integration_s * (t + 0.5);
            const double sh = std::sin(H), ch = std::cos(H), sd = std::sin(dec), cdd = std::cos(dec);
            {
                // M_t = W(H, dec) * E(lat): east-north-up metres to uvw metres.
                const double W[9] = {sh, ch, 0.0, -sd * ch, sd * sh, cdd, cdd * ch, -cdd * sh, sd};
                const double E[9] = {0.0, -std::sin(lat), std::cos(lat), 1.0, 0.0, 0.0, 0.0, std::cos(lat),
                                     std::sin(lat)};
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        double v = 0.0;
                        for (int k = 0; k < 3; ++k) v += W[3 * i + k] * E[3 * k + j];
                        ms->enu_to_uvw[size_t(t) * 9 + 3 * i + j] = v;
                    }
            }
            for (uint32_t a = 0; a < N; ++a) {
                const double X = xyz[3 * a], Y = xyz[3 * a + 1], Z = xyz[3 * a + 2];
                auvw[3 * a] = sh * X + ch * Y;
                auvw[3 * a + 1] = -sd * ch * X + sd * sh * Y + cdd * Z;
                auvw[3 * a + 2] = cdd * ch * X - cdd * sh * Y + sd * Z;
            }
            for (uint32_t b = 0; b < ms->nbl; ++b) {
                const size_t row = size_t(t) * ms->nbl + b;
                for (int k = 0; k < 3; ++k)
                    ms->uvw[3 * row + k] = auvw[3 * ms->a2[b] + k] - auvw[3 * ms->a1[b] + k];
                ms->time[row] = kEpoch0 + cfg.integration_s * t;
            }
        }
        // Sky model.
        SynthParams& sp = ms->params;
        std::memset(&sp, 0, sizeof(sp));
        sp.nbl = ms->nbl;
        sp.C = cfg.n_channels;
        sp.P = cfg.n_pols;
        sp.n_src = cfg.n_sources;
        sp.f0 = cfg.freq_start_hz;
        sp.df = cfg.freq_inc_hz;
        sp.noise_sigma = cfg.noise_sigma;
        sp.nan_fraction = cfg.nan_fraction;
        sp.seed = seed;
        ms->sources.resize(size_t(cfg.n_sources) * 3);
        for (uint32_t s = 0; s < cfg.n_sources; ++s) {
            const double ra = cfg.phase_centre_ra_rad + cfg.source_offset_max_rad * (2.0 * uniform01(seed, 15, s) - 1.0);
            const double de = cfg.phase_centre_dec_rad + cfg.source_offset_max_rad * (2.0 * uniform01(seed, 16, s) - 1.0);
            const double flux = cfg.source_flux_min + (cfg.source_flux_max - cfg.source_flux_min) * uniform01(seed, 17, s);
            ms->sources[3 * s] = ra;
            ms->sources[3 * s + 1] = de;
            ms->sources[3 * s + 2] = flux;
            double dd[9];
            host_dd(cfg.phase_centre_ra_rad, cfg.phase_centre_dec_rad, ra, de, dd);
            sp.src_flux[s] = flux;
            sp.src_dz[s][0] = dd[6];
            sp.src_dz[s][1] = dd[7];
            sp.src_dz[s][2] = dd[8] - 1.0;
        }
        sp.bl_phase = ms->bl_phase.data();
        *out = ms.release();
        return LEAP_OK;
    } catch (const Error& e) {
        g_synth_error = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_synth_error = e.what();
        return LEAP_ERR_INTERNAL;
    }
}

void leap_synth_destroy(leap_synth_ms* ms) { delete ms; }

void leap_synth_view(const leap_synth_ms* ms, leap_ms_view* v) {
    std::memset(v, 0, sizeof(*v));
    v->num_stations = ms->cfg.n_antennas;
    v->total_antennas = ms->cfg.n_antennas;
    v->num_baselines = ms->nbl;
    v->num_rows = ms->n_rows;
    v->num_channels = ms->cfg.n_channels;
    v->num_pols = ms->cfg.n_pols;
    v->freq_start_hz = ms->cfg.freq_start_hz;
    v->freq_inc_hz = ms->cfg.freq_inc_hz;
    v->phase_centre_ra_rad = ms->cfg.phase_centre_ra_rad;
    v->phase_centre_dec_rad = ms->cfg.phase_centre_dec_rad;
    v->antenna1 = ms->a1.data();
    v->antenna2 = ms->a2.data();
    v->time = ms->time.data();
    v->flag = ms->flag.data();
    v->uvw = ms->uvw.data();
    v->vis = nullptr;
    v->vis_dtype = LEAP_VIS_C128;
    v->vis_on_device = 0;
    v->uvw_device = nullptr;
}

void leap_synth_enu_to_uvw(const leap_synth_ms* ms, double* out) {
    std::copy(ms->enu_to_uvw.begin(), ms->enu_to_uvw.end(), out);
}

void leap_synth_sky(const leap_synth_ms* ms, double* sources, double* antenna_phase) {
    if (sources) std::copy(ms->sources.begin(), ms->sources.end(), sources);
    if (antenna_phase) std::copy(ms->antenna_phase.begin(), ms->antenna_phase.end(), antenna_phase);
}

int leap_synth_fill_vis_host(const leap_synth_ms* ms, uint32_t row_start, uint32_t n_rows, int32_t vis_dtype,
                             void* out) {
    if (!ms || !out) return LEAP_ERR_INVALID_ARGUMENT;
    if (uint64_t(row_start) + n_rows > ms->n_rows) return LEAP_ERR_RANGE;
    const uint32_t C = ms->cfg.n_channels, P = ms->cfg.n_pols;
    double v[8];
    for (uint32_t r = 0; r < n_rows; ++r) {
        const uint64_t row = uint64_t(row_start) + r;
        for (uint32_t c = 0; c < C; ++c) {
            synth_visibility(ms->params, row, c, &ms->uvw[3 * row], v);
            const size_t o = (size_t(r) * C + c) * 2 * P;
            if (vis_dtype == LEAP_VIS_C64) {
                float* f = static_cast<float*>(out) + o;
                for (uint32_t q = 0; q < 2 * P; ++q) f[q] = float(v[q]);
            } else {
                double* d = static_cast<double*>(out) + o;
                for (uint32_t q = 0; q < 2 * P; ++q) d[q] = v[q];
            }
        }
    }
    return LEAP_OK;
}

}  // extern "C"
