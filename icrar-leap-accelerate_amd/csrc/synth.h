// Synthetic measurement sets: point sources under per-antenna phase screens,
// with Gaussian noise, in the phase convention of the calibrator itself (a
// source at direction s contributes flux * exp(+2 pi i dw_s / lambda), where
// dw_s = (dd_s * uvw).z - uvw.z, so rotating to s makes it phase-flat).
//
// The per-visibility function is __host__ __device__ so the host fill and the
// device kernel share one definition.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "leap_hip.h"

namespace leap {

constexpr int kMaxSynthSources = 16;

struct SynthParams {
    uint32_t nbl, C, P, n_src;
    double f0, df;
    double src_flux[kMaxSynthSources];
    double src_dz[kMaxSynthSources][3];
    double noise_sigma;
    double nan_fraction;
    uint64_t seed;
    const double* bl_phase;  // [nbl] phase(antenna1) - phase(antenna2), host or device
};

// SplitMix64 finaliser (Steele, Lea & Flood 2014; public-domain reference code by Vigna).
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform in (0, 1] from a counter and a stream key.
__host__ __device__ inline double uniform01(uint64_t seed, uint64_t stream, uint64_t counter) {
    const uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull ^ mix64(counter)));
    return (double((h >> 11) + 1)) * (1.0 / 9007199254740992.0);
}

// Polarisation weights relative to the source model.
__host__ __device__ inline void pol_weight(uint32_t P, uint32_t p, double& wr, double& wi) {
    if (P == 4) {
        const double wre[4] = {1.0, 0.09553364891256060, 0.07648421872844885, 0.9};
        const double wim[4] = {0.0, 0.02955202066613396, -0.06442176872376910, 0.0};
        wr = wre[p];
        wi = wim[p];
    } else {
        wr = (p == 0) ? 1.0 : 0.9;
        wi = 0.0;
    }
}

// Visibilities of one (row, channel): out[2 * p], out[2 * p + 1] = re, im of pol p.
__host__ __device__ inline void synth_visibility(const SynthParams& sp, uint64_t row, uint32_t c,
                                                 const double* uvw, double* out) {
    const uint32_t bl = uint32_t(row % sp.nbl);
    const double kappa = (sp.f0 + double(c) * sp.df) / 299792458.0;  // 1 / lambda
    double vr = 0.0, vi = 0.0;
    for (uint32_t s = 0; s < sp.n_src; ++s) {
        const double dw = sp.src_dz[s][0] * uvw[0] + sp.src_dz[s][1] * uvw[1] + sp.src_dz[s][2] * uvw[2];
        double ph = 6.283185307179586476925286766559005768 * dw * kappa;
        double sn, cs;
        sincos(ph, &sn, &cs);
        vr += sp.src_flux[s] * cs;
        vi += sp.src_flux[s] * sn;
    }
    double gs, gc;
    sincos(sp.bl_phase[bl], &gs, &gc);
    const double br = vr * gc - vi * gs;
    const double bi = vr * gs + vi * gc;
    const uint64_t base = (row * sp.C + c) * sp.P;
    for (uint32_t p = 0; p < sp.P; ++p) {
        double wr, wi;
        pol_weight(sp.P, p, wr, wi);
        double re = br * wr - bi * wi;
        double im = br * wi + bi * wr;
        if (sp.noise_sigma > 0.0) {
            const double u1 = uniform01(sp.seed, 1, base + p);
            const double u2 = uniform01(sp.seed, 2, base + p);
            const double r = sp.noise_sigma * sqrt(-2.0 * log(u1));
            double ns, nc;
            sincos(6.283185307179586476925286766559005768 * u2, &ns, &nc);
            re += r * nc;
            im += r * ns;
        }
        if (sp.nan_fraction > 0.0 && uniform01(sp.seed, 3, base + p) <= sp.nan_fraction) {
            re = NAN;
            im = NAN;
        }
        out[2 * p] = re;
        out[2 * p + 1] = im;
    }
}

}  // namespace leap

struct leap_synth_ms {
    leap_synth_config cfg{};
    uint32_t nbl = 0;
    uint32_t n_rows = 0;
    std::vector<double> enu;        // [N][3]
    std::vector<int32_t> a1, a2;    // [nbl], first epoch
    std::vector<double> uvw;        // [rows][3]
    std::vector<double> time;       // [rows]
    std::vector<uint8_t> flag;      // [nbl]
    std::vector<double> antenna_phase;  // [N]
    std::vector<double> bl_phase;       // [nbl]
    std::vector<double> sources;        // [n_src][3] ra, dec, flux
    std::vector<double> enu_to_uvw;     // [T][3][3]: uvw = M_t * enu, per timestep
    leap::SynthParams params{};
};
