// leap_hip_calibrate_md_noalloc: the FP64 per-direction calibration with every
// intermediate in a fixed device array, for callers that cannot allocate.
//
// The array is a file-scope __device__ variable, so each GPU holds its own copy
// at its own address.  In host code its name is only a stand-in for the
// symbol: the device address comes from hipGetSymbolAddress, resolved once per
// device with that device current (noalloc_init, called by leap_hip_create),
// and only that address is handed to kernels.
//
// Nothing in the array has to start at zero: every kernel of the route writes
// all of its output before the next one reads it (rotate_average_f64 writes
// every (baseline, pol) of avgData, the inversion kernels every phase, cal1,
// dphi and partial entry), so contents left by an earlier call do no harm.
#include <mutex>

#include "capi_common.h"
#include "leap_internal.h"

using namespace leap;

namespace {

constexpr size_t kScratchBytes = LEAP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once resolved

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

void* scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

// Offsets (bytes) of one solution's working set in the scratch array.
struct Layout {
    size_t dirs, dd, avg, phase, cal1, delta, partial, total;
};

Layout layout(uint32_t D, uint32_t nbl, uint32_t P, int32_t K, int32_t N, int32_t K1) {
    Layout l{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t o = at;
        at += align256(bytes);
        return o;
    };
    l.dirs = take(size_t(D) * 2 * sizeof(double));
    l.dd = take(size_t(D) * 9 * sizeof(double));
    l.avg = take(size_t(nbl) * P * 2 * sizeof(double));
    l.phase = take(size_t(K1) * sizeof(double));
    l.cal1 = take(size_t(N) * sizeof(double));
    l.delta = take(size_t(K) * sizeof(double));
    l.partial = take(inversion_partial_doubles(K, N, 1) * sizeof(double));
    l.total = at;
    return l;
}

}  // namespace

namespace leap {

void noalloc_init(int device) {
    require(device >= 0 && device < kMaxDevices, LEAP_ERR_INVALID_ARGUMENT,
            "device index beyond the no-allocation scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return;
    void* p = nullptr;
    LEAP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr, LEAP_ERR_INTERNAL, "no device address for the no-allocation scratch");
    g_scratch[device] = p;
}

}  // namespace leap

extern "C" int leap_hip_calibrate_md_noalloc(leap_ctx* ctx, const leap_metadata* md, const leap_ms_view* ms,
                                             const double* dirs, uint32_t D, int32_t s_start, int32_t s_interval,
                                             int32_t s_end, double* cal_out, double* epochs_out,
                                             int32_t* n_solutions_out) {
    return guarded(ctx, [&] {
        require(ctx && md && ms, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        validate_view(*ms);
        require(md->nbl == ms->num_baselines && md->channels == ms->num_channels && md->pols == ms->num_pols,
                LEAP_ERR_INVALID_ARGUMENT, "metadata was built for a different measurement set");
        require(dirs != nullptr || D == 0, LEAP_ERR_INVALID_ARGUMENT, "null directions");
        const int32_t T = int32_t(ms->num_rows / ms->num_baselines);  // MeasurementSet.cc:115-118
        int32_t range[3];
        evaluate_slice(s_start, s_interval, s_end, T, range);       // CpuLeapCalibrator.cc:90
        const int32_t solutions = (range[2] - range[0]) / range[1];  // Range::GetSize
        if (n_solutions_out) *n_solutions_out = solutions;
        const uint32_t interval = uint32_t(range[1]);
        const uint32_t nbl = ms->num_baselines;
        if (epochs_out) {
            for (int32_t s = 0; s < solutions; ++s) {  // CpuLeapCalibrator.cc:107-109
                epochs_out[2 * s] = ms->time[size_t(s) * interval * nbl];
                epochs_out[2 * s + 1] = ms->time[(size_t(s + 1) * interval - 1) * nbl];
            }
        }
        if (cal_out == nullptr || D == 0 || solutions == 0) return;
        require(ms->vis_on_device && ms->vis != nullptr && ms->uvw_device != nullptr, LEAP_ERR_INVALID_ARGUMENT,
                "the no-allocation calibration needs device-resident visibilities and uvw");
        const int32_t K = md->A.rows, N = md->A.cols, K1 = md->A1.rows;
        const Layout l = layout(D, nbl, md->pols, K, N, K1);
        require(l.total <= kScratchBytes, LEAP_ERR_RANGE,
                "working set of " + std::to_string(l.total) + " bytes exceeds the no-allocation scratch of " +
                    std::to_string(kScratchBytes) + " bytes");
        DeviceScope scope(ctx->device);
        char* base = static_cast<char*>(scratch_on(ctx->device));
        require(base != nullptr, LEAP_ERR_INTERNAL, "no-allocation scratch not set up on this device");
        double* d_dirs = reinterpret_cast<double*>(base + l.dirs);
        double* d_dd = reinterpret_cast<double*>(base + l.dd);
        double* d_avg = reinterpret_cast<double*>(base + l.avg);
        double* d_phase = reinterpret_cast<double*>(base + l.phase);
        double* d_cal1 = reinterpret_cast<double*>(base + l.cal1);
        double* d_delta = reinterpret_cast<double*>(base + l.delta);
        double* d_partial = reinterpret_cast<double*>(base + l.partial);

        LEAP_HIP(hipMemcpyAsync(d_dirs, dirs, size_t(D) * 2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        launch_dd(ctx->stream, md->phase_centre_ra, md->phase_centre_dec, d_dirs, D, d_dd);
        const size_t row_bytes = size_t(ms->num_channels) * ms->num_pols * (ms->vis_dtype == LEAP_VIS_C64 ? 8 : 16);
        for (int32_t s = 0; s < solutions; ++s) {  // CpuLeapCalibrator.cc:104
            // :113-119 — rows [s * interval * nbl, +interval * nbl); the slice start is not added.
            const size_t row_begin = size_t(s) * interval * nbl;
            const void* vis = static_cast<const char*>(ms->vis) + row_begin * row_bytes;
            const double* uvw = ms->uvw_device + row_begin * 3;
            for (uint32_t d = 0; d < D; ++d) {
                launch_rotate_f64(ctx->stream, vis, ms->vis_dtype, uvw, d_dd + 9 * size_t(d), 0, interval, nbl,
                                  md->channels, md->pols, md->d_lambda.as<double>(), d_avg);
                launch_inversion_f64(ctx->stream, *md, d_avg, 1, d_phase, d_cal1, d_delta, d_partial,
                                     cal_out + (size_t(s) * D + d) * size_t(N));
            }
        }
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}
