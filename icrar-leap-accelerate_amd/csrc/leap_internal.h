// Internal declarations of the gfx950 LEAP calibrator library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "leap_hip.h"

namespace leap {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kTwoPi = 6.283185307179586476925286766559005768;
constexpr double kSpeedOfLight = 299792458.0;  // common/constants.h:35 of the reference

// Error carried to the C ABI as a status code plus message.
struct Error : std::runtime_error {
    leap_status status;
    Error(leap_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void hip_check(hipError_t e, const char* what, const char* file, int line);
#define LEAP_HIP(call) ::leap::hip_check((call), #call, __FILE__, __LINE__)

inline void require(bool ok, leap_status s, const std::string& msg) {
    if (!ok) throw Error(s, msg);
}

// Device buffer owned by a context (RAII, frees on the context's device).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer();
    // Grows (never shrinks) to at least n bytes; contents are not kept.
    void reserve(size_t n);
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Pinned host buffer (hipHostMalloc) owned by a context: grows, never shrinks.
struct PinnedBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer();
    // Grows to at least n bytes; contents are not kept.
    void reserve(size_t n);
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Host-side phase matrices, in the reference's layout conventions.
struct PhaseMatrix {
    int32_t rows = 0;  // K: baselines used + the constraint row
    int32_t cols = 0;  // N: max(antenna1, antenna2) + 1
    std::vector<int32_t> I;       // K - 1 baseline indices
    std::vector<int32_t> ant1;    // K - 1: antenna1 of baseline I(k)
    std::vector<int32_t> ant2;    // K - 1: antenna2 of baseline I(k)
    int32_t constraint = 0;       // column of the constraint row
    std::vector<double> dense() const;  // K x N column-major
};

// Builds A (ref_ant < 0) or A1; throws Error with the reference's messages.
PhaseMatrix phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n,
                         int32_t ref_ant);

// Moore-Penrose inverse of the Gram matrix G = A^T A (N x N, column-major)
// for an incidence-plus-constraint matrix A.
std::vector<double> gram_pinv(const PhaseMatrix& pm);

// Dimension of the null space of A (components of the baseline graph without
// the constraint antenna); the reference warns "Ad is degenerate" when > 0.
int32_t null_components(const PhaseMatrix& pm);

// Filtered baselines (flags of the first epoch OR short baselines).
std::vector<uint8_t> filtered_baselines(const leap_ms_view& ms, double threshold);

// icrar::Slice(...).Evaluate(T) -> {start, interval, end}; throws on bad slices.
void evaluate_slice(int32_t start, int32_t interval, int32_t end, int32_t T, int32_t out[3]);

// Validates the ms view (MeasurementSet::Validate, ms/MeasurementSet.cc:64-103).
void validate_view(const leap_ms_view& ms);

uint64_t hash_phase_matrix(const PhaseMatrix& pm);

// Static schedule of the multi-set persistent kernel.
struct MultiShape {
    uint32_t nbl, T, C, D;
};
struct MultiItem {  // the layout of the kernel's uint2 items
    uint32_t set, local;  // local = baseline + nbl * direction block
};
struct MultiPlan {
    std::vector<uint32_t> off;  // n_blocks + 1
    std::vector<MultiItem> items;
    double makespan;
};
bool multi_tile_supported(int32_t tile);
double multi_item_cost(uint32_t T, uint32_t C, int32_t tile);
MultiPlan plan_multi(const std::vector<MultiShape>& sets, int32_t tile, uint32_t n_blocks);

// NUFFT channel sums (nufft_plan.cpp).
constexpr uint32_t kNufftMaxChannels = 512;  // grid N <= 1024
constexpr int kNufftMaxWidth = 16;
constexpr int kNufftQuadratureNodes = 100;
constexpr int kNufftCoefDoubles = (kNufftMaxWidth / 2) * kNufftMaxWidth;
// Automatic choice: fewer directions or channels stay on Horner (per row the FFT
// costs about what 4 C FMAs per lane cost at C = 64..96).  The same bounds
// hold for the width-16 NUFFT of FP64 precision: on the 256-station set (384
// channels) it takes 16.5 / 16.6 / 16.9 / 17.5 / 24.5 ms at 16 / 32 / 64 /
// 128 / 256 directions on the complex128 cube where Horner takes 19.9 / 21.4 /
// 27.8 / 43.6 / 76.9 ms (complex64: 13.2 .. 18.6 against 15.1 .. 69.1 ms).
constexpr uint32_t kNufftMinDirections = 16;
constexpr uint32_t kNufftMinChannels = 96;
struct NufftPlan {
    bool fp32 = false;  // FP32 transform, grid and interpolation (W 8 or 10) instead of FP64 (W 10 or 12)
    bool w16 = false;   // FP64 transform with kernel width 16 (LEAP_CHANNEL_SUM_NUFFT_W16)
    uint32_t N = 0;  // grid size (power of two, >= 2C, >= 256)
    uint32_t W = 0;  // kernel width in grid points
    uint32_t h = 0;  // 64 * round(C / 128): frequency k = c - h
    double beta = 0.0;
    // coef[j][m]: coefficient of s^m in tap j's weight polynomial, j < W / 2
    double coef[kNufftMaxWidth / 2][kNufftMaxWidth] = {};
    std::vector<double> scale;  // C precorrection factors (2 pi / N) / psi^(c - h)
};
bool nufft_supported(uint32_t C);
inline bool is_nufft(int32_t channel_sum) {
    return channel_sum == LEAP_CHANNEL_SUM_NUFFT || channel_sum == LEAP_CHANNEL_SUM_NUFFT_FP32 ||
           channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16;
}
NufftPlan make_nufft_plan(uint32_t C, bool fp32, bool w16 = false);
// The plan's device tables: kNufftCoefDoubles coefficients (coef[j][m]), the C
// precorrection factors, then the coefficients rounded to FP32 (from double
// nufft_coef32_offset(C) on).
std::vector<double> nufft_tables(const NufftPlan& plan);
inline size_t nufft_coef32_offset(uint32_t C) { return size_t(kNufftCoefDoubles) + C; }

}  // namespace leap

// Metadata resident on the device (cpu::MetaData + cuda::ConstantBuffer of the reference).
struct leap_metadata {
    int device = 0;
    uint32_t nbl = 0, channels = 0, pols = 0, rows = 0;
    uint32_t reference_antenna = 0;
    double freq_start_hz = 0, freq_inc_hz = 0;
    double phase_centre_ra = 0, phase_centre_dec = 0;
    leap::PhaseMatrix A, A1;
    // The reference's degeneracy warnings (MetaData.cc:129-136): Ad A != I, Ad1 A1 != I.
    bool ad_degenerate = false, ad1_degenerate = false;
    // device
    leap::DeviceBuffer d_AdT;      // K x N row-major == Ad column-major (N x K) transposed
    leap::DeviceBuffer d_Ad1T;     // K1 x N
    leap::DeviceBuffer d_I1;       // K1 - 1 ints (baseline index of phase angle k)
    leap::DeviceBuffer d_pairs;    // K - 1 int2: (antenna1, antenna2) of baseline I(k)
    // the inversion's adjoint: A's baseline indices I, A1's antenna pairs, and
    // each antenna's rows of A (CSR: offsets N + 1, entries 2 row + (antenna2 ? 1 : 0))
    leap::DeviceBuffer d_I, d_pairs1, d_csr_off, d_csr_ent;
    leap::DeviceBuffer d_lambda;   // channels doubles: wavelength of each channel
    leap::DeviceBuffer d_Gp;       // G+ = pinv(A^T A) then G1+ = pinv(A1^T A1), N x N each
    leap::DeviceBuffer d_AdT_planes;  // split-3 x BF16 planes of AdT in MFMA operand layout
    // NUFFT plans (N == 0 when the channel count is not supported) and their
    // device tables (leap::nufft_tables): FP64 transform (LEAP_CHANNEL_SUM_NUFFT),
    // FP32 transform (LEAP_CHANNEL_SUM_NUFFT_FP32) and FP64 transform with the
    // width-16 kernel (LEAP_CHANNEL_SUM_NUFFT_W16).
    leap::NufftPlan nufft32, nufft64, nufft16;
    leap::DeviceBuffer d_nufft32, d_nufft64, d_nufft16;
    const leap::NufftPlan& nufft_plan(int32_t channel_sum) const {
        return channel_sum == LEAP_CHANNEL_SUM_NUFFT_FP32 ? nufft32
               : channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16 ? nufft16
                                                           : nufft64;
    }
    const double* nufft_tables(int32_t channel_sum) const {
        return (channel_sum == LEAP_CHANNEL_SUM_NUFFT_FP32 ? d_nufft32
                : channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16 ? d_nufft16
                                                            : d_nufft64)
            .as<double>();
    }
    const double* nufft_scale(int32_t channel_sum) const {
        return nufft_tables(channel_sum) + leap::kNufftCoefDoubles;
    }
};

// TuneKey::precision holds the batched tuner's mode code; these bits mark the
// packed path and its NUFFT channel sums.
constexpr int32_t kTunePackedBit = 1024;
constexpr int32_t kTuneNufftBit = 2048;
constexpr int32_t kTuneNufft32Bit = 4096;
constexpr int32_t kTuneNufftW16Bit = 8192;

struct TuneKey {
    int32_t precision, vis_dtype, all_pols;
    uint32_t D, T, C, P;
    bool operator<(const TuneKey& o) const {
        const int64_t a[7] = {precision, vis_dtype, all_pols, D, T, C, P};
        const int64_t b[7] = {o.precision, o.vis_dtype, o.all_pols, o.D, o.T, o.C, o.P};
        for (int i = 0; i < 7; ++i)
            if (a[i] != b[i]) return a[i] < b[i];
        return false;
    }
};

namespace leap {
// Per-set device buffers of leap_hip_calibrate_batch.
struct SetBuffers {
    DeviceBuffer d_dirs, d_dd, d_avg, d_cal, d_vis, d_uvw;
};

// A calibration sequence captured as a HIP graph, with the key of shapes,
// options and device addresses it was captured for.
struct CapturedSequence {
    std::vector<uint64_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DeviceBuffer d_cal;     // cal of every solution interval (and set)
    DeviceBuffer d_tables;  // multi-set descriptor tables of every round
    std::vector<double> h_cal;
    CapturedSequence() = default;
    CapturedSequence(const CapturedSequence&) = delete;
    CapturedSequence& operator=(const CapturedSequence&) = delete;
    ~CapturedSequence();
};
}  // namespace leap

struct leap_ctx {
    int device = 0;
    leap_log_cb log = nullptr;  // diagnostics the reference logs (NULL: dropped)
    void* log_user = nullptr;
    hipStream_t stream = nullptr;
    hipStream_t pack_stream = nullptr;  // pol-0 packs overlapping the Horner launches
    // host -> device staging of host-resident visibilities: pinned double buffer, copy stream
    hipStream_t copy_stream = nullptr;
    void* h_pinned[2] = {nullptr, nullptr};
    size_t pinned_bytes = 0;
    hipEvent_t ev_h2d[2] = {nullptr, nullptr}, ev_widened[2] = {nullptr, nullptr};
    bool h2d_pending[2] = {false, false}, widen_pending[2] = {false, false};
    hipEvent_t ev_copy_fork = nullptr, ev_copy_join = nullptr;
    size_t last_stage_bytes = 0;  // bytes the last call sent host -> device through the pinned buffers
    hipEvent_t ev_fork = nullptr, ev_packed[2] = {nullptr, nullptr}, ev_consumed[2] = {nullptr, nullptr};
    std::string error;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> rotate_events;
    size_t rotate_events_used = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pack_events;
    size_t pack_events_used = 0;
    double last_pack_ms_avg = 0;
    int32_t last_pack_launches = 0;
    double last_rotate_ms_avg = 0, last_total_ms = 0;
    int32_t last_rotate_launches = 0;
    // workspace
    leap::DeviceBuffer d_packed;  // pol-0 pack of the staged rows (FP64 complex per row and channel)
    leap::DeviceBuffer d_widen[2];  // complex64 pol-0 chunks from the host, widened into d_packed
    leap::DeviceBuffer d_uvw, d_vis_stage, d_dirs, d_dd, d_avg, d_phase, d_cal1, d_delta, d_cal,
        d_partial, d_a_planes;
    leap::DeviceBuffer d_nufft_tail;  // partial sums of the NUFFT launch's chunked last round
    // residual-loss gradient: ENU->UVW rotations, baseline antennas, antenna corrections, perturbed
    // uvw, avgData adjoint, uvw / dz gradients, the antennas' baseline lists and their gradient
    leap::DeviceBuffer d_adj_M, d_adj_ants, d_adj_dp, d_adj_uvw, d_adj_abar, d_adj_guvw, d_adj_gdzp, d_adj_gdz,
        d_adj_off, d_adj_bls, d_adj_sgn, d_adj_gant;
    // pol 0 of avgData with its arguments; the inversion adjoint's intermediates; the
    // per-direction losses and offset gradients copied back (pinned)
    leap::DeviceBuffer d_adj_avg0, d_adj_inv;
    // the forward pass's per-row sums (e0 P, e0 z P') for the adjoint, and the key
    // (antennas of every baseline, antenna count) the baseline tables were built for
    leap::DeviceBuffer d_adj_q;
    uint64_t adj_ants_key = 0;
    int32_t adjoint_mode = LEAP_ADJOINT_AUTO;
    leap::PinnedBuffer h_adj_out;
    hipEvent_t ev_grad[4] = {nullptr, nullptr, nullptr, nullptr};  // around the forward and adjoint kernels
    double last_grad_forward_ms = 0.0, last_grad_adjoint_ms = 0.0;
    std::vector<double> h_avg;
    // calibrations of the last solution interval copied back (pinned, so the copy is a direct DMA)
    leap::PinnedBuffer h_cal;
    // pseudo-inverse cache keyed by a hash of A (use_cache)
    std::map<uint64_t, std::vector<double>> gram_cache;
    // autotuned direction blocks of the batched kernel
    std::map<TuneKey, int32_t> tuned;
    int32_t last_direction_block = 0;
    int32_t last_channel_sum = 0, last_direct_cube = 0;  // route of the last batched call
    // multi-set launches
    std::vector<std::unique_ptr<leap::SetBuffers>> batch_sets;
    leap::DeviceBuffer d_tables;     // descriptor tables of the last multi-set launch
    std::vector<uint8_t> h_tables;
    std::map<std::vector<uint32_t>, std::pair<int32_t, int32_t>> tuned_multi;  // -> tile, workgroups per CU
    int32_t last_workgroups_per_cu = 0;
    int32_t cu_count = 0;
    // HIP graphs
    bool capturing = false;
    int32_t last_graph = 0;  // 0 none, 1 captured, 2 replayed
    std::unique_ptr<leap::CapturedSequence> seq_single, seq_batch;
};

namespace leap {
// One measurement set of a multi-set launch (device-resident descriptor).
struct MultiSetDesc {
    const void* vis;    // [rows][C][P] complex, rows = T * nbl
    const double* uvw;  // [rows][3]
    const double* dd;   // [D][3][3]
    double* avg;        // [D][nbl][P] complex128 (pol 0 written, others zeroed)
    uint32_t nbl, T, C, P, D;
    int32_t c64;
    double kappa0, dkappa;
};
void launch_rotate_multi(hipStream_t s, int32_t tile, const MultiSetDesc* descs, const uint2* items,
                         const uint32_t* off, uint32_t n_blocks);
// The autotuners' choices as JSON (tuning.cpp).
std::string tuning_to_json(const leap_ctx& ctx);
void tuning_from_json(leap_ctx& ctx, const char* text);
// Host -> device staging (staging.cpp).  Both order their device writes after
// the work already on ctx->stream and make later work on ctx->stream wait for
// them; host memory is read before they return.
// Copies bytes from host memory to device memory through the pinned buffers.
void stage_copy(leap_ctx* ctx, const void* host_src, size_t bytes, void* device_dst);
// Pol 0 of n (row, channel)s of a host cube (VT = float or double, P pols),
// 0 where any pol is not finite, as FP64 complex into device memory; only
// pol 0 crosses the bus (complex64 sources are widened on the device).
void stage_pol0(leap_ctx* ctx, const void* host_vis, int32_t vis_dtype, uint32_t P, size_t n, double* device_dst);
void launch_widen_c64(hipStream_t s, const void* in, size_t n, double* out);
// Packed pol-0 path: pack the cube, then FP64 Horner fed by scalar loads.  With
// scale non-null every element is multiplied by scale[channel] (n counts whole
// rows of C channels).
void launch_pack_pol0(hipStream_t s, const void* vis, int32_t vis_dtype, size_t n, uint32_t P, double* out,
                      const double* scale = nullptr, uint32_t C = 0);
// x[i] *= scale[i % C] for n packed FP64 complex values (whole rows).
void launch_scale_channels(hipStream_t s, double* x, size_t n, const double* scale, uint32_t C);
// NUFFT channel sums (kernels_nufft.hip) over precorrected packed pol-0 rows
// (cube false) or straight from a cube of vis_dtype with 1, 2 or 4 pols (cube
// true; complex64 for every plan, complex128 for the width-16 plan; scale =
// the plan's precorrection factors on the device); tiles 641 / 1282 / 2564 as
// for launch_rotate_packed.
// tables: the plan on the device, kNufftCoefDoubles coefficients (coef[j][m])
// then the C precorrection factors.
bool nufft_cube_supported(int32_t vis_dtype, uint32_t P, bool w16);
// avg holds P_out pols per (direction, baseline) (pol 0 summed, the others
// set to zero unless accumulating); P is the cube's pol count.
void launch_rotate_nufft(hipStream_t s, int32_t tile, const NufftPlan& plan, const double* tables, const void* src,
                         bool cube, int32_t vis_dtype, const double* uvw, const double* dd, uint32_t D, uint32_t T,
                         uint32_t nbl,
                         uint32_t nbl_grid, uint32_t C, uint32_t P, uint32_t P_out, double f0, double df,
                         double* avg, int32_t accumulate, double* tail = nullptr, size_t tail_bytes = 0);
// Scratch for launch_rotate_nufft's chunked last round (D directions, a single
// direction block): at most twice the resident workgroups x D complex sums.
size_t nufft_tail_bytes(uint32_t D);
// One set of a multi-set NUFFT launch (FP64 transform, complex64 cube).
struct NufftSetDesc {
    const void* vis;      // [T nbl][C][P] complex64 (or complex128 for the width-16 plan)
    const double* uvw;    // [T nbl][3]
    const double* dd;     // [D][3][3]
    double* avg;          // [D][nbl] complex128: pol 0 only
    const double* scale;  // C precorrection factors
    const double* coef;   // weight polynomials (the plan's device tables)
    const float* coef32;
    double kappa_mid, dkappa;  // (f0 + h df) / c_light, df / c_light
    int32_t h;
    uint32_t nbl, T, C, D;
};
// Sets whose keys are equal share one multi-set NUFFT launch (same grid, kernel
// width, channel slots per lane and row source: P the cube's pols, plus 16 for
// complex128).
uint64_t nufft_sets_key(const NufftPlan& plan, uint32_t C, uint32_t P);
// NUFFT channel sums of several sets in one launch: workgroup x takes
// items[x] = (set | direction block << 16, baseline); tile 641 / 1282 / 2564.
// plan, C and P (the row source: the pols, plus 16 for complex128): those of
// any set of the launch (equal keys).
void launch_rotate_nufft_sets(hipStream_t s, int32_t tile, const NufftPlan& plan, uint32_t C, uint32_t P,
                              const NufftSetDesc* sets, const uint32_t* items, uint32_t n_items);
bool packed_tile_supported(int32_t tile);
void launch_rotate_packed(hipStream_t s, int32_t tile, const double* w, const double* uvw, const double* dd,
                          uint32_t D, uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double f0,
                          double df, double* avg, int32_t accumulate);
int32_t multi_blocks_per_cu(int32_t tile);
// Kernel launchers (kernels_f64.hip)
void launch_dd(hipStream_t s, double ra0, double dec0, const double* d_dirs, uint32_t D,
               double* d_dd);
void launch_rotate_f64(hipStream_t s, const void* vis, int32_t vis_dtype, const double* uvw,
                       const double* dd, uint32_t row_start, uint32_t T, uint32_t nbl, uint32_t C,
                       uint32_t P, const double* lambda, double* avg);
// G+ = pinv(A^T A) on the device (kernels_f64.hip); all pointers device memory.
struct GramPinv {
    int32_t N, constraint;
    const int32_t* csr_off;   // N + 1: antenna a's rows of A
    const int32_t* csr_ent;   // 2 k + (a is antenna2 of row k)
    const int2* pairs;        // antenna pairs of A's rows
    int32_t n_components;     // components to invert (unused antennas left out)
    const int32_t* comp_off;  // n_components + 1 offsets into members
    const int32_t* members;   // antennas of each component, ascending
    const uint8_t* has_c;     // 1 where the component holds the constraint antenna
    const int64_t* ws_off;    // offset (doubles) of each component's 2 n^2 workspace
    double* ws;
    double* G;                // N x N scratch (the Gram matrix)
    double* Gp;               // N x N output
    int32_t* status;          // set to 1 when a block is not positive definite
};
void launch_gram_pinv(hipStream_t s, const GramPinv& g);
// Components of the baseline graph of A (plus the constraint) whose block of G+
// is computed: every component but unused antennas (union-find, host).
struct GramComponents {
    std::vector<int32_t> off, members;
    std::vector<uint8_t> has_c;
};
GramComponents gram_components(const PhaseMatrix& pm);
void launch_assemble_ad(hipStream_t s, const double* Gp, int32_t N, const int2* pairs,
                        int32_t n_pairs, int32_t constraint, double* AdT);
void launch_inversion_f64(hipStream_t s, const leap_metadata& md, const double* avg, uint32_t D,
                          double* phase, double* cal1, double* delta, double* partial, double* cal);
void launch_rotate_batched(hipStream_t s, int32_t precision, int32_t direction_block, bool all_pols,
                           const void* vis, int32_t vis_dtype, const double* uvw, const double* dd, uint32_t D,
                           uint32_t T, uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, double f0,
                           double df, double* avg, int32_t fp32_blocks);
size_t inversion_partial_doubles(int32_t K, int32_t N, uint32_t D);
// Batched inversion pieces: cal1 / delta for D directions, then one GEMM.
// phase and cal1 for D directions, and delta unless it is null (the mixed GEMM
// computes its operand from avgData itself).
// avg holds avg_pols pols per (direction, baseline); pol 0 is used.
void launch_inversion_head(hipStream_t s, const leap_metadata& md, const double* avg, uint32_t avg_pols, uint32_t D,
                           double* phase, double* cal1, double* delta);
size_t gemm_f64_partial_doubles(int32_t K, int32_t N, uint32_t D);
void launch_gemm_f64(hipStream_t s, const double* AdT, int32_t K, int32_t N, const double* delta, uint32_t D,
                     const double* cal1, double* partial, double* cal);
size_t bf16_planes_bytes(uint32_t rows, int32_t K);
void launch_split_b(hipStream_t s, const double* B, int32_t K, int32_t N, void* planes);
size_t mfma_partial_floats(int32_t K, uint32_t D, int32_t N);
// Where the mixed GEMM's A operand comes from when it is computed in place of
// the FP64 delta matrix: avgData and the antenna pairs of A's rows.
struct DeltaSource {
    const double* avg;
    uint32_t nbl, pols;
    const int2* pairs;
};
void launch_gemm_bf16x3(hipStream_t s, const void* b_planes, int32_t K, int32_t N, const double* delta, uint32_t D,
                        void* a_planes, const double* cal1, float* partial, double* cal,
                        const DeltaSource* from = nullptr);
void launch_perturb_uvw(hipStream_t s, const double* uvw0, const double* M, const int2* bl_ants, const double* dp,
                        uint32_t nbl, uint32_t rows, double* uvw);
void launch_rotate_adjoint(hipStream_t s, const void* vis, int32_t vis_dtype, const double* uvw, const double* dd,
                           const double* abar, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double f0,
                           double df, double* g_uvw, double* g_dz_partial);
// The residual-loss forward pass that also keeps each row's sums
// q[row][d] = (e0 P(z), e0 z P'(z)) (4 doubles; kernels_batched.hip), pol 0 of
// avgData as launch_rotate_batched in FP64; tiles 162, 321, 641, 1282, 2562.
void launch_rotate_forward_q(hipStream_t s, int32_t tile, const void* vis, int32_t vis_dtype, const double* uvw,
                             const double* dd, uint32_t D, uint32_t T, uint32_t nbl, uint32_t C, uint32_t P, double f0,
                             double df, double* avg, double* q);
// launch_rotate_adjoint's outputs from those row sums (no pass over the cube).
void launch_adjoint_q(hipStream_t s, const double* q, const double* uvw, const double* dd, const double* abar,
                      uint32_t D, uint32_t T, uint32_t nbl, double f0, double df, double* g_uvw,
                      double* g_dz_partial);
// g_dz[d][k] (3 D doubles at out) reduced over baselines, then dL/d(ra, dec) of each
// direction (2 D doubles at out + 3 D); dirs = the directions (ra, dec) on the device.
void launch_reduce_gdz_offsets(hipStream_t s, const double* part, uint32_t nbl, uint32_t D, double ra0, double dec0,
                               const double* dirs, double* out);
void launch_gather_pol0(hipStream_t s, const double* avg, uint32_t n, uint32_t P, double* out);
// The residual loss of D calibrations and the adjoint of the inversion
// (kernels_adjoint.hip); every array [d][...] on the device.
struct InversionAdjoint {
    uint32_t D, nbl;
    int32_t K, N, K1;
    const double* cal;       // D x N calibrations (forward inversion)
    const double* avg0;      // D x nbl complex pol 0 of avgData
    const double* ph;        // D x nbl arguments of avg0
    const int2* pairs;       // K - 1 antenna pairs of A's rows
    const int32_t* I;        // K - 1 baselines of A's rows
    const int2* pairs1;      // K1 - 1 antenna pairs of A1's rows
    const int32_t* I1;       // K1 - 1 baselines of A1's rows
    const int32_t* csr_off;  // N + 1: antenna a's rows are csr_ent[csr_off[a] .. csr_off[a + 1])
    const int32_t* csr_ent;  // 2 (K - 1): 2 row + 1 where a is the row's antenna2 (sign -1)
    const double* Gp;        // N x N pseudo-inverse of A^T A
    const double* G1p;       // N x N pseudo-inverse of A1^T A1
    double* res;             // D x (K - 1) residuals
    double* part;            // D x ceil((K - 1) / 256) partial loss sums
    double* loss;            // D per-direction losses
    // backward (skipped when abar is null)
    double* g_arg;           // D x nbl
    double* g_cal;           // D x N
    double* y;               // D x N
    double* gacal1;          // D x (K - 1)
    double* g_cal1;          // D x N
    double* y1;              // D x N
    double* abar;            // D x nbl complex: dL/dRe avg + i dL/dIm avg
};
void launch_inversion_adjoint(hipStream_t s, const InversionAdjoint& a);
void launch_antenna_grad(hipStream_t s, const double* g_uvw, const double* M, const int32_t* csr_off,
                         const int32_t* csr_bl, const int32_t* csr_sign, int32_t n_ant, uint32_t nbl, uint32_t T,
                         double* out);
void launch_synth_vis(hipStream_t s, const leap_synth_ms* ms, int32_t vis_dtype, void* out,
                      const double* d_uvw);
}  // namespace leap
