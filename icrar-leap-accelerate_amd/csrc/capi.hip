// C ABI of the gfx950 LEAP calibrator: contexts, metadata, the two halves of
// PhaseRotate, the full Calibrate loop, synthetic sets and small helpers.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <sstream>
#include <thread>

#include "capi_common.h"
#include "leap_internal.h"
#include "phasor.h"
#include "synth.h"

using namespace leap;

namespace {

leap_options resolve(const leap_options* o) {
    leap_options r{};
    if (o) r = *o;
    require(r.precision == LEAP_PRECISION_FP64 || r.precision == LEAP_PRECISION_MIXED, LEAP_ERR_INVALID_ARGUMENT,
            "unknown precision");
    require(r.strategy == LEAP_STRATEGY_PER_DIRECTION || r.strategy == LEAP_STRATEGY_BATCHED,
            LEAP_ERR_INVALID_ARGUMENT, "unknown strategy");
    require(!(r.precision == LEAP_PRECISION_MIXED && r.strategy != LEAP_STRATEGY_BATCHED),
            LEAP_ERR_INVALID_ARGUMENT, "mixed precision needs the batched strategy");
    require(r.direction_block == 0 || r.direction_block == 32 || r.direction_block == 64 ||
                r.direction_block == 128 || r.direction_block == 256,
            LEAP_ERR_INVALID_ARGUMENT, "direction block must be 0 (tuned), 32, 64, 128 or 256");
    require(r.lane_directions == 0 || r.lane_directions == 1 || r.lane_directions == 2 || r.lane_directions == 4,
            LEAP_ERR_INVALID_ARGUMENT, "lane directions must be 0 (tuned), 1, 2 or 4");
    require(r.lane_directions <= 1 || !(r.precision == LEAP_PRECISION_MIXED && r.fp32_blocks), LEAP_ERR_INVALID_ARGUMENT,
            "the FP32 blocked kernel runs one direction per lane");
    require(r.lane_directions == 0 || r.direction_block == 0 || r.direction_block >= 32 * r.lane_directions,
            LEAP_ERR_INVALID_ARGUMENT, "direction block too small for the lane directions");
    require(r.graph == 0 || r.graph == 1, LEAP_ERR_INVALID_ARGUMENT, "graph must be 0 or 1");
    require(r.direct_cube == 0 || r.direct_cube == 1, LEAP_ERR_INVALID_ARGUMENT, "direct_cube must be 0 or 1");
    require(r.pack_timesteps >= 0, LEAP_ERR_INVALID_ARGUMENT, "pack_timesteps must be >= 0");
    require(r.workgroups_per_cu >= -1 && r.workgroups_per_cu <= 64, LEAP_ERR_INVALID_ARGUMENT,
            "workgroups per CU must be -1 (one work item each), 0 (tuned) or 1 to 64");
    require(r.channel_sum >= LEAP_CHANNEL_SUM_AUTO && r.channel_sum <= LEAP_CHANNEL_SUM_NUFFT_W16,
            LEAP_ERR_INVALID_ARGUMENT,
            "channel_sum must be 0 (automatic), 1 (Horner), 2 (NUFFT), 3 (NUFFT, FP32 transform) or 4 (NUFFT, "
            "width 16)");
    if (is_nufft(r.channel_sum)) {
        require(r.precision == LEAP_PRECISION_MIXED || r.channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16,
                LEAP_ERR_INVALID_ARGUMENT,
                "the width-10 NUFFT channel sums are a mixed-precision method (FP64: channel_sum 4, width 16)");
        require(!r.all_pols && !r.fp32_blocks, LEAP_ERR_INVALID_ARGUMENT,
                "the NUFFT channel sums run on pol 0 (no all_pols or fp32_blocks)");
    }
    return r;
}

size_t vis_elem_bytes(const leap_ms_view& ms) {
    return (ms.vis_dtype == LEAP_VIS_C64 ? 8 : 16) * size_t(ms.num_pols);
}

void record_rotate_begin(leap_ctx* ctx) {
    if (ctx->capturing) return;  // a captured sequence is timed as a whole
    if (ctx->rotate_events_used == ctx->rotate_events.size()) {
        hipEvent_t a, b;
        LEAP_HIP(hipEventCreate(&a));
        LEAP_HIP(hipEventCreate(&b));
        ctx->rotate_events.emplace_back(a, b);
    }
    LEAP_HIP(hipEventRecord(ctx->rotate_events[ctx->rotate_events_used].first, ctx->stream));
}

void record_rotate_end(leap_ctx* ctx) {
    if (ctx->capturing) return;
    LEAP_HIP(hipEventRecord(ctx->rotate_events[ctx->rotate_events_used].second, ctx->stream));
    ++ctx->rotate_events_used;
}

void record_pack_begin(leap_ctx* ctx, hipStream_t st) {
    if (ctx->capturing) return;
    if (ctx->pack_events_used == ctx->pack_events.size()) {
        hipEvent_t a, b;
        LEAP_HIP(hipEventCreate(&a));
        LEAP_HIP(hipEventCreate(&b));
        ctx->pack_events.emplace_back(a, b);
    }
    LEAP_HIP(hipEventRecord(ctx->pack_events[ctx->pack_events_used].first, st));
}

void record_pack_end(leap_ctx* ctx, hipStream_t st) {
    if (ctx->capturing) return;
    LEAP_HIP(hipEventRecord(ctx->pack_events[ctx->pack_events_used].second, st));
    ++ctx->pack_events_used;
}

void begin_timing(leap_ctx* ctx) {
    ctx->last_stage_bytes = 0;
    ctx->pack_events_used = 0;
    ctx->rotate_events_used = 0;
    LEAP_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
}

// Call after the stream has been synchronised.
void end_timing(leap_ctx* ctx) {
    LEAP_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    LEAP_HIP(hipEventSynchronize(ctx->ev_stop));
    float total = 0.f;
    LEAP_HIP(hipEventElapsedTime(&total, ctx->ev_start, ctx->ev_stop));
    double sum = 0.0;
    for (size_t i = 0; i < ctx->rotate_events_used; ++i) {
        float ms = 0.f;
        LEAP_HIP(hipEventElapsedTime(&ms, ctx->rotate_events[i].first, ctx->rotate_events[i].second));
        sum += ms;
    }
    double pack_sum = 0.0;
    for (size_t i = 0; i < ctx->pack_events_used; ++i) {
        float ms = 0.f;
        LEAP_HIP(hipEventElapsedTime(&ms, ctx->pack_events[i].first, ctx->pack_events[i].second));
        pack_sum += ms;
    }
    ctx->last_pack_launches = int32_t(ctx->pack_events_used);
    ctx->last_pack_ms_avg = ctx->pack_events_used ? pack_sum / double(ctx->pack_events_used) : 0.0;
    ctx->last_total_ms = total;
    ctx->last_rotate_launches = int32_t(ctx->rotate_events_used);
    ctx->last_rotate_ms_avg = ctx->rotate_events_used ? sum / double(ctx->rotate_events_used) : 0.0;
}

// Device pointers of visibilities and uvw covering rows [row_begin, row_end);
// host data is staged into the context's buffers (visibilities through the
// pinned double buffer).  Returned pointers address row row_begin.  With
// with_vis false host visibilities are left where they are: vis is null and
// host_vis addresses row row_begin (the packed path stages pol 0 only).
struct DeviceRows {
    const void* vis;
    const double* uvw;
    const void* host_vis;
};

DeviceRows stage_rows_into(leap_ctx* ctx, const leap_ms_view& ms, uint32_t row_begin, uint32_t row_end,
                           DeviceBuffer& vis_buf, DeviceBuffer& uvw_buf, bool with_vis = true) {
    require(row_end <= ms.num_rows && row_begin <= row_end, LEAP_ERR_RANGE,
            "row selection [" + std::to_string(row_begin) + "," + std::to_string(row_end) +
                "] exceeds total range [0," + std::to_string(ms.num_rows) + "]");
    require(ms.vis != nullptr, LEAP_ERR_INVALID_ARGUMENT, "measurement set view has no visibilities");
    const size_t n = size_t(row_end - row_begin);
    const size_t vb = vis_elem_bytes(ms) * ms.num_channels;
    DeviceRows r{};
    if (ms.vis_on_device) {
        r.vis = static_cast<const char*>(ms.vis) + size_t(row_begin) * vb;
    } else if (!with_vis) {
        r.vis = nullptr;
        r.host_vis = static_cast<const char*>(ms.vis) + size_t(row_begin) * vb;
    } else {
        vis_buf.reserve(n * vb);
        stage_copy(ctx, static_cast<const char*>(ms.vis) + size_t(row_begin) * vb, n * vb, vis_buf.ptr);
        r.vis = vis_buf.ptr;
    }
    if (ms.uvw_device) {
        r.uvw = ms.uvw_device + size_t(row_begin) * 3;
    } else {
        uvw_buf.reserve(n * 3 * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(uvw_buf.ptr, ms.uvw + size_t(row_begin) * 3, n * 3 * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
        r.uvw = uvw_buf.as<double>();
    }
    return r;
}

DeviceRows stage_rows(leap_ctx* ctx, const leap_ms_view& ms, uint32_t row_begin, uint32_t row_end,
                      bool with_vis = true) {
    return stage_rows_into(ctx, ms, row_begin, row_end, ctx->d_vis_stage, ctx->d_uvw, with_vis);
}

void upload_directions(leap_ctx* ctx, const leap_metadata& md, const double* dirs, uint32_t D) {
    ctx->d_dirs.reserve(size_t(D) * 2 * sizeof(double));
    ctx->d_dd.reserve(size_t(D) * 9 * sizeof(double));
    LEAP_HIP(hipMemcpyAsync(ctx->d_dirs.ptr, dirs, size_t(D) * 2 * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream));
    launch_dd(ctx->stream, md.phase_centre_ra, md.phase_centre_dec, ctx->d_dirs.as<double>(), D,
              ctx->d_dd.as<double>());
}

// Whether the NUFFT kernel can read the cube itself: a supported type and pol
// count at an address its loads allow (16 bytes for complex128 and for
// complex64 with 2 or 4 pols, whose channels it loads 16 bytes at a time; 8
// for one complex64 pol).  Row offsets keep that alignment.
bool nufft_cube_ok(const leap_ms_view& ms, bool w16) {
    const uintptr_t align = ms.vis_dtype == LEAP_VIS_C128 || ms.num_pols >= 2 ? 16 : 8;
    return nufft_cube_supported(ms.vis_dtype, ms.num_pols, w16) &&
           (!ms.vis_on_device || reinterpret_cast<uintptr_t>(ms.vis) % align == 0);
}

// The batched pol-0 path through the pol-0 pack and the scalar-load Horner kernel.
bool use_packed(const leap_options& opts) {
    return opts.strategy == LEAP_STRATEGY_BATCHED && !opts.all_pols &&
           !(opts.precision == LEAP_PRECISION_MIXED && opts.fp32_blocks) && !opts.direct_cube;
}

// The batched pol-0 route for a tuned call on a device-resident cube: the pack
// pass (every pol read, pol 0 written as FP64) pays only when the FP64 work is
// large enough.  Rates measured on MI355X at the 256-station set (32,896
// baselines, 128 t, 384 ch): pack 5.2 TB/s; packed Horner 32 / 50 / 55 TFLOP/s
// with 1 / 2 / 4 directions per lane (up to 64, 128, more directions); the
// Horner kernel straight from the cube reads at most 6.4 TB/s and computes
// 30 TFLOP/s up to 128 directions, 51 (complex64) or 45 (complex128) above.
// E.g. 256 directions: from the cube for complex64 (65 vs 72 ms) and
// complex128 (74 vs 82 ms); 64 directions on complex128: from the cube (27 vs
// 50 ms); the pack pays from about 800 (complex64) or 480 (complex128)
// directions.  Host-resident cubes always take the packed route (pol 0 only
// crosses the bus).
leap_options routed(const leap_options& o, const leap_ms_view& ms, uint32_t T, uint32_t D) {
    leap_options r = o;
    if (r.channel_sum == LEAP_CHANNEL_SUM_AUTO) {
        const bool fp64 = r.precision == LEAP_PRECISION_FP64;
        const bool nufft = use_packed(r) && nufft_supported(ms.num_channels) &&
                           ms.num_channels >= kNufftMinChannels && D >= kNufftMinDirections;
        r.channel_sum = !nufft ? LEAP_CHANNEL_SUM_HORNER : fp64 ? LEAP_CHANNEL_SUM_NUFFT_W16 : LEAP_CHANNEL_SUM_NUFFT;
    }
    if (is_nufft(r.channel_sum)) {
        require(nufft_supported(ms.num_channels), LEAP_ERR_INVALID_ARGUMENT,
                "the NUFFT channel sums support 1.." + std::to_string(kNufftMaxChannels) + " channels");
        const bool w16 = r.channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16;
        const bool cube_ok = nufft_cube_ok(ms, w16);
        require(!r.direct_cube || cube_ok, LEAP_ERR_INVALID_ARGUMENT,
                w16 ? "the NUFFT kernel reads the cube directly only with 1, 2 or 4 pols at an aligned address"
                    : "the NUFFT kernel reads the cube directly only for complex64 with 1, 2 or 4 pols at an "
                      "aligned address");
        // A device-resident cube of a supported type is read by the NUFFT kernel itself
        // (no pack pass) unless pack chunks were asked for; host cubes are packed
        // while they are staged (pol 0 only crosses the bus).
        if (cube_ok && ms.vis_on_device && r.pack_timesteps == 0) r.direct_cube = 1;
        return r;
    }
    if (!use_packed(o) || o.direction_block != 0 || !ms.vis_on_device) return r;
    const double rc = double(T) * ms.num_baselines * ms.num_channels;  // (row, channel)s
    const double s = ms.vis_dtype == LEAP_VIS_C64 ? 8.0 : 16.0;
    const double flops = 8.0 * rc * D;
    const double packed = rc * (ms.num_pols * s + 16.0) / 5.2e12 + flops / (D > 128 ? 55e12 : D > 64 ? 50e12 : 32e12);
    const double direct_rate = D > 128 ? (ms.vis_dtype == LEAP_VIS_C64 ? 51e12 : 45e12) : 30e12;
    const double direct = std::max(rc * ms.num_pols * s / 6.4e12, flops / direct_rate);
    if (packed > direct) r.direct_cube = 1;
    return r;
}

// Whether a batched rotate takes the packed path: the caller's tile (or, when
// tuning, the requested one) must be a packed tile.
bool packed_route(const leap_options& opts, int32_t tile) {
    if (!use_packed(opts)) return false;
    if (tile) return packed_tile_supported(tile);
    if (opts.direction_block == 0) return true;
    return packed_tile_supported(opts.direction_block * 10 + (opts.lane_directions ? opts.lane_directions : 1));
}

// Direction block for the batched kernel: the fastest of the candidates on a
// slice of the baselines, timed once per shape and kept in the context.
int32_t tuned_direction_block(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const DeviceRows& rows,
                              uint32_t T, uint32_t D, const leap_options& opts,
                              const std::function<void()>& before_probe = nullptr) {
    const bool fp32 = opts.precision == LEAP_PRECISION_MIXED && opts.fp32_blocks;
    if (opts.direction_block != 0 && (opts.lane_directions != 0 || fp32))
        return opts.direction_block * 10 + (fp32 ? 1 : opts.lane_directions);
    if (opts.direction_block != 0 && opts.lane_directions == 0) return opts.direction_block * 10 + 1;
    const bool packed = use_packed(opts);
    const bool nufft = is_nufft(opts.channel_sum);
    const TuneKey key{opts.precision * 4 + (opts.fp32_blocks ? 2 : 0) + (opts.lane_directions ? 1 : 0) +
                          16 * opts.lane_directions + (packed ? kTunePackedBit : 0) + (nufft ? kTuneNufftBit : 0) +
                          (opts.channel_sum == LEAP_CHANNEL_SUM_NUFFT_FP32 ? kTuneNufft32Bit : 0) +
                          (opts.channel_sum == LEAP_CHANNEL_SUM_NUFFT_W16 ? kTuneNufftW16Bit : 0),
                      ms.vis_dtype, opts.all_pols, D, T, md.channels, md.pols};
    auto it = ctx->tuned.find(key);
    if (it != ctx->tuned.end()) return it->second;
    // Candidate tiles (directions per workgroup * 10 + directions per lane).
    std::vector<int32_t> candidates;
    const int32_t all64[] = {321, 641, 642, 1281, 1282, 1284, 2561, 2562, 2564};
    const int32_t all32[] = {321, 641, 1281, 2561};
    if (packed || nufft) {
        // 64 lanes per row: 64, 128 or 256 directions per workgroup
        for (int32_t t : {641, 1282, 2564})
            if ((t == 641 || uint32_t(t / 10) <= 2 * D - 2) &&
                (opts.lane_directions == 0 || t % 10 == opts.lane_directions))
                candidates.push_back(t);
        if (candidates.empty()) candidates.push_back(opts.lane_directions == 2 ? 1282 : 2564);
    } else if (fp32) {
        candidates.assign(std::begin(all32), std::end(all32));
    } else {
        for (int32_t t : all64)
            if (opts.lane_directions == 0 || t % 10 == opts.lane_directions) candidates.push_back(t);
    }
    // Enough baselines for several rounds of workgroups on every CU.
    const uint32_t probe_bl = std::min<uint32_t>(md.nbl, 8192);
    if (before_probe) before_probe();
    int32_t best = candidates.front();
    float best_ms = 0.f;
    hipEvent_t a, b;
    LEAP_HIP(hipEventCreate(&a));
    LEAP_HIP(hipEventCreate(&b));
    for (int32_t tile : candidates) {
        const uint32_t db = uint32_t(tile / 10);
        if (!packed && !nufft && db > 32 && db >= 2 * D) continue;  // mostly idle lanes
        float best_of = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the first repetition warms the code object
            LEAP_HIP(hipEventRecord(a, ctx->stream));
            if (nufft)
                launch_rotate_nufft(ctx->stream, tile, md.nufft_plan(opts.channel_sum),
                                    md.nufft_tables(opts.channel_sum),
                                    packed ? ctx->d_packed.ptr : rows.vis, !packed, ms.vis_dtype, rows.uvw,
                                    ctx->d_dd.as<double>(),
                                    D, T, md.nbl, probe_bl, md.channels, md.pols, md.pols, md.freq_start_hz,
                                    md.freq_inc_hz, ctx->d_avg.as<double>(), 0);
            else if (packed)
                launch_rotate_packed(ctx->stream, tile, ctx->d_packed.as<double>(), rows.uvw, ctx->d_dd.as<double>(),
                                     D, T, md.nbl, probe_bl, md.channels, md.pols, md.freq_start_hz, md.freq_inc_hz,
                                     ctx->d_avg.as<double>(), 0);
            else
                launch_rotate_batched(ctx->stream, opts.precision, tile, opts.all_pols != 0, rows.vis, ms.vis_dtype,
                                      rows.uvw, ctx->d_dd.as<double>(), D, T, md.nbl, probe_bl, md.channels, md.pols,
                                      md.freq_start_hz, md.freq_inc_hz, ctx->d_avg.as<double>(), opts.fp32_blocks);
            LEAP_HIP(hipEventRecord(b, ctx->stream));
            LEAP_HIP(hipEventSynchronize(b));
            LEAP_HIP(hipEventElapsedTime(&best_of, a, b));
        }
        if (best_ms == 0.f || best_of < best_ms) {
            best_ms = best_of;
            best = tile;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ctx->tuned.emplace(key, best);
    return best;
}

// Timesteps per pack chunk.  Measured on MI355X at 256 stations / 128
// timesteps / 384 channels / 256 directions, one chunk (pack, then Horner)
// and overlapped chunks take the same time: the FP64 Horner kernel runs at the
// power-limited clock, so a concurrent pack slows it by about what it hides.
// Chunks therefore only bound the pack buffer: one chunk up to 32 GiB, else
// the fewest chunks whose two buffers fit in it.
uint32_t pack_chunk(uint32_t T, uint32_t nbl, uint32_t C, const leap_options& opts) {
    if (opts.pack_timesteps > 0) return std::min<uint32_t>(T, uint32_t(opts.pack_timesteps));
    const size_t per_t = size_t(nbl) * C * 2 * sizeof(double);
    const size_t cap = size_t(32) << 30;
    if (size_t(T) * per_t <= cap) return T;
    return uint32_t(std::max<size_t>(1, cap / (2 * per_t)));
}

// The packed pol-0 path over the staged rows.  Chunk 0 is packed first (and
// the tile tuned on it when needed); further chunks of timesteps are packed on
// the pack stream into two alternating buffers while the Horner kernel runs the
// previous chunk on the context stream, accumulating into avgData.  Events
// order each chunk's pack before its Horner launch and each Horner launch
// before the pack that reuses its buffer.  Host-resident visibilities are
// packed by the host staging instead (pol 0 only crosses the bus), which
// orders itself on the context stream.
void rotate_packed_pipeline(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const DeviceRows& rows,
                            uint32_t T, uint32_t D, const leap_options& opts, int32_t tile, uint32_t avg_pols) {
    const uint32_t tc = pack_chunk(T, md.nbl, md.channels, opts);
    const uint32_t n_chunks = (T + tc - 1) / tc;
    const size_t per_t = size_t(md.nbl) * md.channels;  // (row, channel)s per timestep
    const size_t buf_elems = size_t(tc) * per_t;
    ctx->d_packed.reserve(buf_elems * (n_chunks > 1 ? 2 : 1) * 2 * sizeof(double));
    const size_t vis_row_bytes = vis_elem_bytes(ms) * md.channels;
    const bool host = rows.vis == nullptr;
    const bool nufft = is_nufft(opts.channel_sum);
    const double* scale = nufft ? md.nufft_scale(opts.channel_sum) : nullptr;
    if (nufft) ctx->d_nufft_tail.reserve(nufft_tail_bytes(D));
    double* buf[2] = {ctx->d_packed.as<double>(), ctx->d_packed.as<double>() + 2 * buf_elems};
    auto pack = [&](uint32_t k, hipStream_t st) {
        const uint32_t t0 = k * tc, nt = std::min(tc, T - t0);
        const size_t off = size_t(t0) * md.nbl * vis_row_bytes;
        if (host) {
            stage_pol0(ctx, static_cast<const char*>(rows.host_vis) + off, ms.vis_dtype, md.pols,
                       size_t(nt) * per_t, buf[k % 2]);
            if (nufft) launch_scale_channels(ctx->stream, buf[k % 2], size_t(nt) * per_t, scale, md.channels);
            return;
        }
        record_pack_begin(ctx, st);
        launch_pack_pol0(st, static_cast<const char*>(rows.vis) + off, ms.vis_dtype, size_t(nt) * per_t, md.pols,
                         buf[k % 2], scale, md.channels);
        record_pack_end(ctx, st);
    };
    auto horner = [&](uint32_t k) {
        const uint32_t t0 = k * tc, nt = std::min(tc, T - t0);
        record_rotate_begin(ctx);
        if (nufft)
            launch_rotate_nufft(ctx->stream, tile, md.nufft_plan(opts.channel_sum),
                                md.nufft_tables(opts.channel_sum), buf[k % 2], false, ms.vis_dtype,
                                rows.uvw + size_t(t0) * md.nbl * 3, ctx->d_dd.as<double>(), D, nt, md.nbl, md.nbl,
                                md.channels, md.pols, avg_pols, md.freq_start_hz, md.freq_inc_hz,
                                ctx->d_avg.as<double>(), k > 0 ? 1 : 0, ctx->d_nufft_tail.as<double>(),
                                nufft_tail_bytes(D));
        else
            launch_rotate_packed(ctx->stream, tile, buf[k % 2], rows.uvw + size_t(t0) * md.nbl * 3,
                                 ctx->d_dd.as<double>(), D, nt, md.nbl, md.nbl, md.channels, md.pols,
                                 md.freq_start_hz, md.freq_inc_hz, ctx->d_avg.as<double>(), k > 0 ? 1 : 0);
        record_rotate_end(ctx);
    };
    pack(0, ctx->stream);
    // the pack stream may start on chunk 1 once the staged rows are ready
    LEAP_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    if (tile == 0) tile = tuned_direction_block(ctx, md, ms, rows, std::min(tc, T), D, opts);
    ctx->last_direction_block = tile;
    horner(0);
    if (n_chunks == 1) return;
    if (host) {
        // staging chunk k waits for the work queued before it, so buffer k % 2 is free
        for (uint32_t k = 1; k < n_chunks; ++k) {
            pack(k, ctx->stream);
            horner(k);
        }
        return;
    }
    LEAP_HIP(hipEventRecord(ctx->ev_consumed[0], ctx->stream));
    LEAP_HIP(hipStreamWaitEvent(ctx->pack_stream, ctx->ev_fork, 0));
    for (uint32_t k = 1; k < n_chunks; ++k) {
        const int b = int(k % 2);
        if (k >= 2) LEAP_HIP(hipStreamWaitEvent(ctx->pack_stream, ctx->ev_consumed[b], 0));
        pack(k, ctx->pack_stream);
        LEAP_HIP(hipEventRecord(ctx->ev_packed[b], ctx->pack_stream));
        LEAP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_packed[b], 0));
        horner(k);
        LEAP_HIP(hipEventRecord(ctx->ev_consumed[b], ctx->stream));
    }
}

// Tile of the batched kernel for this shape: the caller's, the cached one, or
// tuned now.  The packed path is tuned on its first chunk, packed into the
// first buffer just before the probe launches.
int32_t batched_tile(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const DeviceRows& rows,
                     uint32_t T, uint32_t D, const leap_options& opts) {
    if (!use_packed(opts)) return tuned_direction_block(ctx, md, ms, rows, T, D, opts);
    const uint32_t tc = std::min(T, pack_chunk(T, md.nbl, md.channels, opts));
    const size_t n = size_t(tc) * md.nbl * md.channels;
    const double* scale = is_nufft(opts.channel_sum) ? md.nufft_scale(opts.channel_sum) : nullptr;
    return tuned_direction_block(ctx, md, ms, rows, tc, D, opts, [&] {
        ctx->d_packed.reserve(n * 2 * sizeof(double));
        launch_pack_pol0(ctx->stream, rows.vis, ms.vis_dtype, n, md.pols, ctx->d_packed.as<double>(), scale,
                         md.channels);
    });
}

// avgData of D directions for T timesteps starting at the staged rows, into
// ctx->d_avg; returns the pols it holds per (direction, baseline).  With
// pol0_only (calibrations: only pol 0 reaches cal, CpuLeapCalibrator.cc:161,171)
// the NUFFT route writes pol 0 alone, a quarter of the bytes for 4 pols.
uint32_t rotate_directions(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const DeviceRows& rows,
                           uint32_t T, uint32_t D, const leap_options& opts, int32_t tile = 0,
                           bool pol0_only = false) {
    const bool nufft_route = opts.strategy == LEAP_STRATEGY_BATCHED && is_nufft(opts.channel_sum);
    const uint32_t avg_pols = pol0_only && nufft_route ? 1u : md.pols;
    const size_t per_dir = size_t(md.nbl) * md.pols * 2;  // room for the full layout (autotune probes)
    ctx->d_avg.reserve(per_dir * D * sizeof(double));
    if (opts.strategy == LEAP_STRATEGY_BATCHED && packed_route(opts, tile)) {
        rotate_packed_pipeline(ctx, md, ms, rows, T, D, opts, tile, avg_pols);
        return avg_pols;
    }
    require(rows.vis != nullptr, LEAP_ERR_INTERNAL, "visibilities were not staged");
    if (opts.strategy == LEAP_STRATEGY_BATCHED && is_nufft(opts.channel_sum)) {
        const int32_t db = tile ? tile : batched_tile(ctx, md, ms, rows, T, D, opts);
        ctx->last_direction_block = db;
        record_rotate_begin(ctx);
        ctx->d_nufft_tail.reserve(nufft_tail_bytes(D));
        launch_rotate_nufft(ctx->stream, db, md.nufft_plan(opts.channel_sum), md.nufft_tables(opts.channel_sum),
                            rows.vis, true, ms.vis_dtype, rows.uvw,
                            ctx->d_dd.as<double>(), D, T, md.nbl, md.nbl, md.channels, md.pols, avg_pols,
                            md.freq_start_hz, md.freq_inc_hz, ctx->d_avg.as<double>(), 0,
                            ctx->d_nufft_tail.as<double>(), nufft_tail_bytes(D));
        record_rotate_end(ctx);
        return avg_pols;
    }
    if (opts.strategy == LEAP_STRATEGY_BATCHED) {
        const int32_t db = tile ? tile : batched_tile(ctx, md, ms, rows, T, D, opts);
        ctx->last_direction_block = db;
        record_rotate_begin(ctx);
        launch_rotate_batched(ctx->stream, opts.precision, db, opts.all_pols != 0, rows.vis, ms.vis_dtype, rows.uvw,
                              ctx->d_dd.as<double>(), D, T, md.nbl, md.nbl, md.channels, md.pols, md.freq_start_hz,
                              md.freq_inc_hz, ctx->d_avg.as<double>(), opts.fp32_blocks);
        record_rotate_end(ctx);
        return avg_pols;
    }
    for (uint32_t d = 0; d < D; ++d) {
        record_rotate_begin(ctx);
        launch_rotate_f64(ctx->stream, rows.vis, ms.vis_dtype, rows.uvw, ctx->d_dd.as<double>() + 9 * d, 0, T,
                          md.nbl, md.channels, md.pols, md.d_lambda.as<double>(),
                          ctx->d_avg.as<double>() + per_dir * d);
        record_rotate_end(ctx);
    }
    return avg_pols;
}

// cal[D][N] of D directions from avgData at avg (ctx->d_avg unless given),
// which holds avg_pols pols per (direction, baseline) (md.pols when 0).
void invert_directions(leap_ctx* ctx, const leap_metadata& md, uint32_t D, const leap_options& opts,
                       const double* avg = nullptr, double* cal = nullptr, uint32_t avg_pols = 0) {
    if (avg_pols == 0) avg_pols = md.pols;
    const int32_t K = md.A.rows, N = md.A.cols, K1 = md.A1.rows;
    ctx->d_phase.reserve(size_t(K1) * D * sizeof(double));
    ctx->d_cal1.reserve(size_t(N) * D * sizeof(double));
    ctx->d_delta.reserve(size_t(K) * D * sizeof(double));
    if (cal == nullptr) {
        ctx->d_cal.reserve(size_t(N) * D * sizeof(double));
        cal = ctx->d_cal.as<double>();
    }
    if (avg == nullptr) avg = ctx->d_avg.as<double>();
    if (opts.strategy != LEAP_STRATEGY_BATCHED) {
        // Per-direction GEMVs, the reference's schedule (CudaLeapCalibrator.cu:256-259).
        ctx->d_partial.reserve(inversion_partial_doubles(K, N, D) * sizeof(double));
        launch_inversion_f64(ctx->stream, md, avg, D, ctx->d_phase.as<double>(),
                             ctx->d_cal1.as<double>(), ctx->d_delta.as<double>(), ctx->d_partial.as<double>(),
                             cal);
        return;
    }
    const bool mixed = opts.precision == LEAP_PRECISION_MIXED;
    launch_inversion_head(ctx->stream, md, avg, avg_pols, D, ctx->d_phase.as<double>(), ctx->d_cal1.as<double>(),
                          mixed ? nullptr : ctx->d_delta.as<double>());
    if (mixed) {
        ctx->d_a_planes.reserve(bf16_planes_bytes((D + 63) / 64 * 64, K));
        ctx->d_partial.reserve(mfma_partial_floats(K, D, N) * sizeof(float));
        const DeltaSource from{avg, md.nbl, avg_pols, md.d_pairs.as<int2>()};
        launch_gemm_bf16x3(ctx->stream, md.d_AdT_planes.ptr, K, N, nullptr, D, ctx->d_a_planes.ptr,
                           ctx->d_cal1.as<double>(), static_cast<float*>(ctx->d_partial.ptr), cal, &from);
    } else {
        ctx->d_partial.reserve(gemm_f64_partial_doubles(K, N, D) * sizeof(double));
        launch_gemm_f64(ctx->stream, md.d_AdT.as<double>(), K, N, ctx->d_delta.as<double>(), D,
                        ctx->d_cal1.as<double>(), ctx->d_partial.as<double>(), cal);
    }
}

void check_metadata_matches(const leap_metadata& md, const leap_ms_view& ms) {
    require(md.nbl == ms.num_baselines && md.channels == ms.num_channels && md.pols == ms.num_pols,
            LEAP_ERR_INVALID_ARGUMENT, "metadata was built for a different measurement set");
}

// Copies a host vector into a device buffer (at least one element is reserved).
template <typename T>
void upload(hipStream_t s, DeviceBuffer& buf, const std::vector<T>& v) {
    buf.reserve(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) LEAP_HIP(hipMemcpyAsync(buf.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

// Each antenna's rows of a phase matrix (baseline rows only, in row order):
// entries 2 k + 1 where the antenna is antenna2 of row k, 2 k where antenna1.
void phase_matrix_csr(const PhaseMatrix& pm, std::vector<int32_t>& off, std::vector<int32_t>& ent) {
    const int32_t N = pm.cols, nI = pm.rows - 1;
    off.assign(size_t(N) + 1, 0);
    ent.assign(size_t(2) * size_t(std::max(nI, 0)), 0);
    for (int32_t k = 0; k < nI; ++k) {
        ++off[size_t(pm.ant1[k]) + 1];
        ++off[size_t(pm.ant2[k]) + 1];
    }
    for (int32_t a = 0; a < N; ++a) off[a + 1] += off[a];
    std::vector<int32_t> fill(off.begin(), off.end() - 1);
    for (int32_t k = 0; k < nI; ++k) {
        ent[size_t(fill[pm.ant1[k]]++)] = 2 * k;
        ent[size_t(fill[pm.ant2[k]]++)] = 2 * k + 1;
    }
}

// G+ = pinv(A^T A) of a phase matrix into d_Gp (N x N, device), from the device
// copies of its row pairs and per-antenna rows; the components come from the host.
void gram_pinv_device(leap_ctx* ctx, const PhaseMatrix& pm, const int2* d_pairs, const int32_t* d_off,
                      const int32_t* d_ent, double* d_Gp) {
    hipStream_t s = ctx->stream;
    const int32_t N = pm.cols;
    const GramComponents comps = gram_components(pm);
    std::vector<int64_t> ws_off(comps.has_c.size() + 1, 0);
    for (size_t c = 0; c < comps.has_c.size(); ++c) {
        const int64_t n = comps.off[c + 1] - comps.off[c];
        ws_off[c + 1] = ws_off[c] + 2 * n * n;
    }
    DeviceBuffer d_comp_off, d_members, d_has_c, d_ws_off, d_ws, d_G, d_status;
    upload(s, d_comp_off, comps.off);
    upload(s, d_members, comps.members);
    upload(s, d_has_c, comps.has_c);
    upload(s, d_ws_off, ws_off);
    d_ws.reserve(std::max<size_t>(1, size_t(ws_off.back())) * sizeof(double));
    d_G.reserve(size_t(N) * N * sizeof(double));
    const std::vector<int32_t> zero(1, 0);
    upload(s, d_status, zero);
    GramPinv g{};
    g.N = N;
    g.constraint = pm.constraint;
    g.csr_off = d_off;
    g.csr_ent = d_ent;
    g.pairs = d_pairs;
    g.n_components = int32_t(comps.has_c.size());
    g.comp_off = d_comp_off.as<int32_t>();
    g.members = d_members.as<int32_t>();
    g.has_c = d_has_c.as<uint8_t>();
    g.ws_off = d_ws_off.as<int64_t>();
    g.ws = d_ws.as<double>();
    g.G = d_G.as<double>();
    g.Gp = d_Gp;
    g.status = d_status.as<int32_t>();
    launch_gram_pinv(s, g);
    int32_t status = 0;
    LEAP_HIP(hipMemcpyAsync(&status, d_status.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    LEAP_HIP(hipStreamSynchronize(s));  // before the scratch buffers go
    require(status == 0, LEAP_ERR_INTERNAL, "phase matrix Gram block is not positive definite");
}

leap_metadata* create_metadata(leap_ctx* ctx, const leap_ms_view& ms, int32_t ref_ant, double thr, int use_cache) {
    validate_view(ms);
    auto md = std::make_unique<leap_metadata>();
    md->device = ctx->device;
    md->nbl = ms.num_baselines;                                                       // MetaData.cc:50
    md->reference_antenna = ref_ant >= 0 ? uint32_t(ref_ant) : ms.total_antennas - 1;  // :51
    md->channels = ms.num_channels;                                                   // :58
    md->freq_start_hz = ms.freq_start_hz;                                             // :59
    md->freq_inc_hz = ms.freq_inc_hz;                                                 // :60
    md->rows = ms.num_rows;
    md->pols = ms.num_pols;
    md->phase_centre_ra = ms.phase_centre_ra_rad;
    md->phase_centre_dec = ms.phase_centre_dec_rad;
    const std::vector<uint8_t> fg = filtered_baselines(ms, thr);  // :84
    const int32_t n = int32_t(ms.num_baselines);
    md->A1 = phase_matrix(ms.antenna1, ms.antenna2, fg.data(), n, int32_t(md->reference_antenna));  // :92
    md->A = phase_matrix(ms.antenna1, ms.antenna2, fg.data(), n, -1);                                // :97
    md->ad_degenerate = null_components(md->A) > 0;    // :129-132
    md->ad1_degenerate = null_components(md->A1) > 0;  // :133-136
    if (ctx->log) {
        if (md->ad_degenerate) ctx->log(ctx->log_user, LEAP_LOG_WARNING, "Ad is degenerate");
        if (md->ad1_degenerate) ctx->log(ctx->log_user, LEAP_LOG_WARNING, "Ad1 is degenerate");
    }
    const int32_t K = md->A.rows, N = md->A.cols, K1 = md->A1.rows;
    // Device copies: antenna pairs and baselines of the rows of A and A1, each
    // antenna's rows (CSR), wavelengths.
    std::vector<int2> pairs(size_t(K - 1)), pairs1(size_t(K1 - 1));
    for (int32_t k = 0; k + 1 < K; ++k) pairs[k] = make_int2(md->A.ant1[k], md->A.ant2[k]);
    for (int32_t k = 0; k + 1 < K1; ++k) pairs1[k] = make_int2(md->A1.ant1[k], md->A1.ant2[k]);
    std::vector<int32_t> off, ent, off1, ent1;
    phase_matrix_csr(md->A, off, ent);
    phase_matrix_csr(md->A1, off1, ent1);
    std::vector<double> lambda(md->channels);
    for (uint32_t c = 0; c < md->channels; ++c)
        lambda[c] = kSpeedOfLight / (md->freq_start_hz + int32_t(c) * md->freq_inc_hz);  // MetaData.h:95-98
    hipStream_t s = ctx->stream;
    upload(s, md->d_pairs, pairs);
    upload(s, md->d_pairs1, pairs1);
    upload(s, md->d_I, md->A.I);
    upload(s, md->d_I1, md->A1.I);
    upload(s, md->d_csr_off, off);
    upload(s, md->d_csr_ent, ent);
    DeviceBuffer d_off1, d_ent1;
    upload(s, d_off1, off1);
    upload(s, d_ent1, ent1);
    upload(s, md->d_lambda, lambda);
    // The pseudo-inverses (MetaData.cc:111, :123) as G+ = pinv(A^T A), Ad = G+ A^T, on the
    // device; with use_cache A's G+ is kept in the context by a hash of A (:112-120).
    md->d_Gp.reserve(size_t(N) * N * sizeof(double) * 2);
    double* gp = md->d_Gp.as<double>();
    double* g1p = gp + size_t(N) * N;
    gram_pinv_device(ctx, md->A1, md->d_pairs1.as<int2>(), d_off1.as<int32_t>(), d_ent1.as<int32_t>(), g1p);
    if (use_cache) {
        const uint64_t key = hash_phase_matrix(md->A);
        auto it = ctx->gram_cache.find(key);
        if (it == ctx->gram_cache.end()) {
            gram_pinv_device(ctx, md->A, md->d_pairs.as<int2>(), md->d_csr_off.as<int32_t>(),
                             md->d_csr_ent.as<int32_t>(), gp);
            std::vector<double> host(size_t(N) * N);
            LEAP_HIP(hipMemcpyAsync(host.data(), gp, host.size() * sizeof(double), hipMemcpyDeviceToHost, s));
            LEAP_HIP(hipStreamSynchronize(s));
            ctx->gram_cache.emplace(key, std::move(host));
        } else {
            LEAP_HIP(hipMemcpyAsync(gp, it->second.data(), it->second.size() * sizeof(double), hipMemcpyHostToDevice,
                                    s));
        }
    } else {
        gram_pinv_device(ctx, md->A, md->d_pairs.as<int2>(), md->d_csr_off.as<int32_t>(), md->d_csr_ent.as<int32_t>(),
                         gp);
    }
    md->d_AdT.reserve(size_t(K) * N * sizeof(double));
    md->d_Ad1T.reserve(size_t(K1) * N * sizeof(double));
    launch_assemble_ad(s, gp, N, md->d_pairs.as<int2>(), K - 1, md->A.constraint, md->d_AdT.as<double>());
    launch_assemble_ad(s, g1p, N, md->d_pairs1.as<int2>(), K1 - 1, md->A1.constraint, md->d_Ad1T.as<double>());
    LEAP_HIP(hipStreamSynchronize(s));  // the host vectors above end with this function
    md->d_AdT_planes.reserve(bf16_planes_bytes(uint32_t(N), K));
    launch_split_b(s, md->d_AdT.as<double>(), K, N, md->d_AdT_planes.ptr);
    std::vector<double> tables32, tables64, tables16;
    if (nufft_supported(md->channels)) {
        md->nufft32 = make_nufft_plan(md->channels, true);
        md->nufft64 = make_nufft_plan(md->channels, false);
        md->nufft16 = make_nufft_plan(md->channels, false, true);
        tables32 = nufft_tables(md->nufft32);
        tables64 = nufft_tables(md->nufft64);
        tables16 = nufft_tables(md->nufft16);
        md->d_nufft32.reserve(tables32.size() * sizeof(double));
        md->d_nufft64.reserve(tables64.size() * sizeof(double));
        md->d_nufft16.reserve(tables16.size() * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(md->d_nufft32.ptr, tables32.data(), tables32.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
        LEAP_HIP(hipMemcpyAsync(md->d_nufft64.ptr, tables64.data(), tables64.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
        LEAP_HIP(hipMemcpyAsync(md->d_nufft16.ptr, tables16.data(), tables16.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    }
    LEAP_HIP(hipStreamSynchronize(s));
    return md.release();
}

// ---- HIP graphs ---------------------------------------------------------------

// Device buffers one solution interval uses, reserved before a capture so the
// captured sequence allocates nothing (the sizes rotate_directions and
// invert_directions ask for).
void reserve_solution(leap_ctx* ctx, const leap_metadata& md, uint32_t D, uint32_t T, const leap_options& opts) {
    const int32_t K = md.A.rows, N = md.A.cols, K1 = md.A1.rows;
    ctx->d_avg.reserve(size_t(md.nbl) * md.pols * 2 * D * sizeof(double));
    if (use_packed(opts)) {
        const uint32_t tc = pack_chunk(T, md.nbl, md.channels, opts);
        ctx->d_packed.reserve(size_t(tc) * md.nbl * md.channels * (T > tc ? 2 : 1) * 2 * sizeof(double));
    }
    if (opts.strategy == LEAP_STRATEGY_BATCHED)
        ctx->d_nufft_tail.reserve(nufft_tail_bytes(D));  // the NUFFT route's chunked last round
    ctx->d_phase.reserve(size_t(K1) * D * sizeof(double));
    ctx->d_cal1.reserve(size_t(N) * D * sizeof(double));
    ctx->d_delta.reserve(size_t(K) * D * sizeof(double));
    ctx->d_cal.reserve(size_t(N) * D * sizeof(double));
    if (opts.strategy != LEAP_STRATEGY_BATCHED) {
        ctx->d_partial.reserve(inversion_partial_doubles(K, N, D) * sizeof(double));
    } else if (opts.precision == LEAP_PRECISION_MIXED) {
        ctx->d_a_planes.reserve(bf16_planes_bytes((D + 63) / 64 * 64, K));
        ctx->d_partial.reserve(mfma_partial_floats(K, D, N) * sizeof(float));
    } else {
        ctx->d_partial.reserve(gemm_f64_partial_doubles(K, N, D) * sizeof(double));
    }
}

uint64_t key_bits(double v) {
    uint64_t u;
    std::memcpy(&u, &v, sizeof(u));
    return u;
}

uint64_t key_ptr(const void* p) { return uint64_t(reinterpret_cast<uintptr_t>(p)); }

// What the kernels of a captured sequence take as arguments: options, shapes
// and device addresses.  A sequence is replayed only for an equal key.
void key_common(std::vector<uint64_t>& k, const leap_ctx* ctx, const leap_options& opts) {
    k.insert(k.end(), {uint64_t(opts.precision), uint64_t(opts.strategy), uint64_t(opts.all_pols),
                       uint64_t(opts.fp32_blocks), uint64_t(opts.direct_cube), uint64_t(opts.pack_timesteps),
                       uint64_t(opts.channel_sum),
                       key_ptr(ctx->d_packed.ptr),
                       key_ptr(ctx->d_dd.ptr), key_ptr(ctx->d_avg.ptr),
                       key_ptr(ctx->d_phase.ptr), key_ptr(ctx->d_cal1.ptr), key_ptr(ctx->d_delta.ptr),
                       key_ptr(ctx->d_cal.ptr), key_ptr(ctx->d_partial.ptr), key_ptr(ctx->d_a_planes.ptr)});
}

void key_set(std::vector<uint64_t>& k, const leap_metadata& md, const leap_ms_view& ms, uint32_t D,
             int32_t solutions, uint32_t interval) {
    k.insert(k.end(), {key_ptr(&md), uint64_t(md.A.rows), uint64_t(md.A.cols), uint64_t(md.A1.rows), md.nbl,
                       md.channels, md.pols, key_bits(md.freq_start_hz), key_bits(md.freq_inc_hz),
                       key_ptr(md.d_AdT.ptr), key_ptr(md.d_Ad1T.ptr), key_ptr(md.d_I1.ptr), key_ptr(md.d_pairs.ptr),
                       key_ptr(md.d_lambda.ptr), key_ptr(md.d_AdT_planes.ptr), key_ptr(ms.vis),
                       key_ptr(ms.uvw_device), uint64_t(ms.vis_dtype), ms.num_rows, D, uint64_t(solutions),
                       interval});
}

// Records what enqueue() puts on the context's stream into seq's graph.
template <typename F>
void capture(leap_ctx* ctx, CapturedSequence& seq, F&& enqueue) {
    LEAP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    try {
        enqueue();
    } catch (...) {
        ctx->capturing = false;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    ctx->capturing = false;
    LEAP_HIP(hipStreamEndCapture(ctx->stream, &seq.graph));
    LEAP_HIP(hipGraphInstantiate(&seq.exec, seq.graph, nullptr, nullptr, 0));
}

void require_device_resident(const leap_ms_view& ms) {
    require(ms.vis_on_device && ms.uvw_device != nullptr, LEAP_ERR_INVALID_ARGUMENT,
            "graph capture needs device-resident visibilities and uvw (vis_on_device, uvw_device)");
}

// calibrate_loop with every solution interval's rotate/average and inversion
// captured once as a HIP graph and replayed while the key holds; results of
// all intervals are copied back together and handed to the callback in order.
void calibrate_loop_graph(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const double* dirs,
                          uint32_t D, int32_t solutions, uint32_t interval, const leap_options& opts,
                          leap_solution_cb cb, void* user) {
    require_device_resident(ms);
    const uint32_t nbl = ms.num_baselines;
    const int32_t N = md.A.cols;
    const size_t per_sol = size_t(N) * D;
    begin_timing(ctx);
    CapturedSequence* seq = nullptr;
    if (D > 0 && solutions > 0) {
        upload_directions(ctx, md, dirs, D);
        reserve_solution(ctx, md, D, interval, opts);
        int32_t tile = 0;
        if (opts.strategy == LEAP_STRATEGY_BATCHED) {
            const DeviceRows rows0 = stage_rows(ctx, ms, 0, interval * nbl);
            tile = batched_tile(ctx, md, ms, rows0, interval, D, opts);
            ctx->last_direction_block = tile;
        }
        std::vector<uint64_t> key{uint64_t(tile)};
        key_common(key, ctx, opts);
        key_set(key, md, ms, D, solutions, interval);
        if (!ctx->seq_single || ctx->seq_single->key != key) {
            ctx->seq_single = std::make_unique<CapturedSequence>();
            seq = ctx->seq_single.get();
            seq->key = key;
            seq->d_cal.reserve(per_sol * size_t(solutions) * sizeof(double));
            capture(ctx, *seq, [&] {
                for (int32_t s = 0; s < solutions; ++s) {
                    const DeviceRows rows = stage_rows(ctx, ms, uint32_t(size_t(s) * interval * nbl),
                                                       uint32_t(size_t(s + 1) * interval * nbl));
                    const uint32_t ap = rotate_directions(ctx, md, ms, rows, interval, D, opts, tile, true);
                    invert_directions(ctx, md, D, opts, nullptr, seq->d_cal.as<double>() + per_sol * size_t(s), ap);
                }
            });
            ctx->last_graph = 1;
        } else {
            seq = ctx->seq_single.get();
            ctx->last_graph = 2;
        }
        LEAP_HIP(hipGraphLaunch(seq->exec, ctx->stream));
        seq->h_cal.resize(per_sol * size_t(solutions));
        LEAP_HIP(hipMemcpyAsync(seq->h_cal.data(), seq->d_cal.ptr, seq->h_cal.size() * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    }
    LEAP_HIP(hipStreamSynchronize(ctx->stream));
    end_timing(ctx);
    for (int32_t s = 0; s < solutions; ++s) {
        leap_solution sol{};
        sol.index = uint32_t(s);
        sol.start_epoch = ms.time[size_t(s) * interval * nbl];  // CpuLeapCalibrator.cc:107-109
        sol.end_epoch = ms.time[(size_t(s + 1) * interval - 1) * nbl];
        sol.n_directions = D;
        sol.directions = dirs;
        sol.n_antennas = uint32_t(N);
        sol.calibrations = seq ? seq->h_cal.data() + per_sol * size_t(s) : nullptr;
        if (cb && cb(user, &sol) != 0) throw Error(LEAP_ERR_CALLBACK, "output callback requested a stop");
    }
}

// The Calibrate loop (CpuLeapCalibrator.cc:61-145) over prepared metadata.
void calibrate_loop(leap_ctx* ctx, const leap_metadata& md, const leap_ms_view& ms, const double* dirs, uint32_t D,
                    int32_t s_start, int32_t s_interval, int32_t s_end, const leap_options& opts_in,
                    leap_solution_cb cb, void* user) {
    validate_view(ms);
    check_metadata_matches(md, ms);
    require(dirs != nullptr || D == 0, LEAP_ERR_INVALID_ARGUMENT, "null directions");
    const int32_t T = int32_t(ms.num_rows / ms.num_baselines);  // MeasurementSet.cc:115-118
    int32_t range[3];
    evaluate_slice(s_start, s_interval, s_end, T, range);      // CpuLeapCalibrator.cc:90
    const int32_t solutions = (range[2] - range[0]) / range[1];  // Range::GetSize
    const uint32_t interval = uint32_t(range[1]);
    const uint32_t nbl = ms.num_baselines;
    const int32_t N = md.A.cols;
    const leap_options opts = routed(opts_in, ms, interval, D);
    ctx->last_channel_sum = opts.channel_sum;
    ctx->last_direct_cube = opts.direct_cube;
    if (opts.graph) {
        calibrate_loop_graph(ctx, md, ms, dirs, D, solutions, interval, opts, cb, user);
        return;
    }
    ctx->last_graph = 0;
    begin_timing(ctx);
    if (D > 0) upload_directions(ctx, md, dirs, D);
    ctx->h_cal.reserve(std::max<size_t>(1, size_t(N) * D) * sizeof(double));
    double* h_cal = ctx->h_cal.as<double>();
    for (int32_t s = 0; s < solutions; ++s) {  // :104
        // :113-119 — rows [s * interval * nbl, +interval * nbl); the slice start is not added.
        const uint32_t row_begin = uint32_t(size_t(s) * interval * nbl);
        const uint32_t row_end = uint32_t(size_t(s + 1) * interval * nbl);
        leap_solution sol{};
        sol.index = uint32_t(s);
        sol.start_epoch = ms.time[size_t(s) * interval * nbl];               // :107-109
        sol.end_epoch = ms.time[(size_t(s + 1) * interval - 1) * nbl];
        sol.n_directions = D;
        sol.directions = dirs;
        sol.n_antennas = uint32_t(N);
        sol.calibrations = h_cal;
        if (D > 0) {
            const DeviceRows rows = stage_rows(ctx, ms, row_begin, row_end, !packed_route(opts, 0));
            const uint32_t ap = rotate_directions(ctx, md, ms, rows, interval, D, opts, 0, true);
            invert_directions(ctx, md, D, opts, nullptr, nullptr, ap);
            LEAP_HIP(hipMemcpyAsync(h_cal, ctx->d_cal.ptr, size_t(N) * D * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
            LEAP_HIP(hipStreamSynchronize(ctx->stream));
        }
        if (cb && cb(user, &sol) != 0) throw Error(LEAP_ERR_CALLBACK, "output callback requested a stop");
    }
    LEAP_HIP(hipStreamSynchronize(ctx->stream));
    end_timing(ctx);
}

// ---- several measurement sets in one persistent launch ----------------------

// Workgroups of the multi-set grid: per CU the occupancy limit (per_cu 0), the
// caller's count, or one work item per workgroup (per_cu < 0: the hardware
// dispatcher balances, largest items first).
uint32_t multi_grid(leap_ctx* ctx, int32_t tile, int32_t per_cu, size_t n_items) {
    if (per_cu < 0) return uint32_t(std::max<size_t>(n_items, 1));
    if (ctx->cu_count == 0)
        LEAP_HIP(hipDeviceGetAttribute(&ctx->cu_count, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const size_t g = size_t(std::max(1, ctx->cu_count)) * uint32_t(per_cu > 0 ? per_cu : multi_blocks_per_cu(tile));
    return uint32_t(std::min(g, std::max<size_t>(n_items, 1)));
}

// Device tables of one multi-set launch inside a byte block: descriptors, work
// items and per-workgroup item offsets, each 256-byte aligned.
struct MultiTables {
    size_t descs = 0, items = 0, offsets = 0;  // byte offsets in the block
    uint32_t blocks = 0;
    size_t n_items = 0;
};

// Appends n bytes at the next 256-byte boundary of the block; returns their offset.
size_t put_bytes(std::vector<uint8_t>& bytes, const void* src, size_t n) {
    const size_t at = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(at + std::max<size_t>(n, 1), 0);
    if (n) std::memcpy(bytes.data() + at, src, n);
    return at;
}

// Plans one launch (host LPT schedule) and appends its tables to bytes.
MultiTables append_tables(leap_ctx* ctx, std::vector<uint8_t>& bytes, const std::vector<MultiSetDesc>& descs,
                          const std::vector<MultiShape>& shapes, int32_t tile, int32_t per_cu) {
    MultiTables t;
    const uint32_t db = uint32_t(tile / 10);
    for (const MultiShape& m : shapes) t.n_items += size_t(m.nbl) * ((m.D + db - 1) / db);
    t.blocks = multi_grid(ctx, tile, per_cu, t.n_items);
    const MultiPlan plan = plan_multi(shapes, tile, t.blocks);
    t.descs = put_bytes(bytes, descs.data(), descs.size() * sizeof(MultiSetDesc));
    t.items = put_bytes(bytes, plan.items.data(), plan.items.size() * sizeof(MultiItem));
    t.offsets = put_bytes(bytes, plan.off.data(), plan.off.size() * sizeof(uint32_t));
    return t;
}

// Sets of a multi-set call that take the NUFFT channel sums and share one
// launch (equal nufft_sets_key): their descriptors and the work items
// (set | direction block << 16, baseline), in decreasing rows per item (the
// dispatcher then starts the longest items first), set and block order within.
struct NufftGroup {
    const NufftPlan* plan = nullptr;
    uint64_t key = 0;
    uint32_t C = 0, P = 0;
    int32_t tile = 0;
    std::vector<NufftSetDesc> sets;
    std::vector<uint32_t> items;
};

// The channel sums a set of a multi-set call takes: 0 (Horner) or the NUFFT
// route (LEAP_CHANNEL_SUM_NUFFT in mixed precision, LEAP_CHANNEL_SUM_NUFFT_W16
// in FP64 or when asked for), on a cube the kernel reads (complex64, or
// complex128 for the width-16 kernel, with 1, 2 or 4 pols), by the single-set
// automatic rule (96..512 channels, 16 or more directions) unless the NUFFT is
// asked for.
int32_t set_nufft_route(const leap_options& opts, const leap_metadata& md, const leap_ms_view& ms, uint32_t D) {
    if (D == 0) return 0;
    int32_t cs = opts.channel_sum;
    if (cs == LEAP_CHANNEL_SUM_AUTO)
        cs = opts.precision == LEAP_PRECISION_FP64 ? LEAP_CHANNEL_SUM_NUFFT_W16 : LEAP_CHANNEL_SUM_NUFFT;
    else if (cs != LEAP_CHANNEL_SUM_NUFFT && cs != LEAP_CHANNEL_SUM_NUFFT_W16)
        return 0;
    const bool cube_ok = nufft_cube_ok(ms, cs == LEAP_CHANNEL_SUM_NUFFT_W16) && md.nufft_plan(cs).N != 0;
    if (opts.channel_sum != LEAP_CHANNEL_SUM_AUTO) {
        require(cube_ok, LEAP_ERR_INVALID_ARGUMENT,
                "the multi-set NUFFT route needs 1.." + std::to_string(kNufftMaxChannels) +
                    " channels and a complex64 cube (complex128 too for the width-16 NUFFT) with 1, 2 or 4 "
                    "pols at an aligned address");
        return cs;
    }
    return cube_ok && md.channels >= kNufftMinChannels && D >= kNufftMinDirections ? cs : 0;
}

// NUFFT tile of a multi-set launch: the caller's when it is a NUFFT tile,
// otherwise 64, 128 or 256 directions per workgroup, the fewest that hold the
// largest set's directions.
int32_t nufft_group_tile(const leap_options& opts, uint32_t max_d) {
    if (opts.direction_block != 0) {
        const int32_t t = opts.direction_block * 10 + (opts.lane_directions ? opts.lane_directions : 1);
        if (t == 641 || t == 1282 || t == 2564) return t;
    }
    return max_d <= 64 ? 641 : max_d <= 128 ? 1282 : 2564;
}

void plan_nufft_group(NufftGroup& g, const leap_options& opts) {
    uint32_t max_d = 0;
    for (const NufftSetDesc& d : g.sets) max_d = std::max(max_d, d.D);
    g.tile = nufft_group_tile(opts, max_d);
    const uint32_t per_block = 64u * uint32_t(g.tile % 10);
    require(g.sets.size() <= 0xffff, LEAP_ERR_RANGE, "too many sets for one NUFFT launch");
    std::vector<uint32_t> order(g.sets.size());
    for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return g.sets[a].T > g.sets[b].T; });
    g.items.clear();
    for (uint32_t s : order) {
        const NufftSetDesc& d = g.sets[s];
        const uint32_t blocks = (d.D + per_block - 1) / per_block;
        require(blocks <= 0xffff, LEAP_ERR_RANGE, "too many direction blocks for one NUFFT launch");
        for (uint32_t b = 0; b < blocks; ++b)
            for (uint32_t bl = 0; bl < d.nbl; ++bl) {
                g.items.push_back(s | (b << 16));
                g.items.push_back(bl);
            }
    }
    require(g.items.size() / 2 < (size_t(1) << 31), LEAP_ERR_RANGE, "too many work items for one launch");
}

// A round of a multi-set call: the Horner launch's tables and each NUFFT
// group's descriptors and items, at byte offsets of one uploaded block.
struct RoundTables {
    MultiTables horner;
    struct Nufft {
        size_t descs = 0, items = 0;
        uint32_t n_items = 0, C = 0, P = 0;
        int32_t tile = 0;
        const NufftPlan* plan = nullptr;
    };
    std::vector<Nufft> nufft;
};

RoundTables append_round(leap_ctx* ctx, std::vector<uint8_t>& bytes, const std::vector<MultiSetDesc>& descs,
                         const std::vector<MultiShape>& shapes, int32_t tile, int32_t per_cu,
                         const std::vector<NufftGroup>& groups) {
    RoundTables t;
    if (!descs.empty()) t.horner = append_tables(ctx, bytes, descs, shapes, tile, per_cu);
    for (const NufftGroup& g : groups) {
        RoundTables::Nufft n;
        n.descs = put_bytes(bytes, g.sets.data(), g.sets.size() * sizeof(NufftSetDesc));
        n.items = put_bytes(bytes, g.items.data(), g.items.size() * sizeof(uint32_t));
        n.n_items = uint32_t(g.items.size() / 2);
        n.C = g.C;
        n.P = g.P;
        n.tile = g.tile;
        n.plan = g.plan;
        t.nufft.push_back(n);
    }
    return t;
}


void launch_tables(leap_ctx* ctx, const void* block, const MultiTables& t, int32_t tile) {
    if (t.n_items == 0) return;
    const char* b = static_cast<const char*>(block);
    launch_rotate_multi(ctx->stream, tile, reinterpret_cast<const MultiSetDesc*>(b + t.descs),
                        reinterpret_cast<const uint2*>(b + t.items), reinterpret_cast<const uint32_t*>(b + t.offsets),
                        t.blocks);
}

void launch_round(leap_ctx* ctx, const void* block, const RoundTables& t, int32_t tile) {
    const char* b = static_cast<const char*>(block);
    launch_tables(ctx, block, t.horner, tile);
    for (const RoundTables::Nufft& n : t.nufft)
        launch_rotate_nufft_sets(ctx->stream, n.tile, *n.plan, n.C, n.P,
                                 reinterpret_cast<const NufftSetDesc*>(b + n.descs),
                                 reinterpret_cast<const uint32_t*>(b + n.items), n.n_items);
}

// Plans, uploads and launches one multi-set rotate/average.  The host copy of
// the tables stays in the context until the next launch.
void run_multi(leap_ctx* ctx, const std::vector<MultiSetDesc>& descs, const std::vector<MultiShape>& shapes,
               int32_t tile, int32_t per_cu, hipEvent_t after_upload = nullptr, bool timed = false) {
    LEAP_HIP(hipStreamSynchronize(ctx->stream));  // the previous upload has left h_tables
    ctx->h_tables.clear();
    const MultiTables t = append_tables(ctx, ctx->h_tables, descs, shapes, tile, per_cu);
    ctx->d_tables.reserve(ctx->h_tables.size());
    LEAP_HIP(hipMemcpyAsync(ctx->d_tables.ptr, ctx->h_tables.data(), ctx->h_tables.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    if (after_upload) LEAP_HIP(hipEventRecord(after_upload, ctx->stream));
    if (t.n_items == 0) return;
    if (timed) record_rotate_begin(ctx);
    launch_tables(ctx, ctx->d_tables.ptr, t, tile);
    if (timed) record_rotate_end(ctx);
}

// Plans, uploads and launches one round of a multi-set call (the Horner
// tables and the NUFFT groups in one block), timed as one rotate.
void run_round(leap_ctx* ctx, const std::vector<MultiSetDesc>& descs, const std::vector<MultiShape>& shapes,
               int32_t tile, int32_t per_cu, const std::vector<NufftGroup>& groups) {
    LEAP_HIP(hipStreamSynchronize(ctx->stream));  // the previous upload has left h_tables
    ctx->h_tables.clear();
    const RoundTables t = append_round(ctx, ctx->h_tables, descs, shapes, tile, per_cu, groups);
    ctx->d_tables.reserve(ctx->h_tables.size());
    LEAP_HIP(hipMemcpyAsync(ctx->d_tables.ptr, ctx->h_tables.data(), ctx->h_tables.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    record_rotate_begin(ctx);
    launch_round(ctx, ctx->d_tables.ptr, t, tile);
    record_rotate_end(ctx);
}

// Tile and workgroups per CU of the multi-set kernel: the caller's, or the
// fastest candidates on the launch itself, timed once per combination of shapes.
std::pair<int32_t, int32_t> multi_tile(leap_ctx* ctx, const std::vector<MultiSetDesc>& descs,
                                       const std::vector<MultiShape>& shapes, const leap_options& opts) {
    int32_t fixed_tile = 0;
    if (opts.direction_block != 0) {
        fixed_tile = opts.direction_block * 10 + (opts.lane_directions ? opts.lane_directions : 1);
        require(multi_tile_supported(fixed_tile), LEAP_ERR_INVALID_ARGUMENT,
                "direction block " + std::to_string(opts.direction_block) + " x " +
                    std::to_string(opts.lane_directions ? opts.lane_directions : 1) +
                    " directions per lane is not a multi-set tile");
        if (opts.workgroups_per_cu != 0) return {fixed_tile, opts.workgroups_per_cu};
    }
    std::vector<uint32_t> key{uint32_t(fixed_tile), uint32_t(opts.workgroups_per_cu)};
    uint32_t max_d = 0;
    for (size_t i = 0; i < shapes.size(); ++i) {
        key.insert(key.end(), {shapes[i].nbl, shapes[i].T, shapes[i].C, shapes[i].D, uint32_t(descs[i].c64)});
        max_d = std::max(max_d, shapes[i].D);
    }
    auto it = ctx->tuned_multi.find(key);
    if (it != ctx->tuned_multi.end()) return it->second;
    std::vector<int32_t> tiles;
    if (fixed_tile) {
        tiles.push_back(fixed_tile);
    } else {
        for (int32_t tile : {321, 641, 1281, 1282, 2562}) {
            const uint32_t db = uint32_t(tile / 10);
            if (!(db > 32 && db >= 2 * max_d)) tiles.push_back(tile);  // else mostly idle lanes
        }
    }
    hipEvent_t a, b;
    LEAP_HIP(hipEventCreate(&a));
    LEAP_HIP(hipEventCreate(&b));
    auto time_of = [&](int32_t tile, int32_t per_cu) {
        float best_t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {  // the first repetition warms the code object
            float t = 0.f;
            run_multi(ctx, descs, shapes, tile, per_cu, a);
            LEAP_HIP(hipEventRecord(b, ctx->stream));
            LEAP_HIP(hipEventSynchronize(b));
            LEAP_HIP(hipEventElapsedTime(&t, a, b));
            if (rep == 1 || (rep > 1 && t < best_t)) best_t = t;
        }
        return best_t;
    };
    // Tile at the occupancy limit, then the grid depth for that tile.
    std::pair<int32_t, int32_t> best{tiles.front(), opts.workgroups_per_cu};
    float best_ms = 0.f;
    for (int32_t tile : tiles) {
        const float t = time_of(tile, opts.workgroups_per_cu);
        if (best_ms == 0.f || t < best_ms) {
            best_ms = t;
            best = {tile, opts.workgroups_per_cu};
        }
    }
    if (opts.workgroups_per_cu == 0) {
        const int32_t occ = multi_blocks_per_cu(best.first);
        best.second = occ;
        for (int32_t k : {2, 4, 8, -1}) {
            const int32_t per_cu = k > 0 ? occ * k : -1;
            const float t = time_of(best.first, per_cu);
            if (t < best_ms) {
                best_ms = t;
                best.second = per_cu;
            }
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    ctx->tuned_multi.emplace(key, best);
    return best;
}

struct BatchState {
    const leap_metadata* md;
    const leap_ms_view* ms;
    const double* dirs;
    uint32_t D;
    int32_t solutions;
    uint32_t interval;
    std::vector<double> h_cal;
    int32_t nufft = 0;  // NUFFT route of the channel sums, 0 for Horner (avgData holds pol 0 alone)
};

void calibrate_batch(leap_ctx* ctx, const leap_batch_set* sets, uint32_t n_sets, const leap_options& opts,
                     leap_batch_cb cb, void* user) {
    require(sets != nullptr || n_sets == 0, LEAP_ERR_INVALID_ARGUMENT, "null set list");
    require(opts.strategy == LEAP_STRATEGY_BATCHED, LEAP_ERR_INVALID_ARGUMENT,
            "several measurement sets in one launch need the batched strategy");
    require(!opts.all_pols && !opts.fp32_blocks, LEAP_ERR_NOT_IMPLEMENTED,
            "the multi-set launch accumulates polarisation 0 with the FP64 Horner or NUFFT channel sums only");
    require(opts.channel_sum != LEAP_CHANNEL_SUM_NUFFT_FP32, LEAP_ERR_NOT_IMPLEMENTED,
            "the multi-set launch takes the NUFFT with its FP64 transform (channel_sum 2 or 4)");
    std::vector<BatchState> st(n_sets);
    int32_t rounds = 0;
    for (uint32_t i = 0; i < n_sets; ++i) {
        const leap_batch_set& in = sets[i];
        require(in.metadata != nullptr && in.ms != nullptr, LEAP_ERR_INVALID_ARGUMENT,
                "set " + std::to_string(i) + " has no metadata or measurement set");
        require(in.metadata->device == ctx->device, LEAP_ERR_INVALID_ARGUMENT,
                "metadata belongs to another device");
        validate_view(*in.ms);
        check_metadata_matches(*in.metadata, *in.ms);
        require(in.directions != nullptr || in.n_directions == 0, LEAP_ERR_INVALID_ARGUMENT, "null directions");
        const int32_t T = int32_t(in.ms->num_rows / in.ms->num_baselines);
        int32_t range[3];
        evaluate_slice(in.slice_start, in.slice_interval, in.slice_end, T, range);
        st[i] = BatchState{in.metadata, in.ms, in.directions, in.n_directions,
                           (range[2] - range[0]) / range[1], uint32_t(range[1]), {},
                           set_nufft_route(opts, *in.metadata, *in.ms, in.n_directions)};
        rounds = std::max(rounds, st[i].solutions);
    }
    while (ctx->batch_sets.size() < n_sets) ctx->batch_sets.push_back(std::make_unique<SetBuffers>());
    ctx->last_channel_sum = LEAP_CHANNEL_SUM_HORNER;
    for (const BatchState& b : st)
        if (b.nufft) ctx->last_channel_sum = b.nufft;  // one NUFFT route per call (set by the precision)
    ctx->last_direct_cube = 1;
    begin_timing(ctx);
    for (uint32_t i = 0; i < n_sets; ++i) {
        BatchState& b = st[i];
        SetBuffers& buf = *ctx->batch_sets[i];
        const uint32_t N = uint32_t(b.md->A.cols);
        b.h_cal.assign(size_t(N) * b.D, 0.0);
        if (b.D == 0) continue;
        buf.d_dirs.reserve(size_t(b.D) * 2 * sizeof(double));
        buf.d_dd.reserve(size_t(b.D) * 9 * sizeof(double));
        buf.d_avg.reserve(size_t(b.D) * b.md->nbl * b.md->pols * 2 * sizeof(double));
        buf.d_cal.reserve(size_t(N) * b.D * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(buf.d_dirs.ptr, b.dirs, size_t(b.D) * 2 * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
        launch_dd(ctx->stream, b.md->phase_centre_ra, b.md->phase_centre_dec, buf.d_dirs.as<double>(), b.D,
                  buf.d_dd.as<double>());
    }
    // Descriptors of round r: solution interval r of every set that has one,
    // Horner sets into descs / shapes, NUFFT sets into groups of one launch.
    auto round_descs = [&](int32_t r, std::vector<MultiSetDesc>& descs, std::vector<MultiShape>& shapes,
                           std::vector<NufftGroup>& groups, std::vector<uint32_t>& active) {
        for (uint32_t i = 0; i < n_sets; ++i) {
            BatchState& b = st[i];
            if (r >= b.solutions || b.D == 0) continue;
            SetBuffers& buf = *ctx->batch_sets[i];
            const uint32_t nbl = b.md->nbl;
            // CpuLeapCalibrator.cc:113-119: rows [r * interval * nbl, +interval * nbl).
            const DeviceRows rows = stage_rows_into(ctx, *b.ms, uint32_t(size_t(r) * b.interval * nbl),
                                                    uint32_t(size_t(r + 1) * b.interval * nbl), buf.d_vis, buf.d_uvw);
            active.push_back(i);
            if (b.nufft) {
                const NufftPlan& plan = b.md->nufft_plan(b.nufft);
                const uint32_t C = b.md->channels;
                // row source: the cube's pols, plus 16 for complex128
                const uint32_t src = b.md->pols + (b.ms->vis_dtype == LEAP_VIS_C128 ? 16u : 0u);
                const uint64_t key = nufft_sets_key(plan, C, src);
                auto g = std::find_if(groups.begin(), groups.end(), [&](const NufftGroup& x) { return x.key == key; });
                if (g == groups.end()) {
                    groups.emplace_back();
                    g = groups.end() - 1;
                    g->plan = &plan;
                    g->key = key;
                    g->C = C;
                    g->P = src;
                }
                const double* tables = b.md->nufft_tables(b.nufft);
                const double df = b.md->freq_inc_hz;
                g->sets.push_back(NufftSetDesc{rows.vis, rows.uvw, buf.d_dd.as<double>(), buf.d_avg.as<double>(),
                                               tables + kNufftCoefDoubles, tables,
                                               reinterpret_cast<const float*>(tables + nufft_coef32_offset(C)),
                                               (b.md->freq_start_hz + double(plan.h) * df) / kSpeedOfLight,
                                               df / kSpeedOfLight, int32_t(plan.h), nbl, b.interval, C, b.D});
                continue;
            }
            descs.push_back(MultiSetDesc{rows.vis, rows.uvw, buf.d_dd.as<double>(), buf.d_avg.as<double>(), nbl,
                                         b.interval, b.md->channels, b.md->pols, b.D,
                                         b.ms->vis_dtype == LEAP_VIS_C64 ? 1 : 0,
                                         b.md->freq_start_hz / kSpeedOfLight, b.md->freq_inc_hz / kSpeedOfLight});
            shapes.push_back(MultiShape{nbl, b.interval, b.md->channels, b.D});
        }
        for (NufftGroup& g : groups) plan_nufft_group(g, opts);
    };
    auto avg_pols = [&](uint32_t i) { return st[i].nufft ? 1u : st[i].md->pols; };
    auto deliver = [&](int32_t r, auto&& cal_of) {
        for (uint32_t i = 0; i < n_sets; ++i) {
            const BatchState& b = st[i];
            if (r >= b.solutions) continue;
            const uint32_t nbl = b.md->nbl;
            leap_solution sol{};
            sol.index = uint32_t(r);
            sol.start_epoch = b.ms->time[size_t(r) * b.interval * nbl];  // CpuLeapCalibrator.cc:107-109
            sol.end_epoch = b.ms->time[(size_t(r + 1) * b.interval - 1) * nbl];
            sol.n_directions = b.D;
            sol.directions = b.dirs;
            sol.n_antennas = uint32_t(b.md->A.cols);
            sol.calibrations = cal_of(i);
            if (cb && cb(user, i, &sol) != 0) throw Error(LEAP_ERR_CALLBACK, "output callback requested a stop");
        }
    };
    std::pair<int32_t, int32_t> launch{0, 0};  // tile, workgroups per CU
    if (opts.graph) {
        // Every round's multi-set launch and inversions captured as one graph.
        const size_t n_rounds = static_cast<size_t>(rounds);
        std::vector<std::vector<MultiSetDesc>> rdescs(n_rounds);
        std::vector<std::vector<MultiShape>> rshapes(n_rounds);
        std::vector<std::vector<NufftGroup>> rgroups(n_rounds);
        std::vector<std::vector<uint32_t>> ractive(n_rounds);
        std::vector<size_t> cal_at(n_sets, 0);
        size_t cal_total = 0;
        for (uint32_t i = 0; i < n_sets; ++i) {
            const BatchState& b = st[i];
            cal_at[i] = cal_total;
            if (b.D == 0 || b.solutions == 0) continue;
            require_device_resident(*b.ms);
            leap_options ro = opts;
            ro.direct_cube = 1;  // the multi-set kernel reads the cubes directly
            reserve_solution(ctx, *b.md, b.D, b.interval, ro);
            cal_total += size_t(b.md->A.cols) * b.D * size_t(b.solutions);
        }
        for (int32_t r = 0; r < rounds; ++r) round_descs(r, rdescs[r], rshapes[r], rgroups[r], ractive[r]);
        for (int32_t r = 0; r < rounds && launch.first == 0; ++r)
            if (!rdescs[r].empty()) launch = multi_tile(ctx, rdescs[r], rshapes[r], opts);
        ctx->last_direction_block = launch.first;
        ctx->last_workgroups_per_cu = launch.second;
        std::vector<uint64_t> key{uint64_t(uint32_t(launch.first)), uint64_t(uint32_t(launch.second)), n_sets};
        key_common(key, ctx, opts);
        for (uint32_t i = 0; i < n_sets; ++i) {
            const BatchState& b = st[i];
            key_set(key, *b.md, *b.ms, b.D, b.solutions, b.interval);
            key.insert(key.end(), {key_ptr(ctx->batch_sets[i]->d_dd.ptr), key_ptr(ctx->batch_sets[i]->d_avg.ptr)});
        }
        CapturedSequence* seq = ctx->seq_batch.get();
        if (cal_total > 0 && (!seq || seq->key != key)) {
            ctx->seq_batch = std::make_unique<CapturedSequence>();
            seq = ctx->seq_batch.get();
            seq->key = key;
            std::vector<uint8_t> bytes;
            std::vector<RoundTables> tables(n_rounds);
            for (int32_t r = 0; r < rounds; ++r)
                tables[r] = append_round(ctx, bytes, rdescs[r], rshapes[r], launch.first, launch.second, rgroups[r]);
            seq->d_tables.reserve(bytes.size());
            LEAP_HIP(hipMemcpyAsync(seq->d_tables.ptr, bytes.data(), bytes.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
            LEAP_HIP(hipStreamSynchronize(ctx->stream));
            seq->d_cal.reserve(cal_total * sizeof(double));
            capture(ctx, *seq, [&] {
                for (int32_t r = 0; r < rounds; ++r) {
                    launch_round(ctx, seq->d_tables.ptr, tables[r], launch.first);
                    for (uint32_t i : ractive[r]) {
                        const BatchState& b = st[i];
                        const size_t per_sol = size_t(b.md->A.cols) * b.D;
                        invert_directions(ctx, *b.md, b.D, opts, ctx->batch_sets[i]->d_avg.as<double>(),
                                          seq->d_cal.as<double>() + cal_at[i] + per_sol * size_t(r), avg_pols(i));
                    }
                }
            });
            ctx->last_graph = 1;
        } else {
            ctx->last_graph = cal_total > 0 ? 2 : 0;
        }
        if (cal_total > 0) {
            LEAP_HIP(hipGraphLaunch(seq->exec, ctx->stream));
            seq->h_cal.resize(cal_total);
            LEAP_HIP(hipMemcpyAsync(seq->h_cal.data(), seq->d_cal.ptr, cal_total * sizeof(double),
                                    hipMemcpyDeviceToHost, ctx->stream));
        }
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
        end_timing(ctx);
        for (int32_t r = 0; r < rounds; ++r)
            deliver(r, [&](uint32_t i) -> const double* {
                const BatchState& b = st[i];
                if (b.D == 0) return b.h_cal.data();
                return seq->h_cal.data() + cal_at[i] + size_t(b.md->A.cols) * b.D * size_t(r);
            });
        return;
    }
    ctx->last_graph = 0;
    for (int32_t r = 0; r < rounds; ++r) {
        std::vector<MultiSetDesc> descs;
        std::vector<MultiShape> shapes;
        std::vector<NufftGroup> groups;
        std::vector<uint32_t> active;
        round_descs(r, descs, shapes, groups, active);
        if (!active.empty()) {
            if (launch.first == 0 && !descs.empty()) launch = multi_tile(ctx, descs, shapes, opts);
            ctx->last_direction_block = launch.first;
            ctx->last_workgroups_per_cu = launch.second;
            run_round(ctx, descs, shapes, launch.first, launch.second, groups);
            for (uint32_t i : active) {
                BatchState& b = st[i];
                SetBuffers& buf = *ctx->batch_sets[i];
                invert_directions(ctx, *b.md, b.D, opts, buf.d_avg.as<double>(), buf.d_cal.as<double>(), avg_pols(i));
                LEAP_HIP(hipMemcpyAsync(b.h_cal.data(), buf.d_cal.ptr, b.h_cal.size() * sizeof(double),
                                        hipMemcpyDeviceToHost, ctx->stream));
            }
            LEAP_HIP(hipStreamSynchronize(ctx->stream));
        }
        deliver(r, [&](uint32_t i) -> const double* { return st[i].h_cal.data(); });
    }
    LEAP_HIP(hipStreamSynchronize(ctx->stream));
    end_timing(ctx);
}

}  // namespace

extern "C" {

const char* leap_hip_version(void) { return "leap-hip 0.1.0 (gfx950)"; }

const char* leap_hip_last_error(const leap_ctx* ctx) {
    if (ctx) return ctx->error.c_str();
    return g_error.c_str();
}

int leap_hip_create(leap_ctx** out, int32_t device) {
    return guarded(nullptr, [&] {
        require(out != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null output pointer");
        int count = 0;
        LEAP_HIP(hipGetDeviceCount(&count));
        require(device >= 0 && device < count, LEAP_ERR_INVALID_ARGUMENT,
                "device " + std::to_string(device) + " not available (" + std::to_string(count) + " devices)");
        auto ctx = std::make_unique<leap_ctx>();
        ctx->device = device;
        DeviceScope scope(device);
        noalloc_init(device);
        LEAP_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        LEAP_HIP(hipStreamCreateWithFlags(&ctx->pack_stream, hipStreamNonBlocking));
        LEAP_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_copy_fork, hipEventDisableTiming));
        LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_copy_join, hipEventDisableTiming));
        LEAP_HIP(hipEventCreate(&ctx->ev_start));
        LEAP_HIP(hipEventCreate(&ctx->ev_stop));
        LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        for (auto& e : ctx->ev_grad) LEAP_HIP(hipEventCreate(&e));
        for (int b = 0; b < 2; ++b) {
            LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_packed[b], hipEventDisableTiming));
            LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_consumed[b], hipEventDisableTiming));
            LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_h2d[b], hipEventDisableTiming));
            LEAP_HIP(hipEventCreateWithFlags(&ctx->ev_widened[b], hipEventDisableTiming));
        }
        *out = ctx.release();
    });
}

void leap_hip_destroy(leap_ctx* ctx) {
    if (!ctx) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& e : ctx->rotate_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (ctx->pack_stream) (void)hipStreamSynchronize(ctx->pack_stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    for (int b = 0; b < 2; ++b) {
        if (ctx->h_pinned[b]) (void)hipHostFree(ctx->h_pinned[b]);
        if (ctx->ev_h2d[b]) (void)hipEventDestroy(ctx->ev_h2d[b]);
        if (ctx->ev_widened[b]) (void)hipEventDestroy(ctx->ev_widened[b]);
    }
    if (ctx->ev_copy_fork) (void)hipEventDestroy(ctx->ev_copy_fork);
    if (ctx->ev_copy_join) (void)hipEventDestroy(ctx->ev_copy_join);
    for (auto& e : ctx->pack_events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (auto& e : ctx->ev_grad)
        if (e) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) {
        if (ctx->ev_packed[b]) (void)hipEventDestroy(ctx->ev_packed[b]);
        if (ctx->ev_consumed[b]) (void)hipEventDestroy(ctx->ev_consumed[b]);
    }
    hipStream_t s = ctx->stream;
    hipStream_t ps = ctx->pack_stream;
    hipStream_t cs = ctx->copy_stream;
    delete ctx;  // frees the workspace buffers on this device
    if (s) (void)hipStreamDestroy(s);
    if (ps) (void)hipStreamDestroy(ps);
    if (cs) (void)hipStreamDestroy(cs);
    if (prev >= 0) (void)hipSetDevice(prev);
}

int leap_hip_metadata_create(leap_ctx* ctx, const leap_ms_view* ms, int32_t ref_ant, double thr, int32_t use_cache,
                             leap_metadata** out) {
    return guarded(ctx, [&] {
        require(ctx && ms && out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        *out = create_metadata(ctx, *ms, ref_ant, thr, use_cache);
    });
}

int leap_hip_set_log_callback(leap_ctx* ctx, leap_log_cb cb, void* user) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        ctx->log = cb;
        ctx->log_user = user;
    });
}

int leap_hip_metadata_diagnostics(const leap_metadata* md, int32_t* ad_degenerate, int32_t* ad1_degenerate) {
    return guarded(nullptr, [&] {
        require(md != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null metadata");
        if (ad_degenerate) *ad_degenerate = md->ad_degenerate ? 1 : 0;
        if (ad1_degenerate) *ad1_degenerate = md->ad1_degenerate ? 1 : 0;
    });
}

void leap_hip_metadata_destroy(leap_metadata* md) {
    if (!md) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(md->device);
    delete md;
    if (prev >= 0) (void)hipSetDevice(prev);
}

void leap_hip_metadata_dims(const leap_metadata* md, int32_t* dims) {
    dims[0] = md->A.rows;
    dims[1] = md->A.cols;
    dims[2] = md->A1.rows;
    dims[3] = int32_t(md->nbl);
    dims[4] = int32_t(md->reference_antenna);
}

int leap_hip_metadata_copy(const leap_metadata* md, double* A, int32_t* I, double* Ad, double* A1, int32_t* I1,
                           double* Ad1) {
    return guarded(nullptr, [&] {
        require(md != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null metadata");
        DeviceScope scope(md->device);
        if (A) {
            const auto a = md->A.dense();
            std::copy(a.begin(), a.end(), A);
        }
        if (I) std::copy(md->A.I.begin(), md->A.I.end(), I);
        if (A1) {
            const auto a = md->A1.dense();
            std::copy(a.begin(), a.end(), A1);
        }
        if (I1) std::copy(md->A1.I.begin(), md->A1.I.end(), I1);
        // AdT [K][N] row-major is Ad (N x K) column-major.
        if (Ad)
            LEAP_HIP(hipMemcpy(Ad, md->d_AdT.ptr, size_t(md->A.rows) * md->A.cols * sizeof(double),
                               hipMemcpyDeviceToHost));
        if (Ad1)
            LEAP_HIP(hipMemcpy(Ad1, md->d_Ad1T.ptr, size_t(md->A1.rows) * md->A1.cols * sizeof(double),
                               hipMemcpyDeviceToHost));
    });
}

int leap_hip_rotate_visibilities(leap_ctx* ctx, const leap_metadata* md, const leap_ms_view* ms, uint32_t row_start,
                                 uint32_t n_rows, const double* dirs, uint32_t D, const leap_options* opts,
                                 double* avg_out) {
    return guarded(ctx, [&] {
        require(ctx && md && ms && avg_out && (dirs || D == 0), LEAP_ERR_INVALID_ARGUMENT, "null argument");
        validate_view(*ms);
        check_metadata_matches(*md, *ms);
        require(n_rows % ms->num_baselines == 0, LEAP_ERR_RANGE,
                "integration rows must be a whole number of timesteps");
        const leap_options o = routed(resolve(opts), *ms, n_rows / ms->num_baselines, D);
        ctx->last_channel_sum = o.channel_sum;
        ctx->last_direct_cube = o.direct_cube;
        DeviceScope scope(ctx->device);
        begin_timing(ctx);
        const DeviceRows rows = stage_rows(ctx, *ms, row_start, row_start + n_rows, !packed_route(o, 0));
        upload_directions(ctx, *md, dirs, D);
        rotate_directions(ctx, *md, *ms, rows, n_rows / ms->num_baselines, D, o);
        const size_t n = size_t(md->nbl) * md->pols * 2 * D;
        if (n) LEAP_HIP(hipMemcpyAsync(avg_out, ctx->d_avg.ptr, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
        end_timing(ctx);
    });
}

int leap_hip_apply_inversion(leap_ctx* ctx, const leap_metadata* md, const double* avg, uint32_t D,
                             const leap_options* opts, double* cal_out) {
    return guarded(ctx, [&] {
        require(ctx && md && avg && cal_out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const leap_options o = resolve(opts);
        DeviceScope scope(ctx->device);
        const size_t n = size_t(md->nbl) * md->pols * 2 * D;
        ctx->d_avg.reserve(n * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(ctx->d_avg.ptr, avg, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        invert_directions(ctx, *md, D, o);
        LEAP_HIP(hipMemcpyAsync(cal_out, ctx->d_cal.ptr, size_t(md->A.cols) * D * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int leap_hip_dd_matrix(leap_ctx* ctx, double ra0, double dec0, const double* dirs, uint32_t D, double* dd_out) {
    return guarded(ctx, [&] {
        require(ctx && dirs && dd_out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        ctx->d_dirs.reserve(size_t(D) * 2 * sizeof(double));
        ctx->d_dd.reserve(size_t(D) * 9 * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(ctx->d_dirs.ptr, dirs, size_t(D) * 2 * sizeof(double), hipMemcpyHostToDevice,
                                ctx->stream));
        launch_dd(ctx->stream, ra0, dec0, ctx->d_dirs.as<double>(), D, ctx->d_dd.as<double>());
        LEAP_HIP(hipMemcpyAsync(dd_out, ctx->d_dd.ptr, size_t(D) * 9 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int leap_hip_calibrate_md(leap_ctx* ctx, const leap_metadata* md, const leap_ms_view* ms, const double* dirs,
                          uint32_t D, int32_t s_start, int32_t s_interval, int32_t s_end, const leap_options* opts,
                          leap_solution_cb cb, void* user) {
    return guarded(ctx, [&] {
        require(ctx && md && ms, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const leap_options o = resolve(opts);
        DeviceScope scope(ctx->device);
        calibrate_loop(ctx, *md, *ms, dirs, D, s_start, s_interval, s_end, o, cb, user);
    });
}

int leap_hip_calibrate(leap_ctx* ctx, const leap_ms_view* ms, const double* dirs, uint32_t D, int32_t s_start,
                       int32_t s_interval, int32_t s_end, double thr, int32_t ref_ant, int32_t use_cache,
                       const leap_options* opts, leap_solution_cb cb, void* user) {
    return guarded(ctx, [&] {
        require(ctx && ms, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const leap_options o = resolve(opts);
        DeviceScope scope(ctx->device);
        validate_view(*ms);
        // Slice errors surface before the metadata work, as Evaluate precedes MetaData (CpuLeapCalibrator.cc:90-99).
        int32_t range[3];
        evaluate_slice(s_start, s_interval, s_end, int32_t(ms->num_rows / ms->num_baselines), range);
        std::unique_ptr<leap_metadata> md(create_metadata(ctx, *ms, ref_ant, thr, use_cache));
        calibrate_loop(ctx, *md, *ms, dirs, D, s_start, s_interval, s_end, o, cb, user);
    });
}

int leap_hip_calibrate_batch(leap_ctx* ctx, const leap_batch_set* sets, uint32_t n_sets, const leap_options* opts,
                             leap_batch_cb cb, void* user) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        DeviceScope scope(ctx->device);
        calibrate_batch(ctx, sets, n_sets, resolve(opts), cb, user);
    });
}

int leap_hip_synth_fill_vis_device(leap_ctx* ctx, const leap_synth_ms* ms, int32_t vis_dtype, void* out_device,
                                   double* uvw_device_out) {
    return guarded(ctx, [&] {
        require(ctx && ms && out_device, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        DeviceBuffer d_uvw_tmp, d_blphase;
        double* d_uvw = uvw_device_out;
        const size_t uvw_bytes = ms->uvw.size() * sizeof(double);
        if (!d_uvw) {
            d_uvw_tmp.reserve(uvw_bytes);
            d_uvw = d_uvw_tmp.as<double>();
        }
        LEAP_HIP(hipMemcpyAsync(d_uvw, ms->uvw.data(), uvw_bytes, hipMemcpyHostToDevice, ctx->stream));
        d_blphase.reserve(ms->bl_phase.size() * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(d_blphase.ptr, ms->bl_phase.data(), ms->bl_phase.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream));
        leap_synth_ms dev_view;  // parameters with the device phase table
        dev_view.cfg = ms->cfg;
        dev_view.n_rows = ms->n_rows;
        dev_view.params = ms->params;
        dev_view.params.bl_phase = d_blphase.as<double>();
        launch_synth_vis(ctx->stream, &dev_view, vis_dtype, out_device, d_uvw);
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int leap_hip_malloc(leap_ctx* ctx, uint64_t bytes, void** out) {
    return guarded(ctx, [&] {
        require(ctx && out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        LEAP_HIP(hipMalloc(out, bytes));
    });
}

int leap_hip_free(leap_ctx* ctx, void* ptr) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        DeviceScope scope(ctx->device);
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
        if (ptr) LEAP_HIP(hipFree(ptr));
    });
}

int leap_hip_memcpy_d2h(leap_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&] {
        require(ctx && dst && src, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        LEAP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int leap_hip_memcpy_h2d(leap_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&] {
        require(ctx && dst && src, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        DeviceScope scope(ctx->device);
        LEAP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        LEAP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int leap_hip_last_kernel_times(const leap_ctx* ctx, double* rotate_ms_avg, int32_t* rotate_launches,
                               double* total_ms) {
    if (!ctx) return LEAP_ERR_INVALID_ARGUMENT;
    if (rotate_ms_avg) *rotate_ms_avg = ctx->last_rotate_ms_avg;
    if (rotate_launches) *rotate_launches = ctx->last_rotate_launches;
    if (total_ms) *total_ms = ctx->last_total_ms;
    return LEAP_OK;
}

}  // extern "C"

extern "C" {

int leap_phase_matrix(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n, int32_t ref_ant,
                      double* A_out, int32_t* I_out, int32_t* K, int32_t* N) {
    return guarded(nullptr, [&] {
        require(K && N && (n == 0 || (a1 && a2 && fg)), LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const PhaseMatrix pm = phase_matrix(a1, a2, fg, n, ref_ant);
        *K = pm.rows;
        *N = pm.cols;
        if (A_out) {
            const size_t ld = size_t(n) + 1;
            for (int32_t j = 0; j < pm.cols; ++j)
                for (int32_t i = 0; i < pm.rows; ++i) A_out[size_t(j) * ld + i] = 0.0;
            for (int32_t k = 0; k + 1 < pm.rows; ++k) {
                A_out[size_t(pm.ant1[k]) * ld + k] = 1.0;
                A_out[size_t(pm.ant2[k]) * ld + k] = -1.0;
            }
            A_out[size_t(pm.constraint) * ld + (pm.rows - 1)] = 1.0;
        }
        if (I_out) std::copy(pm.I.begin(), pm.I.end(), I_out);
    });
}

int leap_phase_matrix_pinv(const int32_t* a1, const int32_t* a2, const uint8_t* fg, int32_t n, int32_t ref_ant,
                           double* Ad_out) {
    return guarded(nullptr, [&] {
        require(Ad_out && (n == 0 || (a1 && a2 && fg)), LEAP_ERR_INVALID_ARGUMENT, "null argument");
        const PhaseMatrix pm = phase_matrix(a1, a2, fg, n, ref_ant);
        const std::vector<double> Gp = gram_pinv(pm);
        const int32_t K = pm.rows, Nn = pm.cols;
        // Ad = G+ A^T, column k of Ad = G+ (e_a1 - e_a2), last column G+ e_c.
        for (int32_t k = 0; k < K; ++k) {
            double* col = Ad_out + size_t(k) * Nn;
            if (k + 1 < K) {
                const double* g1 = Gp.data() + size_t(pm.ant1[k]) * Nn;
                const double* g2 = Gp.data() + size_t(pm.ant2[k]) * Nn;
                for (int32_t i = 0; i < Nn; ++i) col[i] = g1[i] - g2[i];
            } else {
                const double* gc = Gp.data() + size_t(pm.constraint) * Nn;
                for (int32_t i = 0; i < Nn; ++i) col[i] = gc[i];
            }
        }
    });
}

int leap_multi_schedule(const uint32_t* nbl, const uint32_t* T, const uint32_t* C, const uint32_t* D,
                        uint32_t n_sets, int32_t tile, uint32_t n_blocks, uint32_t* offsets, uint32_t* item_set,
                        uint32_t* item_local, double* makespan_out) {
    return guarded(nullptr, [&] {
        require(n_sets == 0 || (nbl && T && C && D), LEAP_ERR_INVALID_ARGUMENT, "null shape arrays");
        require(offsets != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null offsets");
        std::vector<MultiShape> shapes(n_sets);
        for (uint32_t i = 0; i < n_sets; ++i) shapes[i] = MultiShape{nbl[i], T[i], C[i], D[i]};
        const MultiPlan plan = plan_multi(shapes, tile, n_blocks);
        require(plan.items.empty() || (item_set && item_local), LEAP_ERR_INVALID_ARGUMENT, "null item arrays");
        std::copy(plan.off.begin(), plan.off.end(), offsets);
        for (size_t i = 0; i < plan.items.size(); ++i) {
            item_set[i] = plan.items[i].set;
            item_local[i] = plan.items[i].local;
        }
        if (makespan_out) *makespan_out = plan.makespan;
    });
}

int leap_phasor64(const double* theta, uint32_t n, double* cos_out, double* sin_out) {
    return guarded(nullptr, [&] {
        require(n == 0 || (theta && cos_out && sin_out), LEAP_ERR_INVALID_ARGUMENT, "null argument");
        for (uint32_t i = 0; i < n; ++i) {
            const Phasor64 p = phasor64(theta[i]);
            cos_out[i] = p.c;
            sin_out[i] = p.s;
        }
    });
}

namespace {
// coef: (width / 2) x width doubles, coef[j * width + m] (width 12 for the
// width-10/12 plans, 16 for the width-16 plan).
int nufft_plan_out(uint32_t C, bool fp32, bool w16, uint32_t* N, uint32_t* W, uint32_t* h, double* beta,
                   double* coef, double* scale) {
    return guarded(nullptr, [&] {
        const NufftPlan p = make_nufft_plan(C, fp32, w16);
        if (N) *N = p.N;
        if (W) *W = p.W;
        if (h) *h = p.h;
        if (beta) *beta = p.beta;
        const int width = w16 ? 16 : 12;
        static_assert(kNufftMaxWidth >= 16, "plan coefficients hold the width-16 kernel");
        if (coef)
            for (int j = 0; j < width / 2; ++j)
                for (int m = 0; m < width; ++m) coef[j * width + m] = p.coef[j][m];
        if (scale) std::copy(p.scale.begin(), p.scale.end(), scale);
    });
}
}  // namespace

int leap_nufft_plan(uint32_t C, uint32_t* N, uint32_t* W, uint32_t* h, double* beta, double* coef, double* scale) {
    return nufft_plan_out(C, false, false, N, W, h, beta, coef, scale);
}

int leap_nufft_plan32(uint32_t C, uint32_t* N, uint32_t* W, uint32_t* h, double* beta, double* coef,
                      double* scale) {
    return nufft_plan_out(C, true, false, N, W, h, beta, coef, scale);
}

int leap_nufft_plan_w16(uint32_t C, uint32_t* N, uint32_t* W, uint32_t* h, double* beta, double* coef,
                        double* scale) {
    return nufft_plan_out(C, false, true, N, W, h, beta, coef, scale);
}

double leap_multi_item_cost(uint32_t T, uint32_t C, int32_t tile) {
    return multi_tile_supported(tile) ? multi_item_cost(T, C, tile) : -1.0;
}

}  // extern "C"

extern "C" int leap_hip_last_direction_block(const leap_ctx* ctx) { return ctx ? ctx->last_direction_block : 0; }

extern "C" int leap_hip_last_route(const leap_ctx* ctx, int32_t* channel_sum, int32_t* direct_cube) {
    if (!ctx) return LEAP_ERR_INVALID_ARGUMENT;
    if (channel_sum) *channel_sum = ctx->last_channel_sum;
    if (direct_cube) *direct_cube = ctx->last_direct_cube;
    return LEAP_OK;
}
extern "C" int leap_hip_last_loss_grad_times(const leap_ctx* ctx, double* forward_ms, double* adjoint_ms) {
    if (!ctx) return LEAP_ERR_INVALID_ARGUMENT;
    if (forward_ms) *forward_ms = ctx->last_grad_forward_ms;
    if (adjoint_ms) *adjoint_ms = ctx->last_grad_adjoint_ms;
    return LEAP_OK;
}
extern "C" int leap_hip_last_workgroups_per_cu(const leap_ctx* ctx) { return ctx ? ctx->last_workgroups_per_cu : 0; }
extern "C" int leap_hip_last_graph(const leap_ctx* ctx) { return ctx ? ctx->last_graph : 0; }
extern "C" uint64_t leap_hip_last_staged_bytes(const leap_ctx* ctx) { return ctx ? ctx->last_stage_bytes : 0; }
extern "C" int leap_hip_tuning_save(const leap_ctx* ctx, char* buf, uint64_t cap, uint64_t* needed) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        const std::string js = tuning_to_json(*ctx);
        if (needed) *needed = js.size() + 1;
        if (buf && cap) {
            const size_t n = std::min<size_t>(size_t(cap) - 1, js.size());
            std::memcpy(buf, js.data(), n);
            buf[n] = '\0';
        }
    });
}
extern "C" int leap_hip_tuning_load(leap_ctx* ctx, const char* json) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        tuning_from_json(*ctx, json);
    });
}
extern "C" int leap_hip_last_pack_times(const leap_ctx* ctx, double* pack_ms_avg, int32_t* pack_launches) {
    if (!ctx) return LEAP_ERR_INVALID_ARGUMENT;
    if (pack_ms_avg) *pack_ms_avg = ctx->last_pack_ms_avg;
    if (pack_launches) *pack_launches = ctx->last_pack_launches;
    return LEAP_OK;
}

namespace {

// Device tables of a set's baselines for the residual-loss gradient, rebuilt
// only when the antennas change: each baseline's antennas, and each antenna's
// baselines (+1 where it is antenna2, -1 where antenna1; autocorrelations,
// whose uvw is 0, left out) as offsets, baselines and signs.
void adjoint_antenna_tables(leap_ctx* ctx, const leap_ms_view& ms, int32_t N) {
    const uint32_t nbl = ms.num_baselines;
    uint64_t key = 1469598103934665603ull;  // FNV-1a over the antenna columns
    auto mix = [&](uint64_t v) {
        for (int k = 0; k < 8; ++k) {
            key ^= (v >> (8 * k)) & 0xffu;
            key *= 1099511628211ull;
        }
    };
    mix(nbl);
    mix(uint64_t(uint32_t(N)));
    for (uint32_t b = 0; b < nbl; ++b) mix((uint64_t(uint32_t(ms.antenna1[b])) << 32) | uint32_t(ms.antenna2[b]));
    if (key == ctx->adj_ants_key && ctx->d_adj_ants.ptr && ctx->d_adj_off.ptr) return;
    hipStream_t s = ctx->stream;
    std::vector<int2> ants(nbl);
    std::vector<int32_t> off(size_t(N) + 1, 0), bls, sgn;
    std::vector<std::vector<std::pair<int32_t, int32_t>>> lists(static_cast<size_t>(N));
    for (uint32_t b = 0; b < nbl; ++b) {
        ants[b] = make_int2(ms.antenna1[b], ms.antenna2[b]);
        if (ms.antenna1[b] == ms.antenna2[b]) continue;
        lists[ms.antenna2[b]].emplace_back(int32_t(b), 1);
        lists[ms.antenna1[b]].emplace_back(int32_t(b), -1);
    }
    for (int32_t a = 0; a < N; ++a) {
        off[a + 1] = off[a] + int32_t(lists[a].size());
        for (auto& e : lists[a]) {
            bls.push_back(e.first);
            sgn.push_back(e.second);
        }
    }
    ctx->d_adj_ants.reserve(nbl * sizeof(int2));
    ctx->d_adj_off.reserve(off.size() * sizeof(int32_t));
    ctx->d_adj_bls.reserve(std::max<size_t>(1, bls.size()) * sizeof(int32_t));
    ctx->d_adj_sgn.reserve(std::max<size_t>(1, sgn.size()) * sizeof(int32_t));
    LEAP_HIP(hipMemcpyAsync(ctx->d_adj_ants.ptr, ants.data(), nbl * sizeof(int2), hipMemcpyHostToDevice, s));
    LEAP_HIP(hipMemcpyAsync(ctx->d_adj_off.ptr, off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    if (!bls.empty()) {
        LEAP_HIP(hipMemcpyAsync(ctx->d_adj_bls.ptr, bls.data(), bls.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        LEAP_HIP(hipMemcpyAsync(ctx->d_adj_sgn.ptr, sgn.data(), sgn.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    LEAP_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    ctx->adj_ants_key = key;
}

}  // namespace

// ---------------------------------------------------------------------------
// Calibration-residual loss and its gradient: forward rotate/average and
// inversion, the loss, the inversion's adjoint and the adjoint of
// rotate/average, all on the device; the host sums the per-direction losses.
// ---------------------------------------------------------------------------
extern "C" int leap_hip_set_adjoint_mode(leap_ctx* ctx, int32_t mode) {
    return guarded(ctx, [&] {
        require(ctx != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null context");
        require(mode == LEAP_ADJOINT_AUTO || mode == LEAP_ADJOINT_RECOMPUTE || mode == LEAP_ADJOINT_ROW_SUMS,
                LEAP_ERR_INVALID_ARGUMENT, "adjoint mode must be 0 (automatic), 1 (recompute) or 2 (row sums)");
        ctx->adjoint_mode = mode;
    });
}

extern "C" int leap_hip_residual_loss_grad(leap_ctx* ctx, const leap_metadata* md, const leap_ms_view* ms,
                                           const double* directions, const double* offsets, uint32_t D,
                                           const double* antenna_delta, const double* enu_to_uvw, double* loss_out,
                                           double* grad_offsets, double* grad_antenna) {
    return guarded(ctx, [&] {
        require(ctx && md && ms && directions && enu_to_uvw && loss_out, LEAP_ERR_INVALID_ARGUMENT, "null argument");
        require(D >= 1 && D <= 256, LEAP_ERR_NOT_IMPLEMENTED, "the adjoint handles 1 to 256 directions per call");
        validate_view(*ms);
        check_metadata_matches(*md, *ms);
        DeviceScope scope(ctx->device);
        hipStream_t s = ctx->stream;
        const uint32_t nbl = ms->num_baselines, R = ms->num_rows, T = R / nbl;
        const int32_t N = md->A.cols;
        const bool want_grad = grad_offsets || grad_antenna;
        // With a gradient the forward pass keeps each row's sums for the adjoint
        // (32 bytes per row and direction) unless they outweigh the cube itself
        // or the context asks for the recomputing adjoint.
        const bool row_sums = want_grad && ctx->adjoint_mode != LEAP_ADJOINT_RECOMPUTE &&
                              (ctx->adjoint_mode == LEAP_ADJOINT_ROW_SUMS ||
                               size_t(32) * D <= vis_elem_bytes(*ms) * md->channels);
        if (antenna_delta || grad_antenna) adjoint_antenna_tables(ctx, *ms, N);
        std::vector<double> dirs(size_t(D) * 2);
        for (uint32_t d = 0; d < 2 * D; ++d) dirs[d] = directions[d] + (offsets ? offsets[d] : 0.0);
        upload_directions(ctx, *md, dirs.data(), D);
        const DeviceRows rows = stage_rows(ctx, *ms, 0, R);
        // perturbed uvw
        DeviceBuffer &d_M = ctx->d_adj_M, &d_ants = ctx->d_adj_ants, &d_dp = ctx->d_adj_dp, &d_uvwp = ctx->d_adj_uvw;
        d_M.reserve(size_t(T) * 9 * sizeof(double));
        LEAP_HIP(hipMemcpyAsync(d_M.ptr, enu_to_uvw, size_t(T) * 9 * sizeof(double), hipMemcpyHostToDevice, s));
        const double* uvw = rows.uvw;
        if (antenna_delta) {
            d_dp.reserve(size_t(N) * 3 * sizeof(double));
            d_uvwp.reserve(size_t(R) * 3 * sizeof(double));
            LEAP_HIP(hipMemcpyAsync(d_dp.ptr, antenna_delta, size_t(N) * 3 * sizeof(double), hipMemcpyHostToDevice, s));
            launch_perturb_uvw(s, rows.uvw, d_M.as<double>(), d_ants.as<int2>(), d_dp.as<double>(), nbl, R,
                               d_uvwp.as<double>());
            uvw = d_uvwp.as<double>();
        }
        // forward avgData (pol 0), FP64
        const size_t per_dir = size_t(nbl) * md->pols * 2;
        ctx->d_avg.reserve(per_dir * D * sizeof(double));
        const int32_t tile = D <= 16 ? 162 : D <= 32 ? 321 : D <= 64 ? 641 : D <= 128 ? 1282 : 2562;
        if (row_sums) ctx->d_adj_q.reserve(size_t(R) * D * 4 * sizeof(double));
        LEAP_HIP(hipEventRecord(ctx->ev_grad[0], s));
        if (row_sums)
            launch_rotate_forward_q(s, tile, rows.vis, ms->vis_dtype, uvw, ctx->d_dd.as<double>(), D, T, nbl,
                                    md->channels, md->pols, md->freq_start_hz, md->freq_inc_hz,
                                    ctx->d_avg.as<double>(), ctx->d_adj_q.as<double>());
        else
            launch_rotate_batched(s, LEAP_PRECISION_FP64, tile, false, rows.vis, ms->vis_dtype, uvw,
                                  ctx->d_dd.as<double>(), D, T, nbl, nbl, md->channels, md->pols, md->freq_start_hz,
                                  md->freq_inc_hz, ctx->d_avg.as<double>(), 0);
        LEAP_HIP(hipEventRecord(ctx->ev_grad[1], s));
        ctx->last_grad_adjoint_ms = 0.0;
        // the inversion (CpuLeapCalibrator.cc:159-178) by the calibration's own kernels
        leap_options inv{};
        invert_directions(ctx, *md, D, inv);
        // pol 0 of avgData and its arguments
        DeviceBuffer& d_avg0 = ctx->d_adj_avg0;
        d_avg0.reserve(size_t(D) * nbl * 3 * sizeof(double));
        launch_gather_pol0(s, ctx->d_avg.as<double>(), D * nbl, md->pols, d_avg0.as<double>());
        // loss and the adjoint of the inversion, on the device
        const int32_t K = md->A.rows, K1 = md->A1.rows, nI = K - 1;
        const size_t rb = size_t(std::max(1, (nI + 255) / 256));
        const size_t n_inv = size_t(D) * (size_t(std::max(nI, 1)) * 2 + rb + 1 + size_t(nbl) + 4 * size_t(N));
        ctx->d_adj_inv.reserve(n_inv * sizeof(double));
        ctx->d_adj_abar.reserve(size_t(D) * nbl * 2 * sizeof(double));
        double* w = ctx->d_adj_inv.as<double>();
        auto take = [&](size_t n) {
            double* p = w;
            w += n;
            return p;
        };
        InversionAdjoint ia{};
        ia.D = D;
        ia.nbl = nbl;
        ia.K = K;
        ia.N = N;
        ia.K1 = K1;
        ia.cal = ctx->d_cal.as<double>();
        ia.avg0 = d_avg0.as<double>();
        ia.ph = d_avg0.as<double>() + 2 * size_t(D) * nbl;
        ia.pairs = md->d_pairs.as<int2>();
        ia.I = md->d_I.as<int32_t>();
        ia.pairs1 = md->d_pairs1.as<int2>();
        ia.I1 = md->d_I1.as<int32_t>();
        ia.csr_off = md->d_csr_off.as<int32_t>();
        ia.csr_ent = md->d_csr_ent.as<int32_t>();
        ia.Gp = md->d_Gp.as<double>();
        ia.G1p = md->d_Gp.as<double>() + size_t(N) * N;
        ia.res = take(size_t(D) * std::max(nI, 1));
        ia.part = take(size_t(D) * rb);
        ia.loss = take(D);
        ia.g_arg = take(size_t(D) * nbl);
        ia.g_cal = take(size_t(D) * N);
        ia.y = take(size_t(D) * N);
        ia.gacal1 = take(size_t(D) * std::max(nI, 1));
        ia.g_cal1 = take(size_t(D) * N);
        ia.y1 = take(size_t(D) * N);
        ia.abar = want_grad ? ctx->d_adj_abar.as<double>() : nullptr;
        launch_inversion_adjoint(s, ia);
        ctx->h_adj_out.reserve((size_t(D) * 3 + size_t(N) * 3) * sizeof(double));
        double* h_out = ctx->h_adj_out.as<double>();
        LEAP_HIP(hipMemcpyAsync(h_out, ia.loss, size_t(D) * sizeof(double), hipMemcpyDeviceToHost, s));
        // The losses are read after the last synchronisation (the gradient's
        // work is queued behind their copy).
        auto finish_loss = [&] {
            float fwd_ms = 0.f;
            LEAP_HIP(hipEventElapsedTime(&fwd_ms, ctx->ev_grad[0], ctx->ev_grad[1]));
            ctx->last_grad_forward_ms = fwd_ms;
            double loss = 0.0;
            for (uint32_t d = 0; d < D; ++d) loss += h_out[d];  // directions in order
            *loss_out = loss;
        };
        if (!want_grad) {
            LEAP_HIP(hipStreamSynchronize(s));
            finish_loss();
            return;
        }
        // adjoint of rotate -> average on the device
        DeviceBuffer &d_abar = ctx->d_adj_abar, &d_guvw = ctx->d_adj_guvw, &d_gdzp = ctx->d_adj_gdzp,
                     &d_gdz = ctx->d_adj_gdz;
        d_guvw.reserve(size_t(R) * 3 * sizeof(double));
        d_gdzp.reserve(size_t(nbl) * D * 3 * sizeof(double));
        d_gdz.reserve(size_t(D) * 5 * sizeof(double));  // g_dz (3 D), then dL/d(ra, dec) (2 D)
        LEAP_HIP(hipEventRecord(ctx->ev_grad[2], s));
        if (row_sums)
            launch_adjoint_q(s, ctx->d_adj_q.as<double>(), uvw, ctx->d_dd.as<double>(), d_abar.as<double>(), D, T, nbl,
                             md->freq_start_hz, md->freq_inc_hz, d_guvw.as<double>(), d_gdzp.as<double>());
        else
            launch_rotate_adjoint(s, rows.vis, ms->vis_dtype, uvw, ctx->d_dd.as<double>(), d_abar.as<double>(), D, T,
                                  nbl, md->channels, md->pols, md->freq_start_hz, md->freq_inc_hz,
                                  d_guvw.as<double>(), d_gdzp.as<double>());
        LEAP_HIP(hipEventRecord(ctx->ev_grad[3], s));
        // dL/d(dz row of dd) reduced over baselines and mapped to dL/d(ra, dec) (MetaData.cc:166-203
        // differentiated) on the device
        launch_reduce_gdz_offsets(s, d_gdzp.as<double>(), nbl, D, md->phase_centre_ra, md->phase_centre_dec,
                                  ctx->d_dirs.as<double>(), d_gdz.as<double>());
        if (grad_offsets)
            LEAP_HIP(hipMemcpyAsync(h_out + D, d_gdz.as<double>() + 3 * size_t(D), size_t(D) * 2 * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        if (grad_antenna) {
            DeviceBuffer& d_gant = ctx->d_adj_gant;
            d_gant.reserve(size_t(N) * 3 * sizeof(double));
            launch_antenna_grad(s, d_guvw.as<double>(), d_M.as<double>(), ctx->d_adj_off.as<int32_t>(),
                                ctx->d_adj_bls.as<int32_t>(), ctx->d_adj_sgn.as<int32_t>(), N, nbl, T,
                                d_gant.as<double>());
            LEAP_HIP(hipMemcpyAsync(h_out + 3 * size_t(D), d_gant.ptr, size_t(N) * 3 * sizeof(double),
                                    hipMemcpyDeviceToHost, s));
        }
        LEAP_HIP(hipStreamSynchronize(s));
        finish_loss();
        if (grad_antenna) std::copy(h_out + 3 * size_t(D), h_out + 3 * size_t(D) + size_t(N) * 3, grad_antenna);
        float adj_ms = 0.f;
        LEAP_HIP(hipEventElapsedTime(&adj_ms, ctx->ev_grad[2], ctx->ev_grad[3]));
        ctx->last_grad_adjoint_ms = adj_ms;
        if (grad_offsets) std::copy(h_out + D, h_out + 3 * size_t(D), grad_offsets);
    });
}
