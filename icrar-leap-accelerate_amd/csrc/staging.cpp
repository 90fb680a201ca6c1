// Host -> device staging of host-resident visibilities.
//
// Pageable host memory is copied in chunks through two pinned buffers: worker
// threads fill one buffer while the DMA engine drains the other on the
// context's copy stream.  For the packed pol-0 path the workers apply the
// reference's skip rule (CpuLeapCalibrator.cc:201-214: a (row, channel) with a
// non-finite polarisation is left out, here written as 0) and keep pol 0
// only, so a quarter of a 4-pol cube crosses the bus.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "leap_internal.h"

namespace leap {

namespace {

constexpr size_t kChunkBytes = size_t(256) << 20;  // per pinned buffer

unsigned worker_count() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hc ? hc : 1u));
}

// Runs f(begin, end) over [0, n) split across the workers.
template <typename F>
void parallel_ranges(size_t n, F&& f) {
    const unsigned w = unsigned(std::min<size_t>(worker_count(), std::max<size_t>(1, n / 4096)));
    if (w <= 1) {
        f(size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(w);
    const size_t step = (n + w - 1) / w;
    for (unsigned i = 0; i < w; ++i) {
        const size_t b = std::min(n, size_t(i) * step), e = std::min(n, b + step);
        if (b < e) pool.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto& t : pool) t.join();
}

void ensure_staging(leap_ctx* ctx) {
    if (ctx->pinned_bytes >= kChunkBytes) return;
    for (int b = 0; b < 2; ++b) {
        if (ctx->h2d_pending[b]) LEAP_HIP(hipEventSynchronize(ctx->ev_h2d[b]));
        ctx->h2d_pending[b] = false;
        if (ctx->h_pinned[b]) LEAP_HIP(hipHostFree(ctx->h_pinned[b]));
        ctx->h_pinned[b] = nullptr;
    }
    ctx->pinned_bytes = 0;
    for (int b = 0; b < 2; ++b) LEAP_HIP(hipHostMalloc(&ctx->h_pinned[b], kChunkBytes, hipHostMallocDefault));
    ctx->pinned_bytes = kChunkBytes;
}

// The copy stream starts after the work already queued on ctx->stream.
void fork_copy_stream(leap_ctx* ctx) {
    LEAP_HIP(hipEventRecord(ctx->ev_copy_fork, ctx->stream));
    LEAP_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_copy_fork, 0));
}

// Later work on ctx->stream starts after the copies.
void join_copy_stream(leap_ctx* ctx) {
    LEAP_HIP(hipEventRecord(ctx->ev_copy_join, ctx->copy_stream));
    LEAP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_copy_join, 0));
}

// The pinned buffer b is free once its previous copy has completed.
void acquire(leap_ctx* ctx, int b) {
    if (ctx->h2d_pending[b]) {
        LEAP_HIP(hipEventSynchronize(ctx->ev_h2d[b]));
        ctx->h2d_pending[b] = false;
    }
}

void send(leap_ctx* ctx, int b, void* device_dst, size_t bytes) {
    LEAP_HIP(hipMemcpyAsync(device_dst, ctx->h_pinned[b], bytes, hipMemcpyHostToDevice, ctx->copy_stream));
    LEAP_HIP(hipEventRecord(ctx->ev_h2d[b], ctx->copy_stream));
    ctx->h2d_pending[b] = true;
    ctx->last_stage_bytes += bytes;
}

template <typename VT>
void pack_pol0_host(const VT* src, uint32_t P, size_t begin, size_t end, VT* out) {
    for (size_t i = begin; i < end; ++i) {
        const VT* p = src + i * 2 * P;
        bool finite = true;
        for (uint32_t k = 0; k < 2 * P; ++k) finite = finite && std::isfinite(p[k]);
        out[2 * i] = finite ? p[0] : VT(0);
        out[2 * i + 1] = finite ? p[1] : VT(0);
    }
}

}  // namespace

void stage_copy(leap_ctx* ctx, const void* host_src, size_t bytes, void* device_dst) {
    if (bytes == 0) return;
    ensure_staging(ctx);
    fork_copy_stream(ctx);
    const char* src = static_cast<const char*>(host_src);
    char* dst = static_cast<char*>(device_dst);
    int b = 0;
    for (size_t off = 0; off < bytes; off += kChunkBytes, b ^= 1) {
        const size_t len = std::min(kChunkBytes, bytes - off);
        acquire(ctx, b);
        char* pin = static_cast<char*>(ctx->h_pinned[b]);
        parallel_ranges(len, [&](size_t lo, size_t hi) { std::memcpy(pin + lo, src + off + lo, hi - lo); });
        send(ctx, b, dst + off, len);
    }
    join_copy_stream(ctx);
}

void stage_pol0(leap_ctx* ctx, const void* host_vis, int32_t vis_dtype, uint32_t P, size_t n, double* device_dst) {
    if (n == 0) return;
    ensure_staging(ctx);
    const bool c64 = vis_dtype == LEAP_VIS_C64;
    const size_t out_elem = c64 ? 8 : 16;  // pol 0 in the source precision
    const size_t per_chunk = kChunkBytes / out_elem;
    if (c64)
        for (int b = 0; b < 2; ++b) ctx->d_widen[b].reserve(kChunkBytes);
    fork_copy_stream(ctx);
    int b = 0;
    for (size_t e0 = 0; e0 < n; e0 += per_chunk, b ^= 1) {
        const size_t m = std::min(per_chunk, n - e0);
        acquire(ctx, b);
        if (c64) {
            const float* src = static_cast<const float*>(host_vis) + e0 * 2 * P;
            float* pin = static_cast<float*>(ctx->h_pinned[b]);
            parallel_ranges(m, [&](size_t lo, size_t hi) { pack_pol0_host(src, P, lo, hi, pin); });
            // the device staging buffer b is free once its previous widen kernel has run
            if (ctx->widen_pending[b]) LEAP_HIP(hipStreamWaitEvent(ctx->copy_stream, ctx->ev_widened[b], 0));
            send(ctx, b, ctx->d_widen[b].ptr, m * out_elem);
            LEAP_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_h2d[b], 0));
            launch_widen_c64(ctx->stream, ctx->d_widen[b].ptr, m, device_dst + 2 * e0);
            LEAP_HIP(hipEventRecord(ctx->ev_widened[b], ctx->stream));
            ctx->widen_pending[b] = true;
        } else {
            const double* src = static_cast<const double*>(host_vis) + e0 * 2 * P;
            double* pin = static_cast<double*>(ctx->h_pinned[b]);
            parallel_ranges(m, [&](size_t lo, size_t hi) { pack_pol0_host(src, P, lo, hi, pin); });
            send(ctx, b, device_dst + 2 * e0, m * out_elem);
        }
    }
    join_copy_stream(ctx);
}

}  // namespace leap
