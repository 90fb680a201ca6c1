// NUFFT channel sums with an FP64 transform, and the entry point that picks
// the transform precision of a plan.  The kernels are in nufft_impl.h.
#include "nufft_impl.h"

namespace leap {

void launch_rotate_nufft_f64(hipStream_t s, const nufft::NufftLaunch& l) {
    const NufftPlan& plan = *l.plan;
#define LEAP_NUFFT(N_, W_)                                                                                   \
    if (plan.N == N_ && plan.W == W_) {                                                                      \
        nufft::launch_nw<N_, W_, double>(s, l.tile, l.nbl_grid, l.src, l.src_pols, l.scale, l.uvw, l.dd, l.D, \
                                         l.T, l.nbl, l.C, l.P_out, l.args, l.avg, l.accumulate, l.tail,      \
                                         l.tail_doubles);                                                    \
        return;                                                                                              \
    }
    LEAP_NUFFT(256, 10)
    LEAP_NUFFT(256, 12)
    LEAP_NUFFT(512, 10)
    LEAP_NUFFT(512, 12)
    LEAP_NUFFT(1024, 10)
    LEAP_NUFFT(1024, 12)
#undef LEAP_NUFFT
    throw Error(LEAP_ERR_INTERNAL, "no FP64 NUFFT kernel for grid " + std::to_string(plan.N) + ", width " +
                                       std::to_string(plan.W));
}

uint64_t nufft_sets_key(const NufftPlan& plan, uint32_t C, uint32_t P) {
    uint32_t slots = plan.N / 128;
    if (!plan.fp32 && plan.N == 1024 && plan.W == 10) slots = uint32_t(nufft::nufft_slots<1024, 10, double>(C));
    if (!plan.fp32 && plan.N == 1024 && plan.W == 16) slots = uint32_t(nufft::nufft_slots<1024, 16, double>(C));
    return (uint64_t(plan.N) << 32) | (uint64_t(plan.W) << 24) | (uint64_t(plan.fp32 ? 1 : 0) << 16) |
           (uint64_t(slots) << 8) | uint64_t(P);
}

void launch_rotate_nufft_sets(hipStream_t s, int32_t tile, const NufftPlan& plan, uint32_t C, uint32_t P,
                              const NufftSetDesc* sets, const uint32_t* items, uint32_t n_items) {
    if (n_items == 0) return;
    if (plan.w16) {
        launch_rotate_nufft_sets_w16(s, tile, plan, C, P, sets, items, n_items);
        return;
    }
    require(!plan.fp32, LEAP_ERR_NOT_IMPLEMENTED, "the multi-set NUFFT launch uses an FP64 transform");
    require(P == 1 || P == 2 || P == 4, LEAP_ERR_INTERNAL, "the width-10 multi-set NUFFT launch reads complex64 cubes");
#define LEAP_NUFFT_SETS(N_, W_)                                                                  \
    if (plan.N == N_ && plan.W == W_) {                                                          \
        nufft::launch_nw_sets<N_, W_, double>(s, tile, int32_t(P), C, sets, items, n_items);     \
        LEAP_HIP(hipGetLastError());                                                             \
        return;                                                                                  \
    }
    LEAP_NUFFT_SETS(256, 10)
    LEAP_NUFFT_SETS(256, 12)
    LEAP_NUFFT_SETS(512, 10)
    LEAP_NUFFT_SETS(512, 12)
    LEAP_NUFFT_SETS(1024, 10)
    LEAP_NUFFT_SETS(1024, 12)
#undef LEAP_NUFFT_SETS
    throw Error(LEAP_ERR_INTERNAL, "no FP64 NUFFT kernel for grid " + std::to_string(plan.N) + ", width " +
                                       std::to_string(plan.W));
}

size_t nufft_tail_bytes(uint32_t D) {
    int dev = 0, cus = 0;
    LEAP_HIP(hipGetDevice(&dev));
    LEAP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // resident workgroups <= 16 per CU for these kernels; the chunked round holds < 2x that
    return 2 * size_t(16) * size_t(std::max(1, cus)) * D * 2 * sizeof(double);
}

bool nufft_cube_supported(int32_t vis_dtype, uint32_t P, bool w16) {
    return (vis_dtype == LEAP_VIS_C64 || (w16 && vis_dtype == LEAP_VIS_C128)) && (P == 1 || P == 2 || P == 4);
}

void launch_rotate_nufft(hipStream_t s, int32_t tile, const NufftPlan& plan, const double* tables, const void* src,
                         bool cube, int32_t vis_dtype, const double* uvw, const double* dd, uint32_t D, uint32_t T,
                         uint32_t nbl, uint32_t nbl_grid, uint32_t C, uint32_t P, uint32_t P_out, double f0,
                         double df, double* avg, int32_t accumulate, double* tail, size_t tail_bytes) {
    if (nbl_grid == 0 || D == 0) return;
    require(P_out >= 1, LEAP_ERR_INTERNAL, "NUFFT output needs at least pol 0");
    require(plan.N != 0 && plan.scale.size() == C, LEAP_ERR_INTERNAL, "NUFFT plan does not match the channel count");
    require(!cube || nufft_cube_supported(vis_dtype, P, plan.w16), LEAP_ERR_INTERNAL,
            "NUFFT cube source needs 1, 2 or 4 pols of complex64 (complex128 for the width-16 kernel)");
    nufft::NufftLaunch l{};
    l.tile = tile;
    l.plan = &plan;
    l.src = src;
    l.src_pols = cube ? int32_t(P) + (vis_dtype == LEAP_VIS_C128 ? 16 : 0) : 0;
    l.scale = tables + kNufftCoefDoubles;
    l.uvw = uvw;
    l.dd = dd;
    l.D = D;
    l.T = T;
    l.nbl = nbl;
    l.nbl_grid = nbl_grid;
    l.C = C;
    l.P_out = P_out;
    l.args.coef = tables;
    l.args.coef32 = reinterpret_cast<const float*>(tables + nufft_coef32_offset(C));
    l.args.kappa_mid = (f0 + double(plan.h) * df) / kSpeedOfLight;
    l.args.dkappa = df / kSpeedOfLight;
    l.args.h = int32_t(plan.h);
    l.avg = avg;
    l.accumulate = accumulate;
    l.tail = tail;
    l.tail_doubles = tail_bytes / sizeof(double);
    if (plan.w16)
        launch_rotate_nufft_w16(s, l);
    else if (plan.fp32)
        launch_rotate_nufft_f32(s, l);
    else
        launch_rotate_nufft_f64(s, l);
    LEAP_HIP(hipGetLastError());
}

}  // namespace leap
