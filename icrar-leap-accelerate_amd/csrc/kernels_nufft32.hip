// NUFFT channel sums with an FP32 transform, grid and interpolation (the
// mixed-precision default).  The kernels are in nufft_impl.h.
#include "nufft_impl.h"

namespace leap {

void launch_rotate_nufft_f32(hipStream_t s, const nufft::NufftLaunch& l) {
    const NufftPlan& plan = *l.plan;
#define LEAP_NUFFT(N_, W_)                                                                                   \
    if (plan.N == N_ && plan.W == W_) {                                                                      \
        nufft::launch_nw<N_, W_, float>(s, l.tile, l.nbl_grid, l.src, l.src_pols, l.scale, l.uvw, l.dd, l.D,  \
                                        l.T, l.nbl, l.C, l.P_out, l.args, l.avg, l.accumulate, l.tail,       \
                                        l.tail_doubles);                                                     \
        return;                                                                                              \
    }
    LEAP_NUFFT(256, 8)
    LEAP_NUFFT(256, 10)
    LEAP_NUFFT(512, 8)
    LEAP_NUFFT(512, 10)
    LEAP_NUFFT(1024, 8)
    LEAP_NUFFT(1024, 10)
#undef LEAP_NUFFT
    throw Error(LEAP_ERR_INTERNAL, "no FP32 NUFFT kernel for grid " + std::to_string(plan.N) + ", width " +
                                       std::to_string(plan.W));
}

}  // namespace leap
