// NUFFT channel sums with an FP64 transform and the width-16 kernel
// (LEAP_CHANNEL_SUM_NUFFT_W16), from packed rows or a complex64 or complex128
// cube.  The kernels are in nufft_impl.h.
#include "nufft_impl.h"

namespace leap {

void launch_rotate_nufft_w16(hipStream_t s, const nufft::NufftLaunch& l) {
    const NufftPlan& plan = *l.plan;
#define LEAP_NUFFT(N_)                                                                                         \
    if (plan.N == N_ && plan.W == 16) {                                                                        \
        nufft::launch_nw<N_, 16, double>(s, l.tile, l.nbl_grid, l.src, l.src_pols, l.scale, l.uvw, l.dd, l.D, \
                                         l.T, l.nbl, l.C, l.P_out, l.args, l.avg, l.accumulate, l.tail,        \
                                         l.tail_doubles);                                                      \
        return;                                                                                                \
    }
    LEAP_NUFFT(256)
    LEAP_NUFFT(512)
    LEAP_NUFFT(1024)
#undef LEAP_NUFFT
    throw Error(LEAP_ERR_INTERNAL, "no width-16 NUFFT kernel for grid " + std::to_string(plan.N) + ", width " +
                                       std::to_string(plan.W));
}

}  // namespace leap
