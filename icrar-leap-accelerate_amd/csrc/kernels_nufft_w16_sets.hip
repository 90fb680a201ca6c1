// The multi-set launch of the width-16 NUFFT channel sums: complex64 or
// complex128 cubes with 1, 2 or 4 pols.  The kernels are in nufft_impl.h.
#include "nufft_impl.h"

namespace leap {

void launch_rotate_nufft_sets_w16(hipStream_t s, int32_t tile, const NufftPlan& plan, uint32_t C, uint32_t src,
                                  const NufftSetDesc* sets, const uint32_t* items, uint32_t n_items) {
    require(src == 1 || src == 2 || src == 4 || src == 17 || src == 18 || src == 20, LEAP_ERR_INTERNAL,
            "the width-16 multi-set NUFFT launch reads complex64 or complex128 cubes with 1, 2 or 4 pols");
#define LEAP_NUFFT_SETS(N_)                                                                          \
    if (plan.N == N_ && plan.W == 16) {                                                              \
        nufft::launch_nw_sets<N_, 16, double>(s, tile, int32_t(src), C, sets, items, n_items);       \
        LEAP_HIP(hipGetLastError());                                                                 \
        return;                                                                                      \
    }
    LEAP_NUFFT_SETS(256)
    LEAP_NUFFT_SETS(512)
    LEAP_NUFFT_SETS(1024)
#undef LEAP_NUFFT_SETS
    throw Error(LEAP_ERR_INTERNAL, "no width-16 multi-set NUFFT kernel for grid " + std::to_string(plan.N) +
                                       ", width " + std::to_string(plan.W));
}

}  // namespace leap
