/*
 * Kernel instantiations: generalized Rodrigues, reference (FP64) form,
 * N = 2..6 levels.
 */
#include "mbh_tile.hpp"

namespace mbh {

#define MBH_ENTRY(LEV, C)                                                     \
    kernel_entry { MBH_ALGO_CVR_RODR, LEV, MBH_FORM_FP64, C,                  \
                   &launch_tile<algo_rodr<LEV>, C, double> }

template __global__ void tile_kernel<algo_rodr<2>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<2>, 2, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<3>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<3>, 2, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<4>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<5>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_rodr<6>, 1, double>(KArgs);

#if !defined(__HIP_DEVICE_COMPILE__)
const kernel_entry k_table_rodr[] = {
    MBH_ENTRY(2, 1), MBH_ENTRY(2, 2), MBH_ENTRY(3, 1), MBH_ENTRY(3, 2),
    MBH_ENTRY(4, 1), MBH_ENTRY(5, 1), MBH_ENTRY(6, 1),
};
const uint32_t k_table_rodr_n = sizeof(k_table_rodr) / sizeof(k_table_rodr[0]);
#endif

} /* namespace mbh */
