/*
 * Kernel instantiations in the step order of solver_cpu_fdtd_red::run
 * (density update first, solver_cpu_fdtd_red.tpp:451-549), reference
 * operation order and FP64 state (MBH_FORM_FP64): noop, regularized Cayley
 * and generalized Rodrigues, N = 2..6.
 */
#include "mbh_tile.hpp"

namespace mbh {

#define MBH_RED_ENTRY(ALGO_ID, LEV, C, A)                                     \
    kernel_entry { ALGO_ID, LEV, MBH_FORM_FP64, C,                            \
                   &launch_tile<A, C, double, true>, MBH_ORDER_RED }

template __global__ void tile_kernel<algo_noop, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_noop, 2, double, true>(KArgs);
template __global__ void tile_kernel<algo_noop, 4, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<2>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<2>, 2, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<3>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<4>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<5>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_cayley<6>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<2>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<2>, 2, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<3>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<3>, 2, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<4>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<5>, 1, double, true>(KArgs);
template __global__ void tile_kernel<algo_rodr<6>, 1, double, true>(KArgs);

#if !defined(__HIP_DEVICE_COMPILE__)
const kernel_entry k_table_red[] = {
    MBH_RED_ENTRY(MBH_ALGO_NOOP, 0, 1, algo_noop),
    MBH_RED_ENTRY(MBH_ALGO_NOOP, 0, 2, algo_noop),
    MBH_RED_ENTRY(MBH_ALGO_NOOP, 0, 4, algo_noop),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 2, 1, algo_cayley<2>),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 2, 2, algo_cayley<2>),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 3, 1, algo_cayley<3>),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 4, 1, algo_cayley<4>),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 5, 1, algo_cayley<5>),
    MBH_RED_ENTRY(MBH_ALGO_REG_CAYLEY, 6, 1, algo_cayley<6>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 2, 1, algo_rodr<2>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 2, 2, algo_rodr<2>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 3, 1, algo_rodr<3>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 3, 2, algo_rodr<3>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 4, 1, algo_rodr<4>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 5, 1, algo_rodr<5>),
    MBH_RED_ENTRY(MBH_ALGO_CVR_RODR, 6, 1, algo_rodr<6>),
};
const uint32_t k_table_red_n = sizeof(k_table_red) / sizeof(k_table_red[0]);
#endif

} /* namespace mbh */
