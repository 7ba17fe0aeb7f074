/*
 * Kernel instantiations: noop (pure Maxwell) and regularized Cayley.
 * Reference operation order: FP64 storage implements MBH_FORM_FP64, FP32
 * storage MBH_FORM_MIXED (three-level Cayley; noop with the folded FDTD
 * constants).  Two-level Cayley in closed
 * Bloch form: FP32 storage MBH_FORM_MIXED, FP64 storage MBH_FORM_ZFORM64,
 * each also specialised for a diagonal Hamiltonian and a real dipole
 * (algo_bloch2r, chosen by the host from the constants), the latter also
 * as tiles of 2 and 4 wavefronts at 4 cells per lane.
 */
#include "mbh_tile.hpp"

namespace mbh {

#define MBH_ENTRY(ALGO_ID, LEV, FORM, C, A, ST)                               \
    kernel_entry { ALGO_ID, LEV, FORM, C, &launch_tile<A, C, ST> }
/* a tile of W wavefronts at C cells per lane: listed as C W cells per lane */
#define MBH_ENTRY_MW(ALGO_ID, LEV, FORM, C, W, A, ST)                         \
    kernel_entry { ALGO_ID, LEV, FORM, (C) * (W), &launch_tile_mw<A, C, W, ST> }

template __global__ void tile_kernel<algo_noop, 1, double>(KArgs);
template __global__ void tile_kernel<algo_noop, 2, double>(KArgs);
template __global__ void tile_kernel<algo_noop, 4, double>(KArgs);
template __global__ void tile_kernel<algo_noop, 8, double>(KArgs);
template __global__ void tile_kernel<algo_noop_fast, 1, float>(KArgs);
template __global__ void tile_kernel<algo_noop_fast, 2, float>(KArgs);
template __global__ void tile_kernel<algo_noop_fast, 4, float>(KArgs);
template __global__ void tile_kernel<algo_noop_fast, 8, float>(KArgs);
template __global__ void tile_kernel<algo_noop_hz0, 1, float>(KArgs);
template __global__ void tile_kernel<algo_noop_hz0, 2, float>(KArgs);
template __global__ void tile_kernel<algo_noop_hz0, 4, float>(KArgs);
template __global__ void tile_kernel<algo_noop_hz0, 8, float>(KArgs);
template __global__ void tile_kernel<algo_cayley<2>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<2>, 2, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<2>, 4, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 1, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 2, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 4, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 1, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 2, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 4, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 8, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2, 8, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 1, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 2, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 4, float>(KArgs);
template __global__ void tile_kernel_mw<algo_bloch2r, 4, 2, float>(KArgs);
template __global__ void tile_kernel_mw<algo_bloch2r, 4, 4, float>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 1, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 2, double>(KArgs);
template __global__ void tile_kernel<algo_bloch2r, 4, double>(KArgs);
template __global__ void tile_kernel_mw<algo_bloch2r, 4, 2, double>(KArgs);
template __global__ void tile_kernel_mw<algo_bloch2r, 4, 4, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<3>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<3>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_cayley<4>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<4>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_cayley<5>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<5>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_cayley<6>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_cayley<6>, 1, float>(KArgs);

#if !defined(__HIP_DEVICE_COMPILE__)

const kernel_entry k_table_basic[] = {
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_FP64, 1, algo_noop, double),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_FP64, 2, algo_noop, double),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_FP64, 4, algo_noop, double),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_FP64, 8, algo_noop, double),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_MIXED, 1, algo_noop_fast, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_MIXED, 2, algo_noop_fast, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_MIXED, 4, algo_noop_fast, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, MBH_FORM_MIXED, 8, algo_noop_fast, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, KFORM_NOOP_HZ0, 1, algo_noop_hz0, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, KFORM_NOOP_HZ0, 2, algo_noop_hz0, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, KFORM_NOOP_HZ0, 4, algo_noop_hz0, float),
    MBH_ENTRY(MBH_ALGO_NOOP, 0, KFORM_NOOP_HZ0, 8, algo_noop_hz0, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_FP64, 1, algo_cayley<2>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_FP64, 2, algo_cayley<2>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_FP64, 4, algo_cayley<2>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_MIXED, 1, algo_bloch2, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_MIXED, 2, algo_bloch2, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_MIXED, 4, algo_bloch2, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_ZFORM64, 1, algo_bloch2, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_ZFORM64, 2, algo_bloch2, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_ZFORM64, 4, algo_bloch2, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_MIXED, 8, algo_bloch2, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, MBH_FORM_ZFORM64, 8, algo_bloch2, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_MIXED, 1, algo_bloch2r, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_MIXED, 2, algo_bloch2r, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_MIXED, 4, algo_bloch2r, float),
    MBH_ENTRY_MW(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_MIXED, 4, 2, algo_bloch2r, float),
    MBH_ENTRY_MW(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_MIXED, 4, 4, algo_bloch2r, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_ZFORM64, 1, algo_bloch2r, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_ZFORM64, 2, algo_bloch2r, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_ZFORM64, 4, algo_bloch2r, double),
    MBH_ENTRY_MW(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_ZFORM64, 4, 2, algo_bloch2r, double),
    MBH_ENTRY_MW(MBH_ALGO_REG_CAYLEY, 2, KFORM_BLOCH2R_ZFORM64, 4, 4, algo_bloch2r, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 3, MBH_FORM_FP64, 1, algo_cayley<3>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 3, MBH_FORM_MIXED, 1, algo_cayley<3>, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 4, MBH_FORM_FP64, 1, algo_cayley<4>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 4, MBH_FORM_MIXED, 1, algo_cayley<4>, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 5, MBH_FORM_FP64, 1, algo_cayley<5>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 5, MBH_FORM_MIXED, 1, algo_cayley<5>, float),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 6, MBH_FORM_FP64, 1, algo_cayley<6>, double),
    MBH_ENTRY(MBH_ALGO_REG_CAYLEY, 6, MBH_FORM_MIXED, 1, algo_cayley<6>, float),
};
const uint32_t k_table_basic_n = sizeof(k_table_basic) / sizeof(k_table_basic[0]);
#endif

} /* namespace mbh */
