/*
 * Kernel instantiations: Rodrigues z-form with the generator product on
 * MFMA, N = 3..6, FP32 state storage (MBH_FORM_MFMA) and FP64 state
 * storage (MBH_FORM_MFMA64).
 */
#include "mbh_zmfma.hpp"

namespace mbh {

template __global__ void zm_kernel<3, float>(KArgs);
template __global__ void zm_kernel<4, float>(KArgs);
template __global__ void zm_kernel<5, float>(KArgs);
template __global__ void zm_kernel<6, float>(KArgs);
template __global__ void zm_kernel<3, double>(KArgs);
template __global__ void zm_kernel<4, double>(KArgs);
template __global__ void zm_kernel<5, double>(KArgs);
template __global__ void zm_kernel<6, double>(KArgs);

#if !defined(__HIP_DEVICE_COMPILE__)
#define MBH_ZMENTRY(LEV, FORM, ST)                                            \
    kernel_entry { MBH_ALGO_CVR_RODR, LEV, FORM, 1, &launch_zm<LEV, ST> }

const kernel_entry k_table_zm[] = {
    MBH_ZMENTRY(3, MBH_FORM_MFMA, float),
    MBH_ZMENTRY(4, MBH_FORM_MFMA, float),
    MBH_ZMENTRY(5, MBH_FORM_MFMA, float),
    MBH_ZMENTRY(6, MBH_FORM_MFMA, float),
    MBH_ZMENTRY(3, MBH_FORM_MFMA64, double),
    MBH_ZMENTRY(4, MBH_FORM_MFMA64, double),
    MBH_ZMENTRY(5, MBH_FORM_MFMA64, double),
    MBH_ZMENTRY(6, MBH_FORM_MFMA64, double),
};
const uint32_t k_table_zm_n = sizeof(k_table_zm) / sizeof(k_table_zm[0]);
#endif

} /* namespace mbh */
