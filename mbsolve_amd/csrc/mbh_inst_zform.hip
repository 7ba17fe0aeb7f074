/*
 * Kernel instantiations: Rodrigues in the rotated basis (z-form) with FP32
 * (MBH_FORM_MIXED) or FP64 (MBH_FORM_ZFORM64) state storage, plus the
 * conversions between the reference coherence vector d and the z-form
 * state y:  y = Q^T A1 (d + d_in) - Q^T d_in,   d = P y + o
 * (FP32 storage of the z-forms holds y - y*, y* = -Q^T d_in).
 */
#include "mbh_tile.hpp"

namespace mbh {

template<int N, typename ST>
__global__ __launch_bounds__(64) void zform_to_y(ZArgs a)
{
    constexpr int K = N * N - 1;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    const uint64_t g = a.g0 + i;
    const bool in = i < a.count;
    const int32_t m = in ? a.mat[g] : -1;
    double d[K], y[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        d[j] = in ? a.dref[(uint64_t)j * a.dstride + g] : 0.0;
        y[j] = 0.0;
    }
    for_each_material(m, [&](int32_t mm) {
        cdp q = (cdp)a.zq + (uint64_t)mm * a.zstride;
        if (q[0] != 0.0) {
            cdp Qt = q + algo_zform<N>::OFF_QT;
            cdp QtA1 = q + algo_zform<N>::OFF_QTA1;
            cdp din = q + algo_zform<N>::OFF_DIN;
            double t[K];
#pragma unroll
            for (int j = 0; j < K; j++) {
                t[j] = d[j] + din[j];
            }
#pragma unroll
            for (int i = 0; i < K; i++) {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < K; j++) {
                    acc = fma(QtA1[i * K + j], t[j], acc);
                    acc = fma(-Qt[i * K + j], din[j], acc);
                }
                y[i] = acc;
            }
            if (std::is_same<ST, float>::value && a.yshift) {
                /* FP32 storage holds y - y* */
#pragma unroll
                for (int i = 0; i < K; i++) {
                    y[i] -= q[algo_zform<N>::OFF_YEQ + i];
                }
            }
        }
    });
    if (in) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            static_cast<ST*>(a.y)[(uint64_t)j * a.dstride + g] = (ST)y[j];
        }
    }
}

template<int N, typename ST>
__global__ __launch_bounds__(64) void zform_from_y(ZArgs a)
{
    constexpr int K = N * N - 1;
    const uint64_t i = (uint64_t)blockIdx.x * 64 + threadIdx.x;
    const uint64_t g = a.g0 + i;
    const bool in = i < a.count;
    const int32_t m = in ? a.mat[g] : -1;
    typename algo_zform<N>::state s;
    double d[K];
#pragma unroll
    for (int j = 0; j < K; j++) {
        s.v[j] = in ? (double)static_cast<const ST*>(a.y)[(uint64_t)j * a.dstride + g]
                    : 0.0;
        d[j] = 0.0;
    }
    for_each_material(m, [&](int32_t mm) {
        cdp q = (cdp)a.zq + (uint64_t)mm * a.zstride;
        if (q[0] != 0.0) {
            if (std::is_same<ST, float>::value && a.yshift) {
#pragma unroll
                for (int j = 0; j < K; j++) {
                    s.v[j] += q[algo_zform<N>::OFF_YEQ + j];
                }
            }
#pragma unroll
            for (int j = 0; j < K; j++) {
                d[j] = algo_zform<N>::dcomp(q, s, j);
            }
        }
    });
    if (in && a.out_cm) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            a.out_cm[i * K + j] = d[j];
        }
    } else if (in) {
#pragma unroll
        for (int j = 0; j < K; j++) {
            const_cast<double*>(a.dref)[(uint64_t)j * a.dstride + g] = d[j];
        }
    }
}

template<int N, typename ST>
hipError_t launch_to_y(const ZArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL((zform_to_y<N, ST>),
                       dim3((unsigned)((a.count + 63) / 64)), dim3(64), 0,
                       stream, a);
    return hipGetLastError();
}
template<int N, typename ST>
hipError_t launch_from_y(const ZArgs& a, hipStream_t stream)
{
    hipLaunchKernelGGL((zform_from_y<N, ST>),
                       dim3((unsigned)((a.count + 63) / 64)), dim3(64), 0,
                       stream, a);
    return hipGetLastError();
}

template __global__ void tile_kernel<algo_zform<2>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_zform<2>, 2, float>(KArgs);
template __global__ void tile_kernel<algo_zform<2>, 4, float>(KArgs);
template __global__ void tile_kernel<algo_zform<3>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_zform<3>, 2, float>(KArgs);
template __global__ void tile_kernel<algo_zform<4>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_zform<5>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_zform<6>, 1, float>(KArgs);
template __global__ void tile_kernel<algo_zform<2>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_zform<2>, 2, double>(KArgs);
template __global__ void tile_kernel<algo_zform<3>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_zform<3>, 2, double>(KArgs);
template __global__ void tile_kernel<algo_zform<4>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_zform<5>, 1, double>(KArgs);
template __global__ void tile_kernel<algo_zform<6>, 1, double>(KArgs);
#define MBH_ZINST(LEV)                                                        \
    template __global__ void zform_to_y<LEV, float>(ZArgs);                   \
    template __global__ void zform_to_y<LEV, double>(ZArgs);                  \
    template __global__ void zform_from_y<LEV, float>(ZArgs);                 \
    template __global__ void zform_from_y<LEV, double>(ZArgs);
MBH_ZINST(2)
MBH_ZINST(3)
MBH_ZINST(4)
MBH_ZINST(5)
MBH_ZINST(6)

#if !defined(__HIP_DEVICE_COMPILE__)
#define MBH_ZENTRY(LEV, FORM, C, ST)                                          \
    kernel_entry { MBH_ALGO_CVR_RODR, LEV, FORM, C,                           \
                   &launch_tile<algo_zform<LEV>, C, ST> }

const kernel_entry k_table_zform[] = {
    MBH_ZENTRY(2, MBH_FORM_MIXED, 1, float),
    MBH_ZENTRY(2, MBH_FORM_MIXED, 2, float),
    MBH_ZENTRY(2, MBH_FORM_MIXED, 4, float),
    MBH_ZENTRY(3, MBH_FORM_MIXED, 1, float),
    MBH_ZENTRY(3, MBH_FORM_MIXED, 2, float),
    MBH_ZENTRY(4, MBH_FORM_MIXED, 1, float),
    MBH_ZENTRY(5, MBH_FORM_MIXED, 1, float),
    MBH_ZENTRY(6, MBH_FORM_MIXED, 1, float),
    MBH_ZENTRY(2, MBH_FORM_ZFORM64, 1, double),
    MBH_ZENTRY(2, MBH_FORM_ZFORM64, 2, double),
    MBH_ZENTRY(3, MBH_FORM_ZFORM64, 1, double),
    MBH_ZENTRY(3, MBH_FORM_ZFORM64, 2, double),
    MBH_ZENTRY(4, MBH_FORM_ZFORM64, 1, double),
    MBH_ZENTRY(5, MBH_FORM_ZFORM64, 1, double),
    MBH_ZENTRY(6, MBH_FORM_ZFORM64, 1, double),
};
const uint32_t k_table_zform_n = sizeof(k_table_zform) / sizeof(k_table_zform[0]);

#define MBH_ZCONV(LEV)                                                        \
    zconv_entry{ LEV, 1, &launch_to_y<LEV, float>, &launch_from_y<LEV, float> }, \
    zconv_entry{ LEV, 0, &launch_to_y<LEV, double>, &launch_from_y<LEV, double> }

const zconv_entry z_table[] = {
    MBH_ZCONV(2), MBH_ZCONV(3), MBH_ZCONV(4), MBH_ZCONV(5), MBH_ZCONV(6),
};
const uint32_t z_table_n = sizeof(z_table) / sizeof(z_table[0]);
#endif

} /* namespace mbh */
