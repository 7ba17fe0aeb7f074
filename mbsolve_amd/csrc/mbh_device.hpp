/*
 * Per-cell density propagators of the gfx950 Maxwell-Bloch backend.
 *
 * Every update works on one grid cell whose state lives in registers and
 * reads its material constants through a wave-uniform pointer, so the
 * compiler turns the constant reads into scalar (s_load) loads.
 *
 * FP64 "mirror" forms follow the reference update expressions operand by
 * operand (dot products accumulate in ascending index order, starting
 * from the first product; no fused multiply-add), which makes them
 * bit-identical to a CPU implementation that uses the same order.
 */
#ifndef MBH_DEVICE_HPP
#define MBH_DEVICE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbsolve_hip.h"

namespace mbh {

/* ---- complex helpers (naive formulas, the libstdc++ fast path) --------- */
struct cplx {
    double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b)
{
    return cplx{ a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b)
{
    return cplx{ a.re + b.re, a.im + b.im };
}
__device__ __forceinline__ cplx csub(cplx a, cplx b)
{
    return cplx{ a.re - b.re, a.im - b.im };
}
/* 1 / z = conj(z) / |z|^2 */
__device__ __forceinline__ cplx cinv(cplx z)
{
    double s = z.re * z.re + z.im * z.im;
    return cplx{ z.re / s, -z.im / s };
}

/* Matrix product C = A B of N x N complex matrices (row-major). */
template<int N>
__device__ __forceinline__ void cmatmul(const cplx* A, const cplx* B, cplx* C)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int j = 0; j < N; j++) {
            cplx acc = cmul(A[i * N], B[j]);
#pragma unroll
            for (int k = 1; k < N; k++) {
                acc = cadd(acc, cmul(A[i * N + k], B[k * N + j]));
            }
            C[i * N + j] = acc;
        }
    }
}

/* C = A B^dagger */
template<int N>
__device__ __forceinline__ void cmatmul_adj(const cplx* A, const cplx* B,
                                            cplx* C)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
#pragma unroll
        for (int j = 0; j < N; j++) {
            cplx b0{ B[j * N].re, -B[j * N].im };
            cplx acc = cmul(A[i * N], b0);
#pragma unroll
            for (int k = 1; k < N; k++) {
                cplx bk{ B[j * N + k].re, -B[j * N + k].im };
                acc = cadd(acc, cmul(A[i * N + k], bk));
            }
            C[i * N + j] = acc;
        }
    }
}

/* Inverse of an N x N complex matrix: cofactor formula for N == 2,
 * Gauss-Jordan with partial pivoting (largest |a|^2) otherwise. */
template<int N>
__device__ __forceinline__ void cinverse(const cplx* A, cplx* R)
{
    if constexpr (N == 2) {
        cplx det = csub(cmul(A[0], A[3]), cmul(A[1], A[2]));
        cplx id = cinv(det);
        R[0] = cmul(A[3], id);
        R[1] = cmul(cplx{ -A[1].re, -A[1].im }, id);
        R[2] = cmul(cplx{ -A[2].re, -A[2].im }, id);
        R[3] = cmul(A[0], id);
    } else {
        cplx W[N * N];
#pragma unroll
        for (int i = 0; i < N * N; i++) {
            W[i] = A[i];
            R[i] = cplx{ (i / N == i % N) ? 1.0 : 0.0, 0.0 };
        }
#pragma unroll
        for (int col = 0; col < N; col++) {
            /* pivot search */
            int piv = col;
            double best = W[col * N + col].re * W[col * N + col].re +
                W[col * N + col].im * W[col * N + col].im;
#pragma unroll
            for (int r = col + 1; r < N; r++) {
                double v = W[r * N + col].re * W[r * N + col].re +
                    W[r * N + col].im * W[r * N + col].im;
                if (v > best) {
                    best = v;
                    piv = r;
                }
            }
            /* row swap (register-friendly: compare-select per row) */
#pragma unroll
            for (int r = col + 1; r < N; r++) {
                if (r == piv) {
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        cplx t = W[col * N + k];
                        W[col * N + k] = W[r * N + k];
                        W[r * N + k] = t;
                        t = R[col * N + k];
                        R[col * N + k] = R[r * N + k];
                        R[r * N + k] = t;
                    }
                }
            }
            cplx ip = cinv(W[col * N + col]);
#pragma unroll
            for (int k = 0; k < N; k++) {
                W[col * N + k] = cmul(W[col * N + k], ip);
                R[col * N + k] = cmul(R[col * N + k], ip);
            }
#pragma unroll
            for (int r = 0; r < N; r++) {
                if (r != col) {
                    cplx f = W[r * N + col];
#pragma unroll
                    for (int k = 0; k < N; k++) {
                        W[r * N + k] = csub(W[r * N + k],
                                            cmul(f, W[col * N + k]));
                        R[r * N + k] = csub(R[r * N + k],
                                            cmul(f, R[col * N + k]));
                    }
                }
            }
        }
    }
}

/* ======================================================================
 * lindblad_noop (algo_lindblad_noop.hpp:35-86): no density, p untouched.
 * ==================================================================== */
struct algo_noop {
    static constexpr int DC = 0; /* stored density components */
    static constexpr int STATE = 1;
    struct state {
        double v[STATE];
    };
    __device__ static void update(const double*, state&, double, double&) {}
    __device__ static double inversion(const double*, const state&)
    {
        return 0.0;
    }
    __device__ static double population(const double*, const state&, int)
    {
        return 0.0;
    }
};

/* ======================================================================
 * lindblad_reg_cayley<N>::update (algo_lindblad_reg_cayley.hpp:228-257)
 *
 * Blob layout: see MBH_CAYLEY_STRIDE in mbsolve_hip.h.  State: the full
 * N x N complex density matrix, row-major, (re, im) interleaved, exactly
 * as the reference keeps it (no Hermitian symmetrisation).
 * ==================================================================== */
template<int N>
struct algo_cayley {
    static constexpr int DC = 2 * N * N;
    static constexpr int STATE = 2 * N * N;
    struct state {
        double v[STATE];
    };

    static constexpr int OFF_H0 = 5;
    static constexpr int OFF_MU = OFF_H0 + 2 * N * N;
    static constexpr int OFF_d1 = OFF_MU + 2 * N * N;
    static constexpr int OFF_d2 = OFF_d1 + N * N;
    static constexpr int OFF_D1 = OFF_d2 + N * N;
    static constexpr int OFF_D2 = OFF_D1 + N * N;

    /* propagate_dissipation (algo_lindblad_reg_cayley.hpp:206-217):
     * coherences *= D2 (elementwise), diagonal <- D1 * diagonal. */
    __device__ static void dissipate(const double* q, const cplx* d, cplx* o)
    {
        const double* D1 = q + OFF_D1;
        const double* D2 = q + OFF_D2;
#pragma unroll
        for (int i = 0; i < N * N; i++) {
            o[i] = cplx{ D2[i] * d[i].re, D2[i] * d[i].im };
        }
#pragma unroll
        for (int i = 0; i < N; i++) {
            cplx acc{ D1[i * N] * d[0].re, D1[i * N] * d[0].im };
#pragma unroll
            for (int k = 1; k < N; k++) {
                acc.re = acc.re + D1[i * N + k] * d[k * N + k].re;
                acc.im = acc.im + D1[i * N + k] * d[k * N + k].im;
            }
            o[i * N + i] = acc;
        }
    }

    __device__ static void update(const double* q, state& s, double e,
                                  double& p)
    {
        if (q[0] == 0.0) {
            return;
        }
        const cplx* H0 = reinterpret_cast<const cplx*>(q + OFF_H0);
        const cplx* mu = reinterpret_cast<const cplx*>(q + OFF_MU);
        const double prop_im = q[3];
        const double rhs_im = q[4];
        cplx* d = reinterpret_cast<cplx*>(s.v);

        /* H = H0 - e mu;  M = (-i/hbar dt) H */
        cplx A[N * N], B[N * N];
#pragma unroll
        for (int i = 0; i < N * N; i++) {
            cplx h{ H0[i].re - e * mu[i].re, H0[i].im - e * mu[i].im };
            cplx m{ -(prop_im * h.im), prop_im * h.re };
            double di = (i / N == i % N) ? 1.0 : 0.0;
            /* A = I - 0.5 M, B = I + 0.5 M */
            A[i] = cplx{ di - 0.5 * m.re, -(0.5 * m.im) };
            B[i] = cplx{ di + 0.5 * m.re, 0.5 * m.im };
        }
        /* Cayley propagator U2 = A^-1 B */
        cplx Ai[N * N], U2[N * N];
        cinverse<N>(A, Ai);
        cmatmul<N>(Ai, B, U2);

        /* dissipation half step, unitary full step, dissipation half step */
        cplx d1[N * N], t[N * N];
        dissipate(q, d, d1);
        cmatmul<N>(U2, d1, t);
        cmatmul_adj<N>(t, U2, d1);
        dissipate(q, d1, d);

        /* polarization derivative: rhs = -i/hbar [H0, d] + D(d);
         * p = n Re tr(rhs mu) */
        const double* r1 = q + OFF_d1;
        const double* r2 = q + OFF_d2;
        cplx c1[N * N], c2[N * N];
        cmatmul<N>(H0, d, c1);
        cmatmul<N>(d, H0, c2);
        cplx rhs[N * N];
#pragma unroll
        for (int i = 0; i < N * N; i++) {
            cplx c = csub(c1[i], c2[i]);
            cplx g{ -(rhs_im * c.im), rhs_im * c.re };
            rhs[i] = cplx{ r2[i] * d[i].re, r2[i] * d[i].im };
            if (i / N == i % N) {
                int ii = i / N;
                cplx acc{ r1[ii * N] * d[0].re, r1[ii * N] * d[0].im };
#pragma unroll
                for (int k = 1; k < N; k++) {
                    acc.re = acc.re + r1[ii * N + k] * d[k * N + k].re;
                    acc.im = acc.im + r1[ii * N + k] * d[k * N + k].im;
                }
                rhs[i] = acc;
            }
            rhs[i] = cadd(g, rhs[i]);
        }
        double tr = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            double di = rhs[i * N].re * mu[i].re - rhs[i * N].im * mu[i].im;
#pragma unroll
            for (int k = 1; k < N; k++) {
                di = di +
                    (rhs[i * N + k].re * mu[k * N + i].re -
                     rhs[i * N + k].im * mu[k * N + i].im);
            }
            tr = (i == 0) ? di : tr + di;
        }
        p = q[1] * tr;
    }

    /* calc_inversion / calc_population (algo_lindblad_reg_cayley.hpp:259-267) */
    __device__ static double inversion(const double*, const state& s)
    {
        return s.v[2 * (N + 1)] - s.v[0];
    }
    __device__ static double population(const double*, const state& s, int m)
    {
        double r = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            if (i == m) {
                r = s.v[2 * (i * N + i)];
            }
        }
        return r;
    }
};

/* ---- coherence-vector population readout ------------------------------
 * cv_representation::calc_population
 * (coherence_vector_representation.hpp:289-336). */
template<int N>
__device__ __forceinline__ double cv_population(const double* d, int m)
{
    if constexpr (N == 2) {
        return (m == 0) ? 0.5 - 0.5 * d[2] : 0.5 + 0.5 * d[2];
    } else if constexpr (N == 3) {
        const double s3 = sqrt(3.0);
        if (m == 0) {
            return 1.0 / 3 + 0.5 * (-1.0 * d[6] - 1.0 / s3 * d[7]);
        } else if (m == 1) {
            return 1.0 / 3 + 0.5 * (+1.0 * d[6] - 1.0 / s3 * d[7]);
        } else if (m == 2) {
            return 1.0 / 3 + 0.5 * (+2.0 / s3 * d[7]);
        }
        return 0.0;
    } else {
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < N - 1; i++) {
            double factor;
            if (m < i + 1) {
                factor = -1.0;
            } else if (m == i + 1) {
                factor = (double)m;
            } else {
                factor = 0.0;
            }
            int k = ((i + 1) * (i + 2)) / 2;
            sum += factor / sqrt((double)k) * d[N * (N - 1) + i];
        }
        return 1.0 / N + 0.5 * sum;
    }
}

/* ======================================================================
 * lindblad_cvr_rodr<N>::update (algo_lindblad_cvr_rodr.hpp:234-268),
 * reference form.  A2 is never materialised: row j of A2 is assembled
 * entry by entry in the reference's accumulation order and consumed
 * immediately, which yields the same floating-point values as building
 * the whole matrix first.
 * ==================================================================== */
template<int N>
struct algo_rodr {
    static constexpr int K = N * N - 1;
    static constexpr int B = K / 2;
    static constexpr int DC = K;
    static constexpr int STATE = K;
    struct state {
        double v[STATE];
    };

    static constexpr int OFF_M = 5;
    static constexpr int OFF_U = OFF_M + K * K;
    static constexpr int OFF_U2 = OFF_U + K * K;
    static constexpr int OFF_A1 = OFF_U2 + K * K;
    static constexpr int OFF_V = OFF_A1 + K * K;
    static constexpr int OFF_DEQ = OFF_V + K;
    static constexpr int OFF_DIN = OFF_DEQ + K;
    static constexpr int OFF_TH = OFF_DIN + K;
    static constexpr int OFF_C1 = OFF_TH + B;
    static constexpr int OFF_C2 = OFF_C1 + B * K * K;

    __device__ static void update(const double* q, state& s, double e,
                                  double& p)
    {
        if (q[0] == 0.0) {
            return;
        }
        const double M_CP = q[1];
        const double d_t = q[2];
        const double* M = q + OFF_M;
        const double* A1 = q + OFF_A1;
        const double* v = q + OFF_V;
        const double* d_eq = q + OFF_DEQ;
        const double* d_in = q + OFF_DIN;
        double* d = s.v;

        /* d1 = A1 (d + d_in) - d_in */
        double tmp[K], d1[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            tmp[k] = d[k] + d_in[k];
        }
#pragma unroll
        for (int i = 0; i < K; i++) {
            double acc = A1[i * K] * tmp[0];
#pragma unroll
            for (int k = 1; k < K; k++) {
                acc = acc + A1[i * K + k] * tmp[k];
            }
            d1[i] = acc - d_in[i];
        }

        /* d2 = A2 d1 */
        double d2[K];
        if constexpr (N == 2) {
            const double* U = q + OFF_U;
            const double* U2 = q + OFF_U2;
            const double th1 = q[3];
            double x = th1 * e * d_t;
            double sa = sin(x) / th1;
            double va = (1 - cos(x)) / (th1 * th1);
#pragma unroll
            for (int i = 0; i < K; i++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    double a2 = sa * U[i * K + k] + va * U2[i * K + k] +
                        ((i == k) ? 1.0 : 0.0);
                    acc = (k == 0) ? a2 * d1[0] : acc + a2 * d1[k];
                }
                d2[i] = acc;
            }
        } else {
            const int nth = (int)q[4];
            const double* th = q + OFF_TH;
            double sv[B], vv[B];
#pragma unroll
            for (int b = 0; b < B; b++) {
                sv[b] = 0.0;
                vv[b] = 0.0;
                if (b < nth) {
                    double x = th[b] * e * d_t;
                    sv[b] = sin(x);
                    vv[b] = 1 - cos(x);
                }
            }
            const double* C1 = q + OFF_C1;
            const double* C2 = q + OFF_C2;
            /* rows: unrolled for small K; for large K one row per iteration
             * with the result routed through a compare-select chain so that
             * no register array is indexed dynamically */
            constexpr int ROW_UNROLL = (K <= 8) ? K : 1;
#pragma unroll ROW_UNROLL
            for (int i = 0; i < K; i++) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    double a2 = (i == k) ? 1.0 : 0.0;
#pragma unroll
                    for (int b = 0; b < B; b++) {
                        if (b < nth) {
                            a2 = a2 +
                                (sv[b] * C1[(b * K + i) * K + k] +
                                 vv[b] * C2[(b * K + i) * K + k]);
                        }
                    }
                    acc = (k == 0) ? a2 * d1[0] : acc + a2 * d1[k];
                }
#pragma unroll
                for (int kk = 0; kk < K; kk++) {
                    if (kk == i) {
                        d2[kk] = acc;
                    }
                }
            }
        }

        /* d = A1 (d2 + d_in) - d_in */
#pragma unroll
        for (int k = 0; k < K; k++) {
            tmp[k] = d2[k] + d_in[k];
        }
#pragma unroll
        for (int i = 0; i < K; i++) {
            double acc = A1[i * K] * tmp[0];
#pragma unroll
            for (int k = 1; k < K; k++) {
                acc = acc + A1[i * K + k] * tmp[k];
            }
            d[i] = acc - d_in[i];
        }

        /* p = M_CP v^T (M d + d_eq) */
        double pacc = 0.0;
#pragma unroll
        for (int i = 0; i < K; i++) {
            double acc = M[i * K] * d[0];
#pragma unroll
            for (int k = 1; k < K; k++) {
                acc = acc + M[i * K + k] * d[k];
            }
            double w = acc + d_eq[i];
            double t = (M_CP * v[i]) * w;
            pacc = (i == 0) ? t : pacc + t;
        }
        p = pacc;
    }

    /* calc_inversion (algo_lindblad_cvr_rodr.hpp:274-277) */
    __device__ static double inversion(const double*, const state& s)
    {
        return s.v[N * (N - 1)];
    }
    __device__ static double population(const double*, const state& s, int m)
    {
        return cv_population<N>(s.v, m);
    }
};

/* ---- sin/cos for the rotated-basis form --------------------------------
 * Taylor series of sin and cos about 0, evaluated in Horner form in x^2,
 * for |x| <= 0.5 (truncation error below 3e-20 relative); larger
 * arguments use the library sincos. */
__device__ __forceinline__ void zsincos(double x, double& s, double& c)
{
    if (fabs(x) <= 0.5) {
        double x2 = x * x;
        /* sin: x (1 - x^2/3! + x^4/5! - ... - x^14/15!) */
        double ps = -7.6471637318198164759e-13; /* -1/15! */
        ps = fma(ps, x2, 1.6059043836821614599e-10);  /* 1/13! */
        ps = fma(ps, x2, -2.5052108385441718775e-08); /* -1/11! */
        ps = fma(ps, x2, 2.7557319223985890653e-06);  /* 1/9! */
        ps = fma(ps, x2, -1.9841269841269841270e-04); /* -1/7! */
        ps = fma(ps, x2, 8.3333333333333333333e-03);  /* 1/5! */
        ps = fma(ps, x2, -1.6666666666666666667e-01); /* -1/3! */
        s = fma(ps * x2, x, x);
        /* cos: 1 - x^2/2! + x^4/4! - ... + x^16/16! */
        double pc = 4.7794773323873852974e-14;         /* 1/16! */
        pc = fma(pc, x2, -1.1470745597729724714e-11);  /* -1/14! */
        pc = fma(pc, x2, 2.0876756987868098979e-09);   /* 1/12! */
        pc = fma(pc, x2, -2.7557319223985890653e-07);  /* -1/10! */
        pc = fma(pc, x2, 2.4801587301587301587e-05);   /* 1/8! */
        pc = fma(pc, x2, -1.3888888888888888889e-03);  /* -1/6! */
        pc = fma(pc, x2, 4.1666666666666666667e-02);   /* 1/4! */
        pc = fma(pc, x2, -0.5);                        /* -1/2! */
        c = fma(pc, x2, 1.0);
    } else {
        sincos(x, &s, &c);
    }
}

/* ======================================================================
 * Rodrigues in the rotated basis (z-form).  Mathematically identical to
 * algo_rodr (A2(e) = Q R(e) Q^T, Strang halves merged into W); see the
 * zform blob description in mbsolve_hip.h.  State: y (K values).
 * ==================================================================== */
template<int N>
struct algo_zform {
    static constexpr int K = N * N - 1;
    static constexpr int B = K / 2;
    static constexpr int DC = K;
    static constexpr int STATE = K;
    struct state {
        double v[STATE];
    };

    static constexpr int OFF_TH = 3;
    static constexpr int OFF_W = OFF_TH + B;
    static constexpr int OFF_C = OFF_W + K * K;
    static constexpr int OFF_A = OFF_C + K;
    static constexpr int OFF_P = OFF_A + K;
    static constexpr int OFF_O = OFF_P + K * K;
    static constexpr int OFF_QT = OFF_O + K;
    static constexpr int OFF_QTA1 = OFF_QT + K * K;
    static constexpr int OFF_DIN = OFF_QTA1 + K * K;

    __device__ static void update(const double* q, state& s, double e,
                                  double& p)
    {
        if (q[0] == 0.0) {
            return;
        }
        const int nth = (int)q[1];
        const double* th = q + OFF_TH;
        const double* W = q + OFF_W;
        const double* c = q + OFF_C;
        const double* a = q + OFF_A;
        double* y = s.v;

        /* r = R(e) y */
        double r[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            r[k] = y[k];
        }
#pragma unroll
        for (int b = 0; b < B; b++) {
            if (b < nth) {
                double sn, cs;
                zsincos(th[b] * e, sn, cs);
                double y0 = y[2 * b], y1 = y[2 * b + 1];
                r[2 * b] = cs * y0 - sn * y1;
                r[2 * b + 1] = fma(sn, y0, cs * y1);
            }
        }
        /* p = a . r + b */
        double pacc = q[2];
#pragma unroll
        for (int k = 0; k < K; k++) {
            pacc = fma(a[k], r[k], pacc);
        }
        p = pacc;
        /* y' = W r + c */
#pragma unroll
        for (int i = 0; i < K; i++) {
            double acc = c[i];
#pragma unroll
            for (int k = 0; k < K; k++) {
                acc = fma(W[i * K + k], r[k], acc);
            }
            y[i] = acc;
        }
    }

    /* component k of the reference coherence vector after the step */
    __device__ static double dcomp(const double* q, const state& s, int k)
    {
        const double* P = q + OFF_P;
        const double* o = q + OFF_O;
        double acc = o[k];
#pragma unroll
        for (int j = 0; j < K; j++) {
            acc = fma(P[k * K + j], s.v[j], acc);
        }
        return acc;
    }
    __device__ static double inversion(const double* q, const state& s)
    {
        if (q[0] == 0.0) {
            return 0.0;
        }
        return dcomp(q, s, N * (N - 1));
    }
    __device__ static double population(const double* q, const state& s,
                                        int m)
    {
        double d[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            d[k] = 0.0;
        }
        if (q[0] != 0.0) {
#pragma unroll
            for (int k = N * (N - 1); k < K; k++) {
                d[k] = dcomp(q, s, k);
            }
        }
        return cv_population<N>(d, m);
    }
};

} /* namespace mbh */

#endif
