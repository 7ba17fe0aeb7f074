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
#include <type_traits>
#include <stdint.h>

#include "../../include/mbsolve_hip.h"

namespace mbh {

/* Material constants are read through the constant address space: with a
 * wave-uniform address these become scalar (s_load) loads even though the
 * kernel stores to global memory. */
typedef const __attribute__((address_space(4))) double* cdp;

/* ---- complex helpers (naive formulas, the libstdc++ fast path) --------- */
struct cplx {
    double re, im;
};
typedef const __attribute__((address_space(4))) cplx* ccp;

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
    static constexpr int STAGE = 0;
    static constexpr bool SHIFT = false;
    static constexpr bool FAST_FDTD = false;
    static constexpr int DC = 0; /* stored density components */
    static constexpr int STATE = 1;
    struct state {
        double v[STATE];
    };
    __device__ static bool has_qm(cdp) { return false; }
    __device__ static void update(cdp, state&, double, double&) {}
    template<int C>
    __device__ static void update_multi(cdp, state*, const double*, double*)
    {
    }
    __device__ static double inversion(cdp, const state&) { return 0.0; }
    __device__ static double population(cdp, const state&, int)
    {
        return 0.0;
    }
};

/* noop with the folded FDTD constants (FP32-storage form: not bound to the
 * reference's operation order) */
struct algo_noop_fast : algo_noop {
    static constexpr bool FAST_FDTD = true;
};
/* the same, for states whose cells with H held at zero start at zero: there
 * fc' = 0 keeps them at zero, so h' = h + fc' (e_c - e_l) needs no factor */
struct algo_noop_hz0 : algo_noop_fast {};

/* ======================================================================
 * lindblad_reg_cayley<N>::update (algo_lindblad_reg_cayley.hpp:228-257)
 *
 * Blob layout: see MBH_CAYLEY_STRIDE in mbsolve_hip.h.  State: the full
 * N x N complex density matrix, row-major, (re, im) interleaved, exactly
 * as the reference keeps it (no Hermitian symmetrisation).
 * ==================================================================== */
template<int N>
struct algo_cayley {
    static constexpr int STAGE = 0;
    static constexpr bool SHIFT = false;
    static constexpr bool FAST_FDTD = false;
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
    __device__ static void dissipate(cdp q, const cplx* d, cplx* o)
    {
        cdp D1 = q + OFF_D1;
        cdp D2 = q + OFF_D2;
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

    __device__ static bool has_qm(cdp q) { return q[0] != 0.0; }

    /* branch-free body; callers check has_qm */
    __device__ static void update(cdp q, state& s, double e, double& p)
    {
        ccp H0 = (ccp)(q + OFF_H0);
        ccp mu = (ccp)(q + OFF_MU);
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
        cplx Ai[N * N], U2[N * N], Hm[N * N];
        cinverse<N>(A, Ai);
        cmatmul<N>(Ai, B, U2);
#pragma unroll
        for (int i = 0; i < N * N; i++) {
            Hm[i] = cplx{ H0[i].re, H0[i].im };
        }

        /* dissipation half step, unitary full step, dissipation half step */
        cplx d1[N * N], t[N * N];
        dissipate(q, d, d1);
        cmatmul<N>(U2, d1, t);
        cmatmul_adj<N>(t, U2, d1);
        dissipate(q, d1, d);

        /* polarization derivative: rhs = -i/hbar [H0, d] + D(d);
         * p = n Re tr(rhs mu) */
        cdp r1 = q + OFF_d1;
        cdp r2 = q + OFF_d2;
        cplx c1[N * N], c2[N * N];
        cmatmul<N>(Hm, d, c1);
        cmatmul<N>(d, Hm, c2);
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

    template<int C>
    __device__ static void update_multi(cdp q, state* s, const double* e,
                                        double* p)
    {
#pragma unroll
        for (int c = 0; c < C; c++) {
            update(q, s[c], e[c], p[c]);
        }
    }

    /* calc_inversion / calc_population (algo_lindblad_reg_cayley.hpp:259-267) */
    __device__ static double inversion(cdp, const state& s)
    {
        return s.v[2 * (N + 1)] - s.v[0];
    }
    __device__ static double population(cdp, const state& s, int m)
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

/* ======================================================================
 * Two-level regularized Cayley in closed Bloch form.  Mathematically the
 * update of algo_cayley<2> (algo_lindblad_reg_cayley.hpp:206-257) for a
 * Hermitian state, rearranged.  With H = h0 I + h.sigma and
 * alpha = -(prop_im / 2) (h0, h_vec), the Cayley factor is proportional to
 *   W = s I - i beta.sigma,  beta = 2 alpha_vec,  s = 1 + alpha0^2 - |alpha|^2
 * so U rho U^+ rotates m = (Re rho01, -Im rho01, (rho00 - rho11) / 2) by
 *   m' = m + lambda (s beta x m + beta x (beta x m)),
 *   lambda = 2 / (s^2 + |beta|^2)
 * and keeps the trace tau.  The dissipation D scales the coherence by g and
 * maps the populations with D1.
 *
 * Carried state: sigma = the state between the unitary step and the second
 * dissipation half step, as (tau, m_x, m_y, m_z).  The reference state
 * after step n is rho_n = D(sigma_n), and the next step needs
 * D(rho_n) = D^2(sigma_n): one merged dissipation per step.  p and the
 * populations are fixed linear forms of sigma.
 *
 * Blob (BLOCH2_STRIDE doubles):
 *   [0] has_qm  [1..3] beta at e = 0  [4..6] d beta / d e
 *   [7] alpha0 at e = 0  [8] d alpha0 / d e  [9] g^2
 *   [10..13] F: D1^2 on (tau, m_z), row-major
 *   [14..17] p = P . (tau, m_x, m_y, m_z)
 *   [18..21] populations: rho00 = R00 tau + R01 m_z, rho11 = R10 tau + R11 m_z
 *   [22..25] D1 (row-major)  [26] g  [27] 1 + alpha0^2 at e = 0
 * ==================================================================== */
template<bool REAL>
struct algo_bloch2_t {
    static constexpr int STAGE = 0;
    static constexpr bool SHIFT = false;
    static constexpr int DC = 4;
    static constexpr int STATE = 4;
    static constexpr bool FAST_FDTD = true;
    struct state {
        double v[STATE];
    };

    __device__ static bool has_qm(cdp q) { return q[0] != 0.0; }

    /* 2 / n2 for n2 near 1: hardware reciprocal and two Newton steps */
    __device__ static double two_over(double n2)
    {
        double r = __builtin_amdgcn_rcp(n2);
        double t = fma(-n2, r, 1.0);
        r = fma(r, t, r);
        t = fma(-n2, r, 1.0);
        r = fma(r, t, r);
        return 2.0 * r;
    }

    template<int C>
    __device__ static void update_multi(cdp q, state* s, const double* e,
                                        double* p)
    {
        if constexpr (REAL) {
            update_real<C>(q, s, e, p);
            return;
        }
        const double bcx = q[1], bcy = q[2], bcz = q[3];
        const double bex = q[4], bey = q[5], bez = q[6];
        const double a0c = q[7], a0e = q[8], g2 = q[9];
        const double F00 = q[10], F01 = q[11], F10 = q[12], F11 = q[13];
        const double P0 = q[14], P1 = q[15], P2 = q[16], P3 = q[17];
#pragma unroll
        for (int c = 0; c < C; c++) {
            double* v = s[c].v;
            /* second half step of the previous step and first half step
             * of this one */
            const double tr = fma(F01, v[3], F00 * v[0]);
            const double mz = fma(F11, v[3], F10 * v[0]);
            const double mx = g2 * v[1];
            const double my = g2 * v[2];
            /* Cayley rotation at field e */
            const double bx = fma(e[c], bex, bcx);
            const double by = fma(e[c], bey, bcy);
            const double bz = fma(e[c], bez, bcz);
            const double a0 = fma(e[c], a0e, a0c);
            const double r2 = fma(bz, bz, fma(by, by, bx * bx));
            const double sc = fma(-0.25, r2, fma(a0, a0, 1.0));
            const double lam = two_over(fma(sc, sc, r2));
            const double c1x = fma(by, mz, -(bz * my));
            const double c1y = fma(bz, mx, -(bx * mz));
            const double c1z = fma(bx, my, -(by * mx));
            const double bm = fma(bz, mz, fma(by, my, bx * mx));
            const double c2x = fma(bx, bm, -(mx * r2));
            const double c2y = fma(by, bm, -(my * r2));
            const double c2z = fma(bz, bm, -(mz * r2));
            v[0] = tr;
            v[1] = fma(lam, fma(sc, c1x, c2x), mx);
            v[2] = fma(lam, fma(sc, c1y, c2y), my);
            v[3] = fma(lam, fma(sc, c1z, c2z), mz);
            p[c] = fma(P3, v[3], fma(P2, v[2], fma(P1, v[1], P0 * tr)));
        }
    }

    /* The same step for a material whose beta_y, d alpha0 / d e and the
     * p coefficients of tau and m_z are exactly zero (a diagonal
     * Hamiltonian and a real off-diagonal dipole, as qm_desc_2lvl builds
     * them, qm_description.hpp:338-432): those terms are dropped and
     * 1 + alpha0^2 is the blob's [27].  Each dropped term is an exact zero
     * in the general form, so the results are the same (up to the sign of
     * zeros). */
    template<int C>
    __device__ static void update_real(cdp q, state* s, const double* e,
                                       double* p)
    {
        const double bcx = q[1], bcz = q[3];
        const double bex = q[4], bez = q[6];
        const double A0 = q[27], g2 = q[9];
        const double F00 = q[10], F01 = q[11], F10 = q[12], F11 = q[13];
        const double P1 = q[15], P2 = q[16];
#pragma unroll
        for (int c = 0; c < C; c++) {
            double* v = s[c].v;
            const double tr = fma(F01, v[3], F00 * v[0]);
            const double mz = fma(F11, v[3], F10 * v[0]);
            const double mx = g2 * v[1];
            const double my = g2 * v[2];
            const double bx = fma(e[c], bex, bcx);
            const double bz = fma(e[c], bez, bcz);
            const double r2 = fma(bz, bz, bx * bx);
            const double sc = fma(-0.25, r2, A0);
            const double lam = two_over(fma(sc, sc, r2));
            const double c1x = -(bz * my);
            const double c1y = fma(bz, mx, -(bx * mz));
            const double c1z = bx * my;
            const double bm = fma(bz, mz, bx * mx);
            const double c2x = fma(bx, bm, -(mx * r2));
            const double c2y = -(my * r2);
            const double c2z = fma(bz, bm, -(mz * r2));
            v[0] = tr;
            v[1] = fma(lam, fma(sc, c1x, c2x), mx);
            v[2] = fma(lam, fma(sc, c1y, c2y), my);
            v[3] = fma(lam, fma(sc, c1z, c2z), mz);
            p[c] = fma(P2, v[2], P1 * v[1]);
        }
    }

    __device__ static void update(cdp q, state& s, double e, double& p)
    {
        update_multi<1>(q, &s, &e, &p);
    }

    __device__ static double population(cdp q, const state& s, int m)
    {
        if (m == 0) {
            return fma(q[19], s.v[3], q[18] * s.v[0]);
        }
        if (m == 1) {
            return fma(q[21], s.v[3], q[20] * s.v[0]);
        }
        return 0.0;
    }
    __device__ static double inversion(cdp q, const state& s)
    {
        return population(q, s, 1) - population(q, s, 0);
    }
};
using algo_bloch2 = algo_bloch2_t<false>;
using algo_bloch2r = algo_bloch2_t<true>;

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
    static constexpr int STAGE = 0;
    static constexpr bool SHIFT = false;
    static constexpr bool FAST_FDTD = false;
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

    __device__ static bool has_qm(cdp q) { return q[0] != 0.0; }

    /* callers check has_qm */
    __device__ static void update(cdp q, state& s, double e, double& p)
    {
        const double M_CP = q[1];
        const double d_t = q[2];
        cdp M = q + OFF_M;
        cdp A1 = q + OFF_A1;
        cdp v = q + OFF_V;
        cdp d_eq = q + OFF_DEQ;
        cdp d_in = q + OFF_DIN;
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
            cdp U = q + OFF_U;
            cdp U2 = q + OFF_U2;
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
            /* planes b >= n_theta have theta = 0 and zero C1, C2 in the
             * blob: their terms are exact zeros and leave A2 unchanged, so
             * every plane is evaluated without a data-dependent branch */
            cdp th = q + OFF_TH;
            double sv[B], vv[B];
#pragma unroll
            for (int b = 0; b < B; b++) {
                double x = th[b] * e * d_t;
                sv[b] = sin(x);
                vv[b] = 1 - cos(x);
            }
            cdp C1 = q + OFF_C1;
            cdp C2 = q + OFF_C2;
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
                        a2 = a2 +
                            (sv[b] * C1[(b * K + i) * K + k] +
                             vv[b] * C2[(b * K + i) * K + k]);
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

    template<int C>
    __device__ static void update_multi(cdp q, state* s, const double* e,
                                        double* p)
    {
#pragma unroll
        for (int c = 0; c < C; c++) {
            update(q, s[c], e[c], p[c]);
        }
    }

    /* calc_inversion (algo_lindblad_cvr_rodr.hpp:274-277) */
    __device__ static double inversion(cdp, const state& s)
    {
        return s.v[N * (N - 1)];
    }
    __device__ static double population(cdp, const state& s, int m)
    {
        return cv_population<N>(s.v, m);
    }
};

/* ---- sin/cos for the rotated-basis form --------------------------------
 * Taylor series of sin (to x^13) and cos (to x^14) about 0 in Horner form
 * in x^2; for |x| <= 0.25 the truncation error is below 1e-21 relative. */
constexpr double TAYLOR_MAX = 0.25;

__device__ __forceinline__ void taylor_sincos(double x, double& s, double& c)
{
    const double x2 = x * x;
    double ps = 1.6059043836821614599e-10;        /* 1/13! */
    ps = fma(ps, x2, -2.5052108385441718775e-08); /* -1/11! */
    ps = fma(ps, x2, 2.7557319223985890653e-06);  /* 1/9! */
    ps = fma(ps, x2, -1.9841269841269841270e-04); /* -1/7! */
    ps = fma(ps, x2, 8.3333333333333333333e-03);  /* 1/5! */
    ps = fma(ps, x2, -1.6666666666666666667e-01); /* -1/3! */
    s = fma(ps * x2, x, x);
    double pc = -1.1470745597729724714e-11;       /* -1/14! */
    pc = fma(pc, x2, 2.0876756987868098979e-09);  /* 1/12! */
    pc = fma(pc, x2, -2.7557319223985890653e-07); /* -1/10! */
    pc = fma(pc, x2, 2.4801587301587301587e-05);  /* 1/8! */
    pc = fma(pc, x2, -1.3888888888888888889e-03); /* -1/6! */
    pc = fma(pc, x2, 4.1666666666666666667e-02);  /* 1/4! */
    pc = fma(pc, x2, -0.5);                       /* -1/2! */
    c = fma(pc, x2, 1.0);
}


/* sin and cos of a small angle x, by wave-uniform tier of max |x|:
 * TIER 0: |x| <= 2^-13, x - x^3/6 and 1 - x^2/2 (relative truncation
 * error below 1e-17); TIER 1: |x| <= 2^-10, Taylor to x^5 / x^4 (below
 * 2e-21 / 3e-19); TIER 2: |x| <= 0.25, Taylor to x^13 / x^14; TIER 3:
 * library. */
template<int TIER>
__device__ __forceinline__ void tier_sincos(double x, double& s, double& c)
{
    if constexpr (TIER == 0) {
        const double x2 = x * x;
        s = fma(x * x2, -1.6666666666666666667e-01, x);
        c = fma(x2, -0.5, 1.0);
    } else if constexpr (TIER == 1) {
        const double x2 = x * x;
        s = fma(x * x2, fma(x2, 8.3333333333333333333e-03,
                            -1.6666666666666666667e-01), x);
        c = fma(x2, fma(x2, 4.1666666666666666667e-02, -0.5), 1.0);
    } else if constexpr (TIER == 2) {
        taylor_sincos(x, s, c);
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
    static constexpr bool FAST_FDTD = true;
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
    /* appended to the device copy of the blob (not part of the C ABI
     * layout): y* = -Qt d_in, the fixed point of the step at zero field
     * (d = -d_in, the relaxation equilibrium).  FP32 storage holds
     * y - y*, so its rounding scales with the distance from equilibrium
     * instead of with |y*|. */
    static constexpr int OFF_YEQ = OFF_DIN + K;
    static constexpr bool SHIFT = true;

    __device__ static bool has_qm(cdp q) { return q[0] != 0.0; }

    /* Small blobs (K <= 8) are staged in LDS by the tile kernel: W, c, a
     * and theta are then read from there every step instead of through
     * the scalar cache (their 600 bytes do not fit the SGPRs). */
    static constexpr int STAGE = (K <= 8) ? OFF_A + K : 0;

    template<int C>
    __device__ static void update_multi(cdp q, state* s, const double* e,
                                        double* p)
    {
        update_multi_p<C>(q, s, e, p);
    }

    /* C cells of one material: straight-line code with the cell index
     * innermost, so the independent chains interleave.  Planes beyond
     * n_theta have angle 0 (identity rotation, exact). */
    template<int C, typename P>
    __device__ static void update_multi_p(P q, state* s, const double* e,
                                          double* p)
    {
        const P th = q + OFF_TH;
        const P W = q + OFF_W;
        const P cv = q + OFF_C;
        const P a = q + OFF_A;

        double sn[C][B], cs[C][B];
        bool t0 = true, t1 = true, t2 = true;
#pragma unroll
        for (int c = 0; c < C; c++) {
#pragma unroll
            for (int b = 0; b < B; b++) {
                const double x = th[b] * e[c];
                const double ax = fabs(x);
                sn[c][b] = x;
                t0 = t0 && (ax <= 0x1p-13);
                t1 = t1 && (ax <= 0x1p-10);
                t2 = t2 && (ax <= TAYLOR_MAX);
            }
        }
        /* wave-uniform tier: the cheapest series accurate for every
         * angle of the wavefront */
        auto all_sincos = [&](auto tier) {
            constexpr int T = decltype(tier)::value;
#pragma unroll
            for (int c = 0; c < C; c++) {
#pragma unroll
                for (int b = 0; b < B; b++) {
                    tier_sincos<T>(sn[c][b], sn[c][b], cs[c][b]);
                }
            }
        };
        if (__all(t0)) {
            all_sincos(std::integral_constant<int, 0>());
        } else if (__all(t1)) {
            all_sincos(std::integral_constant<int, 1>());
        } else if (__all(t2)) {
            all_sincos(std::integral_constant<int, 2>());
        } else {
            all_sincos(std::integral_constant<int, 3>());
        }

        /* r = R(e) y */
        double r[C][K];
#pragma unroll
        for (int c = 0; c < C; c++) {
#pragma unroll
            for (int k = 2 * B; k < K; k++) {
                r[c][k] = s[c].v[k];
            }
#pragma unroll
            for (int b = 0; b < B; b++) {
                const double y0 = s[c].v[2 * b], y1 = s[c].v[2 * b + 1];
                r[c][2 * b] = fma(cs[c][b], y0, -(sn[c][b] * y1));
                r[c][2 * b + 1] = fma(sn[c][b], y0, cs[c][b] * y1);
            }
        }
        /* p = a . r + b, two interleaved partial sums (shorter chain on
         * the step's critical path) */
        double pa[C], pb[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            pa[c] = q[2];
            pb[c] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const double ak = a[k];
#pragma unroll
            for (int c = 0; c < C; c++) {
                if (k & 1) {
                    pb[c] = fma(ak, r[c][k], pb[c]);
                } else {
                    pa[c] = fma(ak, r[c][k], pa[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            p[c] = pa[c] + pb[c];
        }
        /* y' = W r + c.  With one cell per lane and a small blob (K <= 8,
         * read from LDS) the rows are accumulated side by side (k outer):
         * K independent FMA chains instead of one chain per row after
         * another, at the price of holding W in VGPRs during the product.
         * Larger K (the 35 components of six levels) interleaves two rows:
         * half the chain latency without spilling to AGPRs (three rows
         * do).  Several cells per lane already interleave C chains. */
        constexpr int RG = C != 1 ? 1 : (K <= 8 ? K : 2);
#pragma unroll
        for (int i0 = 0; i0 < K; i0 += RG) {
            double acc[C][RG];
#pragma unroll
            for (int g = 0; g < RG; g++) {
                if (i0 + g < K) {
                    const double ci = cv[i0 + g];
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        acc[c][g] = ci;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
#pragma unroll
                for (int g = 0; g < RG; g++) {
                    if (i0 + g < K) {
                        const double w = W[(i0 + g) * K + k];
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            acc[c][g] = fma(w, r[c][k], acc[c][g]);
                        }
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < RG; g++) {
                if (i0 + g < K) {
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        s[c].v[i0 + g] = acc[c][g];
                    }
                }
            }
        }
    }

    __device__ static void update(cdp q, state& s, double e, double& p)
    {
        update_multi<1>(q, &s, &e, &p);
    }

    /* component k of the reference coherence vector after the step */
    __device__ static double dcomp(cdp q, const state& s, int k)
    {
        cdp P = q + OFF_P;
        cdp o = q + OFF_O;
        double acc = o[k];
#pragma unroll
        for (int j = 0; j < K; j++) {
            acc = fma(P[k * K + j], s.v[j], acc);
        }
        return acc;
    }
    __device__ static double inversion(cdp q, const state& s)
    {
        if (q[0] == 0.0) {
            return 0.0;
        }
        return dcomp(q, s, N * (N - 1));
    }
    __device__ static double population(cdp q, const state& s, int m)
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
