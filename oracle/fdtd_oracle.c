/*
 * CPU FP64 ORACLE -- test infrastructure only.
 *
 * Restates the reference's plain FDTD time loop
 *   solver-cpu/include/mbsolve/solver-cpu/solver_cpu_fdtd.tpp:223-338
 * its red (redundant-halo, density-first) time loop
 *   solver-cpu/include/mbsolve/solver-cpu/solver_cpu_fdtd_red.tpp:351-567
 * and its per-cell density updates
 *   lindblad_noop       mbsolve-lib/include/mbsolve/lib/internal/
 *                       algo_lindblad_noop.hpp:53-57
 *   lindblad_reg_cayley algo_lindblad_reg_cayley.hpp:206-274
 *   lindblad_cvr_rodr   algo_lindblad_cvr_rodr.hpp:234-299
 * on the constants of an mbh_problem (include/mbsolve_hip.h), with OpenMP
 * over grid points like the reference.  Dot products accumulate in
 * ascending index order starting from the first product and no
 * multiply-add is fused (build with -ffp-contract=off); with the same
 * constants the GPU FP64 form reproduces these results bit for bit except
 * where libm and the device math library round sin/cos differently.
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * use this file; the product never links it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include "../include/mbsolve_hip.h"

#define MAXN 6
#define MAXK 35

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b)
{
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}
static inline cplx cadd(cplx a, cplx b)
{
    cplx r = { a.re + b.re, a.im + b.im };
    return r;
}
static inline cplx csub(cplx a, cplx b)
{
    cplx r = { a.re - b.re, a.im - b.im };
    return r;
}
static inline cplx cinv(cplx z)
{
    double s = z.re * z.re + z.im * z.im;
    cplx r = { z.re / s, -z.im / s };
    return r;
}

static void cmatmul(int N, const cplx* A, const cplx* B, cplx* C)
{
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            cplx acc = cmul(A[i * N], B[j]);
            for (int k = 1; k < N; k++) {
                acc = cadd(acc, cmul(A[i * N + k], B[k * N + j]));
            }
            C[i * N + j] = acc;
        }
    }
}

/* C = A B^dagger */
static void cmatmul_adj(int N, const cplx* A, const cplx* B, cplx* C)
{
    for (int i = 0; i < N; i++) {
        for (int j = 0; j < N; j++) {
            cplx b0 = { B[j * N].re, -B[j * N].im };
            cplx acc = cmul(A[i * N], b0);
            for (int k = 1; k < N; k++) {
                cplx bk = { B[j * N + k].re, -B[j * N + k].im };
                acc = cadd(acc, cmul(A[i * N + k], bk));
            }
            C[i * N + j] = acc;
        }
    }
}

/* Inverse: 2x2 cofactor formula (as Eigen's size-2 inverse), otherwise
 * Gauss-Jordan with partial pivoting on |a|^2. */
static void cinverse(int N, const cplx* A, cplx* R)
{
    if (N == 2) {
        cplx det = csub(cmul(A[0], A[3]), cmul(A[1], A[2]));
        cplx id = cinv(det);
        cplx m1 = { -A[1].re, -A[1].im }, m2 = { -A[2].re, -A[2].im };
        R[0] = cmul(A[3], id);
        R[1] = cmul(m1, id);
        R[2] = cmul(m2, id);
        R[3] = cmul(A[0], id);
        return;
    }
    cplx W[MAXN * MAXN];
    for (int i = 0; i < N * N; i++) {
        W[i] = A[i];
        R[i].re = (i / N == i % N) ? 1.0 : 0.0;
        R[i].im = 0.0;
    }
    for (int col = 0; col < N; col++) {
        int piv = col;
        double best = W[col * N + col].re * W[col * N + col].re +
            W[col * N + col].im * W[col * N + col].im;
        for (int r = col + 1; r < N; r++) {
            double v = W[r * N + col].re * W[r * N + col].re +
                W[r * N + col].im * W[r * N + col].im;
            if (v > best) {
                best = v;
                piv = r;
            }
        }
        if (piv != col) {
            for (int k = 0; k < N; k++) {
                cplx t = W[col * N + k];
                W[col * N + k] = W[piv * N + k];
                W[piv * N + k] = t;
                t = R[col * N + k];
                R[col * N + k] = R[piv * N + k];
                R[piv * N + k] = t;
            }
        }
        cplx ip = cinv(W[col * N + col]);
        for (int k = 0; k < N; k++) {
            W[col * N + k] = cmul(W[col * N + k], ip);
            R[col * N + k] = cmul(R[col * N + k], ip);
        }
        for (int r = 0; r < N; r++) {
            if (r != col) {
                cplx f = W[r * N + col];
                for (int k = 0; k < N; k++) {
                    W[r * N + k] = csub(W[r * N + k], cmul(f, W[col * N + k]));
                    R[r * N + k] = csub(R[r * N + k], cmul(f, R[col * N + k]));
                }
            }
        }
    }
}

/* ---- lindblad_reg_cayley ----------------------------------------------- */
/* propagate_dissipation (algo_lindblad_reg_cayley.hpp:206-217) */
static void cay_dissipate(int N, const double* D1, const double* D2,
                          const cplx* d, cplx* o)
{
    for (int i = 0; i < N * N; i++) {
        o[i].re = D2[i] * d[i].re;
        o[i].im = D2[i] * d[i].im;
    }
    for (int i = 0; i < N; i++) {
        cplx acc = { D1[i * N] * d[0].re, D1[i * N] * d[0].im };
        for (int k = 1; k < N; k++) {
            acc.re = acc.re + D1[i * N + k] * d[k * N + k].re;
            acc.im = acc.im + D1[i * N + k] * d[k * N + k].im;
        }
        o[i * N + i] = acc;
    }
}

/* update (algo_lindblad_reg_cayley.hpp:228-257) */
static void cay_update(int N, const double* q, double* dv, double e,
                       double* p)
{
    if (q[0] == 0.0) {
        return;
    }
    const int NN = N * N;
    const cplx* H0 = (const cplx*)(q + 5);
    const cplx* mu = (const cplx*)(q + 5 + 2 * NN);
    const double* r1 = q + 5 + 4 * NN;
    const double* r2 = r1 + NN;
    const double* D1 = r2 + NN;
    const double* D2 = D1 + NN;
    const double prop_im = q[3], rhs_im = q[4];
    cplx* d = (cplx*)dv;

    /* H = H0 - e mu; M = -i/hbar dt H; A = I - M/2; B = I + M/2 */
    cplx A[MAXN * MAXN], B[MAXN * MAXN];
    for (int i = 0; i < NN; i++) {
        cplx h = { H0[i].re - e * mu[i].re, H0[i].im - e * mu[i].im };
        cplx m = { -(prop_im * h.im), prop_im * h.re };
        double di = (i / N == i % N) ? 1.0 : 0.0;
        A[i].re = di - 0.5 * m.re;
        A[i].im = -(0.5 * m.im);
        B[i].re = di + 0.5 * m.re;
        B[i].im = 0.5 * m.im;
    }
    /* Cayley transform (algo_lindblad_reg_cayley.hpp:219-225) */
    cplx Ai[MAXN * MAXN], U2[MAXN * MAXN];
    cinverse(N, A, Ai);
    cmatmul(N, Ai, B, U2);

    cplx d1[MAXN * MAXN], t[MAXN * MAXN];
    cay_dissipate(N, D1, D2, d, d1);
    cmatmul(N, U2, d1, t);
    cmatmul_adj(N, t, U2, d1);
    cay_dissipate(N, D1, D2, d1, d);

    /* rhs = -i/hbar (H0 d - d H0) + D(d);  p = n Re tr(rhs mu) */
    cplx c1[MAXN * MAXN], c2[MAXN * MAXN], rhs[MAXN * MAXN];
    cmatmul(N, H0, d, c1);
    cmatmul(N, d, H0, c2);
    for (int i = 0; i < NN; i++) {
        cplx c = csub(c1[i], c2[i]);
        cplx g = { -(rhs_im * c.im), rhs_im * c.re };
        rhs[i].re = r2[i] * d[i].re;
        rhs[i].im = r2[i] * d[i].im;
        if (i / N == i % N) {
            int ii = i / N;
            cplx acc = { r1[ii * N] * d[0].re, r1[ii * N] * d[0].im };
            for (int k = 1; k < N; k++) {
                acc.re = acc.re + r1[ii * N + k] * d[k * N + k].re;
                acc.im = acc.im + r1[ii * N + k] * d[k * N + k].im;
            }
            rhs[i] = acc;
        }
        rhs[i] = cadd(g, rhs[i]);
    }
    double tr = 0.0;
    for (int i = 0; i < N; i++) {
        double di = rhs[i * N].re * mu[i].re - rhs[i * N].im * mu[i].im;
        for (int k = 1; k < N; k++) {
            di = di +
                (rhs[i * N + k].re * mu[k * N + i].re -
                 rhs[i * N + k].im * mu[k * N + i].im);
        }
        tr = (i == 0) ? di : tr + di;
    }
    *p = q[1] * tr;
}

/* ---- lindblad_cvr_rodr ---------------------------------------------------- */
static void rodr_update(int N, const double* q, double* d, double e,
                        double* p)
{
    if (q[0] == 0.0) {
        return;
    }
    const int K = N * N - 1, B = K / 2, KK = K * K;
    const double M_CP = q[1], d_t = q[2];
    const double* M = q + 5;
    const double* U = M + KK;
    const double* U2 = U + KK;
    const double* A1 = U2 + KK;
    const double* v = A1 + KK;
    const double* d_eq = v + K;
    const double* d_in = d_eq + K;
    const double* th = d_in + K;
    const double* C1 = th + B;
    const double* C2 = C1 + B * KK;

    /* d1 = A1 (d + d_in) - d_in  (algo_lindblad_cvr_rodr.hpp:239) */
    double tmp[MAXK], d1[MAXK], d2[MAXK];
    for (int k = 0; k < K; k++) {
        tmp[k] = d[k] + d_in[k];
    }
    for (int i = 0; i < K; i++) {
        double acc = A1[i * K] * tmp[0];
        for (int k = 1; k < K; k++) {
            acc = acc + A1[i * K + k] * tmp[k];
        }
        d1[i] = acc - d_in[i];
    }

    /* d2 = A2 d1 with the Rodrigues propagator (:242-260) */
    if (N == 2) {
        const double th1 = q[3];
        double x = th1 * e * d_t;
        double sa = sin(x) / th1;
        double va = (1 - cos(x)) / (th1 * th1);
        for (int i = 0; i < K; i++) {
            double acc = 0.0;
            for (int k = 0; k < K; k++) {
                double a2 = sa * U[i * K + k] + va * U2[i * K + k] +
                    ((i == k) ? 1.0 : 0.0);
                acc = (k == 0) ? a2 * d1[0] : acc + a2 * d1[k];
            }
            d2[i] = acc;
        }
    } else {
        const int nth = (int)q[4];
        double sv[MAXK / 2], vv[MAXK / 2];
        for (int b = 0; b < nth; b++) {
            double x = th[b] * e * d_t;
            sv[b] = sin(x);
            vv[b] = 1 - cos(x);
        }
        for (int i = 0; i < K; i++) {
            double acc = 0.0;
            for (int k = 0; k < K; k++) {
                double a2 = (i == k) ? 1.0 : 0.0;
                for (int b = 0; b < nth; b++) {
                    a2 = a2 +
                        (sv[b] * C1[(b * K + i) * K + k] +
                         vv[b] * C2[(b * K + i) * K + k]);
                }
                acc = (k == 0) ? a2 * d1[0] : acc + a2 * d1[k];
            }
            d2[i] = acc;
        }
    }

    /* d = A1 (d2 + d_in) - d_in (:263) */
    for (int k = 0; k < K; k++) {
        tmp[k] = d2[k] + d_in[k];
    }
    for (int i = 0; i < K; i++) {
        double acc = A1[i * K] * tmp[0];
        for (int k = 1; k < K; k++) {
            acc = acc + A1[i * K + k] * tmp[k];
        }
        d[i] = acc - d_in[i];
    }

    /* p = M_CP v^T (M d + d_eq) (:266) */
    double pacc = 0.0;
    for (int i = 0; i < K; i++) {
        double acc = M[i * K] * d[0];
        for (int k = 1; k < K; k++) {
            acc = acc + M[i * K + k] * d[k];
        }
        double w = acc + d_eq[i];
        double t = (M_CP * v[i]) * w;
        pacc = (i == 0) ? t : pacc + t;
    }
    *p = pacc;
}

/* cv_representation::calc_population
 * (coherence_vector_representation.hpp:289-336) */
static double cv_population(int N, const double* d, int m)
{
    if (N == 2) {
        return (m == 0) ? 0.5 - 0.5 * d[2] : 0.5 + 0.5 * d[2];
    }
    if (N == 3) {
        const double s3 = sqrt(3.0);
        if (m == 0) {
            return 1.0 / 3 + 0.5 * (-1.0 * d[6] - 1.0 / s3 * d[7]);
        } else if (m == 1) {
            return 1.0 / 3 + 0.5 * (+1.0 * d[6] - 1.0 / s3 * d[7]);
        } else if (m == 2) {
            return 1.0 / 3 + 0.5 * (+2.0 / s3 * d[7]);
        }
        return 0.0;
    }
    double sum = 0.0;
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

/* copy_list_entry_dev::hasto_record (copy_list_entry.hpp:74-80) */
static int hasto_record(double ratio, uint64_t n)
{
    if (n == 0) {
        return 1;
    }
    return floor(ratio * (double)(n - 1)) != floor(ratio * (double)n);
}

static int density_stride(const mbh_problem* P)
{
    if (P->algo == MBH_ALGO_REG_CAYLEY) {
        return 2 * P->num_levels * P->num_levels;
    }
    if (P->algo == MBH_ALGO_CVR_RODR) {
        return P->num_levels * P->num_levels - 1;
    }
    return 0;
}

/* Initial state as solver_cpu_fdtd's constructor leaves it
 * (solver_cpu_fdtd.tpp:113-154). */
void orc_init_state(const mbh_problem* P, double* e, double* h, double* p,
                    double* d)
{
    const uint64_t nx = P->num_gridpoints;
    const int ds = density_stride(P);
    memcpy(e, P->e_init, nx * sizeof(double));
    memcpy(h, P->h_init, (nx + 1) * sizeof(double));
    memset(p, 0, nx * sizeof(double));
    if (ds) {
        memcpy(d, P->d_init, nx * ds * sizeof(double));
    }
}

/* Time steps n in [n_begin, n_end) (solver_cpu_fdtd.tpp:227-326). */
int orc_run(const mbh_problem* P, double* e, double* h, double* p,
            double* d, double* scratch, uint64_t n_begin, uint64_t n_end,
            int threads)
{
    const int64_t nx = (int64_t)P->num_gridpoints;
    const int N = (int)P->num_levels;
    const int ds = density_stride(P);
    const uint64_t qs = P->qm_stride;
    const double dx_inv = P->dx_inv;
    if (threads > 0) {
        omp_set_num_threads(threads);
    }

#pragma omp parallel
    {
        for (uint64_t n = n_begin; n < n_end; n++) {
            /* E update */
#pragma omp for schedule(static)
            for (int64_t i = 0; i < nx; i++) {
                const double* f = P->fdtd_consts + 4 * P->mat_idx[i];
                e[i] = f[0] * e[i] +
                    f[1] * (-f[3] * p[i] + (h[i + 1] - h[i]) * dx_inv);
            }
            /* sources, in list order */
#pragma omp single
            {
                for (uint32_t k = 0; k < P->num_sources; k++) {
                    const double v =
                        P->source_data[(uint64_t)k * P->num_timesteps + n];
                    const uint32_t x = P->sources[k].x_idx;
                    if (P->sources[k].type == MBH_SRC_HARD) {
                        e[x] = v;
                    } else {
                        e[x] = e[x] + v;
                    }
                }
            }
            /* H update and boundary condition */
#pragma omp for schedule(static)
            for (int64_t i = 1; i < nx; i++) {
                const double* f = P->fdtd_consts + 4 * P->mat_idx[i];
                h[i] = h[i] + f[2] * (e[i] - e[i - 1]);
            }
#pragma omp single
            {
                h[0] = 0.0;
                h[nx] = 0.0;
            }
            /* density update */
            if (P->algo != MBH_ALGO_NOOP) {
#pragma omp for schedule(static)
                for (int64_t i = 0; i < nx; i++) {
                    const double* q = P->qm_consts + qs * P->mat_idx[i];
                    if (P->algo == MBH_ALGO_REG_CAYLEY) {
                        cay_update(N, q, d + i * ds, e[i], &p[i]);
                    } else {
                        rodr_update(N, q, d + i * ds, e[i], &p[i]);
                    }
                }
            }
            /* records */
            for (uint32_t r = 0; r < P->num_records; r++) {
                const mbh_record* R = &P->records[r];
                if (!hasto_record(R->interval_ratio, n)) {
                    continue;
                }
                const uint64_t row = (uint64_t)floor(R->interval_ratio * n);
                if (row >= R->rows) {
                    continue;
                }
                if (R->type == MBH_REC_DENSITY && R->row_idx != R->col_idx) {
                    continue;
                }
                double* dst = scratch + R->offset_real + row * R->cols;
                const int64_t pos = (int64_t)R->position_idx;
#pragma omp for schedule(static)
                for (int64_t i = pos; i < pos + (int64_t)R->cols; i++) {
                    double val = 0.0;
                    const double* di = d + i * ds;
                    switch (R->type) {
                    case MBH_REC_ELECTRIC:
                        val = e[i];
                        break;
                    case MBH_REC_POLAR_DT:
                        val = p[i];
                        break;
                    case MBH_REC_MAGNETIC:
                        val = h[i];
                        break;
                    case MBH_REC_INVERSION:
                        if (P->algo == MBH_ALGO_REG_CAYLEY) {
                            val = di[2 * (N + 1)] - di[0];
                        } else if (P->algo == MBH_ALGO_CVR_RODR) {
                            val = di[N * (N - 1)];
                        }
                        break;
                    default:
                        if (P->algo == MBH_ALGO_REG_CAYLEY) {
                            val = di[2 * (R->row_idx * N + R->row_idx)];
                        } else if (P->algo == MBH_ALGO_CVR_RODR) {
                            val = cv_population(N, di, (int)R->row_idx);
                        }
                        break;
                    }
                    dst[i - pos] = val;
                }
            }
        }
    }
    return 0;
}

/* Red time loop of solver_cpu_fdtd_red::run
 * (solver_cpu_fdtd_red.tpp:351-567), steps n in [n_begin, n_end).  Each of
 * T threads owns chunk_base = N_x / T cells (the last one also the
 * remainder) and keeps them with OL halo cells on each side in arrays of
 * its own (:94-201; cells outside the grid hold zeros and are not updated).
 * Every OL steps the halos of d, e and h are copied from the neighbours'
 * arrays (:422-444); then OL sub-steps m run on a range that shrinks by one
 * cell per side and step: density (:451-457), E (:459-461, update_e
 * :288-307 with the border aligned down to VEC = 4), sources (:463-470,
 * apply_sources :326-349, which skips a source at a thread's local index
 * 0), H (:472-473, update_h :309-324), boundary h[0] = h[N_x] = 0
 * (:476-481), records of the owned cells (:484-549).  OL defaults to the
 * reference's 8 (1 for six levels, :45); T is lowered until every chunk
 * holds at least OL cells, which the halo copy needs.  The global arrays
 * hold the state before and after (owned cells gathered). */
#define RED_VEC 4
int orc_run_red(const mbh_problem* P, double* e, double* h, double* p,
                double* d, double* scratch, uint64_t n_begin, uint64_t n_end,
                int threads, int overlap)
{
    const int64_t nx = (int64_t)P->num_gridpoints;
    const int N = (int)P->num_levels;
    const int ds = density_stride(P);
    const uint64_t qs = P->qm_stride;
    const double dx_inv = P->dx_inv;
    const int64_t OL = overlap > 0 ? overlap : (N == 6 ? 1 : 8);
    int T = threads > 0 ? threads : omp_get_max_threads();
    while (T > 1 && nx / T < OL) {
        T--;
    }
    const int64_t chunk_base = nx / T;
    const int64_t chunk_rem = nx % T;

    double** te = calloc(T, sizeof(double*));
    double** th = calloc(T, sizeof(double*));
    double** tp = calloc(T, sizeof(double*));
    double** td = calloc(T, sizeof(double*));
    int32_t** tm = calloc(T, sizeof(int32_t*));
    int ok = te && th && tp && td && tm;
    for (int t = 0; ok && t < T; t++) {
        const int64_t chunk = chunk_base + (t == T - 1 ? chunk_rem : 0);
        const int64_t size = chunk + 2 * OL;
        te[t] = calloc(size, sizeof(double));
        th[t] = calloc(size, sizeof(double));
        tp[t] = calloc(size, sizeof(double));
        td[t] = calloc(size * (ds ? ds : 1), sizeof(double));
        tm[t] = calloc(size, sizeof(int32_t));
        ok = te[t] && th[t] && tp[t] && td[t] && tm[t];
    }
    int team_ok = ok;

    if (ok) {
#pragma omp parallel num_threads(T)
        {
            const int tid = omp_get_thread_num();
            if (omp_get_num_threads() != T) {
#pragma omp atomic write
                team_ok = 0;
            }
#pragma omp barrier
            if (team_ok) {
                const int64_t chunk =
                    chunk_base + (tid == T - 1 ? chunk_rem : 0);
                const int64_t size = chunk + 2 * OL;
                const int64_t base = tid * chunk_base - OL;
                double* E = te[tid];
                double* H = th[tid];
                double* Pp = tp[tid];
                double* D = td[tid];
                int32_t* M = tm[tid];

                /* load this thread's cells, halo included */
                for (int64_t i = 0; i < size; i++) {
                    const int64_t g = base + i;
                    if (g >= 0 && g < nx) {
                        M[i] = (int32_t)P->mat_idx[g];
                        E[i] = e[g];
                        H[i] = h[g];
                        Pp[i] = p[g];
                        for (int j = 0; j < ds; j++) {
                            D[i * ds + j] = d[g * ds + j];
                        }
                    } else {
                        M[i] = -1;
                        H[i] = (g == nx) ? h[nx] : 0.0;
                    }
                }
#pragma omp barrier

                for (uint64_t n = n_begin; n < n_end; n += OL) {
                    const uint64_t sub =
                        (n_end - n < (uint64_t)OL) ? n_end - n : (uint64_t)OL;

                    /* halo exchange (d, e, h) */
                    if (tid > 0) {
                        const double* pe = te[tid - 1];
                        const double* ph = th[tid - 1];
                        const double* pd = td[tid - 1];
                        for (int64_t i = 0; i < OL; i++) {
                            E[i] = pe[chunk_base + i];
                            H[i] = ph[chunk_base + i];
                            for (int j = 0; j < ds; j++) {
                                D[i * ds + j] = pd[(chunk_base + i) * ds + j];
                            }
                        }
                    }
                    if (tid < T - 1) {
                        const double* ne = te[tid + 1];
                        const double* nh = th[tid + 1];
                        const double* nd = td[tid + 1];
                        for (int64_t i = 0; i < OL; i++) {
                            E[OL + chunk_base + i] = ne[OL + i];
                            H[OL + chunk_base + i] = nh[OL + i];
                            for (int j = 0; j < ds; j++) {
                                D[(OL + chunk_base + i) * ds + j] =
                                    nd[(OL + i) * ds + j];
                            }
                        }
                    }
#pragma omp barrier

                    for (uint64_t m = 0; m < sub; m++) {
                        const uint64_t step = n + m;
                        const int64_t border =
                            (int64_t)m - (int64_t)(m % RED_VEC);

                        /* density update */
                        if (P->algo != MBH_ALGO_NOOP) {
                            for (int64_t i = (int64_t)m;
                                 i < size - (int64_t)m - 1; i++) {
                                if (M[i] < 0) {
                                    continue;
                                }
                                const double* q = P->qm_consts + qs * M[i];
                                if (P->algo == MBH_ALGO_REG_CAYLEY) {
                                    cay_update(N, q, D + i * ds, E[i], &Pp[i]);
                                } else {
                                    rodr_update(N, q, D + i * ds, E[i],
                                                &Pp[i]);
                                }
                            }
                        }
                        /* E update */
                        for (int64_t i = border; i < size - border - 1; i++) {
                            if (M[i] < 0) {
                                continue;
                            }
                            const double* f = P->fdtd_consts + 4 * M[i];
                            E[i] = f[0] * E[i] +
                                f[1] * (-f[3] * Pp[i] + (H[i + 1] - H[i]) *
                                        dx_inv);
                        }
                        /* sources, in list order */
                        for (uint32_t k = 0; k < P->num_sources; k++) {
                            const int64_t at = (int64_t)P->sources[k].x_idx -
                                tid * chunk_base + OL;
                            if (at > 0 && at < chunk + 2 * OL) {
                                const double v = P->source_data
                                    [(uint64_t)k * P->num_timesteps + step];
                                if (P->sources[k].type == MBH_SRC_HARD) {
                                    E[at] = v;
                                } else {
                                    E[at] = E[at] + v;
                                }
                            }
                        }
                        /* H update and boundary condition */
                        for (int64_t i = border + 1; i < size - border; i++) {
                            if (M[i] < 0) {
                                continue;
                            }
                            const double* f = P->fdtd_consts + 4 * M[i];
                            H[i] = H[i] + f[2] * (E[i] - E[i - 1]);
                        }
                        if (tid == 0) {
                            H[OL] = 0.0;
                        }
                        if (tid == T - 1) {
                            H[OL + chunk] = 0.0;
                        }
                        /* records of the owned cells */
                        for (uint32_t r = 0; r < P->num_records; r++) {
                            const mbh_record* R = &P->records[r];
                            if (!hasto_record(R->interval_ratio, step)) {
                                continue;
                            }
                            const uint64_t row =
                                (uint64_t)floor(R->interval_ratio * step);
                            if (row >= R->rows) {
                                continue;
                            }
                            if (R->type == MBH_REC_DENSITY &&
                                R->row_idx != R->col_idx) {
                                continue;
                            }
                            double* dst =
                                scratch + R->offset_real + row * R->cols;
                            const int64_t pos = (int64_t)R->position_idx;
                            for (int64_t i = OL; i < chunk + OL; i++) {
                                const int64_t g = base + i;
                                if (g < pos || g >= pos + (int64_t)R->cols) {
                                    continue;
                                }
                                const double* di = D + i * ds;
                                double val = 0.0;
                                switch (R->type) {
                                case MBH_REC_ELECTRIC:
                                    val = E[i];
                                    break;
                                case MBH_REC_POLAR_DT:
                                    val = Pp[i];
                                    break;
                                case MBH_REC_MAGNETIC:
                                    val = H[i];
                                    break;
                                case MBH_REC_INVERSION:
                                    if (P->algo == MBH_ALGO_REG_CAYLEY) {
                                        val = di[2 * (N + 1)] - di[0];
                                    } else if (P->algo == MBH_ALGO_CVR_RODR) {
                                        val = di[N * (N - 1)];
                                    }
                                    break;
                                default:
                                    if (P->algo == MBH_ALGO_REG_CAYLEY) {
                                        val = di[2 * (R->row_idx * N +
                                                      R->row_idx)];
                                    } else if (P->algo == MBH_ALGO_CVR_RODR) {
                                        val = cv_population(N, di,
                                                            (int)R->row_idx);
                                    }
                                    break;
                                }
                                dst[g - pos] = val;
                            }
                        }
                    }
#pragma omp barrier
                }

                /* gather the owned cells (and h[N_x]) */
                for (int64_t i = OL; i < chunk + OL; i++) {
                    const int64_t g = base + i;
                    e[g] = E[i];
                    h[g] = H[i];
                    p[g] = Pp[i];
                    for (int j = 0; j < ds; j++) {
                        d[g * ds + j] = D[i * ds + j];
                    }
                }
                if (tid == T - 1) {
                    h[nx] = H[OL + chunk];
                }
            }
        }
    }

    for (int t = 0; t < T; t++) {
        if (te) free(te[t]);
        if (th) free(th[t]);
        if (tp) free(tp[t]);
        if (td) free(td[t]);
        if (tm) free(tm[t]);
    }
    free(te);
    free(th);
    free(tp);
    free(td);
    free(tm);
    return (ok && team_ok) ? 0 : -1;
}

/* single-cell helpers for unit tests */
void orc_cayley_update(int N, const double* q, double* d, double e,
                       double* p)
{
    cay_update(N, q, d, e, p);
}

void orc_rodr_update(int N, const double* q, double* d, double e, double* p)
{
    rodr_update(N, q, d, e, p);
}

double orc_cv_population(int N, const double* d, int m)
{
    return cv_population(N, d, m);
}
