/*
 * Rodrigues z-form tile kernel with the generator product on the matrix
 * cores (gfx950 MFMA), for N = 3..6 levels (K = N^2 - 1 = 8, 15, 24, 35).
 *
 * One step of the z-form (lindblad_cvr_rodr::update,
 * algo_lindblad_cvr_rodr.hpp:234-268, restated in the rotated basis):
 *   r = R(e) y,  p = a . r + b,  y' = r + c + (W - I) r.
 * The affine part and the rotation stay in FP64; the product (W - I) r,
 * a K x K matrix shared by all cells of a material times the K x 64 block
 * of the tile's rotated states, runs as v_mfma_f32_16x16x32_f16 on FP16
 * pieces: A = W - I as hi + lo (or hi + mid + lo with ZM_A_PIECES = 3),
 * r as hi + lo; products A_hi r_hi + A_hi r_lo (+ A_mid r_hi) + A_lo r_hi
 * with FP32 accumulation.  A is scaled by a power of two into the FP16
 * range; the products are scaled back in the FP64 accumulation.  The
 * relative error of the product is about 2^-22 per step.  The A fragments
 * live in LDS (each lane reads back only its own slots).
 *
 * Layout.  One wavefront owns a tile of 64 cells.  The FDTD part keeps the
 * tile_kernel layout (lane l = cell l).  The density state of the 64 cells
 * is four 16-cell blocks n = 0..3 in the MFMA accumulator layout: lane
 * (q, j) = (l / 16, l % 16) holds ROWS slots of cell 16 n + j: rows
 * 16 t + 4 q + rr of the full M-tiles t < FT (rr < 4) and rows
 * 16 FT + RL q + rr of the last M-tile (rr < RL), where the MFMA output of
 * each M-tile lands on the same lane (the host orders the A rows so).  The
 * next step's B operand is filled from the lane's own slots: K-slot s of
 * chunk kc of lane group q stands for the row of slot u = 8 kc + s, and
 * the host builds the A columns in the same order, so no data crosses
 * lanes between steps.  Rotation planes pair rows (2 b, 2 b + 1), which
 * lie in one lane.
 */
#ifndef MBH_ZMFMA_HPP
#define MBH_ZMFMA_HPP

#include "mbh_tile.hpp"

namespace mbh {

#ifndef ZM_A_PIECES
#define ZM_A_PIECES 2
#endif

typedef _Float16 zm_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 zm_h2 __attribute__((ext_vector_type(2)));
typedef float zm_f4 __attribute__((ext_vector_type(4)));
typedef float zm_f2 __attribute__((ext_vector_type(2)));

/* geometry and per-material constant layout (bytes) of the MFMA form */
template<int N>
struct zm_geom {
    static constexpr int K = N * N - 1;
    static constexpr int B = K / 2;
    static constexpr int MT = (K + 15) / 16; /* M-tiles of 16 rows */
    /* The last M-tile holds the V remaining rows; they go to RL (even)
     * registers per lane group so that no lane carries only padding:
     * row 16 FT + RL q + rr sits in lane group q, register rr < RL. */
    static constexpr int FT = MT - 1;
    static constexpr int V = K - 16 * FT;
    static constexpr int C4 = (V + 3) / 4;
    static constexpr int RL = (C4 + (C4 & 1)) < 4 ? (C4 + (C4 & 1)) : 4;
    static constexpr int ROWS = 4 * FT + RL;  /* state slots per lane */
    static constexpr int KC = (ROWS + 7) / 8; /* K chunks of 32 */
    static constexpr int PL = ROWS / 2;       /* rotation planes per lane */
    static constexpr int NP = N - 1;         /* diagonal generators */
    static constexpr int NB = 4;             /* 16-cell blocks per tile */
    /* with a free K slot (u = ROWS) the constant c rides in the product:
     * B holds 1 there in lane group 0, A holds c in that column */
    static constexpr bool CFOLD = (ROWS % 8) != 0;
    /* header: [0] 2^-s (scale of the product) [1] max theta_dt
     *         [2] b (p offset) [3] has_qm */
    /* FP16 pieces of A = W - I: hi + lo (AP = 2) or hi + mid + lo (3).
     * Three levels at optical frequencies put the free precession in W - I
     * (omega dt ~ 5e-3 per step): its representation error would act as a
     * frequency shift, so it gets three pieces. */
    static constexpr int AP = (N <= 3) ? 3 : ZM_A_PIECES;
    static constexpr int NF = MT * KC * AP;  /* A fragments */
    static constexpr uint64_t FRAG = 32;     /* zm_h8 [MT][KC][AP][64] */
    static constexpr uint64_t TH = FRAG + (uint64_t)NF * 64 * 16;
    static constexpr uint64_t AR = TH + (uint64_t)PL * 64 * 8;   /* a rows */
    static constexpr uint64_t CR = AR + (uint64_t)ROWS * 64 * 8; /* c rows */
    static constexpr uint64_t PR = CR + (uint64_t)ROWS * 64 * 8; /* P rows */
    static constexpr uint64_t OV = PR + (uint64_t)NP * ROWS * 64 * 8; /* o */
    static constexpr uint64_t BYTES = (OV + NP * 8 + 255) & ~uint64_t(255);

    /* state row held in slot u of lane group q (>= K: padding) */
    __host__ __device__ static constexpr int slot_row(int q, int u)
    {
        return u < 4 * FT ? 16 * (u / 4) + 4 * q + u % 4
                          : 16 * FT + RL * q + (u - 4 * FT);
    }
    /* state row of output position i (0..15) of M-tile t (>= K: none) */
    __host__ __device__ static constexpr int out_row(int t, int i)
    {
        return t < FT ? 16 * t + i
                      : ((i % 4) < RL ? 16 * FT + RL * (i / 4) + i % 4 : K);
    }
};

/* v_permlane16_swap / v_permlane32_swap on both 32-bit halves of a double
 * (VALU lane moves, no LDS round trip). With rows of 16 lanes:
 * swap16(a, b): a = [a0 b0 a2 b2], b = [a1 b1 a3 b3];
 * swap32(a, b): a = [a0 a1 b0 b1], b = [a2 a3 b2 b3]. */
__device__ __forceinline__ void swap16(double& a, double& b)
{
    const auto lo = __builtin_amdgcn_permlane16_swap(
        (unsigned)__double2loint(a), (unsigned)__double2loint(b), false,
        false);
    const auto hi = __builtin_amdgcn_permlane16_swap(
        (unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false,
        false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}
__device__ __forceinline__ void swap32(double& a, double& b)
{
    const auto lo = __builtin_amdgcn_permlane32_swap(
        (unsigned)__double2loint(a), (unsigned)__double2loint(b), false,
        false);
    const auto hi = __builtin_amdgcn_permlane32_swap(
        (unsigned)__double2hiint(a), (unsigned)__double2hiint(b), false,
        false);
    a = __hiloint2double((int)hi[0], (int)lo[0]);
    b = __hiloint2double((int)hi[1], (int)lo[1]);
}

/* row r of x, broadcast to all four rows: out[r] */
__device__ __forceinline__ void rows_broadcast(double x, double* out)
{
    double a = x, b = x;
    swap32(a, b); /* a = [x0 x1 x0 x1], b = [x2 x3 x2 x3] */
    double c = a, d = b;
    swap16(a, c); /* a = [x0 x0 x0 x0], c = [x1 x1 x1 x1] */
    swap16(b, d); /* b = [x2 x2 x2 x2], d = [x3 x3 x3 x3] */
    out[0] = a;
    out[1] = c;
    out[2] = b;
    out[3] = d;
}

/* Lane (q, j) holds v[n] for blocks n = 0..3; returns the sum over the
 * four rows of v[q] (block n = q, i.e. the cell of this lane in the
 * one-cell-per-lane layout): halves first, then neighbouring rows, each
 * sum adding the same two operands as an exchange of lanes xor 32, then
 * xor 16, would. */
__device__ __forceinline__ double reduce_scatter4(const double* v)
{
    double a0 = v[0], a2 = v[2], a1 = v[1], a3 = v[3];
    swap32(a0, a2); /* a0: lower halves of v0 and v2; a2: upper halves */
    swap32(a1, a3);
    double s0 = a0 + a2; /* rows 0, 1: v0 over the halves; 2, 3: v2 */
    double s1 = a1 + a3; /* rows 0, 1: v1; rows 2, 3: v3 */
    swap16(s0, s1);
    return s0 + s1;
}

/* v - (float)hi for both halves of a packed FP16 pair, each as one
 * mixed-precision FMA (v_fma_mix_f32 reads the FP16 half directly; the
 * difference is exact in FP32, as the separate conversion and subtraction
 * it replaces) */
__device__ __forceinline__ zm_f2 zm_rest2(zm_h2 hi, zm_f2 v)
{
    const unsigned hp = __builtin_bit_cast(unsigned, hi);
    float r0, r1;
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]"
        : "=v"(r0)
        : "v"(hp), "v"(v[0]));
    asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=v"(r1)
        : "v"(hp), "v"(v[1]));
    zm_f2 r;
    r[0] = r0;
    r[1] = r1;
    return r;
}

/* per-material constants of one lane */
template<int N>
struct zm_consts {
    using Gm = zm_geom<N>;
    double th[Gm::PL], ar[Gm::ROWS], cr[Gm::ROWS];
    double scale, thmax, b;

    /* registers: theta, a, c and the header; the A fragments go to this
     * lane's slots of the LDS array sf ([NF][64]) */
    __device__ void load(const char* base, int lane, zm_h8 (*sf)[64])
    {
        const zm_h8* f = reinterpret_cast<const zm_h8*>(base + Gm::FRAG);
#pragma unroll
        for (int i = 0; i < Gm::NF; i++) {
            sf[i][lane] = f[i * 64 + lane];
        }
        const double* th_ = reinterpret_cast<const double*>(base + Gm::TH);
        const double* ar_ = reinterpret_cast<const double*>(base + Gm::AR);
        const double* cr_ = reinterpret_cast<const double*>(base + Gm::CR);
#pragma unroll
        for (int i = 0; i < Gm::PL; i++) {
            th[i] = th_[i * 64 + lane];
        }
#pragma unroll
        for (int i = 0; i < Gm::ROWS; i++) {
            ar[i] = ar_[i * 64 + lane];
            cr[i] = Gm::CFOLD ? 0.0 : cr_[i * 64 + lane];
        }
        const double* hd = reinterpret_cast<const double*>(base);
        scale = hd[0];
        thmax = hd[1];
        b = hd[2];
    }
};

/* One density step of the tile's four 16-cell blocks.  Phase 1 (VALU,
 * FP64): rotate each block's slots in place (Y := r), this lane's partial
 * a . r into pp, FP16 hi/lo B operands.  Phase 2: the split products into
 * one FP32 accumulator per (block, M-tile), issued chunk by chunk across
 * the twelve independent accumulators.  Phase 3: y' = r + 2^-s acc (+ c
 * when c is not carried in the product). */
template<int N, int TIER>
__device__ __forceinline__ void zm_step(const zm_consts<N>& K_,
                                        const zm_h8 (*sf)[64], int lane,
                                        const double* eb, float one_q0,
                                        double (*Y)[zm_geom<N>::ROWS],
                                        double* pp)
{
    using Gm = zm_geom<N>;
    zm_h8 af[Gm::NF];
#pragma unroll
    for (int f = 0; f < Gm::NF; f++) {
        af[f] = sf[f][lane];
    }
    zm_f4 acc[Gm::NB][Gm::MT];
#pragma unroll
    for (int n = 0; n < Gm::NB; n++) {
#pragma unroll
        for (int pl = 0; pl < Gm::PL; pl++) {
            double sn, cs;
            tier_sincos<TIER>(K_.th[pl] * eb[n], sn, cs);
            const double y0 = Y[n][2 * pl];
            const double y1 = Y[n][2 * pl + 1];
            Y[n][2 * pl] = fma(cs, y0, -(sn * y1));
            Y[n][2 * pl + 1] = fma(sn, y0, cs * y1);
        }
        double pa = 0.0;
#pragma unroll
        for (int i = 0; i < Gm::ROWS; i++) {
            pa = fma(K_.ar[i], Y[n][i], pa);
        }
        pp[n] = pa;
        /* r = hi + lo: hi = f16(r) and lo = f16(r - hi), both rounded to
         * nearest, converted two at a time (v_cvt_pk_f16_f32); r - hi is
         * exact in FP32 and is formed from the packed half directly
         * (zm_rest2) */
        zm_h8 bh[Gm::KC], bl[Gm::KC];
#pragma unroll
        for (int kc = 0; kc < Gm::KC; kc++) {
#pragma unroll
            for (int sl = 0; sl < 8; sl += 2) {
                zm_f2 v;
#pragma unroll
                for (int w = 0; w < 2; w++) {
                    const int u = 8 * kc + sl + w;
                    v[w] = (u < Gm::ROWS)
                        ? (float)Y[n][u]
                        : ((Gm::CFOLD && u == Gm::ROWS) ? one_q0 : 0.0f);
                }
                const zm_h2 hi = __builtin_convertvector(v, zm_h2);
                const zm_h2 lo =
                    __builtin_convertvector(zm_rest2(hi, v), zm_h2);
                bh[kc][sl] = hi[0];
                bh[kc][sl + 1] = hi[1];
                bl[kc][sl] = lo[0];
                bl[kc][sl + 1] = lo[1];
            }
        }
        /* (W - I) r of this block = A_hi r_hi + A_hi r_lo [+ A_mid r_hi]
         * + A_lo r_hi: W - I to about 2^-22 (two pieces) or 2^-33 (three),
         * r to about 2^-22 (a rounding that changes from step to step).
         * Issued here, so the next block's VALU work overlaps them. */
#pragma unroll
        for (int t = 0; t < Gm::MT; t++) {
            zm_f4 a = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int kc = 0; kc < Gm::KC; kc++) {
                const int f = (t * Gm::KC + kc) * Gm::AP;
                a = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[f], bh[kc], a,
                                                           0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_16x16x32_f16(af[f], bl[kc], a,
                                                           0, 0, 0);
                if constexpr (Gm::AP == 3) {
                    a = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        af[f + 1], bh[kc], a, 0, 0, 0);
                }
                a = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                    af[f + Gm::AP - 1], bh[kc], a, 0, 0, 0);
            }
            acc[n][t] = a;
        }
    }
#pragma unroll
    for (int n = 0; n < Gm::NB; n++) {
#pragma unroll
        for (int t = 0; t < Gm::MT; t++) {
#pragma unroll
            for (int rr = 0; rr < (t < Gm::FT ? 4 : Gm::RL); rr++) {
                const int i = (t < Gm::FT ? 4 * t : 4 * Gm::FT) + rr;
                Y[n][i] = Gm::CFOLD
                    ? fma((double)acc[n][t][rr], K_.scale, Y[n][i])
                    : fma((double)acc[n][t][rr], K_.scale,
                          Y[n][i] + K_.cr[i]);
            }
        }
    }
}

/* Wavefronts per SIMD the compiler is asked to fit: two for N = 3 (165
 * registers instead of 212 unasked; 2 % faster on 65536 cells when
 * measured against the unasked build); none for N >= 4
 * (asked for two, four levels fit in 211 registers but ran 30 % slower;
 * five and six levels would spill).  ZM_WAVES_PER_EU (build option)
 * overrides it for every N. */
#ifdef ZM_WAVES_PER_EU
#define ZM_OCCUPANCY \
    __attribute__((amdgpu_waves_per_eu(ZM_WAVES_PER_EU, ZM_WAVES_PER_EU)))
#else
#define ZM_OCCUPANCY \
    __attribute__((amdgpu_waves_per_eu((N <= 3) ? 2 : 1)))
#endif

template<int N, typename ST>
__global__ __launch_bounds__(64) ZM_OCCUPANCY void zm_kernel(KArgs k)
{
    using Gm = zm_geom<N>;
    __shared__ double s_src[TILE_SOURCES][MAX_BLOCK];

    const int lane = threadIdx.x;
    const int q = lane >> 4, j = lane & 15;
    const float one_q0 = (q == 0) ? 1.0f : 0.0f;
    const uint64_t t = k.tile_list ? k.tile_list[blockIdx.x] : blockIdx.x;
    /* a tile whose 64 cells are not all of one material runs the VALU
     * z-form body (same state layout); such tiles come first in the tile
     * list, so they overlap with the MFMA tiles. With the host's table the
     * decision and the tile's material need no load of the cells'
     * materials, and the constant loads start at once. */
    if (!k.tile_mat || blockIdx.x < k.n_mixed) {
        /* (without the table every tile takes the VALU body) */
        tile_body<algo_zform<N>, 1, ST>(k, t);
        return;
    }
    const int32_t tm = as_const(k.tile_mat)[blockIdx.x];
    const int64_t own_lo = (int64_t)(t * k.owned);
    const int64_t own_hi =
        ((own_lo + (int64_t)k.owned) < (int64_t)k.Gown) ? own_lo + k.owned
                                                         : (int64_t)k.Gown;
    const int64_t g0 = own_lo - (int64_t)k.halo;
    const int64_t gl = g0 + lane;
    const uint64_t n0 = *as_const(k.n_base) + k.n_off;
    const uint32_t nsteps = k.nsteps;

    /* FDTD state, one cell per lane (folded constants as in tile_kernel) */
    const bool in = (gl >= 0) && (gl < (int64_t)k.G);
    /* a uniform tile lies inside the grid: every lane has the material
     * (the loads go to a clamped address, without a branch of their own) */
    const int32_t m = tm & ~TILE_MAT_QM;
    const uint64_t glc = in ? (uint64_t)gl : 0u;
    const uint8_t hzv = k.hzero[glc];
    const double e0 = ld<ST>(k.e_in, glc);
    const double h0 = ld<ST>(k.h_in, glc);
    const double p0 = ld<ST>(k.p_in, glc);
    const bool hz = in ? (hzv != 0) : true;
    double e = in ? e0 : 0.0;
    double h = in ? h0 : 0.0;
    double p = in ? p0 : 0.0;
    double fa = 0.0, fg = 0.0, fx = 0.0, fc = 0.0;
    if (m >= 0) {
        const double* f = k.fdtd + (uint64_t)m * FDTD_STRIDE;
        fa = f[0];
        fg = -(f[1] * f[3]);
        fx = f[1] * f[4];
        fc = hz ? 0.0 : f[2];
    }
    const double hm = hz ? 0.0 : 1.0;

    /* density state, MFMA layout.  The loads go to clamped addresses
     * (padding slots read row K - 1) and the padding is zeroed after, so
     * that no load sits in a branch of its own and all of them are in
     * flight together. */
    double Y[Gm::NB][Gm::ROWS];
    int32_t mb[Gm::NB];
#pragma unroll
    for (int n = 0; n < Gm::NB; n++) {
        const int64_t gn = g0 + 16 * n + j;
        const bool inn = (gn >= 0) && (gn < (int64_t)k.G);
        const uint64_t gc = inn ? (uint64_t)gn : 0u;
        mb[n] = m;
#pragma unroll
        for (int u = 0; u < Gm::ROWS; u++) {
            const int row = Gm::slot_row(q, u);
            const int rc = row < Gm::K ? row : Gm::K - 1;
            const double v = ld<ST>(k.d_in, (uint64_t)rc * k.dstride + gc);
            Y[n][u] = (inn && row < Gm::K) ? v : 0.0;
        }
    }

    /* material constants of the tile (one material, from the table) */
    const int32_t m0 = m;
    const char* fbase = static_cast<const char*>(k.frag);
    const bool q0_active = (tm & TILE_MAT_QM) != 0;
    /* FP32 storage holds y - y* (y* = -Q^T d_in, the relaxation
     * equilibrium, at OFF_YEQ of the material's z-form blob): the stored
     * value and its rounding scale with the distance from equilibrium */
    const double* ystar = (const double*)k.qm + (uint64_t)m0 * k.qstride +
        algo_zform<N>::OFF_YEQ;
    const bool shifted =
        std::is_same<ST, float>::value && k.yshift && q0_active;
    if (shifted) {
        /* y* of this lane's slots (the same rows in every block), loaded
         * unconditionally as above */
        double ys[Gm::ROWS];
#pragma unroll
        for (int u = 0; u < Gm::ROWS; u++) {
            const int row = Gm::slot_row(q, u);
            ys[u] = ystar[row < Gm::K ? row : Gm::K - 1];
        }
#pragma unroll
        for (int n = 0; n < Gm::NB; n++) {
#pragma unroll
            for (int u = 0; u < Gm::ROWS; u++) {
                const int row = Gm::slot_row(q, u);
                Y[n][u] = (row < Gm::K) ? fma(1.0, ys[u], Y[n][u]) : Y[n][u];
            }
        }
    }
    __shared__ zm_h8 s_frag[Gm::NF][64];
    zm_consts<N> KC0;
    if (q0_active) {
        KC0.load(fbase + (uint64_t)m0 * k.frag_stride, lane, s_frag);
    }
    const bool my_qm = q0_active; /* this lane's cell has a density */

    /* sources of this tile */
    const uint32_t s_lo = as_const(k.tile_src)[2 * t];
    const uint32_t s_hi = as_const(k.tile_src)[2 * t + 1];
    const uint32_t nsrc = s_hi - s_lo;
    const bool src_staged = nsrc <= TILE_SOURCES && nsteps <= MAX_BLOCK;
    if (src_staged && nsrc > 0) {
        for (uint32_t jj = 0; jj < nsrc; jj++) {
            const double* data = as_const(k.src)[s_lo + jj].data;
            s_src[jj][lane] = (lane < (int)nsteps) ? data[n0 + lane] : 0.0;
        }
        __syncthreads();
    }

    /* record steps of this block */
    const uint32_t r_lo = as_const(k.tile_rec_ptr)[t];
    const uint32_t r_hi = as_const(k.tile_rec_ptr)[t + 1];
    uint64_t rec_steps = 0;
    bool rec_density = false;
    for (uint32_t ri = r_lo; ri < r_hi; ri++) {
        const uint32_t idx = as_const(k.tile_rec_idx)[ri];
        const double ratio = as_const(k.rec)[idx].ratio;
        const uint64_t rows = as_const(k.rec)[idx].rows;
        const uint32_t type = as_const(k.rec)[idx].type;
        rec_steps |= __ballot(lane < (int)nsteps &&
                              record_now(ratio, rows, n0 + lane));
        rec_density = rec_density || type == MBH_REC_DENSITY ||
            type == MBH_REC_INVERSION;
    }

    /* One step in three parts.  The field part (E update, sources, H
     * update and boundary condition, the field's broadcast to the MFMA
     * rows) returns the wave-uniform sin/cos tier of the density step (0
     * without a density); the density part runs zm_step at that tier; the
     * record part copies the step's records.  With ZM_TIER0_LOOP (six
     * levels) the block first runs a loop of tier-0 steps (the tier of
     * weak fields: one code path, so one register assignment for the
     * state, without the copies where four tier paths meet) until a step
     * needs a wider tier, then finishes that step and the rest of the
     * block with the tier chosen step by step.  Each step makes the same
     * tier decision either way, so the results do not depend on the loop
     * that ran it.  (Measured: six levels 13 % faster; three levels, where
     * the register pressure is low, 2.5 % slower.) */
    constexpr bool ZM_TIER0_LOOP = (N >= 6);
    double eb[Gm::NB];
    auto field_step = [&](uint32_t s) __attribute__((always_inline)) -> int {
        const uint64_t n = n0 + s;
        /* E update */
        const double h_right = from_right(h);
        double en = fma(fx, h_right - h, fma(fg, p, fa * e));
        /* sources, in list order */
        if (nsrc > 0) {
            for (uint32_t jj = 0; jj < nsrc; jj++) {
                const uint64_t gx = as_const(k.src)[s_lo + jj].gx;
                const uint32_t type = as_const(k.src)[s_lo + jj].type;
                const double v = src_staged
                    ? s_src[jj][s]
                    : as_const(k.src)[s_lo + jj].data[n];
                if ((int64_t)gx == gl) {
                    en = (type == MBH_SRC_HARD) ? v : en + v;
                }
            }
        }
        /* H update and boundary condition */
        const double e_left = from_left(en);
        h = fma(fc, en - e_left, hm * h);
        e = en;
        static_assert(Gm::NB == 4, "four 16-lane rows per wavefront");
        rows_broadcast(e, eb);
        if (!q0_active) {
            return 0;
        }
        double emax = fabs(eb[0]);
#pragma unroll
        for (int nb = 1; nb < Gm::NB; nb++) {
            emax = fmax(emax, fabs(eb[nb]));
        }
        const double xmax = KC0.thmax * emax;
        return __all(xmax <= 0x1p-13) ? 0
            : (__all(xmax <= 0x1p-10) ? 1
               : (__all(xmax <= TAYLOR_MAX) ? 2 : 3));
    };
    auto density_step = [&](auto tier_c) __attribute__((always_inline)) {
        double pp[Gm::NB];
        zm_step<N, decltype(tier_c)::value>(KC0, s_frag, lane, eb, one_q0, Y,
                                            pp);
        p = reduce_scatter4(pp) + KC0.b;
    };
    auto density_any = [&](int tier) __attribute__((always_inline)) {
        if (tier == 0) {
            density_step(std::integral_constant<int, 0>());
        } else if (tier == 1) {
            density_step(std::integral_constant<int, 1>());
        } else if (tier == 2) {
            density_step(std::integral_constant<int, 2>());
        } else {
            density_step(std::integral_constant<int, 3>());
        }
    };
    auto record_step = [&](uint32_t s) __attribute__((always_inline)) {
        const uint64_t n = n0 + s;
        /* record copy */
        if ((rec_steps >> s) & 1ull) {
            /* diagonal coherence-vector components of this lane's cell:
             * d_k = o_k + sum_i P[k][i] y_i, k = N(N-1) .. K-1 */
            double dd[Gm::K];
#pragma unroll
            for (int kk = 0; kk < Gm::K; kk++) {
                dd[kk] = 0.0;
            }
            if (rec_density) {
#pragma unroll 1
                for (int kk = 0; kk < Gm::NP; kk++) {
                    double part[Gm::NB];
#pragma unroll
                    for (int nb = 0; nb < Gm::NB; nb++) {
                        double acc = 0.0;
                        if (mb[nb] >= 0) {
                            const double* pr = reinterpret_cast<const double*>(
                                fbase + (uint64_t)mb[nb] * k.frag_stride +
                                Gm::PR);
#pragma unroll
                            for (int i = 0; i < Gm::ROWS; i++) {
                                acc = fma(pr[(kk * Gm::ROWS + i) * 64 + lane],
                                          Y[nb][i], acc);
                            }
                        }
                        part[nb] = acc;
                    }
                    double v = reduce_scatter4(part);
                    if (my_qm) {
                        v += reinterpret_cast<const double*>(
                            fbase + (uint64_t)m * k.frag_stride +
                            Gm::OV)[kk];
                    } else {
                        v = 0.0;
                    }
#pragma unroll
                    for (int z = 0; z < Gm::NP; z++) {
                        if (z == kk) {
                            dd[N * (N - 1) + z] = v;
                        }
                    }
                }
            }
            for (uint32_t ri = r_lo; ri < r_hi; ri++) {
                const dev_record R =
                    load_record(k.rec, as_const(k.tile_rec_idx)[ri]);
                if (!record_now(R.ratio, R.rows, n)) {
                    continue;
                }
                if (R.type == MBH_REC_DENSITY && R.row_idx != R.col_idx) {
                    continue; /* coherences are not recorded */
                }
                const uint64_t row = (uint64_t)floor(R.ratio * (double)n);
                if (gl >= own_lo && gl < own_hi &&
                    gl >= (int64_t)R.gbegin &&
                    gl < (int64_t)(R.gbegin + R.cols)) {
                    double v;
                    switch (R.type) {
                    case MBH_REC_ELECTRIC:
                        v = e;
                        break;
                    case MBH_REC_POLAR_DT:
                        v = p;
                        break;
                    case MBH_REC_MAGNETIC:
                        v = h;
                        break;
                    /* without a density (vacuum) d = 0: inversion 0,
                     * populations 1 / N, as calc_population of the zero
                     * coherence vector */
                    case MBH_REC_INVERSION:
                        v = dd[N * (N - 1)];
                        break;
                    default:
                        v = cv_population<N>(dd, R.row_idx);
                        break;
                    }
                    R.dst[row * R.cols + (uint64_t)(gl - (int64_t)R.gbegin)] =
                        v;
                }
            }
        }
    };

    if constexpr (ZM_TIER0_LOOP) {
        uint32_t s = 0;
        int pending = 0;
        for (; s < nsteps; s++) {
            const int tier = field_step(s);
            if (tier != 0) {
                pending = tier;
                break;
            }
            if (q0_active) {
                density_step(std::integral_constant<int, 0>());
            }
            record_step(s);
        }
        if (s < nsteps) {
            /* step s has its field part done and needs a wider tier */
            density_any(pending);
            record_step(s);
            for (s++; s < nsteps; s++) {
                const int tier = field_step(s);
                if (q0_active) {
                    density_any(tier);
                }
                record_step(s);
            }
        }
    } else {
        for (uint32_t s = 0; s < nsteps; s++) {
            const int tier = field_step(s);
            if (q0_active) {
                density_any(tier);
            }
            record_step(s);
        }
    }

    /* write back the owned cells */
    if (gl >= own_lo && gl < own_hi) {
        st<ST>(k.e_out, gl, e);
        st<ST>(k.h_out, gl, h);
        st<ST>(k.p_out, gl, p);
    }
    /* y* again (unconditional loads, then the stores) */
    double ys[Gm::ROWS];
#pragma unroll
    for (int u = 0; u < Gm::ROWS; u++) {
        const int row = Gm::slot_row(q, u);
        ys[u] = shifted ? ystar[row < Gm::K ? row : Gm::K - 1] : 0.0;
    }
#pragma unroll
    for (int n = 0; n < Gm::NB; n++) {
        const int64_t gn = g0 + 16 * n + j;
        if (gn >= own_lo && gn < own_hi) {
#pragma unroll
            for (int u = 0; u < Gm::ROWS; u++) {
                const int row = Gm::slot_row(q, u);
                if (row < Gm::K) {
                    const double v = shifted ? fma(-1.0, ys[u], Y[n][u])
                                             : Y[n][u];
                    st<ST>(k.d_out, (uint64_t)row * k.dstride + gn, v);
                }
            }
        }
    }
}

template<int N, typename ST>
hipError_t launch_zm(const KArgs& k, uint64_t num_tiles, hipStream_t stream)
{
    hipLaunchKernelGGL((zm_kernel<N, ST>), dim3((unsigned)num_tiles),
                       dim3(64), 0, stream, k);
    return hipGetLastError();
}

} /* namespace mbh */

#endif
